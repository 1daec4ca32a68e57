#!/usr/bin/env python3
"""Benchmark of the path stage on one MI355X: paths x k-points per second.

Workload (default, "md"): the batched MD-EXAFS case -- a periodic bcc box of 10 x 10 x 10 conventional cells (2000
atoms, a = 2.87 A), 64 frames of seeded Gaussian displacements (sigma 0.07 A), 16 absorbers per frame, every path of
2..4 legs with half length <= 5 A (about 2.7e5 paths), output grid k = 0..18 A^-1 step 0.05 (361 points), chi summed
per frame. Inputs are synthetic. The K-edge phase file is the shipped V K-edge set of FEFF's validation case A018
(potential 2 = V as the scatterer of every site): no Fe phase file ships with the reference.

A step is one pass of the path stage over the whole batch on the GPU with inputs resident: the fused FP32 kernel
(FP64 geometry and phase) from positions to chi per frame, and -- when the adjoint is built -- its backward to
d loss / d positions. Reported beside it: the fused FP64 kernel on the same workload, the roofline of the dominant
kernel (HIP events on the launch stream), and the FP64 oracle timed on the host cores on a bounded sample.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--workload md|fe] [--cpu-seconds 12]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "paths·k-pts/s fwd+bwd (FP32 fused vs FP64 plain, 1 GPU) at R-factor≤2e-5"
PHASE_K = ROOT / "tests" / "golden" / "validation" / "A018" / "phase.pad"


def build_workload(name: str):
    from feffgpu_amd import workloads as W
    q = np.round(np.arange(0, 18.0 + 1e-9, 0.05), 10)
    if name == "md":
        pos0, box = W.lattice("bcc", 2.87, 10)
        n = len(pos0)
        frames, nabs = 64, 16
        ipot = np.full(n, 2)
        rng = np.random.default_rng(2024)
        pos_all, lists = [], []
        for f in range(frames):
            pos = W.displaced(pos0, 0.07, (2024, f))
            pos_all.append(pos)
            for a in rng.choice(n, nabs, replace=False):
                lists.append(W.enumerate_paths(pos, int(a), ipot, 5.0, 4, box=box, offset=f * n, segment=f, nmax=4))
        pl = W.concat(lists, 4)
        cfg = dict(workload="md-batch bcc 2000 atoms x 64 frames, 16 absorbers/frame, nleg 2-4, Rmax 5 A, k 0-18/0.05",
                   atoms=n, frames=frames, absorbers_per_frame=nabs, rmax_A=5.0, nleg_max=4, nk=len(q),
                   phase="FEFF validation A018 V K edge (substitute for Fe K)", cell_A=float(box[0]))
        return np.concatenate(pos_all), pl, q, float(box[0]), frames, cfg
    if name == "fe":
        cl = W.cluster("bcc", 2.87, 9.0)
        pos = W.displaced(cl, 0.05, 7, keep_first=True)
        pl = W.enumerate_paths(pos, 0, np.full(len(pos), 2), 6.0, 4, nmax=4)
        cfg = dict(workload="bcc cluster 9 A (259 atoms), sigma 0.05 A, nleg <= 4, Rmax 6 A, k 0-18/0.05",
                   atoms=len(pos), rmax_A=6.0, nleg_max=4, nk=len(q), phase="FEFF validation A018 V K edge")
        return pos, pl, q, None, 1, cfg
    raise SystemExit(f"unknown workload {name}")


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def timed_replays(torch, step, steps, warmup, stream):
    """`steps` replays of a captured step, bracketed by device synchronisations (wall clock), each replay also
    between two HIP events on the replay stream. Returns (wall ms per replay, median event ms, all event ms)."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps * 1e3
    per = [a.elapsed_time(b) for a, b in ev]
    return wall, float(np.median(per)), per


def cpu_baseline(pos, pl, q, cell, seconds):
    """The FP64 oracle (OpenMP over paths) on a bounded sample of the same path list."""
    import oracle
    ph = oracle.OraclePhase(PHASE_K)
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or threads))
    rng = np.random.default_rng(5)

    def run(idx):
        t0 = time.perf_counter()
        oracle.chi_paths(ph, pos, pl.atoms[idx], pl.nleg[idx], pl.absorber[idx], pl.ipots[idx], pl.deg[idx], q,
                         cell=cell, threads=threads)
        return time.perf_counter() - t0

    probe = np.sort(rng.choice(len(pl), min(len(pl), 256 * threads), replace=False))
    t = run(probe)
    n = int(min(len(pl), max(len(probe), len(probe) * seconds / max(t, 1e-3))))
    idx = np.sort(rng.choice(len(pl), n, replace=False))
    t = run(idx)
    return dict(value=n * len(q) / t, unit="paths·k-pts/s", cores=threads, kind="port",
                sample=f"{n} of {len(pl)} paths drawn at random (seed 5), all {len(q)} k-points, FP64 oracle "
                       f"(oracle/feffref.c, OpenMP), {t:.1f} s")


def dominant_bucket(plan, pl):
    """The bucket with the most algorithmic work and the per-bucket flops (roofline.path_cmacs x 8)."""
    from feffgpu_amd.roofline import FLOPS_PER_CMAC, path_cmacs
    cm = path_cmacs(plan.tables[0], plan.nleg_h, pl.ipots)
    work = []
    for b in range(len(plan.buckets.kind)):
        sel = plan.buckets.order[plan.buckets.offset[b]:plan.buckets.offset[b + 1]]
        work.append(float(cm[sel].sum() * FLOPS_PER_CMAC))
    return int(np.argmax(work)), work


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary (profiles/<kernel>_md_batch.md, a line
    'traffic_bytes_per_launch: N' written from rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by a plain copy's
    readings), or None when there is none."""
    f = ROOT / "profiles" / f"{kernel}_md_batch.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_only(torch, steps):
    """A device-to-device copy of 1 GiB, `steps` times: the known traffic against which the FETCH_SIZE and
    WRITE_SIZE counters are calibrated."""
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"[bench] copy of {4 * n} bytes x {steps}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="md")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-only", choices=["fwd", "bwd", "copy"], default=None,
                    help="run only the dominant bucket's forward or adjoint kernel (or a plain device copy of known "
                         "size, for calibrating counters) --steps times, for rocprofv3 --pmc passes")
    ap.add_argument("--lds-kib", type=int, nargs=2, default=None, metavar=("FWD", "BWD"),
                    help="launch policy (LDS budget per workgroup) instead of tuning")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py measures one GPU: the path stage runs on a single MI355X (no multi-GPU path)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.kernel_only == "copy":
        return copy_only(torch, args.steps)
    from feffgpu_amd.paths import PathPlan
    from feffgpu_amd.roofline import MI355X_PEAK_TFLOPS

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a ROCm GPU")
    dev = torch.device("cuda", 0)
    t_setup = time.perf_counter()
    pos, pl, q, cell, nseg, cfg = build_workload(args.workload)
    print(f"[bench] workload {args.workload}: {len(pl)} paths, {len(q)} k-points, setup {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)
    common = dict(segment=pl.segment, nseg=nseg, cell=cell, device=dev)
    plan32 = PathPlan(str(PHASE_K), q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, precision=32, **common)
    plan64 = PathPlan(str(PHASE_K), q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, precision=64, **common)
    pos_t = torch.tensor(pos, device=dev, dtype=torch.float64)
    sync = torch.cuda.synchronize
    units = len(pl) * len(q)
    # launch policy (paths per workgroup) measured on this workload before anything is timed
    from feffgpu_amd.tune import set_budget, tune
    if args.lds_kib:
        set_budget(0, args.lds_kib[0])
        set_budget(1, args.lds_kib[1])
        policy = dict(chosen=dict(forward=args.lds_kib[0], backward=args.lds_kib[1]), forward={}, backward={})
    else:
        policy = tune(plan32, pos_t)
        policy64 = tune(plan64, pos_t, budgets=(80, 160))
        print(f"[bench] FP64 launch policy {policy64['chosen']}", file=sys.stderr, flush=True)
    print(f"[bench] launch policy {policy}", file=sys.stderr, flush=True)
    bdom, b_work = dominant_bucket(plan32, pl)
    gseg_one = torch.ones(nseg, len(q), dtype=torch.float64, device=dev)
    if args.kernel_only:
        run = (lambda: plan32.forward_bucket(pos_t, bdom)) if args.kernel_only == "fwd" else \
              (lambda: plan32.backward_bucket(pos_t, gseg_one, bdom))
        for _ in range(args.steps):
            run()
        sync()
        print(f"[bench] {args.kernel_only} of bucket {bdom} ({int(plan32.buckets.nleg[bdom])} legs, "
              f"{int(plan32.buckets.offset[bdom + 1] - plan32.buckets.offset[bdom])} paths) x {args.steps}",
              file=sys.stderr, flush=True)
        return

    # parity of this workload: FP32 fused against FP64 fused (equal to the FP64 oracle to ~1e-14)
    s32 = plan32.forward(pos_t)[0].cpu().numpy()
    s64 = plan64.forward(pos_t)[0].cpu().numpy()
    from feffgpu_amd.compare import r_factor
    Rs = [r_factor(q, s64[i], s32[i], kmin=3.0, kmax=17.0, rmin=1.0, rmax=5.0) for i in range(nseg)]
    maxrel = float(np.abs(s32 - s64).max() / np.abs(s64).max())

    ms_fwd32 = timed(lambda: plan32.forward(pos_t), args.steps, args.warmup, sync) * 1e3
    print(f"[bench] fp32 fused forward {ms_fwd32:.3f} ms/step", file=sys.stderr, flush=True)
    # forward + backward: the k^2-weighted L2 loss against a target spectrum per frame, sum_s,q (q^2 (chi - t))^2,
    # whose gradient dL/dchi_seg = 2 q^4 (chi - t) is formed on the GPU between the two kernels; the whole step
    # (bucket kernels, segment sum, loss gradient, adjoint kernels, per-atom sum) captured as one hipGraph
    from feffgpu_amd.graph import CapturedStep
    target = torch.tensor(s64 * 0.97, device=dev)
    q4 = torch.tensor(q ** 4, device=dev)

    def loss_grad(seg):
        return 2.0 * q4 * (seg - target)
    stream = torch.cuda.current_stream(dev)
    step = CapturedStep(plan32, len(pos), loss_grad=loss_grad)
    step.pos.copy_(pos_t)
    ms_fb32, med_fb32, per = timed_replays(torch, step, args.steps, args.warmup, stream)
    print(f"[bench] fp32 fused forward+backward (hipGraph replay) {ms_fb32:.3f} ms/step, median {med_fb32:.3f} ms",
          file=sys.stderr, flush=True)
    # gradient parity of this workload: FP32 against FP64 adjoint
    g32 = step()[1].cpu().numpy()
    del step
    seg64 = plan64.forward(pos_t)[0]
    g64 = plan64.backward(pos_t, 2.0 * q4 * (seg64 - target)).cpu().numpy()
    grad_rel = float(np.abs(g32 - g64).max() / np.abs(g64).max())
    ms_fwd64 = timed(lambda: plan64.forward(pos_t), max(2, args.steps // 2), 1, sync) * 1e3
    print(f"[bench] fp64 fused forward {ms_fwd64:.3f} ms/step", file=sys.stderr, flush=True)
    # the comparator: the plain FP64 pipeline (the reference's op sequence, one kernel per stage, HBM intermediates)
    plain_seg = plan64.forward_plain(pos_t)[0].cpu().numpy()
    plain_rel = float(np.abs(plain_seg - s64).max() / np.abs(s64).max())
    ms_plain64 = timed(lambda: plan64.forward_plain(pos_t), 2, 1, sync) * 1e3
    print(f"[bench] fp64 plain forward {ms_plain64:.3f} ms/step (vs fused fp64 max rel {plain_rel:.1e})",
          file=sys.stderr, flush=True)

    # roofline of the dominant kernel of the step, the adjoint of the bucket with the most work (algorithmic flops
    # = 2 x the forward's, recomputation not counted), and of that bucket's forward; HIP events on the launch stream
    nrep = max(3, args.steps)

    def launch_ms(fn):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        for _ in range(2):
            fn()
        ev[0].record(stream)
        for _ in range(nrep):
            fn()
        ev[1].record(stream)
        sync()
        return ev[0].elapsed_time(ev[1]) / nrep

    kind_names = {0: "K multiple scattering", 1: "L/M multiple scattering", 2: "single scattering"}
    what = f"{kind_names.get(int(plan32.buckets.kind[bdom]))}, {int(plan32.buckets.nleg[bdom])} legs"
    t_bwd = launch_ms(lambda: plan32.backward_bucket(pos_t, gseg_one, bdom))
    t_fwd = launch_ms(lambda: plan32.forward_bucket(pos_t, bdom))
    peak = MI355X_PEAK_TFLOPS["fp32"]
    note = ("FP32 vector pipes (no MFMA); 157.3 is MI355X's FP32 peak with packed math, which these kernels are "
            "compiled without (non-packed VALU ceiling 78.6)")
    ach_b = 2 * b_work[bdom] / (t_bwd * 1e-3) / 1e12
    ach_f = b_work[bdom] / (t_fwd * 1e-3) / 1e12
    nl = int(plan32.buckets.nleg[bdom])
    kset = int(plan32.buckets.kind[bdom]) if nl > 2 else 3 + plan32.tables[int(plan32.buckets.edge[bdom])].ilinit - 1
    roof = dict(bound="mfma", achieved=ach_b, peak=peak, unit="TFLOP/s", frac=ach_b / peak,
                traffic=measured_traffic("k_adjoint") if (kset, nl) == (0, 4) else None,
                kernel=f"k_adjoint<float,{kset},{nl},*> + k_adjoint_tail<float,{nl}> ({what})", launch_ms=t_bwd,
                flops_per_launch=2 * b_work[bdom], note=note)
    roof_fwd = dict(achieved=ach_f, peak=peak, unit="TFLOP/s", frac=ach_f / peak,
                    kernel=f"k_forward<float,{kset}> ({what})", launch_ms=t_fwd,
                    flops_per_launch=b_work[bdom])

    # the misfit after the sum: R-factor of every frame and its gradient (fm_r_factor), the alternative loss_grad of
    # the step; timed on its own
    from feffgpu_amd.transform import FitWindow, RFactorLoss
    rloss = RFactorLoss(target, q, FitWindow(3.0, 16.0, 1.0, 2, 1.0, 5.0), device=dev)
    seg32 = plan32.forward(pos_t)[0].clone()
    t_rf = launch_ms(lambda: rloss(seg32))
    misfit = dict(kernel="k_r_factor", segments=int(nseg), r_points=rloss.nm, launch_ms=t_rf,
                  note="R-factor per frame and d R / d chi_seg in one launch (FP64)")

    # single scattering of whole frames from pair tables: every atom of every frame an absorber, pairs within 5 A
    pairs = None
    if cell is not None:
        from feffgpu_amd.pairs import PairTables
        pt = PairTables(str(PHASE_K), q, box=cell, rmax=5.0, r_min=1.5, dr=0.002, ms=False, device=dev)
        Pf = pos_t.reshape(nseg, -1, 3).clone().requires_grad_(True)
        gk = torch.tensor(q ** 2, device=dev)
        t_ss = launch_ms(lambda: pt.single_scattering(Pf.detach()))
        t_ssb = launch_ms(lambda: torch.autograd.grad((pt.single_scattering(Pf) * gk).sum(), Pf))
        npair = int(Pf.shape[0]) * int(Pf.shape[1]) ** 2
        pairs = dict(kernels="k_pair_hist_lds (forward), k_pair_grad (adjoint)", frames=int(Pf.shape[0]),
                     atoms=int(Pf.shape[1]), pair_evaluations=npair, forward_ms=t_ss, forward_backward_ms=t_ssb,
                     note="every atom an absorber, minimum image, rmax 5 A; forward includes the histogram-table product")

    out = dict(metric=METRIC, value=units / (ms_fb32 * 1e-3), unit="paths·k-pts/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=ms_fb32, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="fp32 (fp64 geometry and phase)", data="synthetic",
               config=dict(cfg, paths=len(pl), parallelism="1 GPU"),
               mode="fused FP32 forward + adjoint (amplitudes recomputed in the adjoint; k^2-weighted L2 loss per "
                    "frame -> dL/dpositions), one hipGraph per step",
               timing=f"{args.steps} graph replays between device synchronisations (wall clock); median of the "
                      "per-replay HIP events beside it",
               median_event_ms_per_step=med_fb32, min_event_ms_per_step=float(min(per)),
               fp32_fused_fwd=dict(value=units / (ms_fwd32 * 1e-3), ms_per_step=ms_fwd32),
               fp64_fused_fwd=dict(value=units / (ms_fwd64 * 1e-3), ms_per_step=ms_fwd64),
               fp64_plain_fwd=dict(value=units / (ms_plain64 * 1e-3), ms_per_step=ms_plain64,
                                   max_rel_vs_fused_fp64=plain_rel),
               speedup_fp32_fused_fwd_bwd_vs_fp64_plain_fwd=ms_plain64 / ms_fb32,
               ratio_fp64_fused_over_fp32_fused_fwd=ms_fwd64 / ms_fwd32,
               parity_vs_fp64=dict(r_factor_max=float(max(Rs)), max_rel=maxrel, grad_max_rel=grad_rel),
               launch_policy=dict(lds_budget_kib=policy["chosen"], forward_ms=policy["forward"],
                                  backward_ms=policy["backward"]),
               roofline=roof, roofline_forward=roof_fwd, misfit=misfit, pair_tables=pairs)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pos, pl, q, cell, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
