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

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--workload md|fe|cu|refine|multiedge] [--cpu-seconds 12]

The other workloads are the remaining configurations: "cu" the fcc Cu cluster (k 0-16, sigma2 0.003), "fe" the bcc
cluster, "refine" the 4000-atom NiO box (every 4th Ni an absorber), "multiedge" the 8000-atom Pt30Ni70 packing (L3
and K edges, 2-6 legs). Each prints its own line with its roofline and its parity against the oracle.
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
GOLD = ROOT / "tests" / "golden"
PHASE_K = GOLD / "validation" / "A018" / "phase.pad"
PHASE_L3 = GOLD / "validation" / "A014" / "phase.pad"
PHASE_W = GOLD / "W" / "phase.pad"


class Workload:
    """A configuration of the benchmark: positions, the path list with its edges, the output grid, the cell, the
    segments (frames or absorbers) and what the configuration is called."""

    def __init__(self, pos, pl, q, cell, nseg, cfg, phases, edge=None, sigma2=None):
        self.pos, self.pl, self.q, self.cell, self.nseg, self.cfg = pos, pl, q, cell, nseg, cfg
        self.phases, self.edge, self.sigma2 = phases, edge, sigma2


def build_workload(name: str) -> Workload:
    from feffgpu_amd import workloads as W
    q = np.round(np.arange(0, 18.0 + 1e-9, 0.05), 10)
    if name == "md":
        # the path lists of every frame from the GPU path finder with FEFF's plane-wave filter (critpw 2 %, the
        # reference's MultipleScattering default); the finder's time is reported beside the stage's
        import torch
        from feffgpu_amd.pathfinder import GPUPathFinder
        pos0, box = W.lattice("bcc", 2.87, 10)
        n = len(pos0)
        frames, nabs, critpw = 64, 16, 2.0
        ipot = np.full(n, 2)
        rng = np.random.default_rng(2024)
        frames_pos = [W.displaced(pos0, 0.07, (2024, f)) for f in range(frames)]
        absorbers = [rng.choice(n, nabs, replace=False) for _ in range(frames)]
        finder = GPUPathFinder(str(PHASE_K), device="cuda", rmax=5.0, nlegmax=4, critpw=critpw,
                               merge_degenerate=False)
        stack = np.stack(frames_pos)
        ab_all = np.concatenate([absorbers[f] + f * n for f in range(frames)])
        finder.enumerate_many(stack, float(box[0]), ab_all, ipot)        # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kept, _ = finder.enumerate_many(stack, float(box[0]), ab_all, ipot)
        torch.cuda.synchronize()
        finder_s = time.perf_counter() - t0
        lists = []
        for nleg, v in sorted(kept.items()):
            P = v["deg"].numel()
            if P == 0:
                continue
            atoms = np.zeros((P, 3), np.int64)
            atoms[:, :nleg - 1] = v["atoms"].cpu().numpy()
            ipots = np.zeros((P, 4), np.int64)
            ipots[:, :nleg] = v["ipots"].cpu().numpy()
            ab = ab_all[v["slot"].cpu().numpy()]
            lists.append(W.PathList(atoms, np.full(P, nleg), ab, ipots, v["deg"].double().cpu().numpy(), ab // n,
                                    np.zeros(P)))
        pl = W.concat(lists, 4)
        cfg = dict(workload="md-batch bcc 2000 atoms x 64 frames, 16 absorbers/frame, nleg 2-4, Rmax 5 A, "
                            f"FEFF importance filter {critpw} %, k 0-18/0.05",
                   atoms=n, frames=frames, absorbers_per_frame=nabs, rmax_A=5.0, nleg_max=4, nk=len(q),
                   critpw_percent=critpw, phase="FEFF validation A018 V K edge (substitute for Fe K)",
                   cell_A=float(box[0]),
                   path_finder=dict(ms_for_64_frames=finder_s * 1e3, paths_kept=len(pl),
                                    note="one GPUPathFinder.enumerate_many call over the 64 frames (1024 absorbers): "
                                         "candidates, enumeration, float32 importance and filter on the GPU, wall "
                                         "clock including its host synchronisations"))
        return Workload(np.concatenate(frames_pos), pl, q, float(box[0]), frames, cfg, [PHASE_K])
    if name == "fe":
        cl = W.cluster("bcc", 2.87, 9.0)
        pos = W.displaced(cl, 0.05, 7, keep_first=True)
        pl = W.enumerate_paths(pos, 0, np.full(len(pos), 2), 6.0, 4, nmax=4)
        cfg = dict(workload="bcc cluster 9 A (259 atoms), sigma 0.05 A, nleg <= 4, Rmax 6 A, k 0-18/0.05",
                   atoms=len(pos), rmax_A=6.0, nleg_max=4, nk=len(q), phase="FEFF validation A018 V K edge")
        return Workload(pos, pl, q, None, 1, cfg, [PHASE_K])
    if name == "cu":
        cl = W.cluster("fcc", 3.61, 6.0)
        pos = W.displaced(cl, 0.03, 101)
        pos = pos - pos[0]
        qc = np.round(np.arange(0, 16.0 + 1e-9, 0.05), 10)
        pl = W.enumerate_paths(pos, 0, np.ones(len(pos), int), 5.0, 4, nmax=4)
        cfg = dict(workload="fcc Cu cluster 6 A (79 atoms), sigma 0.03 A, nleg <= 4, Rmax 5 A, k 0-16/0.05, "
                            "sigma2 0.003", atoms=len(pos), rmax_A=5.0, nleg_max=4, nk=len(qc),
                   phase="FEFF W K edge (substitute for Cu K)")
        return Workload(pos, pl, qc, None, 1, cfg, [PHASE_W], sigma2=0.003)
    if name == "refine":
        pos0, species, box = W.rocksalt(4.18, (10, 10, 5))
        ipot = np.where(species == 0, 2, 1)
        pos = W.displaced(pos0, 0.1, 41)
        absorbers = np.nonzero(species == 0)[0][::4]
        pl = W.concat([W.enumerate_paths(pos, int(a), ipot, 5.0, 4, box=box, segment=i, nmax=4)
                       for i, a in enumerate(absorbers)], 4)
        cfg = dict(workload=f"rocksalt NiO 4000 atoms (10x10x5 cells, a 4.18 A), sigma 0.1 A, {len(absorbers)} Ni "
                            "absorbers (every 4th), nleg <= 4, Rmax 5 A, k 0-18/0.05, k^2 L2 loss",
                   atoms=len(pos), absorbers=len(absorbers), rmax_A=5.0, nleg_max=4, nk=len(q),
                   phase="FEFF validation A018 V K edge in Ge-V rock salt (substitute for Ni K in NiO)")
        return Workload(pos, pl, q, np.diag(box), len(absorbers), cfg, [PHASE_K])
    if name == "multiedge":
        n = 8000
        box = (n * (0.3 * 195.08 + 0.7 * 58.69) / 6.02214076e23 / 8.9 * 1e24) ** (1 / 3)
        pos = W.random_packing(n, box, 2.3, 11)
        heavy = np.zeros(n, bool)
        heavy[np.random.default_rng(12).choice(n, 2400, replace=False)] = True
        rng = np.random.default_rng(13)
        pt_abs = rng.choice(np.nonzero(heavy)[0], 32, replace=False)
        ni_abs = rng.choice(np.nonzero(~heavy)[0], 32, replace=False)
        ipot_l3 = np.ones(n, int)                       # A014 has one scatterer potential
        ipot_k = np.where(heavy, 1, 2)                  # A018: 1 for the heavy species, 2 for the light one
        lists, edges = [], []
        for e, (absr, ipot) in enumerate(((pt_abs, ipot_l3), (ni_abs, ipot_k))):
            for a in absr:
                lst = W.enumerate_paths(pos, int(a), ipot, 7.0, 6, box=np.full(3, box), segment=len(lists), nmax=6)
                lists.append(lst)
                edges.append(np.full(len(lst), e))
        pl = W.concat(lists, 6)
        cfg = dict(workload="random-packed Pt30Ni70 8000 atoms (8.9 g/cm3, box 52.98 A, d_min 2.3 A), Pt L3 + Ni K, "
                            "32 absorbers per edge, nleg <= 6, Rmax 7 A, k 0-18/0.05", atoms=n, rmax_A=7.0,
                   nleg_max=6, nk=len(q), cell_A=float(box),
                   phase="FEFF validation A014 Zn L3 (substitute for Pt L3) and A018 V K (substitute for Ni K)")
        return Workload(pos, pl, q, float(box), len(lists), cfg, [PHASE_L3, PHASE_K], edge=np.concatenate(edges))
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


def _oracle_args(wl, idx):
    pl = wl.pl
    s2 = None if wl.sigma2 is None else np.broadcast_to(np.asarray(wl.sigma2, float), (len(pl),))[idx]
    return dict(atoms=pl.atoms[idx], nleg=pl.nleg[idx], absorber=pl.absorber[idx], ipots=pl.ipots[idx],
                deg=pl.deg[idx], segment=pl.segment[idx], nseg=wl.nseg, sigma2=s2,
                edge=None if wl.edge is None else wl.edge[idx], cell=wl.cell)


def cpu_baseline(wl, seconds, threads):
    """The FP64 oracle (OpenMP over paths) on a bounded sample of the workload's path list: its rate, and its
    per-path chi of the sample (returned for the parity check)."""
    import oracle
    phs = [oracle.OraclePhase(p) for p in wl.phases]
    rng = np.random.default_rng(5)
    pl, q = wl.pl, wl.q

    def run(idx, want=False):
        a = _oracle_args(wl, idx)
        t0 = time.perf_counter()
        seg, path, _ = oracle.chi_paths(phs, wl.pos, a["atoms"], a["nleg"], a["absorber"], a["ipots"], a["deg"], q,
                                        segment=a["segment"], nseg=a["nseg"], edge=a["edge"], sigma2=a["sigma2"],
                                        cell=a["cell"], threads=threads, want_paths=want)
        return time.perf_counter() - t0, seg, path

    probe = np.sort(rng.choice(len(pl), min(len(pl), 256 * threads), replace=False))
    t, _, _ = run(probe)
    n = int(min(len(pl), max(len(probe), len(probe) * seconds / max(t, 1e-3))))
    idx = np.sort(rng.choice(len(pl), n, replace=False))
    t, seg, path = run(idx, want=True)
    base = dict(value=n * len(q) / t, unit="paths·k-pts/s", cores=threads, kind="port",
                sample=f"{n} of {len(pl)} paths drawn at random (seed 5), all {len(q)} k-points, FP64 oracle "
                       f"(oracle/feffref.c, OpenMP), {t:.1f} s")
    return base, idx, seg, path


def oracle_parity(torch, wl, plan32, idx, seg_ref, path_ref, rmax, dev):
    """The FP32 fused kernel against the FP64 oracle on the sampled paths: per-path chi (max|dchi_p| over the
    sample relative to the sample's largest |chi_p|), the per-segment sums of the sample (R-factor, max|dchi|/max|chi|), and the gradient of a
    k^2-weighted loss of a sub-sample of 1000 paths along two random directions (five-point differences of the
    oracle) and FP32 against FP64 (max over atoms)."""
    import oracle
    from feffgpu_amd.compare import r_factor
    from feffgpu_amd.paths import PathPlan
    q = wl.q
    path32 = plan32.forward(torch.tensor(wl.pos, device=dev))[1][torch.tensor(idx, device=dev)].double().cpu().numpy()
    # per path: max|dchi_p| relative to the largest path of the sample (a path's own scale would weigh paths whose
    # signal is a millionth of the largest, whose FP32 error is then relatively large and irrelevant)
    per_path = float(np.abs(path32 - path_ref).max() / np.abs(path_ref).max())
    pl = wl.pl
    seg32 = np.zeros_like(seg_ref)
    np.add.at(seg32, pl.segment[idx], pl.deg[idx][:, None] * path32)
    live = np.abs(seg_ref).max(1) > 0
    kmax = min(17.0, float(q[-1]) - 1.0)
    R = max(r_factor(q, seg_ref[s], seg32[s], kmin=3.0, kmax=kmax, rmin=1.0, rmax=rmax) for s in np.nonzero(live)[0])
    seg_rel = float(np.abs(seg32 - seg_ref).max() / np.abs(seg_ref).max())
    # gradient: a sub-sample, its own plans (FP32 and FP64), the loss sum_s,q (q^2 chi_s(q) w_s(q))
    rng = np.random.default_rng(6)
    sub = np.sort(rng.choice(idx, min(1000, len(idx)), replace=False))
    a = _oracle_args(wl, sub)
    Wt = rng.normal(size=(wl.nseg, len(q))) * q ** 2
    grads = {}
    for prec in (64, 32):
        plan = PathPlan([str(p) for p in wl.phases], q, a["atoms"], a["nleg"], a["absorber"], a["ipots"], a["deg"],
                        segment=a["segment"], nseg=wl.nseg, edge=a["edge"], sigma2=a["sigma2"], cell=wl.cell,
                        precision=prec, device=dev)
        p = torch.tensor(wl.pos, device=dev, requires_grad=True)
        seg, _ = plan(p)
        (g,) = torch.autograd.grad((seg * torch.tensor(Wt, device=dev)).sum(), p)
        grads[prec] = g.cpu().numpy()
    phs = [oracle.OraclePhase(p) for p in wl.phases]

    def L(x):
        s, _, _ = oracle.chi_paths(phs, x, a["atoms"], a["nleg"], a["absorber"], a["ipots"], a["deg"], q,
                                   segment=a["segment"], nseg=a["nseg"], edge=a["edge"], sigma2=a["sigma2"],
                                   cell=a["cell"])
        return float((s * Wt).sum())
    h, fd_err = 1e-4, 0.0
    for _ in range(2):
        d = rng.normal(size=wl.pos.shape)
        d /= np.linalg.norm(d)
        fd = (-L(wl.pos + 2 * h * d) + 8 * L(wl.pos + h * d) - 8 * L(wl.pos - h * d) + L(wl.pos - 2 * h * d)) / (12 * h)
        fd_err = max(fd_err, abs(float((grads[32] * d).sum()) - fd) / np.abs(grads[64]).sum())
    g32 = float(np.abs(grads[32] - grads[64]).max() / np.abs(grads[64]).max())
    return dict(sample_paths=int(len(idx)), per_path_max_abs_rel_to_largest=per_path, sample_segments_r_factor_max=float(R),
                sample_segments_max_rel=seg_rel, grad_paths=int(len(sub)),
                grad_fp32_vs_oracle_fd_directional=float(fd_err), grad_fp32_vs_fp64_max_rel=g32)


def dominant_bucket(plan, pl):
    """The bucket with the most algorithmic work and the per-bucket flops (roofline.path_cmacs x 8)."""
    from feffgpu_amd.roofline import FLOPS_PER_CMAC, path_cmacs
    work = []
    for b in range(len(plan.buckets.kind)):
        sel = plan.buckets.order[plan.buckets.offset[b]:plan.buckets.offset[b + 1]]
        T = plan.tables[int(plan.buckets.edge[b])]
        work.append(float(path_cmacs(T, plan.nleg_h[sel], pl.ipots[sel]).sum() * FLOPS_PER_CMAC))
    return int(np.argmax(work)), work


def measured_traffic(summary: str, kernel: str):
    """HBM bytes per launch of `kernel` (e.g. "k_adjoint<float,0,3>") from the committed counter summary
    profiles/<summary>.md, a line 'traffic_bytes_per_launch <kernel>: N' written from rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes scaled by a plain copy's readings; None when there is none."""
    f = ROOT / "profiles" / f"{summary}.md"
    if not f.exists():
        return None
    key = f"traffic_bytes_per_launch {kernel}:"
    for line in f.read_text().splitlines():
        if line.strip().startswith(key):
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
    ap.add_argument("--policy", choices=["tuned", "per-group", "persistent", "persistent-queue"], default="tuned",
                    help="launch policy of every bucket: tuned per bucket (default), or one fixed policy with the "
                         "paths per workgroup from the LDS budget (for profiles whose launches must all match)")
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
    wl = build_workload(args.workload)
    pos, pl, q, cell, nseg, cfg = wl.pos, wl.pl, wl.q, wl.cell, wl.nseg, wl.cfg
    counts = {int(n): int((pl.nleg == n).sum()) for n in np.unique(pl.nleg)}
    print(f"[bench] workload {args.workload}: {len(pl)} paths {counts}, {len(q)} k-points, setup "
          f"{time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)
    common = dict(segment=pl.segment, nseg=nseg, cell=cell, edge=wl.edge, sigma2=wl.sigma2, device=dev)
    phases = [str(p) for p in wl.phases]
    plan32 = PathPlan(phases, q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, precision=32, **common)
    plan64 = PathPlan(phases, q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, precision=64, **common)
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
    from feffgpu_amd.tune import apply_buckets, set_all, tune_buckets
    fixed = {"per-group": (0, False), "persistent": (-1, False), "persistent-queue": (-1, True)}
    if args.kernel_only and args.policy == "tuned":
        raise SystemExit("--kernel-only needs a fixed --policy, so that every launch of the kernel is the same")
    if args.policy != "tuned":
        set_all(plan32, *fixed[args.policy])
    bdom, b_work = dominant_bucket(plan32, pl)
    gseg_one = torch.ones(nseg, len(q), dtype=torch.float64, device=dev)
    if args.kernel_only:
        keep = plan32.keep_amplitudes
        plan32.forward_bucket(pos_t, bdom, keep=keep)
        run = (lambda: plan32.forward_bucket(pos_t, bdom, keep=keep)) if args.kernel_only == "fwd" else \
              (lambda: plan32.backward_bucket(pos_t, gseg_one, bdom, kept=keep))
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
    rmax = float(cfg["rmax_A"])
    kmax_r = min(17.0, float(q[-1]) - 1.0)
    Rs = [r_factor(q, s64[i], s32[i], kmin=3.0, kmax=kmax_r, rmin=1.0, rmax=rmax) for i in range(nseg)
          if np.abs(s64[i]).max() > 0]
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

    def step_median_ms():
        st = CapturedStep(plan32, len(pos), loss_grad=loss_grad)
        st.pos.copy_(pos_t)
        med = timed_replays(torch, st, args.steps, args.warmup, stream)[1]
        del st
        return med
    # per-bucket launch policy (paths per workgroup, one workgroup per group or a persistent grid), and the whole
    # step under each uniform policy beside the tuned one
    buckets, step_by_policy = [], {}
    if args.policy == "tuned":
        buckets = tune_buckets(plan32, pos_t)
        for name, (grid, queue) in fixed.items():
            set_all(plan32, grid, queue)
            step_by_policy[name] = step_median_ms()
        apply_buckets(plan32, buckets)
        step_by_policy["tuned-per-bucket"] = step_median_ms()
        if plan32.keep_amplitudes:
            # the same step with the adjoint recomputing the amplitudes (nothing kept from the forward)
            plan32.keep_amplitudes = False
            step_by_policy["tuned-per-bucket, amplitudes recomputed"] = step_median_ms()
            plan32.keep_amplitudes = True
    print(f"[bench] step median ms by launch policy {step_by_policy}", file=sys.stderr, flush=True)
    for r in buckets:
        print(f"[bench] bucket kind {r['kind']} edge {r['edge']} nleg {r['nleg']} ({r['paths']} paths): forward "
              f"{r['forward']['chosen']} {r['forward']['ms'][r['forward']['chosen']]:.3f} ms (default "
              f"{r['forward']['ms']['budget/per-group']:.3f}), adjoint {r['adjoint']['chosen']} "
              f"{r['adjoint']['ms'][r['adjoint']['chosen']]:.3f} ms (default "
              f"{r['adjoint']['ms']['budget/per-group']:.3f})", file=sys.stderr, flush=True)
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
    # (two-leg paths of L and M edges have no plain kernel: such a workload reports no comparator)
    try:
        plain_seg = plan64.forward_plain(pos_t)[0].cpu().numpy()
        plain_rel = float(np.abs(plain_seg - s64).max() / np.abs(s64).max())
        ms_plain64 = timed(lambda: plan64.forward_plain(pos_t), 2, 1, sync) * 1e3
        print(f"[bench] fp64 plain forward {ms_plain64:.3f} ms/step (vs fused fp64 max rel {plain_rel:.1e})",
              file=sys.stderr, flush=True)
    except NotImplementedError as e:
        plain_rel = ms_plain64 = None
        print(f"[bench] fp64 plain forward not available: {e}", file=sys.stderr, flush=True)

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
    what = f"{kind_names.get(int(plan32.buckets.kind[bdom]))}, {int(plan32.buckets.nleg[bdom])} legs, " \
           f"{int(plan32.buckets.offset[bdom + 1] - plan32.buckets.offset[bdom])} paths"
    keep = plan32.keep_amplitudes       # as in the step: the forward keeps the scattering matrices, the adjoint reads
    t_fwd = launch_ms(lambda: plan32.forward_bucket(pos_t, bdom, keep=keep))
    t_bwd = launch_ms(lambda: plan32.backward_bucket(pos_t, gseg_one, bdom, kept=keep))
    peak = MI355X_PEAK_TFLOPS["fp32"]
    note = ("FP32 vector pipes (no MFMA); 157.3 is MI355X's FP32 peak with packed math, which these kernels are "
            "compiled without (non-packed VALU ceiling 78.6). " +
            ("The adjoint reads the scattering matrices its forward kept; its algorithmic work is counted as twice "
             "the forward's either way" if keep else "The adjoint recomputes the amplitudes"))
    ach_b = 2 * b_work[bdom] / (t_bwd * 1e-3) / 1e12
    ach_f = b_work[bdom] / (t_fwd * 1e-3) / 1e12
    nl = int(plan32.buckets.nleg[bdom])
    kset = int(plan32.buckets.kind[bdom]) if nl > 2 else 3 + plan32.tables[int(plan32.buckets.edge[bdom])].ilinit - 1
    summary = f"k_adjoint_{args.workload}_batch" if args.workload == "md" else f"k_adjoint_{args.workload}"
    traffic = measured_traffic(summary, f"k_adjoint<float,{kset},{nl}>")
    roof = dict(bound="valu", achieved=ach_b, peak=peak, unit="TFLOP/s", frac=ach_b / peak, traffic=traffic,
                traffic_source=None if traffic is None else
                f"profiles/{summary}.md (rocprofv3 FETCH_SIZE / WRITE_SIZE passes of the sweep and the tail, "
                "copy-calibrated)",
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
    if args.workload == "md":
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
               mode="fused FP32 forward + adjoint (" + ("the forward keeps the scattering matrices of the 3-6-leg "
                    "paths for the adjoint" if plan32.keep_amplitudes else "amplitudes recomputed in the adjoint") +
                    "; k^2-weighted L2 loss per frame -> dL/dpositions), one hipGraph per step",
               timing=f"{args.steps} graph replays between device synchronisations (wall clock); median of the "
                      "per-replay HIP events beside it",
               median_event_ms_per_step=med_fb32, min_event_ms_per_step=float(min(per)),
               fp32_fused_fwd=dict(value=units / (ms_fwd32 * 1e-3), ms_per_step=ms_fwd32),
               fp64_fused_fwd=dict(value=units / (ms_fwd64 * 1e-3), ms_per_step=ms_fwd64),
               fp64_plain_fwd=None if ms_plain64 is None else dict(value=units / (ms_plain64 * 1e-3),
                                                                   ms_per_step=ms_plain64,
                                                                   max_rel_vs_fused_fp64=plain_rel),
               speedup_fp32_fused_fwd_bwd_vs_fp64_plain_fwd=None if ms_plain64 is None else ms_plain64 / ms_fb32,
               ratio_fp64_fused_over_fp32_fused_fwd=ms_fwd64 / ms_fwd32,
               parity_vs_fp64=dict(r_factor_max=float(max(Rs)), max_rel=maxrel, grad_max_rel=grad_rel),
               launch_policy=dict(lds_budget_kib=policy["chosen"], forward_ms=policy["forward"],
                                  backward_ms=policy["backward"],
                                  step_median_ms_by_policy=step_by_policy,
                                  per_bucket=[dict(kind=r["kind"], edge=r["edge"], nleg=r["nleg"], paths=r["paths"],
                                                   forward=r["forward"]["chosen"], adjoint=r["adjoint"]["chosen"],
                                                   forward_ms={k: round(v, 4) for k, v in r["forward"]["ms"].items()},
                                                   adjoint_ms={k: round(v, 4) for k, v in r["adjoint"]["ms"].items()})
                                              for r in buckets],
                                  bucket_policy=args.policy,
                                  note="per bucket: G paths per workgroup and grid (per-group = one workgroup per "
                                       "group of G paths; persistent = as many workgroups as the CUs hold, looping "
                                       "over the groups in a fixed order; persistent-queue = the same taking groups "
                                       "from an atomic counter), the fastest of each bucket's launch timed alone; "
                                       "step_median_ms_by_policy: the captured step with every bucket on one policy "
                                       "(G by the LDS budget) and with the tuned per-bucket choice, which the "
                                       "measured step uses"),
               roofline=roof, roofline_forward=roof_fwd, misfit=misfit, pair_tables=pairs)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        threads = max(1, min(threads, os.cpu_count() or threads))
        base, idx, seg_ref, path_ref = cpu_baseline(wl, args.cpu_seconds, threads)
        out["cpu_baseline"] = base
        print(f"[bench] cpu baseline {base['value']:.3e} {base['unit']} ({base['sample']})", file=sys.stderr, flush=True)
        out["parity_vs_oracle"] = oracle_parity(torch, wl, plan32, idx, seg_ref, path_ref, rmax, dev)
        print(f"[bench] parity vs oracle {out['parity_vs_oracle']}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
