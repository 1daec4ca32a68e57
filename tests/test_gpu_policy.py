"""Every launch policy the autotuner can choose is checked against the oracle: the LDS budget per kernel and
precision (40, 80 and 160 KiB) decides the paths per workgroup (G = 1..8), the band plan and, for the adjoint, the
register budget (one or two wavefronts per SIMD). A mixed list of 2-4-leg paths on a K and an L3 edge runs the
forward and the adjoint under each budget: chi against the FP64 oracle, the FP64 gradient against five-point
differences of the oracle along two directions, the FP32 gradient against the FP64 one."""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
Q = np.arange(0.0, 18.0001, 0.05)
H = 1e-4


def _list(case, npot):
    pos0, box = W.lattice("bcc", 2.87, 4)
    pos = W.displaced(pos0, 0.07, 31)
    ipot = np.full(len(pos), npot)
    ipot[::4] = 1
    pl = W.concat([W.enumerate_paths(pos, a, ipot, 4.6, 4, box=box, segment=i, nmax=4)
                   for i, a in enumerate([5, 77])], 4)
    return GOLDEN / "validation" / case / "phase.pad", pos, box, pl


@pytest.mark.parametrize("case,npot", [("A018", 2), ("A014", 1)], ids=["K", "L3"])
@pytest.mark.parametrize("kib", [40, 80, 160])
def test_every_lds_budget_against_oracle(case, npot, kib):
    from feffgpu_amd.paths import PathPlan
    from feffgpu_amd.tune import set_budget
    phase, pos, box, pl = _list(case, npot)
    assert {2, 3, 4} <= set(pl.nleg.tolist())
    ph = oracle.OraclePhase(phase)
    s2 = np.full(len(pl), 0.002)
    ref, _, _ = oracle.chi_paths(ph, pos, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, segment=pl.segment,
                                 nseg=2, sigma2=s2, cell=box[0])
    rng = np.random.default_rng(kib)
    Wt = rng.normal(size=ref.shape) * Q ** 2
    prev = {}
    grads = {}
    try:
        for prec in (64, 32):
            for kernel in (0, 1):
                prev[(kernel, prec)] = set_budget(kernel, kib, prec)
            plan = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment,
                            nseg=2, sigma2=s2, cell=box[0], precision=prec, device="cuda")
            p = torch.tensor(pos, device="cuda", requires_grad=True)
            seg, _ = plan(p)
            err = float(np.abs(seg.detach().cpu().numpy() - ref).max() / np.abs(ref).max())
            (g,) = torch.autograd.grad((seg * torch.tensor(Wt, device="cuda")).sum(), p)
            grads[prec] = g.cpu().numpy()
            print(f"[policy] {case} {kib} KiB FP{prec}: {len(pl)} paths, chi max rel {err:.2e}", flush=True)
            assert err < (1e-10 if prec == 64 else 1e-5)
    finally:
        for (kernel, prec), kb in prev.items():
            set_budget(kernel, kb, prec)

    def L(x):
        s, _, _ = oracle.chi_paths(ph, x, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, segment=pl.segment,
                                   nseg=2, sigma2=s2, cell=box[0])
        return float((s * Wt).sum())
    for _ in range(2):
        d = rng.normal(size=pos.shape)
        d /= np.linalg.norm(d)
        fd = (-L(pos + 2 * H * d) + 8 * L(pos + H * d) - 8 * L(pos - H * d) + L(pos - 2 * H * d)) / (12 * H)
        an = float((grads[64] * d).sum())
        assert abs(an - fd) < 1e-6 * np.abs(grads[64]).sum(), (an, fd)
    err32 = float(np.abs(grads[32] - grads[64]).max() / np.abs(grads[64]).max())
    print(f"[policy] {case} {kib} KiB: FP32 vs FP64 gradient {err32:.2e}", flush=True)
    assert err32 < 1e-4


@pytest.mark.parametrize("case,npot", [("A018", 2), ("A014", 1)], ids=["K", "L3"])
def test_per_bucket_paths_per_workgroup_and_persistent_grids(case, npot):
    """Every bucket launched with G = 1, 2, 4, 8 paths per workgroup (where the edge's band plan and one CU's LDS
    allow it) and each grid: one workgroup per group, a persistent grid as large as the CUs hold and a small one of
    3 workgroups, in a fixed order or taking the groups from a counter. At one G the persistent grids give the
    per-group launch's chi and gradient bit for bit; every G agrees with the oracle (chi) and with the default
    policy (gradient). Calls in a row and hipGraph replays check that the counters are back at zero after each
    launch (a counter left behind would skip groups)."""
    from feffgpu_amd.graph import CapturedStep
    from feffgpu_amd.paths import PathPlan
    phase, pos, box, pl = _list(case, npot)
    ph = oracle.OraclePhase(phase)
    ref, _, _ = oracle.chi_paths(ph, pos, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, segment=pl.segment,
                                 nseg=2, cell=box[0])
    plan = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=2,
                    cell=box[0], precision=32, device="cuda")
    p = torch.tensor(pos, device="cuda")
    gseg = torch.tensor(np.random.default_rng(5).normal(size=ref.shape) * Q ** 2, device="cuda")
    nb = len(plan.buckets.kind)

    def run():
        seg, path, _ = plan.forward(p)
        return seg.clone(), path.clone(), plan.backward(p, gseg)
    seg0, _, g0 = run()
    gscale = float(g0.abs().max())
    grids = [(0, False), (-1, False), (-1, True), (3, False), (3, True)]
    tried = 0
    for G in (1, 2, 4, 8):
        base = None
        for grid, queue in grids:
            for b in range(nb):
                for kernel in (0, 1):
                    plan.set_launch(kernel, b, G, grid, queue)
            try:
                out = run()
            except ValueError:      # this G has no band plan on the edge, or its paths do not fit one CU
                break
            again = run()
            for x, y in zip(out, again):
                assert torch.equal(x, y), (G, grid, queue)
            if base is None:
                base = out
                err = float(np.abs(out[0].cpu().numpy() - ref).max() / np.abs(ref).max())
                gerr = float((out[2] - g0).abs().max()) / gscale
                print(f"[policy] {case} G={G}: chi max rel {err:.2e}, gradient vs default {gerr:.2e}", flush=True)
                assert err < 1e-5 and gerr < 1e-5
            else:
                for x, y in zip(out, base):
                    assert torch.equal(x, y), (G, grid, queue)
            tried += 1
    assert tried >= 10
    assert int(plan._tickets.abs().sum()) == 0
    # persistent grids with counters inside a graph: replays equal the eager step
    for b in range(nb):
        for kernel in (0, 1):
            plan.set_launch(kernel, b, 0, -1, True)
    def run_kept():           # as the captured step runs: the forward keeps the scattering matrices
        seg, path, _ = plan.forward(p, keep=True)
        return seg.clone(), path.clone(), plan.backward(p, gseg, kept=True)
    eager_seg, _, eager_g = run_kept()
    step = CapturedStep(plan, len(pos))
    for _ in range(3):
        seg, g = step(p, gseg)
        torch.cuda.synchronize()
        assert torch.equal(seg, eager_seg) and torch.equal(g, eager_g)
        run_kept()
    assert int(plan._tickets.abs().sum()) == 0
