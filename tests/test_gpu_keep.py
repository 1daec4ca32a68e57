"""Kept scattering matrices (PathPlan keep_amplitudes; fm_launch.keep): the forward writes the matrices of the 3-6-leg
multiple-scattering paths, the adjoint at the same positions reads them instead of rebuilding them.

* The adjoint from kept matrices equals the recomputing adjoint (FP32 and FP64, K and L3 edges, 2-4 legs of a
  periodic box; FEFF's L001 list of 2-8 legs, whose 7-8-leg buckets keep nothing), also when a small workspace
  splits the buckets into chunks; the forward's outputs do not change when it keeps.
* The differentiable call: an adjoint whose forward's matrices were overwritten by a later keeping forward at other
  positions recomputes, and gives the gradient at its own positions.
* A captured step that keeps equals the eager step on every replay.
* The gradient of the kept path against five-point differences of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
Q = np.arange(0.0, 18.0001, 0.05)


def _box(case, npot):
    pos0, box = W.lattice("bcc", 2.87, 4)
    pos = W.displaced(pos0, 0.07, 31)
    ipot = np.full(len(pos), npot)
    ipot[::4] = 1
    pl = W.concat([W.enumerate_paths(pos, a, ipot, 5.0, 4, box=box, segment=i, nmax=4)
                   for i, a in enumerate([5, 77, 20, 100])], 4)
    return GOLDEN / "validation" / case / "phase.pad", pos, float(box[0]), pl


def _l001():
    from feffgpu_amd.feffio import read_paths_dat
    paths = read_paths_dat(GOLDEN / "validation" / "L001" / "paths.dat")
    nmax = max(p.nleg for p in paths)
    pos, atoms, nleg, ipots, deg = [np.zeros(3)], [], [], [], []
    for p in paths:
        row = []
        for c in p.coords[1:]:
            pos.append(c)
            row.append(len(pos) - 1)
        atoms.append(row + [0] * (nmax - 1 - len(row)))
        nleg.append(p.nleg)
        ipots.append(list(p.ipots) + [0] * (nmax - p.nleg))
        deg.append(p.deg)
    return GOLDEN / "validation" / "L001" / "phase.pad", np.array(pos), np.array(atoms), np.array(nleg), \
        np.array(ipots), np.array(deg, float)


def _close(a, b, prec):
    return float((a - b).abs().max()) <= (1e-12 if prec == 64 else 2e-6) * float(b.abs().max())


@pytest.mark.parametrize("case,npot", [("A018", 2), ("A014", 1), ("L001", 0)], ids=["K-box", "L3-box", "L001"])
@pytest.mark.parametrize("prec", [32, 64])
def test_kept_equals_recomputed(case, npot, prec):
    from feffgpu_amd import _native as N
    from feffgpu_amd.paths import PathPlan
    if case == "L001":
        phase, pos, atoms, nleg, ipots, deg = _l001()
        plan = PathPlan(str(phase), Q, atoms, nleg, np.zeros(len(nleg), int), ipots, deg, precision=prec,
                        device="cuda")
    else:
        phase, pos, box, pl = _box(case, npot)
        plan = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=4,
                        cell=box, precision=prec, device="cuda")
    kinds = {(int(k), int(n)) for k, n in zip(plan.buckets.kind, plan.buckets.nleg)}
    p = torch.tensor(pos, device="cuda")
    gseg = torch.tensor(np.random.default_rng(2).normal(size=(plan.nseg, len(Q))) * Q ** 2, device="cuda")
    seg0 = plan.forward(p)[0].clone()
    path0 = plan.chi_path.clone()
    g_re = plan.backward(p, gseg)
    seg1 = plan.forward(p, keep=True)[0].clone()
    assert torch.equal(seg1, seg0) and torch.equal(plan.chi_path, path0)
    g_kept = plan.backward(p, gseg, kept=True)
    err = float((g_kept - g_re).abs().max() / g_re.abs().max())
    print(f"[keep] {case} FP{prec}: buckets {sorted(kinds)}, kept vs recomputed gradient {err:.2e}", flush=True)
    assert _close(g_kept, g_re, prec)
    # a workspace that holds only part of a bucket: the kept matrices follow the chunks
    full = plan.workspace
    per = max((4 * n + 1) * T.E * (8 if prec == 32 else 16) for n, T in
              ((int(n), plan.tables[int(e)]) for n, e in zip(plan.buckets.nleg, plan.buckets.edge)))
    small = max(int(N.lib().fm_backward_workspace_bytes(prec, plan.nmax, 16)), 40 * per)
    assert small < full.numel()
    plan.workspace = full[:small]
    try:
        g_chunk = plan.backward(p, gseg, kept=True)
    finally:
        plan.workspace = full
    assert _close(g_chunk, g_re, prec)


def test_generation_and_graph_and_differences():
    from feffgpu_amd.graph import CapturedStep
    from feffgpu_amd.paths import PathPlan
    phase, pos, box, pl = _box("A018", 2)
    plan = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=4,
                    cell=box, precision=64, device="cuda")
    assert plan.keep_amplitudes
    rng = np.random.default_rng(9)
    w = torch.tensor(rng.normal(size=(4, len(Q))) * Q ** 2, device="cuda")
    p1 = torch.tensor(pos, device="cuda", requires_grad=True)
    p2 = torch.tensor(pos + 0.01 * rng.normal(size=pos.shape), device="cuda", requires_grad=True)
    s1, _ = plan(p1)
    gen1 = plan.kept_generation
    s2, _ = plan(p2)                       # a later keeping forward at other positions
    assert plan.kept_generation == gen1 + 1
    (g1,) = torch.autograd.grad((s1 * w).sum(), p1)
    g1_ref = plan.backward(p1.detach(), w)
    assert _close(g1, g1_ref, 64)
    (g2,) = torch.autograd.grad((s2 * w).sum(), p2)     # its own matrices are the kept ones
    plan.forward(p2.detach())
    assert _close(g2, plan.backward(p2.detach(), w), 64)
    # against five-point differences of the FP64 oracle
    ph = oracle.OraclePhase(phase)
    wn = w.cpu().numpy()

    def L(x):
        s, _, _ = oracle.chi_paths(ph, x, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, segment=pl.segment,
                                   nseg=4, cell=box)
        return float((s * wn).sum())
    x0 = p2.detach().cpu().numpy()
    h = 1e-4
    gk = g2.cpu().numpy()
    for _ in range(2):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        fd = (-L(x0 + 2 * h * d) + 8 * L(x0 + h * d) - 8 * L(x0 - h * d) + L(x0 - 2 * h * d)) / (12 * h)
        assert abs(float((gk * d).sum()) - fd) < 1e-6 * np.abs(gk).sum()
    # captured step: forward keeps, adjoint reads, every replay equal to the eager step
    step = CapturedStep(plan, len(pos))
    for _ in range(3):
        seg, g = step(pos, w)
        torch.cuda.synchronize()
        eager_seg = plan.forward(torch.tensor(pos, device="cuda"))[0]
        assert torch.equal(seg, eager_seg)
        assert _close(g, plan.backward(torch.tensor(pos, device="cuda"), w), 64)
    print("[keep] generations, graph replays and differences", flush=True)


def test_plan_over_its_keep_limit_recomputes():
    """A plan whose kept matrices would exceed keep_limit_bytes keeps nothing: the differentiable call and a captured
    step recompute the amplitudes in the adjoint, and give the recomputing adjoint's gradient."""
    from feffgpu_amd.graph import CapturedStep
    from feffgpu_amd.paths import PathPlan
    phase, pos, box, pl = _box("A018", 2)
    kw = dict(segment=pl.segment, nseg=4, cell=box, precision=32, device="cuda")
    plan = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, keep_limit_bytes=1, **kw)
    need = plan.keep_bytes()
    assert need > 1 and not plan.keep_amplitudes
    w = torch.tensor(np.random.default_rng(4).normal(size=(4, len(Q))) * Q ** 2, device="cuda")
    p = torch.tensor(pos, device="cuda", requires_grad=True)
    seg, _ = plan(p)
    (g,) = torch.autograd.grad((seg * w).sum(), p)
    assert plan._keep is None and plan.kept_generation == 0
    assert _close(g, plan.backward(p.detach(), w), 32)
    step = CapturedStep(plan, len(pos))
    _, g2 = step(pos, w)
    torch.cuda.synchronize()
    assert _close(g2, g, 32) and plan._keep is None
    roomy = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, keep_limit_bytes=need, **kw)
    assert roomy.keep_amplitudes
