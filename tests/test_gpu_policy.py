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
