"""The step captured as a hipGraph (feffgpu_amd.graph.CapturedStep): every replay equals the eager forward + adjoint
at the same positions, bit for bit, with eager work on the same plan and host synchronisations between replays. The
eager reference runs as the step does: the forward keeps the scattering matrices, the adjoint reads them."""
import numpy as np
import pytest
import torch

from feffgpu_amd import workloads as W
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
PHASE = GOLDEN / "validation" / "A018" / "phase.pad"
Q = np.arange(0, 18.0001, 0.05)


def _setup():
    from feffgpu_amd.paths import PathPlan
    pos0, box = W.lattice("bcc", 2.87, 4)
    ipot = np.full(len(pos0), 2)
    ipot[::5] = 1
    pos = W.displaced(pos0, 0.07, 31)
    pl = W.concat([W.enumerate_paths(pos, a, ipot, 4.8, 4, box=box, segment=i, nmax=4) for i, a in enumerate([0, 9, 77])], 4)
    plan = PathPlan(str(PHASE), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=3,
                    cell=box[0], precision=32, device="cuda")
    return pos0, plan


def test_captured_step_replays_equal_eager():
    from feffgpu_amd.graph import CapturedStep
    pos0, plan = _setup()
    q4 = torch.tensor(Q ** 4, device="cuda")
    target = torch.tensor(np.random.default_rng(1).normal(size=(3, len(Q))) * 1e-3, device="cuda")

    def loss_grad(seg):
        return 2.0 * q4 * (seg - target)

    def eager(p):
        seg, _, _ = plan.forward(p, keep=True)
        seg = seg.clone()
        return seg, plan.backward(p, loss_grad(seg), kept=True).clone()

    poss = [torch.tensor(W.displaced(pos0, 0.07, s), device="cuda") for s in (40, 41)]
    ref = [eager(p) for p in poss]
    torch.cuda.synchronize()
    step = CapturedStep(plan, len(pos0), loss_grad=loss_grad)
    for rep, i in enumerate((0, 1, 0, 1)):
        seg, g = step(poss[i])
        torch.cuda.synchronize()
        print(f"[graph] replay {rep} positions {i}", flush=True)
        assert torch.equal(seg, ref[i][0]), (rep, i)
        assert torch.equal(g, ref[i][1]), (rep, i)
        eager(poss[1 - i])                       # eager work on the same plan between replays
        torch.cuda.synchronize()


def test_captured_forward_only_and_static_gseg():
    from feffgpu_amd.graph import CapturedStep
    pos0, plan = _setup()
    p = torch.tensor(W.displaced(pos0, 0.07, 50), device="cuda")
    seg_ref = plan.forward(p, keep=True)[0].clone()
    gs = torch.tensor(np.random.default_rng(2).normal(size=(3, len(Q))), device="cuda")
    g_ref = plan.backward(p, gs, kept=True).clone()
    fwd = CapturedStep(plan, len(pos0), adjoint=False)
    both = CapturedStep(plan, len(pos0))
    for _ in range(3):
        seg, g = fwd(p)
        torch.cuda.synchronize()
        assert g is None and torch.equal(seg, seg_ref)
        seg, g = both(p, gseg=gs)
        torch.cuda.synchronize()
        assert torch.equal(seg, seg_ref) and torch.equal(g, g_ref)
