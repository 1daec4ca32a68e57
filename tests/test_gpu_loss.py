"""The misfit after the path sum on the GPU (fm_r_factor through feffgpu_amd.transform): R-factors against the
oracle's FFT form (FP64, 1e-12), gradients against central differences of the oracle, r_factor_torch's autograd,
and a captured step whose loss gradient is the R-factor's, replay by replay equal to the eager step."""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
KSTEP = 0.05


def _windows():
    from feffgpu_amd.transform import FitWindow
    return [FitWindow(3.0, 14.0, 1.0, 2, 1.0, 6.0), FitWindow(2.0, 16.0, 2.0, 3, 1.2, 5.3),
            FitWindow(4.0, 12.0, 0.0, 1, 0.5, 9.0), FitWindow(0.0, 17.5, 1.0, 0, 0.0, 31.0)]


@pytest.mark.parametrize("per_segment_data", [False, True])
def test_r_factor_matches_oracle(per_segment_data):
    from feffgpu_amd.transform import r_factor_raw, r_range
    rng = np.random.default_rng(3)
    S, nk = 5, 361
    data = rng.normal(size=(S, nk) if per_segment_data else nk)
    chi = (data if per_segment_data else data[None]) * 1.1 + 0.2 * rng.normal(size=(S, nk))
    for win in _windows():
        for scale, f32 in ((1.0, True), (0.25, False)):
            rf, g = r_factor_raw(torch.tensor(chi, device="cuda"), torch.tensor(data, device="cuda"), win, KSTEP,
                                 scale=scale, grad=True, float32_grid=f32)
            m0, nm = r_range(win, KSTEP, 2048, f32)
            ref = oracle.r_factor(chi, data, KSTEP, win.kmin, win.kmax, win.dk, win.kweight, m0, nm, scale=scale)
            err = float(np.abs(rf.cpu().numpy() - ref).max() / np.abs(ref).max())
            print(f"[loss] {win} scale {scale}: R-factors {ref[:2]}, max rel {err:.1e}", flush=True)
            assert err < 1e-12
            # gradient: central differences of the oracle along random directions, per segment
            g = g.cpu().numpy()
            for s in (0, S - 1):
                d = rng.normal(size=nk)
                h = 1e-4
                cp, cm = chi.copy(), chi.copy()
                cp[s] += h * d
                cm[s] -= h * d
                fd = (oracle.r_factor(cp, data, KSTEP, win.kmin, win.kmax, win.dk, win.kweight, m0, nm, scale=scale)[s] -
                      oracle.r_factor(cm, data, KSTEP, win.kmin, win.kmax, win.dk, win.kweight, m0, nm, scale=scale)[s]) / (2 * h)
                an = float(g[s] @ d)
                assert abs(an - fd) < 1e-7 * max(abs(fd), np.abs(g[s]).sum() * 1e-3), (an, fd)


def test_r_factor_torch_autograd_and_closed_form():
    from feffgpu_amd.transform import FitWindow, k_grid, r_factor_torch
    k = torch.tensor(k_grid(18.0), device="cuda")
    rng = np.random.default_rng(4)
    data = torch.tensor(rng.normal(size=len(k)), device="cuda")
    win = FitWindow(3.0, 16.0, 2.0, 2, 1.0, 5.0)
    # a model 1/0.8 times the data: (1/0.8 - 1)^2 for every window
    rf = r_factor_torch((data / 0.8)[None], data, k, win)
    assert abs(float(rf[0]) - 0.0625) < 1e-12
    model = (data[None] + 0.1 * torch.tensor(rng.normal(size=(3, len(k))), device="cuda")).requires_grad_(True)
    rf = r_factor_torch(model, data, k, win)
    wts = torch.tensor([1.0, -2.0, 0.5], device="cuda", dtype=torch.float64)
    (g,) = torch.autograd.grad((rf * wts).sum(), model)
    from feffgpu_amd.transform import r_factor_raw
    _, g1 = r_factor_raw(model.detach(), data, win, grad=True)
    assert torch.allclose(g, wts[:, None] * g1, rtol=0, atol=1e-15 * float(g1.abs().max()))
    with pytest.raises(ValueError):
        r_factor_torch(model, data, k + 0.01, win)          # not a grid from zero
    with pytest.raises(ValueError):
        FitWindow(kmin=5.0, kmax=4.0)


def test_r_factor_rejects_bad_arguments():
    import ctypes
    from feffgpu_amd import _native as N
    L = N.lib()
    w = N.FmWindow(3.0, 14.0, 1.0, 2, 1.0, 6.0)
    x = torch.zeros(2, 361, dtype=torch.float64, device="cuda")
    r = torch.zeros(2, dtype=torch.float64, device="cuda")
    for nfft, m0, nm, kw in ((1000, 33, 163, 2), (2048, 1000, 100, 2), (256, 0, 10, 2)):
        w.kweight = kw
        rc = L.fm_r_factor(2, 361, KSTEP, nfft, ctypes.byref(w), m0, nm, x.data_ptr(), 1.0, x.data_ptr(), 361,
                           r.data_ptr(), None, None)
        assert rc == N.FM_EINVAL, (nfft, m0, nm)
    w.kweight = -1
    assert L.fm_r_factor(2, 361, KSTEP, 2048, ctypes.byref(w), 33, 163, x.data_ptr(), 1.0, x.data_ptr(), 361,
                         r.data_ptr(), None, None) == N.FM_EINVAL


def test_captured_step_with_r_factor_loss():
    """forward -> R-factor gradient of every frame against its target (fm_r_factor, in the graph) -> adjoint, replayed
    against the same sequence run eagerly; the gradient also against central differences of the oracle's chain."""
    from feffgpu_amd.graph import CapturedStep
    from feffgpu_amd.paths import PathPlan
    from feffgpu_amd.transform import FitWindow, RFactorLoss
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    q = np.arange(0, 18.0001, 0.05)
    pos0, box = W.lattice("bcc", 2.87, 4)
    ipot = np.full(len(pos0), 2)
    ipot[::5] = 1
    pos = W.displaced(pos0, 0.07, 31)
    pl = W.concat([W.enumerate_paths(pos, a, ipot, 4.8, 4, box=box, segment=i, nmax=4) for i, a in enumerate([0, 9, 77])], 4)
    plan = PathPlan(str(phase), q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=3,
                    cell=box[0], precision=64, device="cuda")
    hidden = torch.tensor(W.displaced(pos0, 0.07, 32), device="cuda")
    target = plan.forward(hidden)[0].clone() * 0.5
    win = FitWindow(3.0, 16.0, 1.0, 2, 1.0, 5.0)
    loss = RFactorLoss(target, q, win, scale=0.5)
    p = torch.tensor(pos, device="cuda")
    seg = plan.forward(p, keep=True)[0].clone()                         # as the captured step runs it
    g_eager = plan.backward(p, loss(seg).clone(), kept=True).clone()
    rf_eager = loss.rfac.clone()
    torch.cuda.synchronize()
    step = CapturedStep(plan, len(pos0), loss_grad=loss)
    for rep in range(3):
        s2, g2 = step(p)
        torch.cuda.synchronize()
        print(f"[loss] replay {rep}: R-factors {loss.rfac.cpu().numpy()}", flush=True)
        assert torch.equal(s2, seg) and torch.equal(g2, g_eager) and torch.equal(loss.rfac, rf_eager)
        plan.forward(hidden)                        # eager work on the same plan between replays
        torch.cuda.synchronize()
    # d (sum of R-factors) / d positions against central differences of the oracle's chi and R-factor
    from feffgpu_amd.transform import r_range
    m0, nm = r_range(win, 0.05, 2048, float32_grid=False)
    orc = oracle.OraclePhase(phase)
    tgt = target.cpu().numpy()

    def L(x):
        s = oracle.chi_paths(orc, x, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, q, segment=pl.segment, nseg=3,
                             cell=box[0])[0]
        return float(oracle.r_factor(s, tgt, 0.05, win.kmin, win.kmax, win.dk, win.kweight, m0, nm, scale=0.5).sum())
    d = np.random.default_rng(9).normal(size=pos.shape)
    d /= np.linalg.norm(d)
    h = 1e-4
    fd = (-L(pos + 2 * h * d) + 8 * L(pos + h * d) - 8 * L(pos - h * d) + L(pos - 2 * h * d)) / (12 * h)
    an = float((g_eager.cpu().numpy() * d).sum())
    print(f"[loss] directional derivative of the R-factor sum: adjoint {an:.10e}, FD {fd:.10e}", flush=True)
    assert abs(an - fd) < 1e-6 * np.abs(g_eager.cpu().numpy()).sum() + 1e-8 * abs(fd)


def test_shape_checks():
    from feffgpu_amd.transform import FitWindow, RFactorLoss, r_factor_raw
    x = torch.zeros(2, 361, dtype=torch.float64, device="cuda")
    for bad in (torch.zeros(3, 361, dtype=torch.float64, device="cuda"), torch.zeros(360, dtype=torch.float64, device="cuda"),
                torch.zeros(1, 2, 361, dtype=torch.float64, device="cuda")):
        with pytest.raises(ValueError):
            r_factor_raw(x, bad, FitWindow())
    q = np.arange(0, 18.0001, 0.05)
    loss = RFactorLoss(np.zeros((3, len(q))), q, FitWindow())
    with pytest.raises(ValueError):
        loss(x)
