"""The adjoint kernels against finite differences of the FP64 oracle and against each other.

Gradient bar: relative error <= 1e-4 of the FP32 adjoint against the FP64 reference gradient. The FP64
adjoint is first pinned against five-point central differences of the FP64 oracle (h = 1e-4 A, generic geometry: no
collinear legs, where FEFF's regularised angles are not differentiable), to 1e-6 of the largest component.
"""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN, load_case

pytestmark = pytest.mark.gpu
H = 1e-4


def _d(L, x, dx):
    """Five-point central difference of L along dx (error O(h^4): chi oscillates like sin(2 q reff) with 2q up to
    36 1/A, so the three-point rule's O(h^2) term would be ~1e-6 of the derivative)."""
    return (-L(x + 2 * H * dx) + 8 * L(x + H * dx) - 8 * L(x - H * dx) + L(x - 2 * H * dx)) / (12 * H)


def _paths(case, nlegs, count, seed):
    phase, paths, _, _ = load_case(case)
    rng = np.random.default_rng(seed)
    out = []
    for n in nlegs:
        grp = [p for p in paths if p.nleg == n][:count]
        for p in grp:
            c = p.coords + np.concatenate([np.zeros((1, 3)), rng.normal(0, 0.08, (n - 1, 3))])
            out.append((c, p.ipots))
    return phase, out


def _cchi_loss_grad_fd(orc, coords, ipots, w, wr):
    def L(c):
        cc, reff, _ = orc.cchi(c, ipots)
        return float(np.sum(w.real * cc.real + w.imag * cc.imag) + wr * reff)
    g = np.zeros_like(coords)
    for i in range(coords.shape[0]):
        for k in range(3):
            e = np.zeros_like(coords)
            e[i, k] = 1.0
            g[i, k] = _d(L, coords, e)
    return g


@pytest.mark.parametrize("case,nlegs", [("validation/L001", (2, 3, 4, 5, 6)), ("validation/A018", (2, 3, 4)),
                                        ("validation/A014", (2, 3, 4)), ("validation/A009", (2, 3))])
def test_cchi_adjoint_matches_fd_and_fp32(case, nlegs):
    from feffgpu_amd.stage import GPUPathStage
    phase, paths = _paths(case, nlegs, 3, 1)
    orc = oracle.OraclePhase(phase)
    s64 = GPUPathStage(str(phase), dtype=torch.complex128)
    s32 = GPUPathStage(str(phase), dtype=torch.complex64)
    rng = np.random.default_rng(2)
    worst64 = worst32 = 0.0
    for coords, ipots in paths:
        n = len(ipots)
        w = rng.normal(size=s64.E) + 1j * rng.normal(size=s64.E)
        w *= np.exp(-0.1 * np.arange(s64.E))
        wr = float(rng.normal())
        ct = torch.tensor(coords[None], device="cuda", dtype=torch.float64)
        it = torch.tensor(ipots[None], device="cuda")
        g64 = s64._cchi_backward_raw(ct, it, torch.tensor(w[None], device="cuda"),
                                     torch.tensor([wr], device="cuda", dtype=torch.float64)).cpu().numpy()[0]
        g32 = s32._cchi_backward_raw(ct, it, torch.tensor(w[None], device="cuda", dtype=torch.complex64),
                                     torch.tensor([wr], device="cuda", dtype=torch.float64)).cpu().numpy()[0]
        gfd = _cchi_loss_grad_fd(orc, coords, ipots, w, wr)
        scale = np.abs(gfd).max()
        e64 = np.abs(g64 - gfd).max() / scale
        e32 = np.abs(g32 - g64).max() / np.abs(g64).max()
        worst64, worst32 = max(worst64, e64), max(worst32, e32)
        assert e64 < 1e-6, (case, n, e64)
        assert e32 < 1e-4, (case, n, e32)
    print(f"[grad] {case} cchi adjoint: FP64 vs FD {worst64:.2e}, FP32 vs FP64 {worst32:.2e}", flush=True)


def test_autograd_through_cchi_and_chi_fine():
    """Autograd through the drop-in API (cchi -> chi_fine -> loss) against five-point differences of the FP64 oracle.
    In complex64 mode cchi returns reff in float32, as the reference does, and chi_fine evaluates 2 q reff (hundreds
    of radians) with the FP64 value behind that tensor, so the FP32 gradient meets the 1e-4 bar."""
    from feffgpu_amd.stage import GPUPathStage
    phase, paths = _paths("validation/A018", (3, 4), 4, 3)
    orc = oracle.OraclePhase(phase)
    q = np.arange(0.0, 18.0001, 0.05)
    rng = np.random.default_rng(4)
    wq = rng.normal(size=len(q)) * q ** 2
    for dtype, tol in ((torch.complex128, 1e-6), (torch.complex64, 1e-4)):
        st = GPUPathStage(str(phase), dtype=dtype)
        for coords, ipots in paths:
            ct = torch.tensor(coords[None], device="cuda", dtype=torch.float64, requires_grad=True)
            cc, reff = st.cchi(ct, torch.tensor(ipots[None], device="cuda"))
            chi = st.chi_fine(cc, reff, q)
            loss = (chi.double() * torch.tensor(wq, device="cuda")).sum()
            (g,) = torch.autograd.grad(loss, ct)
            g = g.cpu().numpy()[0]

            def L(c):
                c1, r1, _ = orc.cchi(c, ipots)
                return float(orc.chi_fine(c1, r1, q)[0] @ wq)
            gfd = np.zeros_like(coords)
            for i in range(coords.shape[0]):
                for k in range(3):
                    e = np.zeros_like(coords)
                    e[i, k] = 1.0
                    gfd[i, k] = _d(L, coords, e)
            err = np.abs(g - gfd).max() / np.abs(gfd).max()
            print(f"[grad] drop-in autograd {dtype}: {len(ipots)} legs, max rel {err:.2e}", flush=True)
            assert err < tol, (dtype, err)
            assert reff.dtype == (torch.float32 if dtype == torch.complex64 else torch.float64)


def _box(seed, absorbers, rmax=4.6, npot=2):
    pos0, box = W.lattice("bcc", 2.87, 4)
    pos = W.displaced(pos0, 0.07, seed)
    ipot = np.full(len(pos), npot)
    ipot[::4] = 1
    pl = W.concat([W.enumerate_paths(pos, a, ipot, rmax, 4, box=box, segment=i, nmax=4)
                   for i, a in enumerate(absorbers)], 4)
    return pos, box, pl


@pytest.mark.parametrize("case,npot", [("A018", 2), ("A014", 1), ("A009", 2)], ids=["K", "L3", "M5"])
def test_chi_paths_gradient_fd_and_precisions(case, npot):
    """Path lists of 2-4 legs on a K, an L3 and an M5 edge (two-leg paths run the single-scattering kernels)."""
    from feffgpu_amd.paths import PathPlan
    phase = GOLDEN / "validation" / case / "phase.pad"
    q = np.arange(0.0, 18.0001, 0.05)
    pos, box, pl = _box(8, [0, 33], npot=npot)
    assert {2, 3, 4} <= set(pl.nleg.tolist())
    rng = np.random.default_rng(9)
    Wt = rng.normal(size=(2, len(q))) * q ** 2
    s2 = np.full(len(pl), 0.003)
    grads = {}
    for prec in (64, 32):
        plan = PathPlan(str(phase), q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=2,
                        sigma2=s2, cell=box[0], precision=prec, device="cuda")
        p = torch.tensor(pos, device="cuda", requires_grad=True)
        seg, _ = plan(p)
        loss = (seg * torch.tensor(Wt, device="cuda")).sum()
        (g,) = torch.autograd.grad(loss, p)
        grads[prec] = g.cpu().numpy()
    ph = oracle.OraclePhase(phase)

    def L(x):
        s, _, _ = oracle.chi_paths(ph, x, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, q, segment=pl.segment,
                                   nseg=2, sigma2=s2, cell=box[0])
        return float((s * Wt).sum())
    for _ in range(3):
        d = rng.normal(size=pos.shape)
        d /= np.linalg.norm(d)
        fd = _d(L, pos, d)
        an = float((grads[64] * d).sum())
        assert abs(an - fd) < 1e-6 * np.abs(grads[64]).sum() + 1e-9 * abs(fd), (an, fd)
    err32 = np.abs(grads[32] - grads[64]).max() / np.abs(grads[64]).max()
    print(f"[grad] chi_paths {case}: FP32 vs FP64 gradient {err32:.2e}", flush=True)
    assert err32 < 1e-4
    # a periodic box has no net force from a translation of every atom
    assert np.abs(grads[64].sum(0)).max() < 1e-9 * np.abs(grads[64]).sum()
