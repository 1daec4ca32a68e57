"""The drop-in GPUPathStage at the edges of its domain, against the FP64 oracle.

* Zero-signal energies. The reference's chi_fine masks the phase where amp < 1e-16 (src/feffgpu/stage.py:380-384),
  with amp = sqrt(|cchi|^2 + 1e-30) >= 1e-15: the mask never applies, in the reference and in the oracle alike, and
  the kernels take atan2 of every cchi. Energies whose cchi is exactly zero or tiny must still give the oracle's
  chi (forward), and exactly zero gradient where cchi = 0 (torch's angle and the amp's epsilon give 0 there).
* Potentials outside the phase file, given on the device: no host check (that would synchronise), the kernel reads
  potential 0's tables and writes NaN for the path, and the other paths are unaffected.
* complex64: reff is returned in float32 as the reference does; chi_fine given that tensor evaluates with its FP64
  value, another float32 tensor is taken as given.
"""
import numpy as np
import pytest
import torch

import oracle
from fixtures import GOLDEN, load_case

pytestmark = pytest.mark.gpu
PHASE = GOLDEN / "validation" / "A018" / "phase.pad"
Q = np.arange(0.0, 18.0001, 0.05)


def _path_cchi(stage, case="validation/A018", nleg=3):
    phase, paths, _, _ = load_case(case)
    p = [p for p in paths if p.nleg == nleg][0]
    c, r = stage.cchi(torch.tensor(p.coords[None], device="cuda", dtype=torch.float64),
                      torch.tensor(p.ipots[None], device="cuda"))
    return c.detach().to(torch.complex128), float(r.detach().double().cpu()[0])


def _zeroed(cc):
    """cchi with energies set exactly to zero (one isolated, a run of three, the first) and three tiny values."""
    c = cc.clone()
    c[0, [0, 9, 21, 22, 23]] = 0
    c[0, 14] = 1e-25 * (1 - 2j)
    c[0, 30] = -3e-20 + 0j
    c[0, 31] = 1e-30j                  # normal in float32 (denormals are not exercised)
    return c


@pytest.mark.parametrize("convention", ["feff", "larch"])
def test_chi_fine_zero_and_tiny_energies(convention):
    from feffgpu_amd.stage import GPUPathStage
    s64 = GPUPathStage(str(PHASE), dtype=torch.complex128)
    s32 = GPUPathStage(str(PHASE), dtype=torch.complex64)
    cc, reff = _path_cchi(s64)
    c = _zeroed(cc)
    orc = oracle.OraclePhase(PHASE)
    ch = c.cpu().numpy()[0]
    if convention == "feff":
        ref = orc.chi_fine(ch, reff, Q)[0]
    else:
        ref = oracle.OraclePhase.chi_fine_dense(ch, reff, Q, orc.spline_matrix(Q))[0]
    scale = np.abs(ref).max()
    for st, tol in ((s64, 1e-12), (s32, 1e-5)):
        got = st.chi_fine(c.to(st.cdtype), torch.tensor([reff], device="cuda", dtype=torch.float64), Q,
                          convention=convention)
        err = np.abs(got.double().cpu().numpy()[0] - ref).max() / scale
        print(f"[dropin] zero-signal energies, {convention}, FP{st.precision}: max rel {err:.2e}", flush=True)
        assert np.isfinite(got.cpu().numpy()).all() and err < tol


@pytest.mark.parametrize("convention", ["feff", "larch"])
def test_chi_fine_adjoint_at_zero_energies(convention):
    """d loss / d cchi: zero at the energies where cchi = 0 (the reference's autograd gives 0 there); elsewhere
    central differences of the oracle's chi_fine in the real and imaginary parts (FP64), FP32 within 1e-4."""
    from feffgpu_amd.stage import GPUPathStage
    s64 = GPUPathStage(str(PHASE), dtype=torch.complex128)
    s32 = GPUPathStage(str(PHASE), dtype=torch.complex64)
    cc, reff = _path_cchi(s64)
    c = cc.clone()
    zeros = [0, 9, 21, 22, 23]
    c[0, zeros] = 0
    rng = np.random.default_rng(2)
    wq = rng.normal(size=len(Q)) * Q ** 2
    orc = oracle.OraclePhase(PHASE)
    S = orc.spline_matrix(Q) if convention == "larch" else None

    def L(ch):
        if S is None:
            return float(orc.chi_fine(ch, reff, Q)[0] @ wq)
        return float(oracle.OraclePhase.chi_fine_dense(ch, reff, Q, S)[0] @ wq)
    grads = {}
    for st in (s64, s32):
        x = c.to(st.cdtype).clone().requires_grad_(True)
        chi = st.chi_fine(x, torch.tensor([reff], device="cuda", dtype=torch.float64), Q, convention=convention)
        (g,) = torch.autograd.grad((chi.double() * torch.tensor(wq, device="cuda")).sum(), x)
        grads[st.precision] = g.to(torch.complex128).cpu().numpy()[0]
        assert np.isfinite(grads[st.precision].real).all() and np.isfinite(grads[st.precision].imag).all()
        assert (grads[st.precision][zeros] == 0).all()
    g64 = grads[64]
    ch = c.cpu().numpy()[0]
    scale = np.abs(g64).max()
    for e in (3, 10, 20, 24, 35, 45):
        h = 1e-6 * abs(ch[e])
        for part, unit in (("re", 1.0), ("im", 1j)):
            d = np.zeros_like(ch)
            d[e] = unit * h
            fd = (-L(ch + 2 * d) + 8 * L(ch + d) - 8 * L(ch - d) + L(ch - 2 * d)) / (12 * h)
            an = g64[e].real if part == "re" else g64[e].imag
            assert abs(an - fd) < 1e-6 * scale, (e, part, an, fd)
    err32 = np.abs(grads[32] - g64).max() / scale
    print(f"[dropin] chi_fine adjoint with zero energies, {convention}: FP32 vs FP64 {err32:.2e}", flush=True)
    assert err32 < 1e-4


def test_out_of_range_potentials_on_device_give_nan_rows():
    from feffgpu_amd.stage import GPUPathStage
    phase, paths, _, _ = load_case("validation/A018")
    grp = [p for p in paths if p.nleg == 3][:3]
    coords = torch.tensor(np.stack([p.coords for p in grp]), device="cuda", dtype=torch.float64)
    ip = np.stack([p.ipots for p in grp])
    for dtype in (torch.complex128, torch.complex64):
        st = GPUPathStage(str(PHASE), dtype=dtype)
        good, _ = st.cchi(coords, torch.tensor(ip, device="cuda"))
        bad = ip.copy()
        bad[1, 2] = st.tables.nph1          # one past the last potential
        bad[2, 0] = -1
        with pytest.raises(ValueError):      # on the host: checked there
            st.cchi(coords, torch.tensor(bad))
        got, reff = st.cchi(coords, torch.tensor(bad, device="cuda"))
        got = got.cpu().numpy()
        assert np.isnan(got[1]).all() and np.isnan(got[2]).all()
        np.testing.assert_array_equal(got[0], good.cpu().numpy()[0])
        assert np.isnan(reff.cpu().numpy()[1:]).all()
        # the adjoint of such a path is NaN too, the others unaffected
        x = coords.clone().requires_grad_(True)
        c2, _ = st.cchi(x, torch.tensor(bad, device="cuda"))
        (g,) = torch.autograd.grad(c2[0].real.sum(), x)
        g = g.cpu().numpy()
        assert np.isfinite(g[0]).all() and np.isnan(g[1:]).all()


def test_complex64_reff_keeps_its_fp64_value():
    from feffgpu_amd.stage import GPUPathStage
    st = GPUPathStage(str(PHASE), dtype=torch.complex64)
    s64 = GPUPathStage(str(PHASE), dtype=torch.complex128)
    phase, paths, _, _ = load_case("validation/A018")
    p = [p for p in paths if p.nleg == 4][0]
    ct = torch.tensor(p.coords[None], device="cuda", dtype=torch.float64)
    it = torch.tensor(p.ipots[None], device="cuda")
    c32, r32 = st.cchi(ct, it)
    _, r64 = s64.cchi(ct, it)
    assert r32.dtype == torch.float32
    q = np.arange(0.0, 40.0, 0.05)                       # 2 q reff up to ~400 rad
    a = st.chi_fine(c32, r32, q)                          # the FP64 reff behind r32
    b = st.chi_fine(c32, r64, q)                          # given in FP64
    c = st.chi_fine(c32, r32.clone(), q)                  # another tensor: taken as the float32 it holds
    assert float((a - b).abs().max()) <= 1e-6 * float(a.abs().max())
    assert not torch.equal(a, c)
    # the float32 rounding of reff moves the phase by 2 q (reff32 - reff64), up to ~1e-5 rad here
    dphi = 2 * q.max() * abs(float(r32.double()[0]) - float(r64[0]))
    assert float((a - c).abs().max()) <= 1.5 * dphi * float(a.abs().max()) + 1e-6
