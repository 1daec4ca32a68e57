"""The path-list forward models (feffgpu_amd.forward: PathSpectra, MultipleScattering; the reference's
src/feffgpu/exafs/forward.py:33-175, 323-730) with an energy shift e0 = 7 eV, against the FP64 oracle.

The grid is q = sign(k^2 - e0 ETOK) sqrt|k^2 - e0 ETOK| (forward.py:79-82): negative below k = 1.36 1/A, where the
interpolation clamps to the first phase-grid energy and 2 q reff keeps the sign. The held path lists come from the GPU
path finder (tested against its own oracle in test_gpu_finder.py); here the oracle evaluates the same lists:
* PathSpectra.spectra: chi per copy, the degeneracy-weighted sum over the copy's paths divided by the absorbers per
  copy, FP32 within 1e-5 and FP64 within 1e-10 of the largest value; its gradient along two directions against
  five-point differences of the oracle; the misfit against the R-factor of the oracle's numpy transform.
* MultipleScattering: signals of explicit paths and path_signals of a slice of a list against the oracle's cchi and
  fine grid; full_pass rows per absorber and chi_absorbers of a subset against the oracle's per-absorber sums;
  full_gradient of a loss of the mean spectrum against differences of the oracle.
"""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from feffgpu_amd.transform import r_range
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
PHASE = GOLDEN / "validation" / "A018" / "phase.pad"
ETOK = 0.2624684236075176
E0 = 7.0
K = np.round(np.arange(0.0, 18.0 + 1e-9, 0.05), 10)
H = 1e-4


def _q():
    en = K ** 2 - E0 * ETOK
    return np.sign(en) * np.sqrt(np.abs(en))


def _copies():
    pos0, box = W.lattice("bcc", 2.87, 4)
    frames = np.stack([W.displaced(pos0, 0.06, (77, c)) for c in range(2)])
    ipot = np.full(len(pos0), 2)
    ipot[::4] = 1
    return frames, float(box[0]), ipot


def _oracle_lists(lists, nmax):
    rows = [L for L in lists.values() if L["atoms"].shape[0]]
    pad = lambda a, w: np.pad(a, ((0, 0), (0, w - a.shape[1])))
    atoms = np.concatenate([pad(L["atoms"].cpu().numpy(), nmax - 1) for L in rows])
    ipots = np.concatenate([pad(L["ipots"].cpu().numpy(), nmax) for L in rows])
    nleg = np.concatenate([np.full(L["atoms"].shape[0], L["atoms"].shape[1] + 1) for L in rows])
    absorber = np.concatenate([L["absorber"].cpu().numpy() for L in rows])
    deg = np.concatenate([L["deg"].double().cpu().numpy() for L in rows])
    return atoms, ipots, nleg, absorber, deg


@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-10), (torch.complex64, 1e-5)], ids=["fp64", "fp32"])
def test_path_spectra_with_energy_shift(dtype, tol):
    from feffgpu_amd.forward import PathSpectra
    frames, box, ipot = _copies()
    C, n, _ = frames.shape
    ps = PathSpectra(str(PHASE), K, np.zeros(len(K)), E0, window=(3.0, 15.0), r_range=(1.0, 4.0), box=box, rmax=5.5,
                     nlegmax=4, critpw=0.5, device="cuda", dtype=dtype)
    np.testing.assert_array_equal(ps.q_np, _q())
    absorbers = PathSpectra.sample_absorbers(C, n, 3, np.random.default_rng(8))
    counts = ps.refresh(frames, absorbers, ipot)
    assert {2, 3, 4} <= set(counts)
    pos_flat = torch.tensor(frames.reshape(-1, 3), device="cuda", requires_grad=True)
    chi = ps.spectra(pos_flat)
    atoms, ipots, nleg, absorber, deg = _oracle_lists(ps.lists, 4)
    copy = np.concatenate([L["copy"].cpu().numpy() for L in ps.lists.values()])
    ph = oracle.OraclePhase(PHASE)

    def spectra_oracle(x):
        s, _, _ = oracle.chi_paths(ph, x, atoms, nleg, absorber, ipots, deg, _q(), segment=copy, nseg=C, cell=box)
        return s / absorbers.shape[1]
    ref = spectra_oracle(frames.reshape(-1, 3))
    err = float(np.abs(chi.detach().double().cpu().numpy() - ref).max() / np.abs(ref).max())
    print(f"[forward] PathSpectra {dtype}: {counts} paths, e0 {E0} eV: max rel {err:.2e}", flush=True)
    assert err < tol
    # gradient of a weighted sum of the spectra
    w = np.random.default_rng(3).normal(size=ref.shape) * K ** 2
    (g,) = torch.autograd.grad((chi.double() * torch.tensor(w, device="cuda")).sum(), pos_flat)
    g = g.cpu().numpy()
    rng = np.random.default_rng(4)
    x0 = frames.reshape(-1, 3)
    for _ in range(2):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        L = lambda x: float((spectra_oracle(x) * w).sum())
        fd = (-L(x0 + 2 * H * d) + 8 * L(x0 + H * d) - 8 * L(x0 - H * d) + L(x0 - 2 * H * d)) / (12 * H)
        assert abs(float((g * d).sum()) - fd) < (1e-6 if dtype == torch.complex128 else 1e-4) * np.abs(g).sum()
    # the misfit of the mean against a target: the oracle's numpy transform
    target = ref.mean(0) * 1.1
    ps.set_target(target)
    mean, mis, _ = ps.spectrum_only(frames)
    m0, nm = r_range(ps.fit_window)
    want = oracle.r_factor(mean.double().cpu().numpy(), target, 0.05, 3.0, 15.0, 1.0, 2, m0, nm)[0]
    assert abs(float(mis) - want) < 1e-6 * want + 1e-12


def test_multiple_scattering_signals_rows_and_gradient():
    from feffgpu_amd.forward import MultipleScattering
    frames, box, ipot = _copies()
    C, n, _ = frames.shape
    ms = MultipleScattering(str(PHASE), K, E0, box, rmax=5.5, critpw=1.0, device="cuda", nlegmax=4)
    absorbers = np.array([0, 5, 70])
    counts = ms.enumerate(frames, ipot, absorbers)
    assert set(counts) == {3, 4}
    ph = oracle.OraclePhase(PHASE)
    q = _q()
    pos_flat = torch.tensor(frames.reshape(-1, 3), device="cuda")
    # explicit paths and a slice of a list: the oracle's amplitude and fine grid on the shifted grid
    L3 = ms.lists[3]
    chi_s, reff = ms.path_signals(pos_flat.float(), L3, 2, 9)
    for b, p in enumerate(range(2, 9)):
        a = int(L3["absorber"][p])
        rel = frames.reshape(-1, 3)[L3["atoms"][p].cpu().numpy()] - frames.reshape(-1, 3)[a]
        rel -= box * np.round(rel / box)
        coords = np.concatenate([np.zeros((1, 3)), rel])
        cc, rf, _ = ph.cchi(coords, L3["ipots"][p].cpu().numpy())
        want = ph.chi_fine(cc, rf, q)[0]
        assert np.abs(chi_s[b].double().cpu().numpy() - want).max() < 2e-5 * np.abs(want).max()
    # rows of every absorber, and a subset with its gradient
    rows = ms.full_pass(pos_flat).double().cpu().numpy()
    atoms, ipots, nleg, absorber, deg = _oracle_lists(ms.lists, 4)
    row = ms.row_of[torch.tensor(absorber, device="cuda")].cpu().numpy()

    def rows_oracle(x):
        return oracle.chi_paths(ph, x, atoms, nleg, absorber, ipots, deg, q, segment=row, nseg=ms.n_abs, cell=box)[0]
    ref = rows_oracle(frames.reshape(-1, 3))
    err = float(np.abs(rows - ref).max() / np.abs(ref).max())
    print(f"[forward] MultipleScattering: {counts} paths, {ms.n_abs} absorbers: rows max rel {err:.2e}", flush=True)
    assert err < 1e-5
    sub = ms.abs_atoms[torch.tensor([1, 4], device="cuda")]
    part = ms.chi_absorbers(pos_flat, sub).double().cpu().numpy()
    assert np.abs(part - rows[[1, 4]]).max() <= 1e-6 * np.abs(rows).max()
    loss = lambda chi: ((chi * torch.tensor(K ** 2, device="cuda", dtype=chi.dtype)) ** 2).sum()
    g = ms.full_gradient(pos_flat, loss).cpu().numpy()
    k2 = K ** 2
    Lo = lambda x: float(((rows_oracle(x).mean(0) * k2) ** 2).sum())
    x0 = frames.reshape(-1, 3)
    d = np.random.default_rng(6).normal(size=x0.shape)
    d /= np.linalg.norm(d)
    fd = (-Lo(x0 + 2 * H * d) + 8 * Lo(x0 + H * d) - 8 * Lo(x0 - H * d) + Lo(x0 - 2 * H * d)) / (12 * H)
    assert abs(float((g * d).sum()) - fd) < 1e-4 * np.abs(g).sum(), (float((g * d).sum()), fd)
