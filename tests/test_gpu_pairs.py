"""Single scattering from pair tables on the GPU (feffgpu_amd.pairs.PairTables; src/feffgpu/exafs/forward.py:178-320):
the histogram kernel against the oracle's numpy histogram (cubic, general and no cell; absorber subsets; several
potentials), the tables against the oracle's two-leg path, the tabulated single scattering against the two-leg paths
of every absorber through the oracle (the 0.002 A grid's interpolation error, as the reference's own test bounds it),
and the gradient in the positions against central differences."""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from feffgpu_amd.pairs import _ONE
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
PHASE = GOLDEN / "validation" / "A018" / "phase.pad"
Q = np.arange(0, 18.0001, 0.05)


@pytest.fixture(scope="module")
def tables():
    from feffgpu_amd.pairs import PairTables
    pos0, box = W.lattice("bcc", 2.87, 5)
    return PairTables(str(PHASE), Q, box=float(box[0]), rmax=5.6, r_min=1.5, dr=0.002, ms=False,
                      dtype=torch.complex128), pos0, float(box[0])


def test_histogram_matches_oracle(tables):
    pt, pos0, L = tables
    rng = np.random.default_rng(1)
    pos = W.displaced(pos0, 0.1, 5)
    n = len(pos)
    ipot = rng.integers(0, pt.npot + 1, n)             # includes a value outside 0..npot-1 (never a neighbor)
    for cell in (L, np.eye(3) * L, None):
        ab = rng.choice(n, 23, replace=False)
        h, short = pt.histogram(torch.tensor(pos, device="cuda"), torch.tensor(ab, device="cuda"),
                                torch.tensor(ipot, device="cuda"), box=cell)
        ref, rshort = oracle.pair_histogram(pos, ab, ipot, cell, pt.r0, pt.dr, pt.nr, pt.rmax, pt.npot)
        err = float(np.abs(h.cpu().numpy() - ref).max())
        print(f"[pairs] histogram cell {None if cell is None else np.shape(cell)}: {ref.sum():.0f} pairs, max abs {err:.1e}, "
              f"short {short}/{rshort}", flush=True)
        assert err < 1e-8 and short == rshort          # 2^-32 fixed point: <= 2^-33 per pair and bin
    # pairs closer than the grid's start land on its first bin and are counted
    close = pos.copy()
    close[1] = close[0] + np.array([0.0, 0.0, 1.2])
    h, short = pt.histogram(torch.tensor(close, device="cuda"), torch.tensor([0, 1], device="cuda"),
                            torch.ones(n, dtype=torch.int64, device="cuda"), box=L)
    ref, rshort = oracle.pair_histogram(close, [0, 1], np.ones(n, int), L, pt.r0, pt.dr, pt.nr, pt.rmax, pt.npot)
    assert short == rshort >= 2 and float(np.abs(h.cpu().numpy() - ref).max()) < 1e-8
    with pytest.raises(ValueError):
        pt.histogram(torch.tensor(pos, device="cuda"), torch.tensor([n], device="cuda"), torch.tensor(ipot, device="cuda"))


def test_tables_are_the_two_leg_path(tables):
    pt, _, _ = tables
    orc = oracle.OraclePhase(PHASE)
    for ipot in sorted(pt.tables):
        for i in (0, 700, pt.nr - 1):
            r = np.float64(np.float32(pt.r0) + np.float32(pt.dr) * np.float32(i))   # the grid in float32, as the reference
            c, reff, _ = orc.cchi(np.array([[0.0, 0.0, 0.0], [0.0, 0.0, r]]), np.array([0, ipot]))
            ref = orc.chi_fine(c, reff, Q)[0]
            got = pt.tables[ipot][i].double().cpu().numpy()
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            assert err < 1e-6, (ipot, i, err)


def test_single_scattering_equals_two_leg_paths(tables):
    """Every atom of one configuration an absorber: the histogram times the tables against the two-leg paths of every
    absorber through the oracle, divided by the number of absorbers (2e-3 on 3 < k < 17: the grid's interpolation)."""
    pt, pos0, L = tables
    pos = W.displaced(pos0, 0.05, 7)
    n = len(pos)
    P = torch.tensor(pos, device="cuda")[None]
    chi = pt.single_scattering(P).cpu().numpy()
    pl = W.concat([W.enumerate_paths(pos, a, np.ones(n, int), pt.rmax, 2, box=np.full(3, L), segment=0, nmax=2)
                   for a in range(0, n, 25)], 2)
    sub = np.arange(0, n, 25)
    orc = oracle.OraclePhase(PHASE)
    ref = oracle.chi_paths(orc, pos, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, cell=L)[0][0] / len(sub)
    # the same absorbers through the tables: the histogram of the subset times the tables
    h, _ = pt.histogram(P[0], torch.tensor(sub, device="cuda"), torch.ones(n, dtype=torch.int64, device="cuda"), box=L)
    tab = (h[1] @ pt.tables[1].double()).cpu().numpy() / len(sub)
    sel = (Q > 3) & (Q < 17)
    err = float(np.abs(tab - ref)[sel].max() / np.abs(ref[sel]).max())
    print(f"[pairs] tabulated vs two-leg paths ({len(pl)} paths): max rel {err:.1e}", flush=True)
    assert err < 2e-3
    # every atom an absorber: the mean over all absorbers equals the subset's on a perfect lattice; here compare the
    # all-absorber spectrum with the tables times the all-absorber histogram
    h_all, _ = pt.histogram(P[0], torch.arange(n, device="cuda"), torch.ones(n, dtype=torch.int64, device="cuda"), box=L)
    want = (h_all[1] @ pt.tables[1].double()).cpu().numpy() / n
    assert np.abs(chi - want).max() < 1e-12 * np.abs(want).max()


def test_single_scattering_gradient(tables):
    pt, pos0, L = tables
    pos = W.displaced(pos0, 0.05, 9)
    n = len(pos)
    rng = np.random.default_rng(3)
    w = rng.normal(size=len(Q)) * Q ** 2
    T = pt.tables[1].double().cpu().numpy()
    P = torch.tensor(pos[None], device="cuda", requires_grad=True)
    chi = pt.single_scattering(P)
    (g,) = torch.autograd.grad((chi * torch.tensor(w, device="cuda")).sum(), P)
    g = g.cpu().numpy()[0]

    def loss(x):
        h, _ = oracle.pair_histogram(x, np.arange(n), np.ones(n, int), L, pt.r0, pt.dr, pt.nr, pt.rmax, pt.npot)
        return float((h[1] @ T / n) @ w)
    for _ in range(2):
        d = rng.normal(size=pos.shape)
        d /= np.linalg.norm(d)
        h = 1e-6
        fd = (loss(pos + h * d) - loss(pos - h * d)) / (2 * h)
        an = float((g * d).sum())
        print(f"[pairs] directional derivative: adjoint {an:.8e}, FD {fd:.8e}", flush=True)
        assert abs(an - fd) < 1e-5 * np.abs(g).sum()
    # float32 positions, as the reference passes them: same spectrum to float32 rounding, gradient in float32
    P32 = torch.tensor(pos[None], device="cuda", dtype=torch.float32, requires_grad=True)
    c32 = pt.single_scattering(P32)
    assert c32.dtype == torch.float32
    (g32,) = torch.autograd.grad((c32 * torch.tensor(w, device="cuda", dtype=torch.float32)).sum(), P32)
    assert g32.dtype == torch.float32 and torch.isfinite(g32).all()


def test_cell_must_exceed_twice_the_cutoff():
    from feffgpu_amd.pairs import PairTables
    with pytest.raises(ValueError):
        PairTables(str(PHASE), Q, box=10.0, rmax=5.6, ms=False)


def test_histogram_without_lds_privatisation(tables):
    """A grid too fine for a workgroup's LDS (3 x 8201 bins) takes the kernel with global atomics: same histogram."""
    import ctypes
    from feffgpu_amd import _native as N
    pt, pos0, L = tables
    pos = W.displaced(pos0, 0.1, 6)
    n = len(pos)
    g = pt._grid()
    g.dr = 0.0005
    g.nr = int(round((pt.rmax - pt.r0) / g.dr)) + 1
    ipot = np.random.default_rng(5).integers(0, 3, n)
    ab = np.arange(0, n, 3)
    hist = torch.zeros(3, g.nr, dtype=torch.int64, device="cuda")
    short = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = torch.tensor(pos, device="cuda")
    abt = torch.tensor(ab, dtype=torch.int32, device="cuda")
    ipt = torch.tensor(ipot, dtype=torch.int32, device="cuda")
    N.check(N.lib().fm_pair_histogram(1, n, P.data_ptr(), len(ab), abt.data_ptr(), ipt.data_ptr(), 3, ctypes.byref(g),
                                      hist.data_ptr(), short.data_ptr(), None), "fm_pair_histogram")
    ref, rshort = oracle.pair_histogram(pos, ab, ipot, L, pt.r0, g.dr, g.nr, pt.rmax, 3)
    err = float(np.abs(hist.cpu().numpy() / _ONE - ref).max())
    print(f"[pairs] histogram of {g.nr} bins per potential: max abs {err:.1e}", flush=True)
    assert err < 1e-8 and int(short.item()) == rshort


def test_edge_cases(tables):
    """No absorber gives an empty histogram (not every atom's); an empty configuration is refused."""
    pt, pos0, L = tables
    n = len(pos0)
    P = torch.tensor(pos0, device="cuda")
    h, short = pt.histogram(P, torch.zeros(0, dtype=torch.int64, device="cuda"), torch.ones(n, dtype=torch.int64, device="cuda"))
    assert float(h.abs().sum()) == 0.0 and short == 0
    with pytest.raises(ValueError):
        pt.single_scattering(torch.zeros(1, 0, 3, device="cuda", dtype=torch.float64))
