"""Pins of the FP64 oracle (oracle/feffref.c) before it is trusted as the parity reference.

Each function of the oracle is checked against data or mathematics it does not share code with:
  * phase.pad decoding + lmax/xk/ck: against the product's independent numpy reader (two parsers) and, end to end,
    through chi.dat;
  * geometry (pathgeom): against the rleg/beta columns FEFF printed in paths.dat;
  * rotation matrices: "exact" against Legendre / associated Legendre closed forms (scipy) and Wigner's explicit
    sum; "feff" (rot3i) against "exact" where FEFF's recursion is still accurate (l <= 15);
  * the whole amplitude (Hankel factors, scattering matrices, termination, chain, cfac) and the fine grid: against
    FEFF 8.5L's chi.dat for the W set and the six validation cases (R-factor, the reference's own bound
    tests/test_xsph.py:125 is 1e-9), and against the per-path feffNNNN.dat columns (|F_eff| and phase);
  * invariants: rigid rotation, cubic-cell minimum image equals the cluster result.
"""
import math

import numpy as np
import pytest

import oracle
from fixtures import CASES, GOLDEN, load_case, r_factor_vs, rmax_of
from feffgpu_amd.feffio import read_feff_dat
from feffgpu_amd.pad import BOHR, read_phase_pad


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("rotation", ["exact", "feff"])
def test_oracle_reproduces_feff_chi_dat(case, rotation):
    phase, paths, k, chi_ref = load_case(case)
    ph = oracle.OraclePhase(phase)
    acc = np.zeros(len(k))
    for p in paths:
        cc, reff, _ = ph.cchi(p.coords, p.ipots, rotation=rotation)
        acc += p.deg * ph.chi_fine(cc, reff, k)[0]
    R = r_factor_vs(k, chi_ref, acc, rmax_of(case))
    assert R < 1e-9, R


@pytest.mark.parametrize("case", CASES)
def test_two_independent_phase_readers_agree(case):
    phase = GOLDEN / case / "phase.pad"
    ph = oracle.OraclePhase(phase)
    P = read_phase_pad(phase)
    lm, xk, ck = ph.tables()
    assert (ph.ne, ph.ne1, ph.nph, ph.ihole) == (P.ne, P.ne1, P.nph, P.ihole)
    np.testing.assert_array_equal(lm, P.lmax[:P.ne1])
    np.testing.assert_allclose(xk, P.xk[:P.ne1], rtol=0, atol=1e-13)
    np.testing.assert_allclose(ck, P.ck[:P.ne1], rtol=1e-13)


@pytest.mark.parametrize("case", ["W", "validation/L001", "validation/A009"])
def test_geometry_matches_paths_dat(case):
    phase, paths, _, _ = load_case(case)
    ph = oracle.OraclePhase(phase)
    for p in paths:
        _, reff, g = ph.cchi(p.coords, p.ipots)
        # paths.dat lists the scatterers then the absorber: row j holds the leg ending at that atom and the angle
        # there; the oracle's leg j ends at atom j (scatterers 1..n-1, then the absorber)
        np.testing.assert_allclose(g["ri"], p.rleg, atol=6e-5)
        np.testing.assert_allclose(np.degrees(g["beta"]), p.beta, atol=6e-3)
        assert abs(reff - g["ri"].sum() / 2) < 1e-12


def _wigner_explicit(l, m, n, beta):
    """Wigner's explicit sum (exact rational coefficients, float powers); fine for small l."""
    f = math.factorial
    c2, s2 = math.cos(beta / 2), math.sin(beta / 2)
    pref = math.sqrt(f(l + m) * f(l - m) * f(l + n) * f(l - n))
    tot = 0.0
    for k in range(0, 2 * l + 1):
        a, b, c, d = l + n - k, k, m - n + k, l - m - k
        if min(a, b, c, d) < 0:
            continue
        tot += (-1) ** (m - n + k) * pref / (f(a) * f(b) * f(c) * f(d)) * c2 ** (2 * l + n - m - 2 * k) * s2 ** (m - n + 2 * k)
    return tot


def test_rotation_matrices_closed_forms():
    from scipy.special import eval_legendre, lpmv
    rng = np.random.default_rng(3)
    for beta in list(rng.uniform(0, math.pi, 5)) + [0.0, math.pi, 1e-7]:
        x = math.cos(beta)
        for l in (0, 1, 2, 5, 13, 30, 56):
            assert abs(oracle.wigner_d(l, 0, 0, beta) - eval_legendre(l, x)) < 1e-13
            for m in (1, 2, 3):
                if m > l or math.sin(beta) < 1e-3:      # scipy's P_l^m(cos b) loses digits through cos b near b = 0
                    continue
                # d^l_{m0} = (-1)^m sqrt((l-m)!/(l+m)!) P_l^m with scipy's Condon-Shortley P_l^m ... Edmonds sign
                ref = math.sqrt(math.exp(math.lgamma(l - m + 1) - math.lgamma(l + m + 1))) * lpmv(m, l, x)
                got = oracle.wigner_d(l, m, 0, beta)
                assert abs(abs(got) - abs(ref)) < 1e-12 * max(1.0, abs(ref))
        for l in range(0, 9):
            for m in range(-min(l, 3), min(l, 3) + 1):
                for n in range(-min(l, 3), min(l, 3) + 1):
                    assert abs(oracle.wigner_d(l, m, n, beta) - _wigner_explicit(l, m, n, beta)) < 1e-12
    # FEFF's recursion agrees with the exact values at low l (it drifts above l ~ 20)
    for l in range(0, 16):
        for m, n in ((0, 0), (1, 0), (2, 1), (1, -1), (2, -2), (0, 2)):
            if max(abs(m), abs(n)) <= l:
                b = 1.234
                assert abs(oracle.wigner_d(l, m, n, b, "feff") - oracle.wigner_d(l, m, n, b)) < 1e-10


def test_per_path_feff_columns():
    """|F_eff| and the phase FEFF printed per path (feffNNNN.dat, 4-5 digits), inverted through fdtarr's
    definitions (src/feffgpu/genfmt.py:449-461) to the oracle's cchi."""
    phase, paths, _, _ = load_case("W")
    ph = oracle.OraclePhase(phase)
    P = read_phase_pad(phase)
    E = P.ne1
    xk, ck = P.xk[:E], P.ck[:E]
    kinit = -1
    caps = P.ph[:E, -(1) + P.ltot, 0, 0]            # l = 1, kappa < 0 -> ll = -(linit + 1)
    redfac = np.exp(-2 * caps.imag)
    xlam = np.where(np.abs(ck.imag) > 1e-16, 1 / ck.imag, 1e10)
    byidx = {p.index: p for p in paths}
    files = sorted((GOLDEN / "W").glob("feff0*.dat"))
    assert len(files) >= 10
    for f in files:
        d = read_feff_dat(f)
        p = byidx[int(f.stem[4:])]
        cc, reff, _ = ph.cchi(p.coords, p.ipots)
        rb = reff / BOHR
        mag = np.abs(cc) * xk * rb ** 2 * np.exp(2 * rb / xlam) / redfac * BOHR
        n = min(len(mag), len(d["mag"]))
        sel = d["k"][:n] > 2.0
        rel = np.abs(mag[:n] - d["mag"][:n])[sel].max() / np.abs(d["mag"][:n]).max()
        assert rel < 2e-4, (f.name, rel)
        # phase: phff = phase + real_phc (mod 2 pi)
        ph_feff = d["phase"][:n] + d["real_phc"][:n]
        dph = np.angle(np.exp(1j * (np.angle(cc[:n]) - ph_feff)))[sel]
        assert np.abs(dph).max() < 2e-3, (f.name, np.abs(dph).max())


def test_invariant_rigid_rotation():
    """cchi is unchanged by a rigid rotation of a generic path. (Time reversal is not an exact invariant of the
    Rehr-Albers expansion truncated at iorder 2: a path and its reverse differ at the 1e-3 level, which FEFF
    accepts when it merges them with degeneracy 2.)"""
    phase, paths, _, _ = load_case("validation/L000")
    ph = oracle.OraclePhase(phase)
    rng = np.random.default_rng(11)
    from scipy.spatial.transform import Rotation
    for p in [q for q in paths if q.nleg >= 3][:12]:
        # displaced off the lattice: on exactly collinear legs FEFF's regularised Euler angles (trig/arg thresholds
        # at 1e-6) depend on the orientation, so the invariant is checked on generic geometry
        coords = p.coords + np.concatenate([np.zeros((1, 3)), rng.normal(0, 0.05, (p.nleg - 1, 3))])
        c0, _, _ = ph.cchi(coords, p.ipots)
        Rm = Rotation.random(random_state=rng).as_matrix()
        c1, _, _ = ph.cchi(coords @ Rm.T, p.ipots)
        assert np.abs(c1 - c0).max() < 1e-10 * np.abs(c0).max()


def test_chi_paths_cell_and_segments():
    """The batched oracle entry equals the per-path sum; a cubic cell's minimum image equals the explicit cluster."""
    phase, paths, k, _ = load_case("validation/A000")
    ph = oracle.OraclePhase(phase)
    sel = paths[:40]
    nmax = max(p.nleg for p in sel)
    pos = [np.zeros(3)]
    atoms, nleg, ipots, deg = [], [], [], []
    for p in sel:
        row = []
        for c in p.coords[1:]:
            pos.append(c)
            row.append(len(pos) - 1)
        atoms.append(row + [0] * (nmax - 1 - len(row)))
        nleg.append(p.nleg)
        ipots.append(list(p.ipots) + [0] * (nmax - p.nleg))
        deg.append(p.deg)
    pos = np.array(pos)
    seg = np.arange(len(sel)) % 3
    s1, cp1, r1 = ph_call = oracle.chi_paths(ph, pos, atoms, nleg, np.zeros(len(sel)), ipots, deg, k, segment=seg, nseg=3,
                                            want_paths=True)
    manual = np.zeros((3, len(k)))
    for i, p in enumerate(sel):
        cc, reff, _ = ph.cchi(p.coords, p.ipots)
        manual[seg[i]] += p.deg * ph.chi_fine(cc, reff, k)[0]
    np.testing.assert_allclose(s1, manual, rtol=0, atol=1e-12 * np.abs(manual).max())
    L = 40.0
    shifted = pos + L * np.round(np.random.default_rng(2).uniform(-2, 2, pos.shape))
    s2, _, _ = oracle.chi_paths(ph, shifted, atoms, nleg, np.zeros(len(sel)), ipots, deg, k, segment=seg, nseg=3, cell=L)
    np.testing.assert_allclose(s2, s1, rtol=0, atol=1e-9 * np.abs(s1).max())


def test_dense_fine_grid_matches_linear_and_spline_interpolates():
    """The oracle's dense-matrix fine grid (chi_fine_dense, any interpolation matrix S) with FEFF's linear matrix
    (src/feffgpu/stage.py:353-363) equals the oracle's own C fine grid; its spline matrix has rows summing to 1 and
    reproduces cubic polynomials exactly (an interpolating cubic spline does)."""
    phase, paths, _, _ = load_case("validation/A014")
    orc = oracle.OraclePhase(phase)
    q = np.arange(0.0, 18.0001, 0.05)
    kf = orc.tables()[1] / oracle.BOHR
    qn = np.clip(q, kf[0], kf[-1])
    j = np.clip(np.searchsorted(kf, qn, side="right") - 1, 0, orc.ne1 - 2)
    f = (qn - kf[j]) / (kf[j + 1] - kf[j])
    S = np.zeros((len(q), orc.ne1))
    S[np.arange(len(q)), j] = 1 - f
    S[np.arange(len(q)), j + 1] = f
    for p in paths[:20]:
        c, r, _ = orc.cchi(p.coords, p.ipots)
        a = orc.chi_fine(c, r, q)[0]
        b = orc.chi_fine_dense(c, r, q, S)[0]
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    Ssp = orc.spline_matrix(q)
    assert np.abs(Ssp.sum(1) - 1.0).max() < 1e-12
    for deg in (1, 2, 3):
        assert np.abs(Ssp @ (kf / 10) ** deg - (qn / 10) ** deg).max() < 1e-10
    # linearity: the matrix applied to random data is scipy's spline of that data (a transposed or shifted column fails)
    from scipy.interpolate import UnivariateSpline
    y = np.random.default_rng(4).normal(size=orc.ne1)
    assert np.abs(Ssp @ y - UnivariateSpline(kf, y, s=0, k=3)(qn)).max() < 1e-9 * np.abs(y).max()
    # the product's host-built matrices (feffgpu_amd.tables) are these
    from feffgpu_amd.tables import interpolation, spline_matrix
    assert np.abs(spline_matrix(kf, q) - Ssp).max() == 0.0
    jj, ff = interpolation(kf, q)
    assert np.array_equal(jj, j) and np.abs(ff - f).max() < 1e-15


def test_misfit_oracle_pins():
    """The oracle's window, transform and R-factor (oracle.hanning / chi_r / r_factor) against what fixes them: the
    reference's closed form for a scaled spectrum ((1/0.8 - 1)^2 = 0.0625, as tests/test_mcp.py:83 of the reference
    has it), the Hanning window at its sill centres, on its plateau and outside, the transform of a single sample
    (normalisation kstep / sqrt(pi), the e^{-i} sign, the R grid), Parseval's identity over the half spectrum, and
    the product's numpy form (feffgpu_amd.compare.r_factor, src/feffgpu/compare.py:17-27)."""
    from feffgpu_amd.compare import r_factor as r_np
    from feffgpu_amd.transform import FitWindow, r_range
    kstep, N = 0.05, 2048
    k = kstep * np.arange(361)
    w = oracle.hanning(k, 3.0, 14.0, 1.0)
    assert w[k < 2.5 - 1e-9].max() == 0.0 and w[k > 14.5 + 1e-9].max() == 0.0
    assert np.abs(w[(k > 3.5 + 1e-9) & (k < 13.5 - 1e-9)] - 1.0).max() < 1e-15
    assert np.abs(oracle.hanning(np.array([3.0, 14.0]), 3.0, 14.0, 1.0) - 0.5).max() < 1e-15
    assert np.array_equal(oracle.hanning(k, 3.0, 14.0, 0.0), ((k >= 3.0) & (k <= 14.0)).astype(float))
    rng = np.random.default_rng(11)
    data = rng.normal(size=361)
    assert abs(oracle.r_factor(data / 0.8, data, kstep, 3, 14, 1, 2, 33, 163)[0] - 0.0625) < 1e-12
    j0 = 150                                   # k = 7.5, on the plateau
    delta = np.zeros(361)
    delta[j0] = 1.0
    m = np.arange(N // 2)
    want = kstep / np.sqrt(np.pi) * k[j0] ** 2 * np.exp(-2j * np.pi * j0 * m / N)
    assert np.abs(oracle.chi_r(delta, kstep, 3, 14, 1, 2) - want).max() < 1e-13
    x = k ** 3 * oracle.hanning(k, 2.0, 16.0, 2.0) * data
    X0, XN = x.sum(), (x * (-1.0) ** np.arange(361)).sum()
    lhs = (np.abs(oracle.chi_r(data, kstep, 2, 16, 2, 3)) ** 2).sum()
    rhs = kstep ** 2 / np.pi * (N * (x ** 2).sum() + X0 ** 2 - XN ** 2) / 2
    assert abs(lhs - rhs) < 1e-12 * rhs
    model = data + 0.1 * rng.normal(size=361)
    for win in (FitWindow(3, 14, 1, 2, 1, 6), FitWindow(2, 16, 2, 3, 1.2, 5.3), FitWindow(4, 12, 0, 1, 0.5, 9)):
        m0, nm = r_range(win, kstep, N, float32_grid=False)
        a = r_np(k, data, model, kmin=win.kmin, kmax=win.kmax, rmin=win.rmin, rmax=win.rmax, kweight=win.kweight,
                 nfft=N, dk=win.dk)
        b = oracle.r_factor(model, data, kstep, win.kmin, win.kmax, win.dk, win.kweight, m0, nm, N)[0]
        assert abs(a - b) < 1e-12 * a


def test_pair_histogram_oracle_pins():
    """oracle.pair_histogram against what fixes it: every pair within the cutoff adds a mass of one; linear weights
    keep the first moment (sum_i hist_i r_i = sum of the pair distances, for pairs on the grid); a perfect bcc
    lattice puts its shells (8 at a sqrt(3)/2, 6 at a, 12 at a sqrt(2) per atom) on their bins; a cubic cell given as
    3 x 3 gives the cubic result."""
    from feffgpu_amd import workloads as W
    a = 3.0
    pos0, box = W.lattice("bcc", a, 4)
    L = float(box[0])
    n = len(pos0)
    r0, dr, rmax = 1.5, 0.002, 4.5
    nr = int(round((rmax - r0) / dr)) + 1
    h, short = oracle.pair_histogram(pos0, np.arange(n), np.ones(n, int), L, r0, dr, nr, rmax, 2)
    assert short == 0 and abs(h[0]).max() == 0.0
    for r, cnt in ((a * np.sqrt(3) / 2, 8), (a, 6), (a * np.sqrt(2), 12)):
        u = (r - r0) / dr
        i0 = int(np.floor(u))
        assert abs(h[1, i0] + h[1, i0 + 1] - cnt * n) < 1e-9 * n
    pos = W.displaced(pos0, 0.08, 3)
    rng = np.random.default_rng(2)
    ipot = rng.integers(0, 3, n)
    ab = rng.choice(n, 17, replace=False)
    h, short = oracle.pair_histogram(pos, ab, ipot, L, r0, dr, nr, rmax, 3)
    rel = np.concatenate([oracle.min_image(pos - pos[i], L) for i in ab])
    nb = np.concatenate([ipot for _ in ab])
    r = np.sqrt((rel ** 2).sum(1))
    sel = (r > 1e-3) & (r <= rmax)
    assert abs(h.sum() - sel.sum()) < 1e-9 * sel.sum()
    grid = r0 + dr * np.arange(nr)
    for p in range(3):
        on = sel & (nb == p) & (r >= r0)
        assert abs((h[p] * grid).sum() - r[on].sum()) < 1e-9 * max(r[on].sum(), 1.0)
    h2, _ = oracle.pair_histogram(pos, ab, ipot, np.eye(3) * L, r0, dr, nr, rmax, 3)
    assert np.abs(h2 - h).max() < 1e-12


def test_oracle_debye_waller_closed_form():
    """The oracle's Debye-Waller factor (not in the reference; SURVEY's a19 form chi_p(q) exp(-2 sigma_p^2 q^2)):
    per path of a list with a different sigma^2 for every path, the oracle's chi with sigma^2 equals its chi without
    times the closed form evaluated here in numpy, on a grid with negative q (an e0 shift), and the segment sums are
    the degeneracy-weighted sums of those rows. Also through the per-path fine grid (chi_fine)."""
    from feffgpu_amd import workloads as W
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    ph = oracle.OraclePhase(phase)
    pos0, box = W.lattice("bcc", 2.87, 3)
    pos = W.displaced(pos0, 0.05, 4)
    pl = W.concat([W.enumerate_paths(pos, a, np.full(len(pos), 2), 4.0, 4, box=box, segment=i, nmax=4)
                   for i, a in enumerate([0, 9])], 4)
    k = np.arange(0.0, 18.0001, 0.05)
    en = k ** 2 - 5.0 * 0.2624684236075176                    # e0 = 5 eV: q < 0 below k = 1.15
    q = np.sign(en) * np.sqrt(np.abs(en))
    s2 = np.random.default_rng(1).uniform(0.0, 0.012, len(pl))
    seg0, path0, _ = oracle.chi_paths(ph, pos, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, q,
                                      segment=pl.segment, nseg=2, cell=box[0], want_paths=True)
    seg, path, _ = oracle.chi_paths(ph, pos, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, q,
                                    segment=pl.segment, nseg=2, sigma2=s2, cell=box[0], want_paths=True)
    want = path0 * np.exp(-2.0 * s2[:, None] * q[None, :] ** 2)
    assert np.abs(path - want).max() <= 1e-13 * np.abs(path0).max()
    want_seg = np.zeros_like(seg)
    np.add.at(want_seg, pl.segment, pl.deg[:, None] * want)
    assert np.abs(seg - want_seg).max() <= 1e-12 * np.abs(want_seg).max()
    # the per-path fine grid
    from fixtures import load_case
    _, paths, _, _ = load_case("validation/A018")
    p = paths[3]
    cc, reff, _ = ph.cchi(p.coords, p.ipots)
    a = ph.chi_fine(cc, reff, q)[0]
    b = ph.chi_fine(cc, reff, q, sigma2=0.0071)[0]
    assert np.abs(b - a * np.exp(-2 * 0.0071 * q ** 2)).max() <= 1e-14 * np.abs(a).max()


def test_oracle_debye_waller_is_the_gaussian_average():
    """Physics pin of the same factor: a single-scattering path averaged over Gaussian bond lengths r ~ N(R, s^2)
    (24-point Gauss-Hermite quadrature of the oracle's sigma^2 = 0 chi) is chi(R) exp(-2 s^2 k^2), up to the
    amplitude's dependence on r (1/r^2, mean free path), which adds a relative ~2 k s^2 |d ln A / dr|. With
    s = 0.03 A and k^2 weighting over k 4..16 the closed form is within 1.1% of the average; the plausible
    mistakes (exp(-s^2 k^2), exp(-3 s^2 k^2), exp(-4 s^2 k^2), exp(+2 s^2 k^2), no factor) miss by 7% or more."""
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    ph = oracle.OraclePhase(phase)
    R, s = 2.62, 0.03
    k = np.arange(0.0, 18.0001, 0.05)
    x, w = np.polynomial.hermite_e.hermegauss(24)
    w = w / w.sum()

    def chi_at(r, sigma2=None):
        cc, reff, _ = ph.cchi(np.array([[0.0, 0.0, 0.0], [0.3 * r, -0.4 * r, np.sqrt(0.75) * r]]), np.array([0, 2]))
        return ph.chi_fine(cc, reff, k, sigma2=sigma2)[0]
    avg = sum(wi * chi_at(R + s * xi) for xi, wi in zip(x, w))
    m = (k >= 4.0) & (k <= 16.0)
    kw = k[m] ** 2

    def miss(y):
        return np.abs(kw * (avg[m] - y[m])).max() / np.abs(kw * avg[m]).max()
    assert miss(chi_at(R, sigma2=s * s)) < 0.025
    base = chi_at(R)
    for wrong in (1.0, 3.0, 4.0, -2.0, 0.0):
        assert miss(base * np.exp(-wrong * s * s * k ** 2)) > 0.05, wrong


def _refgrad():
    return np.load(GOLDEN / "refgrad" / "ref_autograd.npz")


def test_oracle_matches_reference_vectors_at_arbitrary_geometry():
    """Values written by the reference's own stage in complex128 (tests/golden/make_ref_autograd.py): exactly
    collinear and axis-aligned paths, FEFF's W list, and the 2-8-leg random walks of the reference's
    tests/test_pathfinder.py:193-213. The oracle's cchi within 1e-8 of the largest below k = 20 (the reference's own
    bound against its literal port, tests/test_pathfinder.py:229; both keep every partial wave there), its reff to
    1e-12 and its chi on the fine grid to 1e-8."""
    d = _refgrad()
    ph = oracle.OraclePhase(GOLDEN / "W" / "phase.pad")
    lm, xk, ck = ph.tables()
    low = xk / BOHR < 20.0
    q = d["q"]
    for g in d["groups"]:
        C, I = d[f"coords_{g}"], d[f"ipots_{g}"]
        for b in range(len(C)):
            cc, reff, _ = ph.cchi(C[b], I[b])
            ref = d[f"cchi_{g}"][b]
            assert np.abs(cc - ref)[low].max() <= 1e-8 * np.abs(ref[low]).max(), (g, b)
            assert abs(reff - d[f"reff_{g}"][b]) <= 1e-12 * reff
            chi = ph.chi_fine(cc, reff, q)[0]
            assert np.abs(chi - d[f"chi_{g}"][b]).max() <= 1e-8 * np.abs(d[f"chi_{g}"][b]).max(), (g, b)


def test_reference_autograd_vectors_against_oracle_differences():
    """Where the reference's autograd gradient of the cchi-level loss is finite, it equals five-point differences of
    the oracle (1e-5 of the path's largest component: at exactly collinear legs the reference's regularised
    beta = atan2(sqrt(|u1 x u2|^2 + 1e-30), b) leaves its smallest components a few 1e-7 off the differences); the
    fixture's non-finite rows are the reference's own
    (legs parallel to z: _trig's torch.where passes the 0/0 of d sqrt(x^2 + y^2) on, src/feffgpu/stage.py:147-157)
    and are counted here so that a regenerated fixture that changes them is noticed."""
    d = _refgrad()
    ph = oracle.OraclePhase(GOLDEN / "W" / "phase.pad")
    wc, wr = d["wc"], float(d["wr"])
    h = 1e-4

    def L(c, ip):
        cc, r, _ = ph.cchi(c, ip)
        return float((np.conj(wc) * cc).real.sum() + wr * r)
    nonfinite = 0
    for g in d["groups"]:
        C, I, G = d[f"coords_{g}"], d[f"ipots_{g}"], d[f"gc_{g}"]
        for b in range(0, len(C), 3):
            if not np.isfinite(G[b]).all():
                nonfinite += 1
                continue
            fd = np.zeros_like(C[b])
            for i in range(C.shape[1]):
                for k in range(3):
                    e = np.zeros_like(C[b])
                    e[i, k] = h
                    fd[i, k] = (-L(C[b] + 2 * e, I[b]) + 8 * L(C[b] + e, I[b]) - 8 * L(C[b] - e, I[b])
                                + L(C[b] - 2 * e, I[b])) / (12 * h)
            assert np.abs(fd - G[b]).max() <= 1e-5 * np.abs(fd).max(), (g, b)
    total = sum(int((~np.isfinite(d[f"gc_{g}"]).all((1, 2))).sum()) for g in d["groups"])
    assert total == 33 and nonfinite > 0
