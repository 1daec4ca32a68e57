"""Host tables of the kernels: lambda sets, final-state channels, normalisation, rotation-recursion coefficients and
the interpolation weights, each against an independent statement."""
import math

import numpy as np
import pytest

import oracle
from feffgpu_amd import tables as T
from feffgpu_amd.pad import BOHR, read_phase_pad
from fixtures import GOLDEN


def test_setlam_sets_by_hand():
    # K edge, three or more legs (iorder 2): 2n + m <= 2, |m| <= 2, n <= 1; first the 4 with n <= 1, |m| <= 1
    m, n, lamx, laml0x, mmaxp1, nmax = T.setlam(2, 3, 1)
    assert (lamx, laml0x, mmaxp1, nmax) == (6, 4, 3, 1)
    assert sorted(zip(m[:4], n[:4])) == sorted([(0, 0), (-1, 0), (1, 0), (0, 1)])
    assert sorted(zip(m[4:], n[4:])) == [(-2, 0), (2, 0)]
    # K edge, single scattering: mmax = nmax = 1, iord 3 -> all (m, n) with |m| <= 1, n <= 1
    m, n, lamx, laml0x, _, _ = T.setlam(2, 1, 1)
    assert (lamx, laml0x) == (6, 6) and sorted(zip(m, n)) == sorted((a, b) for a in (-1, 0, 1) for b in (0, 1))
    # L3, single scattering: 15 entries (LAMTOT), m up to 2, n up to 2
    m, n, lamx, laml0x, _, _ = T.setlam(2, 1, 2)
    assert lamx == 15 and laml0x == 15 and max(n) == 2 and max(abs(v) for v in m) == 2
    # M5, single scattering: truncated at LAMTOT = 15
    m, n, lamx, _, _, _ = T.setlam(2, 1, 3)
    assert lamx == 15
    # L/M edges, three or more legs: all six in the first part
    assert T.setlam(2, 2, 2)[3] == 6


def test_kernel_lambda_sets_match_setlam():
    """The kernels' internal lambda orders (csrc/fm_device.h lamdef) hold the same sets, laml0x entries first."""
    sets = {0: ([0, -1, 1, 0, -2, 2], [0, 0, 0, 1, 0, 0], 4), 1: ([0, -1, 1, 0, -2, 2], [0, 0, 0, 1, 0, 0], 6),
            2: ([0, -1, 1, 0, -1, 1], [0, 0, 0, 1, 1, 1], 6)}
    for kind, (nsc, ilinit) in {0: (3, 1), 1: (3, 2), 2: (1, 1)}.items():
        m, n, lamx, laml0x, _, _ = T.setlam(2, nsc, ilinit)
        km, kn, kl0 = sets[kind]
        assert kl0 == laml0x and lamx == 6
        assert sorted(zip(m[:laml0x], n[:laml0x])) == sorted(zip(km[:kl0], kn[:kl0]))
        assert sorted(zip(m, n)) == sorted(zip(km, kn))
    src = (GOLDEN.parents[1] / "feffgpu_amd/csrc/fm_device.h").read_text()
    assert "LamDef{6, 6, {0, -1, 1, 0, -1, 1}, {0, 0, 0, 1, 1, 1}}" in src
    assert "LamDef{6, 4, {0, -1, 1, 0, -2, 2}, {0, 0, 0, 1, 0, 0}}" in src


def test_bcoef_channels():
    lind, bdiag = T.bcoef_unpolarized(-1)                 # K: two p channels (j = 1/2, 3/2), both l = 1
    assert [l for l in lind[1:] if l >= 0] == [1, 1]
    assert bdiag[1] == pytest.approx(-1 / 9) and bdiag[2] == pytest.approx(-1 / 9)
    lind, _ = T.bcoef_unpolarized(-2)                     # L3: d, d, s
    assert sorted(l for l in lind[1:] if l >= 0) == [0, 2, 2]
    assert T.setkap(1) == (-1, 0) and T.setkap(4) == (-2, 1) and T.setkap(9) == (-3, 2)


def test_xnlm_closed_form():
    for il in range(1, 60):
        l = il - 1
        for m in range(0, min(4, il)):
            ref = math.sqrt((2 * l + 1) * math.factorial(l - m) / math.factorial(l + m))
            assert T.xnlm(il, m) == pytest.approx(ref, rel=1e-12)
        assert T.xnlm(il, il) == 0.0


def test_rotation_recursion_reproduces_exact_d():
    """The kernels' recursion d^{l+1} = (P c - Q) d^l - R d^{l-1} with the host coefficients, seeded at l <= 2,
    against the oracle's Jacobi-polynomial values (an independent formula) up to l = 60."""
    L1 = 61
    rotc, _ = T.rotation_coefficients(L1)
    for beta in (0.3, 1.234, 2.9, math.pi - 1e-9):
        c = math.cos(beta)
        for m1 in range(-2, 3):
            for m2 in range(-2, 3):
                d = [oracle.wigner_d(l, m1, m2, beta) for l in range(3)]
                for l in range(3, L1):
                    P, Q, R = rotc[l - 1, (m1 + 3) * 7 + m2 + 3]
                    d.append((P * c - Q) * d[-1] - R * d[-2])
                for l in range(L1):
                    assert abs(d[l] - oracle.wigner_d(l, m1, m2, beta)) < 1e-12, (beta, m1, m2, l)


def test_interpolation_matches_dense_linear_matrix():
    P = read_phase_pad(GOLDEN / "W" / "phase.pad")
    kf = P.xk[:P.ne1] / BOHR
    q = np.linspace(-1.0, 45.0, 977)
    j, f = T.interpolation(kf, q)
    # the dense construction of linear_matrix (src/feffgpu/stage.py:356-363), restated
    qn = np.clip(q, kf[0], kf[-1])
    S = np.zeros((len(q), len(kf)))
    jj = np.clip(np.searchsorted(kf, qn, side="right") - 1, 0, len(kf) - 2)
    ff = (qn - kf[jj]) / (kf[jj + 1] - kf[jj])
    S[np.arange(len(q)), jj] = 1 - ff
    S[np.arange(len(q)), jj + 1] = ff
    y = np.sin(kf) * kf
    np.testing.assert_allclose((1 - f) * y[j] + f * y[j + 1], S @ y, rtol=0, atol=1e-13)
    q18 = np.arange(0, 18.0001, 0.05)
    assert T.energies_needed(P, q18) == int(T.interpolation(kf, q18)[0].max()) + 2


def test_edge_tables_shapes_and_t_matrix():
    P = read_phase_pad(GOLDEN / "validation" / "A018" / "phase.pad")
    E = T.energies_needed(P, np.arange(0, 18.0001, 0.05))
    tab = T.edge_tables(P, E)
    assert tab.tl.shape == (P.nph + 1, tab.L1 + 1, E) and tab.lmax.shape == (P.nph + 1, E)
    assert tab.chan_il == [2] and tab.ilinit == 1
    ie, iph, l = 20, 2, 3
    tm = (np.exp(2j * P.ph[ie, -l + P.ltot, 0, iph]) - 1) / 2j
    tp = (np.exp(2j * P.ph[ie, l + P.ltot, 0, iph]) - 1) / 2j
    assert tab.tl[iph, l + 1, ie] == pytest.approx(tm * (l + 1) + tp * l)
    assert (tab.lmax1 == tab.lmax.max(1) + 1).all()
