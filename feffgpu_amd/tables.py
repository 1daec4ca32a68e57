"""Per-edge tables of the path stage, built on the host in FP64 from a phase file.

What each table restates (reference file:line):
  * setkap: initial-state kappa and l from the hole index (src/feffgpu/genfmt.py:22-38);
  * bcoef_unpolarized: final-state channels and the spin-traced diagonal coefficients (src/feffgpu/genfmt.py:230-264);
  * setlam: the lambda = (m, n) index sets (src/feffgpu/genfmt.py:115-153);
  * xnlm: Legendre normalisation sqrt((2l+1)(l-m)!/(l+m)!) (snlm, src/feffgpu/genfmt.py:41-53; FEFF's 1/64 scaling
    cancels between numerator and denominator, so it is computed here from log-factorials);
  * the j-averaged t-matrix (l+1) t(-l) + l t(l), t = (exp(2i delta) - 1)/2i (src/feffgpu/stage.py:111-119);
  * the three-term recursion coefficients of d^l_{m1 m2}(beta) (src/feffgpu/stage.py:80-86);
  * the linear interpolation from the phase grid to q (linear_matrix, src/feffgpu/stage.py:353-363).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .pad import BOHR, PhaseData

MTOT, NTOT, LAMTOT = 4, 2, 15
NREP = 20


def setkap(ihole: int) -> tuple[int, int]:
    """(kinit, linit) of the core hole (FEFF's setkap table)."""
    table = (((1, 2, 5, 10, 17, 24, 27), (-1, 0)), ((3, 6, 11, 18, 25, 30), (1, 1)), ((4, 7, 12, 19, 26), (-2, 1)),
             ((8, 13, 20, 27), (2, 2)), ((9, 14, 21, 28), (-3, 2)), ((15, 22), (3, 3)), ((16, 23), (-4, 3)))
    for holes, kl in table:
        if ihole in holes:
            return kl
    raise ValueError(f"invalid hole number {ihole}")


def bcoef_unpolarized(kinit: int, le2: int = 0, lx: int = 4):
    """lind[1..8] (final-state l per channel, -1 unphysical; index 0 unused) and the diagonal bdiag[1..8] for
    ipol = 0 traced over spin (factor 2)."""
    lind = [-1] * 9
    for k in (-1, 0, 1):
        kap = -(kinit + k) if k == 0 else kinit + k
        lk = kap if kap > 0 else -kap - 1
        lind[k + 2] = lk if lk <= lx else -1
    for k in (-2, -1, 0, 1, 2):
        jk = max(abs(kinit) + k, 0)
        kap = jk
        if kinit < 0 and abs(k) != 1:
            kap = -jk
        if kinit > 0 and abs(k) == 1:
            kap = -jk
        lk = kap if kap > 0 else -kap - 1
        if lk > lx or le2 == 0 or (le2 == 1 and abs(k) == 2):
            lk = -1
        lind[k + 6] = lk
    bdiag = [0.0] * 9
    for k in range(1, 9):
        if lind[k] < 0:
            continue
        i2 = 5 if (le2 == 2 and k > 3) else 3
        v = 0.5 / (2 * lind[k] + 1.0) / i2
        bdiag[k] = 2.0 * (-v if k <= 3 else v)
    return lind, bdiag


def setlam(iorder: int, nsc: int, ilinit: int):
    """lambda set: (m list, n list, lamx, laml0x, mmaxp1, nmax), lists 0-based, first laml0x entries first."""
    if nsc == 1:
        mmax = nmax = ilinit
        iord = 2 * nmax + mmax
    elif iorder < 10:
        iord, mmax, nmax = iorder, iorder, iorder // 2
    else:
        raise NotImplementedError("icalc >= 10 not ported")
    ms, ns = [], []
    for n in range(nmax + 1):
        for m in range(mmax + 1):
            if 2 * n + m > iord:
                continue
            for sm in ((0,) if m == 0 else (-m, m)):
                if len(ms) < LAMTOT:
                    ms.append(sm)
                    ns.append(n)
    first = [i for i in range(len(ms)) if ns[i] <= ilinit and abs(ms[i]) <= ilinit]
    rest = [i for i in range(len(ms)) if i not in first]
    order = first + rest
    m = [ms[i] for i in order]
    n = [ns[i] for i in order]
    mmaxp1 = max(v + 1 for v in m)
    if max(n) > NTOT or mmaxp1 > MTOT + 1:
        raise ValueError("lambda set exceeds mtot/ntot")
    return m, n, len(m), len(first), mmaxp1, max(n)


def xnlm(il: int, m: int) -> float:
    """sqrt((2l+1)(l-m)!/(l+m)!), l = il - 1, for 0 <= m <= l; 0 otherwise."""
    l = il - 1
    if m < 0 or m > l:
        return 0.0
    return math.sqrt((2 * l + 1) * math.exp(math.lgamma(l - m + 1) - math.lgamma(l + m + 1)))


def rotation_coefficients(L1: int):
    """rotc[l, (m1+3)*7 + m2+3, :] = (P, Q, R) with d^{l+1} = (P cos(beta) - Q) d^l - R d^{l-1}; rotn[il, |m1|*4+|m2|]
    = xnlm(il, |m1|) (-1)^|m1| / xnlm(il, |m2|) (0 where either is 0)."""
    rotc = np.zeros((L1 + 1, 49, 3))
    for l in range(L1 + 1):
        for m1 in range(-3, 4):
            for m2 in range(-3, 4):
                den = l * math.sqrt(max(((l + 1) ** 2 - m1 ** 2) * ((l + 1) ** 2 - m2 ** 2), 0))
                if den == 0:
                    den = 1.0
                a = (l + 1) * math.sqrt(max(l * l - m1 * m1, 0) * max(l * l - m2 * m2, 0))
                k = (m1 + 3) * 7 + m2 + 3
                rotc[l, k] = ((2 * l + 1) * l * (l + 1) / den, (2 * l + 1) * m1 * m2 / den, a / den)
    rotn = np.zeros((L1 + 1, 16))
    for il in range(1, L1 + 1):
        for ma in range(4):
            for mb in range(4):
                xa, xb = xnlm(il, ma), xnlm(il, mb)
                if xa != 0.0 and xb != 0.0:
                    rotn[il, ma * 4 + mb] = xa * (-1) ** ma / xb
    return rotc, rotn


def spline_matrix(kf: np.ndarray, q) -> np.ndarray:
    """Cubic-spline interpolation matrix (len(q), len(kf)) from the phase grid kf (1/A) to q, q clamped to the grid:
    column j is scipy's interpolating UnivariateSpline (s = 0, k = 3, as Larch) of the j-th unit vector
    (src/feffgpu/stage.py:341-351)."""
    from scipy.interpolate import UnivariateSpline
    kf = np.asarray(kf, float)
    qn = np.clip(np.asarray(q, float), kf[0], kf[-1])
    S = np.zeros((len(qn), len(kf)))
    for j in range(len(kf)):
        y = np.zeros(len(kf))
        y[j] = 1.0
        S[:, j] = UnivariateSpline(kf, y, s=0, k=3)(qn)
    return S


def interpolation(kf_full: np.ndarray, q: np.ndarray):
    """Lower index j (clipped to 0..E-2) and upper weight f of the linear interpolation of q on the ascending grid
    kf_full (linear_matrix's rule, src/feffgpu/stage.py:356-362: q clamped to the grid, searchsorted right - 1)."""
    kf = np.asarray(kf_full, float)
    q = np.asarray(q, float)
    qn = np.clip(q, kf[0], kf[-1])
    j = np.clip(np.searchsorted(kf, qn, side="right") - 1, 0, len(kf) - 2)
    f = (qn - kf[j]) / (kf[j + 1] - kf[j])
    return j.astype(np.int32), f


@dataclass
class EdgeTables:
    """Host (numpy) form of the device tables of one edge, for the first E energies of the phase grid."""
    E: int
    L1: int
    nph1: int
    ilinit: int
    kinit: int
    lmax: np.ndarray            # (nph1, E) int32
    lmax1: np.ndarray           # (nph1,) int32
    ck: np.ndarray              # (E,) complex
    ckmxk: np.ndarray           # (E,) complex
    kf: np.ndarray              # (E,) 1/A
    kf_full: np.ndarray         # (ne1,) 1/A
    tl: np.ndarray              # (nph1, L1+1, E) complex
    chan_il: list = field(default_factory=list)
    chw: np.ndarray = None      # (nchan, E) complex
    rotc: np.ndarray = None
    rotn: np.ndarray = None
    bands: list = field(default_factory=list)   # band plans for 1, 2, 4, 8 paths per workgroup (see energy_bands)


MAX_BANDS = 16
WAVES = 4
BAND_OVERHEAD = 6.0   # per-lane work of an energy outside the partial-wave rows (termination, chain, cfac), in rows


def _splits(cost, cap: int, nbmax: int):
    """Contiguous splits of the energies into nb = 1..nbmax bands of at most cap energies, each minimising the sum
    over bands of the band's largest cost (a wavefront runs a band for as long as its most expensive lane).
    Returns {nb: starts} for the nb that admit a split."""
    E = len(cost)
    # M[i, j] = max(cost[i:j]) for 0 < j - i <= cap, inf otherwise
    M = np.full((E + 1, E + 1), np.inf)
    for i in range(E):
        hi = min(E, i + cap)
        M[i, i + 1:hi + 1] = np.maximum.accumulate(cost[i:hi])
    best = np.full(E + 1, np.inf)
    best[0] = 0.0
    cuts, out = [], {}
    for nb in range(1, nbmax + 1):
        tot = best[:, None] + M                       # (i, j): bands so far ending at i, then [i, j)
        cut = np.argmin(tot, axis=0)                  # first minimum: deterministic
        best = tot[cut, np.arange(E + 1)]
        cuts.append(cut)
        if np.isfinite(best[E]):
            starts, j = [E], E
            for c in reversed(cuts):
                j = int(c[j])
                starts.append(j)
            out[nb] = starts[::-1]
    return out


def _assign(costs, nwave: int = WAVES):
    """Longest-processing-time-first assignment of the bands to the wavefronts of a workgroup: (wave of each band,
    the largest wave total). Ties go to the lowest wave index (deterministic)."""
    load = [0.0] * nwave
    wave = [0] * len(costs)
    for b in sorted(range(len(costs)), key=lambda b: (-costs[b], b)):
        w = min(range(nwave), key=lambda w: (load[w], w))
        wave[b] = w
        load[w] += costs[b]
    return wave, max(load)


def energy_bands(profile: np.ndarray):
    """Band plans of a workgroup of WAVES wavefronts holding gp = 1, 2, 4 or 8 paths: the energies are split into at
    most MAX_BANDS contiguous bands of at most 64 / gp energies (lane (path j, energy i) = j * (64 / gp) + i), and
    each band is run by one wavefront, chosen so that the wavefronts' totals are balanced. profile[e] is the number
    of partial-wave rows at energy e. Returns [(gp, starts, waves)] for gp = 1, 2, 4, 8; starts = [] where the
    energies do not fit."""
    E = len(profile)
    cost = np.asarray(profile, float) + BAND_OVERHEAD
    plans = []
    for gp in (1, 2, 4, 8):
        cap = 64 // gp
        best = None
        for nb, st in sorted(_splits(cost, cap, MAX_BANDS).items()):
            bc = [float(cost[st[b]:st[b + 1]].max()) for b in range(nb)]
            wave, span = _assign(bc)
            if best is None or span < best[0] - 1e-9:
                best = (span, st, wave)
        plans.append((gp, best[1], best[2]) if best else (gp, [], []))
    return plans


def edge_tables(P: PhaseData, E: int | None = None) -> EdgeTables:
    if P.nsp != 1:
        raise NotImplementedError("spin-dependent phase files are not ported")
    E = int(P.ne1 if E is None else E)
    if not 2 <= E <= P.ne1:
        raise ValueError(f"E must be within 2..{P.ne1}")
    lm = P.lmax[:E].astype(np.int32)                     # (E, nph1)
    L1 = int(lm.max()) + 1
    nph1 = P.nph + 1
    kinit, linit = setkap(P.ihole)
    if linit + 1 > 3:
        raise NotImplementedError("edges with an initial state l > 2 (N edges) are not supported")
    ck = P.ck[:E]
    xk = P.xk[:E]
    tl = np.zeros((nph1, L1 + 1, E), complex)
    for iph in range(nph1):
        for il in range(1, L1 + 1):
            l = il - 1
            tm = (np.exp(2j * P.ph[:E, -l + P.ltot, 0, iph]) - 1) / 2j
            tp = (np.exp(2j * P.ph[:E, l + P.ltot, 0, iph]) - 1) / 2j
            tl[iph, il] = tm * (l + 1) + tp * l
    lind, bdiag = bcoef_unpolarized(kinit)
    ls = sorted({lind[k] for k in range(1, 9) if lind[k] >= 0})
    chw = np.zeros((len(ls), E), complex)
    for c, l in enumerate(ls):
        for k in range(1, 9):
            if lind[k] == l:
                chw[c] += bdiag[k] * P.rkk[:E, k - 1, 0] ** 2
    rotc, rotn = rotation_coefficients(L1)
    return EdgeTables(E=E, L1=L1, nph1=nph1, ilinit=linit + 1, kinit=kinit, lmax=np.ascontiguousarray(lm.T),
                      lmax1=(lm.max(0) + 1).astype(np.int32), ck=ck, ckmxk=ck - xk, kf=xk / BOHR,
                      kf_full=P.xk[:P.ne1] / BOHR, tl=tl, chan_il=[l + 1 for l in ls], chw=chw, rotc=rotc, rotn=rotn,
                      bands=energy_bands(lm.max(1) + 1))


def energies_needed(P: PhaseData, q) -> int:
    """Phase-grid energies the fine-grid evaluation on q reads: through the upper interpolation point of the
    largest q (the unwrapped phase is a running sum, so every lower energy is needed too)."""
    j, _ = interpolation(P.xk[:P.ne1] / BOHR, q)
    return int(min(P.ne1, int(j.max()) + 2))
