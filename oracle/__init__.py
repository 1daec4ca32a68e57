"""FP64 CPU oracle of the path stage: TEST INFRASTRUCTURE ONLY.

Only the test suite, the smoke check and the benchmark's CPU-baseline leg may import this package, and only as the
checker (or the timed CPU baseline), never as the thing measured or shipped. The product (``feffgpu_amd``) never
imports it.

``liborc.so`` (built from ``feffref.c`` by ``make -C oracle``) is an FP64 restatement of feffgpu's literal FEFF
port ``genfmt.PathStage`` (src/feffgpu/genfmt.py:279-461) with the leg geometry and fine-grid evaluation of
``GPUPathStage`` (src/feffgpu/stage.py:147-206, 353-388). It parses phase.pad itself, so it shares no code with the
product. What pins each function is listed in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"
_lib = None
ROTMODE = {"feff": 0, "exact": 1}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build(force: bool = False) -> Path:
    """Compile liborc.so with gcc (make -C oracle)."""
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "feffref.c").stat().st_mtime:
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_info.argtypes = [ctypes.c_void_p, _ip]
        L.orc_tables.argtypes = [ctypes.c_void_p, _ip, _dp, _dp]
        L.orc_cchi.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _lp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_wigner.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_phase_shifts.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip]
        L.orc_chi_fine.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.orc_chi_paths.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _ip,
                                    _ip, _ip, _lp, _dp, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


BOHR = 0.52917721067   # A, as FEFF and the reference (src/feffgpu/pad.py:13)


class OraclePhase:
    """A phase.pad file loaded by the oracle's own parser."""

    def __init__(self, path):
        self.path = str(path)
        h = ctypes.c_void_p()
        _check(lib().orc_load(self.path.encode(), ctypes.byref(h)))
        self.h = h
        info = np.zeros(8, np.int32)
        lib().orc_info(self.h, _p(info, _ip))
        self.ne, self.ne1, self.nph, self.ihole, self.ltot, self.lmaxp1, self.kinit, self.linit = (int(v) for v in info)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_free(self.h)
            self.h = None

    def tables(self):
        """lmax (ne1, nph+1), xk (ne1,) 1/bohr, ck (ne1,) complex, as the oracle derives them."""
        lm = np.zeros((self.ne1, self.nph + 1), np.int32)
        xk = np.zeros(self.ne1)
        ck = np.zeros(2 * self.ne1)
        lib().orc_tables(self.h, _p(lm, _ip), _p(xk, _dp), _p(ck, _dp))
        return lm, xk, ck[0::2] + 1j * ck[1::2]

    def phase_shifts(self, sign=-1):
        """Phase shifts of spin 0, (ne1, ltot + 1, nph + 1) complex: ph(ie, -l, iph) for sign < 0 (as FEFF's path
        finder tabulates them), ph(ie, l, iph) otherwise; and ik0 (1-based, as in phase.pad)."""
        out = np.zeros((self.ne1, self.ltot + 1, self.nph + 1, 2))
        ik0 = np.zeros(1, np.int32)
        lib().orc_phase_shifts(self.h, int(sign), _p(out, _dp), _p(ik0, _ip))
        return out[..., 0] + 1j * out[..., 1], int(ik0[0])

    def spline_matrix(self, q):
        """Cubic-spline interpolation matrix (len(q), ne1) from the phase grid (1/A) to q, q clamped to the grid:
        scipy's interpolating UnivariateSpline (s = 0, k = 3) of each unit vector, as Larch and the reference's
        spline_matrix build it (src/feffgpu/stage.py:341-351)."""
        from scipy.interpolate import UnivariateSpline
        kf = self.tables()[1] / BOHR
        qn = np.clip(np.asarray(q, float), kf[0], kf[-1])
        S = np.zeros((len(qn), self.ne1))
        for j in range(self.ne1):
            y = np.zeros(self.ne1)
            y[j] = 1.0
            S[:, j] = UnivariateSpline(kf, y, s=0, k=3)(qn)
        return S

    @staticmethod
    def chi_fine_dense(cchi, reff, q, S):
        """chi on the points q with a dense interpolation matrix S (len(q), E) of amplitude and unwrapped phase
        (src/feffgpu/stage.py:372-388): amp = sqrt(|cchi|^2 + 1e-30), the phase unwrapped with its successive
        differences brought to (-pi, pi] by rounding half to even, chi = (amp S^T) sin(phase S^T + 2 q reff).
        cchi (B, E) or (E,), reff (B,) or a number; returns (B, len(q))."""
        c = np.atleast_2d(np.asarray(cchi))
        r = np.broadcast_to(np.asarray(reff, float), (c.shape[0],))
        amp = np.sqrt(c.real ** 2 + c.imag ** 2 + 1e-30)
        ang = np.arctan2(c.imag, c.real)
        d = np.diff(ang, axis=-1)
        d = d - 2 * np.pi * np.round(d / (2 * np.pi))
        pha = np.concatenate([ang[:, :1], ang[:, :1] + np.cumsum(d, -1)], -1)
        S = np.asarray(S, float)
        return (amp @ S.T) * np.sin(pha @ S.T + 2 * np.asarray(q, float)[None, :] * r[:, None])

    def cchi(self, coords, ipots, iorder=2, rotation="exact"):
        """One path: coords (nleg, 3) A, absorber first at the origin; ipots (nleg,). Returns (cchi (ne1,) complex,
        reff A, geometry dict with ri (A), beta, eta per leg). rotation "exact" (Wigner d through Jacobi
        polynomials: the batched stage's values) or "feff" (FEFF's rot3i recursion, as in chi.dat)."""
        c = np.ascontiguousarray(coords, dtype=np.float64)
        ip = np.ascontiguousarray(ipots, dtype=np.int64)
        n = c.shape[0]
        out = np.zeros(2 * self.ne1)
        reff = np.zeros(1)
        geom = np.zeros(3 * n)
        _check(lib().orc_cchi(self.h, n, _p(c, _dp), _p(ip, _lp), int(iorder), ROTMODE[rotation], _p(out, _dp), _p(reff, _dp),
                             _p(geom, _dp)))
        return out[0::2] + 1j * out[1::2], float(reff[0]), dict(ri=geom[:n], beta=geom[n:2 * n], eta=geom[2 * n:])

    def chi_fine(self, cchi, reff, q, sigma2=None):
        cc = np.ascontiguousarray(np.stack([np.asarray(cchi).real, np.asarray(cchi).imag], -1), dtype=np.float64)
        cc = cc.reshape(-1, self.ne1, 2)
        B = cc.shape[0]
        rf = np.ascontiguousarray(np.broadcast_to(np.asarray(reff, np.float64), (B,)))
        s2 = None if sigma2 is None else np.ascontiguousarray(np.broadcast_to(np.asarray(sigma2, np.float64), (B,)))
        qq = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros((B, len(qq)))
        _check(lib().orc_chi_fine(self.h, B, _p(cc, _dp), _p(rf, _dp), _p(s2, _dp), len(qq), _p(qq, _dp), _p(out, _dp)))
        return out


def wigner_d(l, m1, m2, beta, rotation="exact"):
    out = np.zeros(1)
    _check(lib().orc_wigner(ROTMODE[rotation], int(l), int(m1), int(m2), float(beta), _p(out, _dp)))
    return float(out[0])


def chi_paths(phases, pos, atoms, nleg, absorber, ipots, deg, q, segment=None, nseg=1, edge=None, sigma2=None,
              cell=None, iorder=2, threads=None, want_paths=False, rotation="exact"):
    """The fused path-list evaluation in FP64 (see orc_chi_paths in feffref.c). Returns (chi_seg (nseg, NK),
    chi_path (P, NK) or None, reff (P,))."""
    if not isinstance(phases, (list, tuple)):
        phases = [phases]
    P = int(len(nleg))
    pos = np.ascontiguousarray(pos, np.float64)
    atoms = np.ascontiguousarray(atoms, np.int32)
    nmax = atoms.shape[1] + 1
    nleg = np.ascontiguousarray(nleg, np.int32)
    absorber = np.ascontiguousarray(absorber, np.int32)
    ipots = np.ascontiguousarray(ipots, np.int64)
    deg = np.ascontiguousarray(deg, np.float64)
    segment = np.zeros(P, np.int32) if segment is None else np.ascontiguousarray(segment, np.int32)
    edge = None if edge is None else np.ascontiguousarray(edge, np.int32)
    sigma2 = None if sigma2 is None else np.ascontiguousarray(np.broadcast_to(np.asarray(sigma2, np.float64), (P,)))
    q = np.ascontiguousarray(q, np.float64)
    if cell is None:
        cellv, ncell = None, 0
    else:
        cellv = np.ascontiguousarray(np.asarray(cell, np.float64).reshape(-1))
        ncell = cellv.size
        if ncell not in (1, 9):
            raise ValueError("cell must be a cubic length or a 3 x 3 matrix")
    handles = (ctypes.c_void_p * len(phases))(*[ph.h.value for ph in phases])
    chi_seg = np.zeros((nseg, len(q)))
    chi_path = np.zeros((P, len(q))) if want_paths else None
    reff = np.zeros(P)
    nthr = int(threads or os.cpu_count() or 1)
    _check(lib().orc_chi_paths(handles, len(phases), P, nmax, _p(pos, _dp), _p(atoms, _ip), _p(nleg, _ip), _p(absorber, _ip),
                               _p(ipots, _lp), _p(deg, _dp), _p(segment, _ip), _p(edge, _ip), _p(sigma2, _dp), int(nseg),
                               len(q), _p(q, _dp), _p(cellv, _dp), ncell, int(iorder), ROTMODE[rotation], nthr, _p(chi_seg, _dp),
                               _p(chi_path, _dp), _p(reff, _dp)))
    return chi_seg, chi_path, reff


# ------------------------------------------------------------------------------------------------ misfit after the sum
def hanning(k, kmin, kmax, dk=1.0):
    """xraylarch's Hanning window (src/feffgpu/compare.py:6-14, src/feffgpu/exafs/transform.py:71-94): sills centered
    on kmin and kmax, dk wide; dk <= 0 the rectangular window."""
    k = np.asarray(k, float)
    if dk <= 0:
        return ((k >= kmin) & (k <= kmax)).astype(float)
    a = np.clip((k - (kmin - dk / 2)) / dk, 0.0, 1.0)
    b = np.clip((k - (kmax - dk / 2)) / dk, 0.0, 1.0)
    return 0.5 * (1.0 - np.cos(np.pi * a)) * 0.5 * (1.0 + np.cos(np.pi * b))


def chi_r(chi, kstep, kmin, kmax, dk, kweight, nfft=2048):
    """Complex chi(R) at R_m = m pi / (nfft kstep), m < nfft / 2, of chi (..., NK) on k_j = j kstep: numpy's FFT of
    k^w chi window zero-padded to nfft, times kstep / sqrt(pi) (src/feffgpu/exafs/transform.py:104-122)."""
    chi = np.asarray(chi, float)
    k = kstep * np.arange(chi.shape[-1])
    x = chi * k ** kweight * hanning(k, kmin, kmax, dk)
    return np.fft.fft(x, n=nfft, axis=-1)[..., : nfft // 2] * kstep / np.sqrt(np.pi)


def r_factor(chi, data, kstep, kmin, kmax, dk, kweight, m0, nm, nfft=2048, scale=1.0):
    """Per row of chi (S, NK): sum |M - D|^2 / sum |D|^2 over R points m0 .. m0 + nm - 1, M the transform of
    scale * chi, D of data ((NK,) or (S, NK)) (src/feffgpu/exafs/transform.py:130-145, src/feffgpu/compare.py:17-27)."""
    M = chi_r(scale * np.atleast_2d(chi), kstep, kmin, kmax, dk, kweight, nfft)[:, m0:m0 + nm]
    D = chi_r(np.atleast_2d(data), kstep, kmin, kmax, dk, kweight, nfft)[:, m0:m0 + nm]
    return (np.abs(M - D) ** 2).sum(1) / (np.abs(D) ** 2).sum(1)


# ------------------------------------------------------------------------------------------------ pair histograms
def min_image(rel, cell):
    """Minimum image of displacements rel (..., 3) (src/feffgpu/cell.py:26-35): cell None, a cubic length or 3 x 3
    (rows the lattice vectors)."""
    rel = np.asarray(rel, float)
    if cell is None:
        return rel
    c = np.asarray(cell, float)
    if c.ndim == 0:
        return rel - float(c) * np.round(rel / float(c))
    f = rel @ np.linalg.inv(c)
    return (f - np.round(f)) @ c


def pair_histogram(pos, absorbers, ipot, cell, r0, dr, nr, rmax, npot):
    """PairTables.histogram (src/feffgpu/exafs/forward.py:233-280) in numpy: every pair (absorber a, neighbor j)
    within rmax and not coincident adds 1 - f to bin i0 and f to bin i0 + 1 of row ipot[j]. Returns (hist
    (npot, nr), pairs closer than r0)."""
    pos = np.asarray(pos, float)
    ipot = np.asarray(ipot)
    hist = np.zeros((npot, nr))
    short = 0
    for a in np.asarray(absorbers).reshape(-1):
        rel = min_image(pos - pos[a], cell)
        r2 = (rel ** 2).sum(1)
        m = (r2 <= rmax ** 2) & (r2 > 1e-6)
        j = np.nonzero(m)[0]
        r = np.sqrt(r2[m])
        short += int((r < r0).sum())
        u = (r - r0) / dr
        i0 = np.clip(np.floor(u), 0, nr - 2).astype(int)
        f = np.clip(u - i0, 0.0, 1.0)
        p = ipot[j]
        keep = (p >= 0) & (p < npot)
        np.add.at(hist, (p[keep], i0[keep]), 1.0 - f[keep])
        np.add.at(hist, (p[keep], i0[keep] + 1), f[keep])
    return hist, short
