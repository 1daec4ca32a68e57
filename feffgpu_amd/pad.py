"""FEFF 8.5L ``phase.pad`` reader: the phase shifts, energy grid and matrix elements the path stage consumes.

Follows the record layout that feffgpu's reader documents (src/feffgpu/pad.py:124-151) and the packed-ASCII
number format of FEFF's padlib (src/feffgpu/pad.py:12-27): each number is ``npack`` printable characters, the
first carrying the decimal exponent, the second the sign and the leading base-90 digit, the rest further base-90
digits. Lines of real numbers start with ``!``, lines of complex numbers with ``$`` (real and imaginary part
back to back). Decoding here is vectorised with numpy over whole records.

Units are FEFF's: energies in Hartree, momenta in 1/bohr.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

BOHR = 0.52917721067          # A per bohr (src/feffgpu/pad.py:13)
HART = 2.0 * 13.60569301      # eV per Hartree (src/feffgpu/pad.py:14)
PHMIN = 1.0e-7                # |sin(delta_l)| below this is a negligible phase (src/feffgpu/pad.py:15)

_BASE, _OFFSET, _EXPBIAS = 90, 37, 45


def _decode(codes: np.ndarray) -> np.ndarray:
    """Decode packed numbers. ``codes`` is (n, npack) uint8 of printable characters; returns (n,) float64."""
    c = codes.astype(np.int64) - _OFFSET
    npack = c.shape[1]
    expo = c[:, 0] - _EXPBIAS
    sign = np.where(c[:, 1] % 2 == 1, 1.0, -1.0)
    mant = (c[:, 1] // 2).astype(np.float64) / float(_BASE) ** 2
    for i in range(3, npack + 1):                                  # digit at position i-1 weighs 90^-i
        mant = mant + c[:, i - 1].astype(np.float64) / float(_BASE) ** i
    return 2.0 * sign * _BASE * mant * np.power(10.0, expo.astype(np.float64))


class _Records:
    """Sequential access to the lines of a pad file."""

    def __init__(self, text: str):
        self._lines = text.splitlines()
        self._pos = 0

    def line(self) -> str:
        if self._pos >= len(self._lines):
            raise ValueError("phase file ends early")
        s = self._lines[self._pos]
        self._pos += 1
        return s

    def _packed(self, marker: str, width: int, count: int) -> np.ndarray:
        chunks, have = [], 0
        while have < count:
            s = self.line().rstrip()
            if not s or s[0] != marker:
                raise ValueError(f"expected a packed line starting with {marker!r}, got {s[:20]!r}")
            body = s[1:].encode("latin-1")
            n = min(len(s) // width, count - have)
            if n <= 0:
                continue
            chunks.append(np.frombuffer(body[: n * width], dtype=np.uint8).reshape(n, width))
            have += n
        return np.concatenate(chunks) if chunks else np.zeros((0, width), np.uint8)

    def reals(self, npack: int, count: int) -> np.ndarray:
        return _decode(self._packed("!", npack, count))

    def complexes(self, npack: int, count: int) -> np.ndarray:
        raw = self._packed("$", 2 * npack, count)
        return _decode(raw[:, :npack]) + 1j * _decode(raw[:, npack:])


@dataclass
class PhaseData:
    """Content of a phase.pad file; field names as feffgpu's ``PhaseData`` (src/feffgpu/pad.py:69-91)."""
    nsp: int
    ne: int
    ne1: int
    ne3: int
    nph: int
    ihole: int
    ik0: int
    ixc: int
    rs: float
    vint: float
    rnrmav: float
    xmu: float
    edge: float
    em: np.ndarray            # (ne,) complex energies, Hartree
    eref: np.ndarray          # (ne, nsp) complex energy reference
    lmax0: np.ndarray         # (nph+1,) partial waves stored per potential
    iz: np.ndarray            # (nph+1,) atomic numbers
    potlbl: list = field(default_factory=list)
    ph: np.ndarray = None     # (ne, 2 ltot + 1, nsp, nph+1) complex phase shifts, index l + ltot
    rkk: np.ndarray = None    # (ne, 8, nsp) complex reduced matrix elements
    ltot: int = 0

    @property
    def lmax(self) -> np.ndarray:
        """(ne, nph+1): the largest l <= lmax0 whose phase shift has |sin| > PHMIN in either spin, else 0
        (the per-energy partial-wave limit of src/feffgpu/pad.py:93-106)."""
        out = np.zeros((self.ne, self.nph + 1), dtype=np.int64)
        for iph in range(self.nph + 1):
            l0 = int(self.lmax0[iph])
            ls = np.arange(l0 + 1)
            s0 = np.abs(np.sin(self.ph[:, ls + self.ltot, 0, iph]))
            s1 = np.abs(np.sin(self.ph[:, ls + self.ltot, self.nsp - 1, iph]))
            big = (s0 > PHMIN) | (s1 > PHMIN)                        # (ne, l0+1)
            # the largest l with a big phase; l = 0 when none is big
            last = np.where(big.any(1), l0 - np.argmax(big[:, ::-1], axis=1), 0)
            out[:, iph] = last
        return out

    def phase(self, ie: int, l: int, iph: int, isp: int = 0) -> complex:
        return self.ph[ie, l + self.ltot, isp, iph]

    @property
    def xk(self) -> np.ndarray:
        """Real momentum relative to the edge, sign(E - edge) sqrt(2 |E - edge|), 1/bohr (src/feffgpu/pad.py:111-115)."""
        e = self.em.real - self.edge
        return np.sign(e) * np.sqrt(np.abs(2.0 * e))

    @property
    def ck(self) -> np.ndarray:
        """Complex momentum sqrt(2 (E - E_ref)), spin-averaged reference, 1/bohr (src/feffgpu/pad.py:117-121)."""
        eref = self.eref[:, 0] if self.nsp == 1 else 0.5 * (self.eref[:, 0] + self.eref[:, -1])
        return np.sqrt(2.0 * (self.em - eref))


def read_phase_pad(filename) -> PhaseData:
    """Read a FEFF 8.5L phase.pad (record order of src/feffgpu/pad.py:124-151)."""
    text = Path(filename).read_text(encoding="latin-1")
    rec = _Records(text)
    head = rec.line()
    nsp, ne, ne1, ne3, nph, ihole, ik0, npack, ixc = (int(head[1 + 5 * j:5 + 5 * j]) for j in range(9))
    rs, vint = float(head[45:56]), float(head[56:67])
    rnrmav, xmu, edge = rec.reals(npack, 3)
    em = rec.complexes(npack, ne)
    eref = rec.complexes(npack, ne * nsp).reshape(nsp, ne).T.copy()
    lmax0, iz, labels, per_pot = [], [], [], []
    for _ in range(nph + 1):
        s = rec.line()
        l0 = int(s[1:4])
        lmax0.append(l0)
        iz.append(int(s[5:8]))
        labels.append(s[9:15].strip())
        spins = [rec.complexes(npack, ne * (2 * l0 + 1)).reshape(ne, 2 * l0 + 1) for _ in range(nsp)]
        per_pot.append(np.stack(spins, -1))                              # (ne, 2 l0 + 1, nsp)
    ltot = max(lmax0)
    ph = np.zeros((ne, 2 * ltot + 1, nsp, nph + 1), complex)
    for iph, block in enumerate(per_pot):
        lo = ltot - lmax0[iph]
        ph[:, lo:lo + block.shape[1], :, iph] = block
    rkk = rec.complexes(npack, ne * 8 * nsp).reshape(nsp, 8, ne).transpose(2, 1, 0).copy()
    return PhaseData(nsp, ne, ne1, ne3, nph, ihole, ik0, ixc, rs, vint, float(rnrmav), float(xmu), float(edge),
                     em, eref, np.array(lmax0), np.array(iz), labels, ph, rkk, ltot)


def load_phase(phase) -> PhaseData:
    """A ``PhaseData`` from itself or from a phase.pad path. Pickled phase files are not loaded: they can run code."""
    if isinstance(phase, PhaseData) or (hasattr(phase, "ph") and hasattr(phase, "iz")):
        return phase
    p = Path(phase)
    if not p.is_file():
        raise FileNotFoundError(f"no such phase file: {p}")
    return read_phase_pad(p)
