"""Readers for FEFF 8.5L's text outputs that carry path-stage data: ``paths.dat`` (the path list), ``chi.dat``
(the summed chi(k)) and ``feffNNNN.dat`` (per-path amplitude and phase columns).

``paths.dat`` lists each path's scatterers in order with the absorber last (at the origin, potential 0); the path
stage wants the absorber first, so :func:`read_paths_dat` returns coordinates ``[0, s_1, ..., s_{n-1}]`` and
potentials ``[0, ipot_1, ...]`` exactly as feffgpu's own W test builds them (tests/test_xsph.py:93-105 of the
reference).
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class FeffPath:
    index: int
    nleg: int
    deg: float
    coords: np.ndarray        # (nleg, 3) A, absorber first at the origin
    ipots: np.ndarray         # (nleg,) absorber first (0)
    rleg: np.ndarray          # (nleg,) A, leg ending at each listed atom (FEFF's order: scatterers, then absorber)
    beta: np.ndarray          # (nleg,) degrees, scattering angle at each listed atom
    eta: np.ndarray           # (nleg,) degrees


def read_paths_dat(filename) -> list[FeffPath]:
    lines = Path(filename).read_text(encoding="latin-1").splitlines()
    out, i = [], 0
    while i < len(lines):
        if "index, nleg" not in lines[i]:
            i += 1
            continue
        w = lines[i].split()
        index, nleg, deg = int(w[0]), int(w[1]), float(w[2])
        rows = [lines[i + 2 + j].split() for j in range(nleg)]
        xyz = np.array([[float(v) for v in r[:3]] for r in rows])
        ip = np.array([int(r[3]) for r in rows])
        tail = np.array([[float(v) for v in r[-3:]] for r in rows])
        coords = np.concatenate([np.zeros((1, 3)), xyz[:-1]])
        ipots = np.concatenate([[0], ip[:-1]]).astype(np.int64)
        out.append(FeffPath(index, nleg, deg, coords, ipots, tail[:, 0], tail[:, 1], tail[:, 2]))
        i += 2 + nleg
    return out


def read_chi_dat(filename):
    """(k, chi, mag, phase) columns of chi.dat, header lines skipped."""
    rows = []
    for s in Path(filename).read_text(encoding="latin-1").splitlines():
        t = s.split()
        if not t or s.lstrip().startswith("#"):
            continue
        try:
            rows.append([float(v) for v in t[:4]])
        except ValueError:
            continue
    a = np.array(rows)
    return a[:, 0], a[:, 1], a[:, 2], a[:, 3]


def read_feff_dat(filename):
    """feffNNNN.dat: dict with nleg, deg, reff (A) and the columns k, real_phc, mag, phase, red_fact, lam, rep."""
    lines = Path(filename).read_text(encoding="latin-1").splitlines()
    head = next(i for i, s in enumerate(lines) if "nleg, deg, reff" in s)
    w = lines[head].split()
    nleg, deg, reff = int(w[0]), float(w[1]), float(w[2])
    start = next(i for i, s in enumerate(lines) if s.strip().startswith("k") and "real[2*phc]" in s) + 1
    data = np.array([[float(v) for v in s.split()] for s in lines[start:] if s.strip()])
    names = ["k", "real_phc", "mag", "phase", "red_fact", "lam", "rep"]
    out = {n: data[:, j] for j, n in enumerate(names)}
    out.update(nleg=nleg, deg=deg, reff=reff)
    return out
