"""Loaders for the FEFF 8.5L golden files kept under tests/golden (copied from the reference's tests/data/W and
data/validation: phase.pad, paths.dat, chi.dat, feffNNNN.dat)."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from feffgpu_amd.feffio import read_chi_dat, read_feff_dat, read_paths_dat

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["W"] + [f"validation/{c}" for c in ("A000", "A009", "A014", "A018", "L000", "L001")]


def case_dir(case: str) -> Path:
    return GOLDEN / case


def rmax_of(case: str) -> float:
    if case == "W":
        return 9.3          # the reference's W test window (tests/test_xsph.py:124)
    txt = (GOLDEN / case / "cards.txt").read_text()
    return float(re.search(r"RPATH\s+([\d.]+)", txt).group(1))


def load_case(case: str):
    d = case_dir(case)
    paths = read_paths_dat(d / "paths.dat")
    k, chi, _, _ = read_chi_dat(d / "chi.dat")
    return d / "phase.pad", paths, k, chi


def r_factor_vs(k, ref, got, rmax):
    from feffgpu_amd.compare import r_factor
    if k[0] > 0:
        k = np.concatenate([[0.0], k])
        ref = np.concatenate([[0.0], ref])
        got = np.concatenate([[0.0], got])
    kmax = min(23.0, float(k[-1]) - 1.0)
    return r_factor(k, ref, got, kmin=3.0, kmax=kmax, rmin=1.0, rmax=rmax)


def group_by_nleg(paths):
    out = {}
    for p in paths:
        out.setdefault(p.nleg, []).append(p)
    return out
