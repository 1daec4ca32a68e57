"""Inputs of the path-finder tests: FEFF's feff.inp of the golden cases (ATOMS, RPATH, NLEG, CRITERIA) and the
degeneracy-weighted number of paths per number of legs that FEFF's paths.dat lists."""
from __future__ import annotations

import re

import numpy as np


def read_feff_inp(path):
    """(xyz (n, 3) A, ipot (n,), rmax, nlegmax, critpw) of a feff.inp."""
    lines = open(path).read().splitlines()
    txt = "\n".join(lines)
    rmax = float(re.search(r"RPATH\s+([\d.]+)", txt).group(1))
    nleg = int(re.search(r"NLEG\s+(\d+)", txt).group(1))
    critpw = float(re.search(r"CRITERIA\s+[\d.]+\s+([\d.]+)", txt).group(1))
    i = [k for k, l in enumerate(lines) if l.strip().upper().startswith("ATOMS")][0]
    xyz, ip = [], []
    for l in lines[i + 1:]:
        w = l.split()
        if not w or w[0].upper().startswith("END"):
            break
        try:
            xyz.append([float(w[0]), float(w[1]), float(w[2])])
            ip.append(int(w[3]))
        except ValueError:
            break
    return np.array(xyz), np.array(ip), rmax, nleg, critpw


def feff_degeneracies(path):
    out = {}
    for l in open(path).read().splitlines():
        if "index, nleg" in l:
            w = l.split()
            out[int(w[1])] = out.get(int(w[1]), 0.0) + float(w[2])
    return out


def compact_cluster(seed=3):
    """The absorber and six atoms 1.7 A from it along the axes, displaced by 0.05 A (seeded): paths of up to eight
    legs fit inside 7.4 A (the shape of the reference's tests/test_pathfinder.py:117-122 case)."""
    pos = np.vstack([np.zeros(3), 1.7 * np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])])
    pos[1:] += np.random.default_rng(seed).normal(0, 0.05, (6, 3))
    return np.round(pos, 5)
