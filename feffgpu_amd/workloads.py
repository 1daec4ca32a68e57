"""Synthetic structures and path lists for the benchmark and parity workloads.

Structures: fcc / bcc / rocksalt lattices (clusters or periodic boxes), seeded Gaussian displacements, and random
sequential packing with a minimum separation. Path lists: every closed path of 2..nlegmax legs from an absorber
whose half length is at most rmax, with the definitions of feffgpu's finder (src/feffgpu/pathfinder.py:211-279):
scatterers are the atoms within rmax of the absorber; consecutive atoms differ; the absorber may appear in the
middle of a path of four or more legs but neither first nor last; a path and its time reverse are one path of
degeneracy 2 (1 if it is its own reverse). The finder's plane-wave importance filter is optional here
(``max_paths`` keeps the shortest paths instead, to size a workload).

The path finder itself is not part of the GPU path stage; this module only produces inputs of the right shape.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def lattice(kind: str, a: float, ncell: int | tuple):
    """Atom positions of an ncell^3 (or (nx, ny, nz)) supercell of a cubic lattice with constant a, and the box."""
    basis = {"sc": [[0, 0, 0]], "bcc": [[0, 0, 0], [0.5, 0.5, 0.5]],
             "fcc": [[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]}[kind]
    n = (ncell, ncell, ncell) if np.isscalar(ncell) else tuple(ncell)
    g = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3)
    pos = (g[:, None, :] + np.asarray(basis)[None]).reshape(-1, 3) * a
    return pos, np.asarray(n, float) * a


def rocksalt(a: float, ncell):
    """Rocksalt (NaCl) supercell: species 0 on the fcc sites, species 1 shifted by a/2."""
    fcc, box = lattice("fcc", a, ncell)
    pos = np.concatenate([fcc, fcc + np.array([0.5 * a, 0, 0])]) % box
    species = np.concatenate([np.zeros(len(fcc), int), np.ones(len(fcc), int)])
    return pos, species, box


def cluster(kind: str, a: float, radius: float):
    """Lattice sites within radius of the origin (an atom at the origin first)."""
    n = int(math.ceil(radius / a)) + 1
    pos, _ = lattice(kind, a, 2 * n)
    pos = pos - n * a
    d = np.linalg.norm(pos, axis=1)
    pos = pos[d <= radius + 1e-9]
    d = np.linalg.norm(pos, axis=1)
    return pos[np.argsort(d, kind="stable")]


def displaced(pos, sigma, seed, keep_first=False):
    rng = np.random.default_rng(seed)
    out = pos + rng.normal(0.0, sigma, pos.shape)
    if keep_first:
        out[0] = pos[0]
    return out


def random_packing(n: int, box: float, dmin: float, seed: int, max_rounds: int = 100000):
    """Random sequential addition of n points in a cubic periodic box with minimum image separation dmin: uniform
    candidates in batches, each kept if it is at least dmin from every point placed before it (a periodic k-d tree
    for the points of earlier batches, a direct check within the batch, in the candidates' order)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pts = np.zeros((0, 3))
    tree = None
    for _ in range(max_rounds):
        if len(pts) == n:
            return pts
        m = min(4 * (n - len(pts)) + 16, 20000)
        cand = rng.uniform(0.0, box, (m, 3))
        if tree is not None:
            d, _ = tree.query(cand, k=1)
            cand = cand[d >= dmin]
        new = []
        for c in cand:
            if len(pts) + len(new) == n:
                break
            if new:
                v = np.asarray(new) - c
                v -= box * np.round(v / box)
                if (np.einsum("ij,ij->i", v, v) < dmin * dmin).any():
                    continue
            new.append(c)
        if new:
            pts = np.concatenate([pts, np.asarray(new)])
            tree = cKDTree(np.mod(pts, box), boxsize=box)
    raise RuntimeError(f"random packing placed only {len(pts)} of {n} atoms")


def _wrap(v, box):
    if box is None:
        return v
    b = np.asarray(box, float)
    return v - b * np.round(v / b)


@dataclass
class PathList:
    atoms: np.ndarray        # (P, nmax-1) global atom indices (padded with 0)
    nleg: np.ndarray         # (P,)
    absorber: np.ndarray     # (P,)
    ipots: np.ndarray        # (P, nmax)
    deg: np.ndarray          # (P,)
    segment: np.ndarray      # (P,)
    half_length: np.ndarray  # (P,) A

    def __len__(self):
        return len(self.nleg)

    def coords(self, pos, box=None):
        """(P, nmax, 3) coordinates relative to the absorber (absorber first at the origin)."""
        rel = _wrap(pos[self.atoms] - pos[self.absorber][:, None, :], box)
        return np.concatenate([np.zeros((len(self), 1, 3)), rel], 1)


def enumerate_paths(pos, absorber: int, ipot, rmax: float, nlegmax: int = 4, box=None, offset: int = 0,
                    segment: int = 0, nmax: int | None = None, max_paths: int | None = None):
    """Paths of absorber `absorber` (index into pos) with half length <= rmax. ipot (N,) potential per atom (the
    absorber gets 0). Atom indices in the result are shifted by `offset` (frames stacked in one position array)."""
    nmax = nmax or nlegmax
    rel_all = _wrap(pos - pos[absorber], box)
    d0 = np.linalg.norm(rel_all, axis=1)
    cl = np.nonzero((d0 <= rmax) & (np.arange(len(pos)) != absorber))[0]
    inter = np.concatenate([cl, [absorber]])
    rc, ri_ = rel_all[cl], rel_all[inter]
    dc0 = d0[cl]
    two_r = 2 * rmax
    out = []                                       # (atoms tuple array (B, nleg-1), deg, total length)
    out.append((cl[:, None], np.ones(len(cl)), 2 * dc0))
    if nlegmax >= 3 and len(cl) > 1:
        dij = np.linalg.norm(rc[:, None] - rc[None], axis=-1)
        tot = dc0[:, None] + dij + dc0[None, :]
        ii, jj = np.nonzero((tot <= two_r) & (np.arange(len(cl))[:, None] < np.arange(len(cl))[None, :]))
        out.append((np.stack([cl[ii], cl[jj]], 1), np.full(len(ii), 2.0), tot[ii, jj]))
    if nlegmax >= 4 and len(cl) > 1:
        dci = np.linalg.norm(rc[:, None] - ri_[None], axis=-1)        # (nc, ni)
        tot = dc0[:, None, None] + dci[:, :, None] + dci.T[None, :, :] + dc0[None, None, :]
        ok = tot <= two_r
        ok &= np.arange(len(cl))[:, None, None] <= np.arange(len(cl))[None, None, :]
        ok &= (cl[:, None, None] != inter[None, :, None]) & (inter[None, :, None] != cl[None, None, :])
        a, b, c = np.nonzero(ok)
        atoms4 = np.stack([cl[a], inter[b], cl[c]], 1)
        out.append((atoms4, np.where(atoms4[:, 0] == atoms4[:, 2], 1.0, 2.0), tot[a, b, c]))
    if nlegmax >= 5 and len(cl) > 1:
        D = np.linalg.norm(ri_[:, None] - ri_[None], axis=-1)
        back = np.linalg.norm(ri_, axis=1)
        seq = np.arange(len(cl))[:, None]
        length = back[:len(cl)].copy()
        for m in range(2, nlegmax):
            tot = length[:, None] + D[seq[:, -1]]
            ok = (tot + back[None, :] <= two_r) & (np.arange(len(inter))[None, :] != seq[:, -1:])
            i, j = np.nonzero(ok)
            seq, length = np.concatenate([seq[i], j[:, None]], 1), tot[i, j]
            if len(seq) == 0:
                break
            if m + 1 < 5:
                continue
            done = seq[seq[:, -1] != len(cl)]
            dl = length[seq[:, -1] != len(cl)] + back[done[:, -1]]
            rev = done[:, ::-1]
            differ = done != rev
            first = np.argmax(differ, 1)
            pal = ~differ.any(1)
            keep = pal | (done[np.arange(len(done)), first] < rev[np.arange(len(done)), first])
            out.append((inter[done[keep]], np.where(pal[keep], 1.0, 2.0), dl[keep]))
    atoms_l, nleg_l, deg_l, tot_l = [], [], [], []
    for atoms, deg, tot in out:
        if len(atoms) == 0:
            continue
        pad = np.zeros((len(atoms), nmax - 1), np.int64)
        pad[:, :atoms.shape[1]] = atoms
        atoms_l.append(pad)
        nleg_l.append(np.full(len(atoms), atoms.shape[1] + 1))
        deg_l.append(deg)
        tot_l.append(tot)
    if not atoms_l:
        z = np.zeros(0, np.int64)
        return PathList(np.zeros((0, nmax - 1), np.int64), z, z, np.zeros((0, nmax), np.int64), np.zeros(0), z, np.zeros(0))
    atoms = np.concatenate(atoms_l)
    nleg = np.concatenate(nleg_l)
    deg = np.concatenate(deg_l)
    half = np.concatenate(tot_l) / 2
    if max_paths is not None and len(nleg) > max_paths:
        keep = np.argsort(half, kind="stable")[:max_paths]
        atoms, nleg, deg, half = atoms[keep], nleg[keep], deg[keep], half[keep]
    ip = np.asarray(ipot)
    ipots = np.zeros((len(nleg), nmax), np.int64)
    for j in range(1, nmax):
        live = j < nleg
        a = atoms[:, j - 1]
        ipots[:, j] = np.where(live, np.where(a == absorber, 0, ip[a]), 0)
    atoms = np.where(np.arange(nmax - 1)[None, :] < (nleg[:, None] - 1), atoms + offset, 0)
    P = len(nleg)
    return PathList(atoms, nleg, np.full(P, absorber + offset), ipots, deg, np.full(P, segment), half)


def concat(lists, nmax):
    """Concatenate PathLists (padding atoms/ipots to nmax)."""
    def pad(a, w):
        out = np.zeros((a.shape[0], w), a.dtype)
        out[:, :a.shape[1]] = a
        return out
    return PathList(np.concatenate([pad(l.atoms, nmax - 1) for l in lists]), np.concatenate([l.nleg for l in lists]),
                    np.concatenate([l.absorber for l in lists]), np.concatenate([pad(l.ipots, nmax) for l in lists]),
                    np.concatenate([l.deg for l in lists]), np.concatenate([l.segment for l in lists]),
                    np.concatenate([l.half_length for l in lists]))
