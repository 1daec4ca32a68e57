"""Workloads shaped like the refinement and multi-edge configurations, at sizes the FP64 oracle evaluates in seconds.

* Refinement: a rocksalt box (a = 4.18 A, two species on two potentials, validation phase A000 as the substitute
  for Ni K in NiO), a structure perturbed by sigma = 0.1 A, a target chi from a hidden structure, the k^2-weighted L2
  loss sum_s,q (q^2 (chi - chi_t))^2; the fused FP32 forward + adjoint against the FP64 oracle (values) and five-point
  finite differences of the oracle's loss (gradient), FP32 against FP64 gradient <= 1e-4.
* Multi-edge: a random packing (d_min = 2.3 A) with 30 % of a heavy species, an L3 edge (validation phase A014, the
  substitute for Pt L3) and a K edge (A018, for Ni K) in one plan, paths of 2-6 legs: every fused adjoint variant
  (single scattering, 3-6 legs, K and L/M lambda sets) against the oracle and finite differences.
"""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
Q = np.arange(0.0, 18.0001, 0.05)
H = 1e-4


def _d(L, x, dx):
    return (-L(x + 2 * H * dx) + 8 * L(x + H * dx) - 8 * L(x - H * dx) + L(x - 2 * H * dx)) / (12 * H)


def _plan(phases, pl, prec, nseg, cell, edge=None):
    from feffgpu_amd.paths import PathPlan
    return PathPlan([str(p) for p in phases], Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment,
                    nseg=nseg, edge=edge, cell=cell, precision=prec, device="cuda")


def _subset(pl, keep):
    return W.PathList(pl.atoms[keep], pl.nleg[keep], pl.absorber[keep], pl.ipots[keep], pl.deg[keep], pl.segment[keep],
                      pl.half_length[keep])


def _check(phases, pos, pl, nseg, cell, edge, target, label):
    """Values against the oracle; the loss gradient of FP64 against FD of the oracle, FP32 against FP64."""
    orc = [oracle.OraclePhase(p) for p in phases]

    def seg_oracle(x):
        return oracle.chi_paths(orc, x, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, segment=pl.segment,
                                nseg=nseg, edge=edge, cell=cell)[0]
    q2 = Q ** 2

    def loss(seg):
        return float(((q2 * (seg - target)) ** 2).sum())
    ref = seg_oracle(pos)
    grads = {}
    for prec, tol in ((64, 1e-10), (32, 1e-5)):
        plan = _plan(phases, pl, prec, nseg, cell, edge)
        p = torch.tensor(pos, device="cuda", requires_grad=True)
        seg, _ = plan(p)
        err = float(np.abs(seg.detach().cpu().numpy() - ref).max() / np.abs(ref).max())
        print(f"[{label}] FP{prec}: {len(pl)} paths, max|dchi|/max|chi| vs oracle {err:.2e}", flush=True)
        assert err < tol, (prec, err)
        t = torch.tensor(target, device="cuda")
        qq = torch.tensor(q2, device="cuda")
        L = ((qq * (seg - t)) ** 2).sum()
        (g,) = torch.autograd.grad(L, p)
        grads[prec] = g.cpu().numpy()
    rng = np.random.default_rng(5)
    scale = np.abs(grads[64]).sum()
    for _ in range(2):
        d = rng.normal(size=pos.shape)
        d /= np.linalg.norm(d)
        fd = _d(lambda x: loss(seg_oracle(x)), pos, d)
        an = float((grads[64] * d).sum())
        print(f"[{label}] directional derivative: adjoint {an:.10e}, FD {fd:.10e}", flush=True)
        assert abs(an - fd) < 1e-6 * scale + 1e-9 * abs(fd), (an, fd)
    err32 = np.abs(grads[32] - grads[64]).max() / np.abs(grads[64]).max()
    print(f"[{label}] FP32 vs FP64 gradient {err32:.2e}", flush=True)
    assert err32 < 1e-4


def test_refinement_rocksalt_two_potentials():
    phase = GOLDEN / "validation" / "A000" / "phase.pad"
    pos0, species, box = W.rocksalt(4.18, 3)
    ipot = np.where(species == 0, 1, 2)
    pos = W.displaced(pos0, 0.1, 1)
    hidden = W.displaced(pos0, 0.1, 2)
    absorbers = np.nonzero(species == 0)[0][[0, 13, 40]]
    pl = W.concat([W.enumerate_paths(pos, int(a), ipot, 4.2, 4, box=box, segment=i, nmax=4)
                   for i, a in enumerate(absorbers)], 4)
    orc = oracle.OraclePhase(phase)
    target = oracle.chi_paths(orc, hidden, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, segment=pl.segment,
                              nseg=len(absorbers), cell=box[0])[0]
    _check([phase], pos, pl, len(absorbers), box[0], None, target, "refinement")


def test_multi_edge_two_to_six_legs():
    ph = [GOLDEN / "validation" / "A014" / "phase.pad", GOLDEN / "validation" / "A018" / "phase.pad"]
    box = 16.0
    pos = W.random_packing(200, box, 2.3, 7)
    heavy = np.random.default_rng(8).random(len(pos)) < 0.3
    absorbers = [int(np.nonzero(heavy)[0][0]), int(np.nonzero(~heavy)[0][0])]   # L3 edge, K edge
    ipots_by_edge = [np.ones(len(pos), int), np.where(heavy, 2, 1)]
    rng = np.random.default_rng(9)
    lists, edges = [], []
    for e, a in enumerate(absorbers):
        pl = W.enumerate_paths(pos, a, ipots_by_edge[e], 7.6, 6, box=np.full(3, box), segment=e, nmax=6)
        keep = (pl.nleg != 4) | (rng.random(len(pl)) < 0.25)
        pl = _subset(pl, keep)
        lists.append(pl)
        edges.append(np.full(len(pl), e))
    pl = W.concat(lists, 6)
    edge = np.concatenate(edges)
    counts = {(e, n): int(((edge == e) & (pl.nleg == n)).sum()) for e in (0, 1) for n in range(2, 7)}
    print(f"[multi-edge] paths per (edge, nleg): {counts}", flush=True)
    assert all(counts[(e, n)] > 0 for e in (0, 1) for n in range(2, 7))
    orc = [oracle.OraclePhase(p) for p in ph]
    hidden = W.displaced(pos, 0.05, 3)
    target = oracle.chi_paths(orc, hidden, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, Q, segment=pl.segment,
                              nseg=2, edge=edge, cell=box)[0]
    _check(ph, pos, pl, 2, box, edge, target, "multi-edge")
