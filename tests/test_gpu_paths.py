"""The fused path-list entry (PathPlan / fm_chi_paths) against the FP64 oracle's batched entry: gather + minimum
image, per-path chi, Debye-Waller factor, per-segment sums, edge cases (empty list, one path, long paths, a general
cell), and bit-for-bit repeatability (the segment sum is an ordered reduction)."""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
PHASE = GOLDEN / "validation" / "A018" / "phase.pad"
Q = np.arange(0, 18.0001, 0.05)


def _plan(pl, prec, **kw):
    from feffgpu_amd.paths import PathPlan
    return PathPlan(str(PHASE), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment,
                    precision=prec, device="cuda", **kw)


def _box_paths(seed, absorbers, rmax=5.0, nleg=4, ncell=4):
    pos0, box = W.lattice("bcc", 2.87, ncell)
    pos = W.displaced(pos0, 0.07, seed)
    ipot = np.full(len(pos), 2)
    ipot[::5] = 1                                      # a second scatterer species
    pl = W.concat([W.enumerate_paths(pos, a, ipot, rmax, nleg, box=box, segment=i, nmax=nleg)
                   for i, a in enumerate(absorbers)], nleg)
    return pos, box, pl


@pytest.mark.parametrize("prec,tol", [(64, 1e-10), (32, 1e-5)], ids=["fp64", "fp32"])
def test_chi_paths_matches_oracle_periodic(prec, tol):
    pos, box, pl = _box_paths(3, [0, 7, 40, 101])
    print(f"[paths] periodic box {len(pos)} atoms, {len(pl)} paths, prec {prec}", flush=True)
    s2 = np.linspace(0.0, 0.006, len(pl))
    ref_seg, ref_path, ref_reff = oracle.chi_paths(oracle.OraclePhase(PHASE), pos, pl.atoms, pl.nleg, pl.absorber,
                                                   pl.ipots, pl.deg, Q, segment=pl.segment, nseg=4, sigma2=s2,
                                                   cell=box[0], want_paths=True)
    plan = _plan(pl, prec, nseg=4, sigma2=s2, cell=box[0])
    seg, path, reff = plan.forward(torch.tensor(pos, device="cuda"))
    seg, path, reff = seg.cpu().numpy(), path.double().cpu().numpy(), reff.cpu().numpy()
    err_seg = np.abs(seg - ref_seg).max() / np.abs(ref_seg).max()
    err_path = (np.abs(path - ref_path).max(1) / np.abs(ref_path).max(1)).max()
    print(f"[paths] max rel: segments {err_seg:.2e}, per path {err_path:.2e}", flush=True)
    assert err_seg < tol and err_path < 5 * tol
    np.testing.assert_allclose(reff, ref_reff, rtol=1e-12)


@pytest.mark.parametrize("case,npot", [("A014", 1), ("A009", 2)], ids=["L3", "M5"])
@pytest.mark.parametrize("prec,tol", [(64, 1e-10), (32, 1e-5)], ids=["fp64", "fp32"])
def test_chi_paths_l_and_m_edges(case, npot, prec, tol):
    """L3 and M5 edges: multiple scattering (lambda0 = 6) and single scattering (the 15-entry lambda set, M
    truncated at 15 entries as FEFF's setlam) against the oracle; the plain FP64 pipeline has no such two-leg
    kernel and says so."""
    from feffgpu_amd.paths import PathPlan
    phase = GOLDEN / "validation" / case / "phase.pad"
    pos0, box = W.lattice("bcc", 2.87, 4)
    pos = W.displaced(pos0, 0.07, 21)
    ipot = np.full(len(pos), npot)
    ipot[::5] = 1
    pl = W.concat([W.enumerate_paths(pos, a, ipot, 4.8, 4, box=box, segment=i, nmax=4)
                   for i, a in enumerate([2, 50])], 4)
    assert {2, 3, 4} <= set(pl.nleg.tolist())
    print(f"[paths] {case} edge: {len(pl)} paths, prec {prec}", flush=True)
    s2 = np.linspace(0.0, 0.004, len(pl))
    ref_seg, ref_path, _ = oracle.chi_paths(oracle.OraclePhase(phase), pos, pl.atoms, pl.nleg, pl.absorber, pl.ipots,
                                            pl.deg, Q, segment=pl.segment, nseg=2, sigma2=s2, cell=box[0],
                                            want_paths=True)
    plan = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=2,
                    sigma2=s2, cell=box[0], precision=prec, device="cuda")
    seg, path, _ = plan.forward(torch.tensor(pos, device="cuda"))
    seg, path = seg.cpu().numpy(), path.double().cpu().numpy()
    err_seg = np.abs(seg - ref_seg).max() / np.abs(ref_seg).max()
    rel = np.abs(path - ref_path).max(1) / np.abs(ref_path).max(1)
    for n in (2, 3, 4):
        print(f"[paths] {case} nleg {n}: per path {rel[pl.nleg == n].max():.2e}", flush=True)
    print(f"[paths] {case} segments {err_seg:.2e}", flush=True)
    assert err_seg < tol and rel.max() < 5 * tol
    if prec == 64:
        with pytest.raises(NotImplementedError):
            plan.forward_plain(torch.tensor(pos, device="cuda"))


def test_general_cell_equals_cubic():
    pos, box, pl = _box_paths(4, [3, 9], rmax=4.6)
    plan_c = _plan(pl, 64, nseg=2, cell=box[0])
    plan_g = _plan(pl, 64, nseg=2, cell=np.diag(box))
    p = torch.tensor(pos, device="cuda")
    a = plan_c.forward(p)[0].cpu().numpy()
    b = plan_g.forward(p)[0].cpu().numpy()
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_long_paths_cluster():
    """Paths of up to 8 legs (L001 list) through the fused entry: positions = FEFF's listed atoms."""
    from feffgpu_amd.feffio import read_paths_dat
    from feffgpu_amd.paths import PathPlan
    paths = read_paths_dat(GOLDEN / "validation" / "L001" / "paths.dat")
    phase = GOLDEN / "validation" / "L001" / "phase.pad"
    nmax = max(p.nleg for p in paths)
    pos, atoms, nleg, ipots, deg = [np.zeros(3)], [], [], [], []
    for p in paths:
        row = []
        for c in p.coords[1:]:
            pos.append(c)
            row.append(len(pos) - 1)
        atoms.append(row + [0] * (nmax - 1 - len(row)))
        nleg.append(p.nleg)
        ipots.append(list(p.ipots) + [0] * (nmax - p.nleg))
        deg.append(p.deg)
    pos = np.array(pos)
    k = np.arange(0.05, 18.0001, 0.05)
    ref, _, _ = oracle.chi_paths(oracle.OraclePhase(phase), pos, atoms, nleg, np.zeros(len(paths)), ipots, deg, k)
    for prec, tol in ((64, 1e-10), (32, 1e-5)):
        plan = PathPlan(str(phase), k, atoms, nleg, np.zeros(len(paths), int), ipots, deg, precision=prec, device="cuda")
        got = plan.forward(torch.tensor(pos, device="cuda"))[0].cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"[paths] L001 {len(paths)} paths up to {nmax} legs, prec {prec}: {err:.2e}", flush=True)
        assert err < tol


def test_empty_and_single_path_and_repeatability():
    pos, box, pl = _box_paths(5, [2], rmax=4.0)
    one = W.PathList(pl.atoms[:1], pl.nleg[:1], pl.absorber[:1], pl.ipots[:1], pl.deg[:1], pl.segment[:1],
                     pl.half_length[:1])
    p = torch.tensor(pos, device="cuda")
    plan1 = _plan(one, 32, nseg=1, cell=box[0])
    r1 = plan1.forward(p)[0].cpu().numpy()
    ref, _, _ = oracle.chi_paths(oracle.OraclePhase(PHASE), pos, one.atoms, one.nleg, one.absorber, one.ipots, one.deg,
                                 Q, cell=box[0])
    assert np.abs(r1 - ref).max() < 1e-5 * np.abs(ref).max()
    empty = W.PathList(pl.atoms[:0], pl.nleg[:0], pl.absorber[:0], pl.ipots[:0], pl.deg[:0], pl.segment[:0],
                       pl.half_length[:0])
    plan0 = _plan(empty, 32, nseg=3, cell=box[0])
    assert (plan0.forward(p)[0] == 0).all()
    planr = _plan(pl, 32, nseg=1, cell=box[0])
    a = planr.forward(p)[0].clone()
    b = planr.forward(p)[0].clone()
    assert torch.equal(a, b)


def test_bad_inputs_raise_before_launch():
    from feffgpu_amd.paths import PathPlan
    pos, box, pl = _box_paths(6, [1], rmax=4.0)
    bad = pl.ipots.copy()
    bad[0, 1] = 7
    with pytest.raises(ValueError):
        PathPlan(str(PHASE), Q, pl.atoms, pl.nleg, pl.absorber, bad, pl.deg, device="cuda")
    plan = _plan(pl, 32, nseg=1, cell=box[0])
    with pytest.raises(ValueError):
        plan.forward(torch.zeros(5, 3, device="cuda", dtype=torch.float64))
