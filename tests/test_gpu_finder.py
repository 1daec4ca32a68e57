"""The HIP path finder (feffgpu_amd.pathfinder.GPUPathFinder) against the path-finder oracle and FEFF's own lists.

* FEFF 8.5L's paths.dat of every golden case: the degeneracy-weighted number of kept paths per number of legs (the
  reference's tests/test_validation_cases.py:55-57), with FEFF's phase file and its CRITERIA.
* The oracle (float32 numpy, oracle/finder.py) on a disordered periodic box (general and cubic cell) and a cluster:
  with no filter the same paths in the same order, the same degeneracies and importance to 1e-5; with a filter the
  same kept paths, apart from paths whose degeneracy-weighted importance ties the threshold to float32 rounding.
* Paths of up to eight legs of a compact cluster (the 5+-leg extension) equal to the oracle's.
* Symmetry totals on an ideal simple cubic lattice (shells of 6, 8, 12; triangles of 24 and 48).
* Many absorbers in one pass, of one configuration or of several, equal to the absorbers one by one.
"""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from finder_cases import compact_cluster, feff_degeneracies, read_feff_inp
from fixtures import CASES, GOLDEN
from oracle.finder import OracleFinder

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES)
def test_finder_reproduces_feff_path_lists(case):
    from feffgpu_amd.pathfinder import GPUPathFinder
    xyz, ip, rmax, nleg, critpw = read_feff_inp(GOLDEN / case / "feff.inp")
    a = int(np.nonzero(ip == 0)[0][0])
    f = GPUPathFinder(str(GOLDEN / case / "phase.pad"), device="cuda", rmax=rmax, nlegmax=nleg, critpw=critpw)
    kept, ref = f.enumerate(xyz, None, a, ip)
    mine = {n: round(float(v[2].sum())) for n, v in kept.items() if v[2].numel()}
    feff = {n: round(d) for n, d in feff_degeneracies(GOLDEN / case / "paths.dat").items()}
    print(f"[finder] {case}: {mine}", flush=True)
    assert mine == feff, (mine, feff)


def _box():
    pos0, box = W.lattice("bcc", 2.87, 4)
    pos = W.displaced(pos0, 0.07, 5)
    ipot = np.full(len(pos), 2)
    ipot[::3] = 1
    return pos, box, ipot


@pytest.mark.parametrize("cell", ["cubic", "general", "cluster"])
def test_finder_equals_oracle_without_filter(cell):
    from feffgpu_amd.pathfinder import GPUPathFinder
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    pos, box, ipot = _box()
    bx = {"cubic": float(box[0]), "general": np.diag(box), "cluster": None}[cell]
    orc = OracleFinder(oracle.OraclePhase(phase), 5.0, 4, 0.0, merge_degenerate=False)
    gpu = GPUPathFinder(str(phase), device="cuda", rmax=5.0, nlegmax=4, critpw=0.0, merge_degenerate=False)
    for a in (0, 37, 101):
        ko, ro = orc.enumerate(pos, bx, a, ipot)
        kg, rg = gpu.enumerate(pos, bx, a, ipot)
        assert set(ko) == set(kg)
        for n in ko:
            co, io, do, xo, ao = ko[n]
            cg, ig, dg, xg, ag = (t.cpu().numpy() for t in kg[n])
            np.testing.assert_array_equal(ag, ao)
            np.testing.assert_array_equal(ig, io)
            np.testing.assert_array_equal(dg, do)
            np.testing.assert_allclose(cg, co, rtol=0, atol=2e-6)
            np.testing.assert_allclose(xg, xo, rtol=2e-5, atol=0)
        assert abs(rg - ro) <= 2e-5 * ro
    print(f"[finder] {cell} cell: the oracle's paths, order and importance", flush=True)


def test_finder_filter_matches_oracle():
    from feffgpu_amd.pathfinder import GPUPathFinder
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    pos, box, ipot = _box()
    for merge in (False, True):
        orc = OracleFinder(oracle.OraclePhase(phase), 5.5, 4, 1.0, merge_degenerate=merge)
        gpu = GPUPathFinder(str(phase), device="cuda", rmax=5.5, nlegmax=4, critpw=1.0, merge_degenerate=merge)
        for a in (3, 64):
            ko, ro = orc.enumerate(pos, box[0], a, ipot)
            kg, rg = gpu.enumerate(pos, box[0], a, ipot)
            for n in ko:
                so = {tuple(r) for r in ko[n][4].tolist()}
                sg = {tuple(r) for r in kg[n][4].cpu().tolist()}
                # a path may differ only where its weighted importance ties the threshold to float32 rounding
                odd = so ^ sg
                if odd:
                    full = OracleFinder(oracle.OraclePhase(phase), 5.5, 4, 0.0, merge_degenerate=merge)
                    kall, rall = full.enumerate(pos, box[0], a, ipot)
                    co, _, dd, xo, ao = kall[n]
                    dt = OracleFinder.degenerate_total(co, dd) if (merge and n > 2) else dd
                    frac = {tuple(r): 100 * t * x / ro for r, t, x in zip(ao.tolist(), dt, xo)}
                    assert all(abs(frac[p] - 1.0) < 1e-4 for p in odd), odd
                print(f"[finder] filter merge={merge} absorber {a} nleg {n}: {len(sg)} kept, {len(odd)} ties",
                      flush=True)


def test_finder_long_paths_of_a_compact_cluster():
    from feffgpu_amd.pathfinder import GPUPathFinder
    pos = compact_cluster()
    ip = np.ones(len(pos), int)
    orc = OracleFinder(oracle.OraclePhase(GOLDEN / "W" / "phase.pad"), 7.4, 9, 0.0, merge_degenerate=False)
    gpu = GPUPathFinder(str(GOLDEN / "W" / "phase.pad"), device="cuda", rmax=7.4, nlegmax=9, critpw=0.0,
                        merge_degenerate=False)
    ko, _ = orc.enumerate(pos, None, 0, ip)
    kg, _ = gpu.enumerate(pos, None, 0, ip)
    assert set(ko) == set(kg) and 8 in kg
    for n in ko:
        np.testing.assert_array_equal(kg[n][4].cpu().numpy(), ko[n][4])
        np.testing.assert_array_equal(kg[n][2].cpu().numpy(), ko[n][2])
        np.testing.assert_allclose(kg[n][3].cpu().numpy(), ko[n][3], rtol=2e-5)
    print(f"[finder] compact cluster: {({n: len(v[4]) for n, v in ko.items()})}", flush=True)


def test_finder_symmetry_totals_and_many_absorbers(monkeypatch):
    from feffgpu_amd.pathfinder import GPUPathFinder
    a, m = 2.823, 3
    cells = np.stack(np.meshgrid(*[np.arange(-m, m + 1)] * 3, indexing="ij"), -1).reshape(-1, 3) * a
    pos = cells[np.argsort(np.linalg.norm(cells, axis=1), kind="stable")]
    f = GPUPathFinder(str(GOLDEN / "W" / "phase.pad"), device="cuda", rmax=5.32, nlegmax=3, critpw=0.0)
    kept, _ = f.enumerate(pos, None, 0, np.ones(len(pos), int))
    assert sorted(kept[2][2].cpu().tolist()) == [6.0, 8.0, 12.0]
    coords, deg = kept[3][0], kept[3][2]
    tot = f.degenerate_total(coords, deg).cpu().numpy()
    c = coords.cpu().numpy()
    d1, d2 = np.linalg.norm(c[:, 1], axis=1), np.linalg.norm(c[:, 2], axis=1)
    ff = (np.abs(d1 - a) < 1e-3) & (np.abs(d2 - a) < 1e-3)
    assert ff.sum() == 12 and np.all(tot[ff] == 24.0)
    # many absorbers of a box in one pass = one by one
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    pos, box, ipot = _box()
    g = GPUPathFinder(str(phase), device="cuda", rmax=5.0, nlegmax=5, critpw=0.5)
    absorbers = [0, 9, 55, 127]
    many, refs = g.enumerate_many(pos, box[0], absorbers, ipot)
    for s, a in enumerate(absorbers):
        one, r1 = g.enumerate(pos, box[0], a, ipot)
        assert abs(refs[s] - r1) == 0
        for n, v in one.items():
            sel = many[n]["slot"] == s
            assert torch.equal(many[n]["atoms"][sel], v[4]) and torch.equal(many[n]["deg"][sel], v[2])
    # configurations one after the other: each absorber's paths run over its own configuration's atoms
    pos2 = W.displaced(pos, 0.05, 9)
    frames = np.stack([pos, pos2])
    n = len(pos)
    ab2 = [9, n + 9, n + 127, 55]
    many, refs = g.enumerate_many(frames, box[0], ab2, ipot)
    for s, a in enumerate(ab2):
        one, r1 = g.enumerate(frames[a // n], box[0], a % n, ipot)
        assert abs(refs[s] - r1) == 0
        for nl, v in one.items():
            sel = many[nl]["slot"] == s
            assert torch.equal(many[nl]["atoms"][sel], v[4] + (a // n) * n)
            assert torch.equal(many[nl]["deg"][sel], v[2]) and torch.equal(many[nl]["ipots"][sel], v[1])
            assert torch.equal(many[nl]["coords"][sel], v[0])
    # absorbers taken a few at a time: the same result as one pass
    import feffgpu_amd.pathfinder as PF
    monkeypatch.setattr(PF, "ABSORBERS_PER_PASS", 3)
    split, rs = g.enumerate_many(frames, box[0], ab2, ipot)
    np.testing.assert_array_equal(rs, refs)
    assert set(split) == set(many)
    for nl in many:
        for k in many[nl]:
            assert torch.equal(split[nl][k], many[nl][k]), (nl, k)
