"""The plain FP64 pipeline (fm_chi_paths_plain: the reference's op sequence, one kernel per stage, intermediates in
HBM) against the FP64 oracle and the fused kernels on the same path lists; chunking; and the launch contract that a
bucket's paths all have its leg count (a path that does not gets NaN, the others are unaffected)."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
Q = np.arange(0, 18.0001, 0.05)


def _box(seed, absorbers, phase, rmax=4.8, nleg=4, ipot2=True, legs=None):
    pos0, box = W.lattice("bcc", 2.87, 4)
    pos = W.displaced(pos0, 0.06, seed)
    ipot = np.full(len(pos), 2 if ipot2 else 1)
    if ipot2:
        ipot[::5] = 1
    pl = W.concat([W.enumerate_paths(pos, a, ipot, rmax, nleg, box=box, segment=i, nmax=nleg)
                   for i, a in enumerate(absorbers)], nleg)
    if legs is not None:
        keep = np.isin(pl.nleg, legs)
        pl = W.PathList(pl.atoms[keep], pl.nleg[keep], pl.absorber[keep], pl.ipots[keep], pl.deg[keep],
                        pl.segment[keep], pl.half_length[keep])
    return pos, box, pl


def _plan(phase, pl, prec, box, nseg, **kw):
    from feffgpu_amd.paths import PathPlan
    return PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=nseg,
                    precision=prec, cell=box[0], device="cuda", **kw)


def test_plain_matches_oracle_and_fused_k_edge():
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    pos, box, pl = _box(11, [0, 33], phase)
    s2 = np.linspace(0.0, 0.005, len(pl))
    print(f"[plain] K edge: {len(pl)} paths, legs {sorted(set(pl.nleg.tolist()))}", flush=True)
    ref_seg, ref_path, ref_reff = oracle.chi_paths(oracle.OraclePhase(phase), pos, pl.atoms, pl.nleg, pl.absorber,
                                                   pl.ipots, pl.deg, Q, segment=pl.segment, nseg=2, sigma2=s2,
                                                   cell=box[0], want_paths=True)
    plan = _plan(phase, pl, 64, box, 2, sigma2=s2)
    p = torch.tensor(pos, device="cuda")
    seg, path, reff = (t.cpu().numpy() for t in plan.forward_plain(p))
    err_path = (np.abs(path - ref_path).max(1) / np.abs(ref_path).max(1)).max()
    err_seg = np.abs(seg - ref_seg).max() / np.abs(ref_seg).max()
    print(f"[plain] vs oracle: per path {err_path:.2e}, segments {err_seg:.2e}", flush=True)
    assert err_path < 1e-10 and err_seg < 1e-10
    np.testing.assert_allclose(reff, ref_reff, rtol=1e-12)
    fseg = plan.forward(p)[0].cpu().numpy()
    assert np.abs(fseg - seg).max() < 1e-10 * np.abs(seg).max()


def test_plain_l3_edge_multiple_scattering():
    """The L/M lambda set (lambda0 = 6) through the plain pipeline and the fused FP64 / FP32 kernels."""
    phase = GOLDEN / "validation" / "A014" / "phase.pad"
    pos, box, pl = _box(12, [5], phase, rmax=4.6, ipot2=False, legs=[3, 4])
    print(f"[plain] L3 edge: {len(pl)} paths", flush=True)
    ref_seg, _, _ = oracle.chi_paths(oracle.OraclePhase(phase), pos, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg,
                                     Q, segment=pl.segment, nseg=1, cell=box[0], want_paths=True)
    p = torch.tensor(pos, device="cuda")
    plain = _plan(phase, pl, 64, box, 1).forward_plain(p)[0].cpu().numpy()
    f64 = _plan(phase, pl, 64, box, 1).forward(p)[0].cpu().numpy()
    f32 = _plan(phase, pl, 32, box, 1).forward(p)[0].cpu().numpy()
    scale = np.abs(ref_seg).max()
    errs = [np.abs(x - ref_seg).max() / scale for x in (plain, f64, f32)]
    print(f"[plain] L3 max rel: plain {errs[0]:.2e}, fused fp64 {errs[1]:.2e}, fused fp32 {errs[2]:.2e}", flush=True)
    assert errs[0] < 1e-10 and errs[1] < 1e-10 and errs[2] < 1e-5


def test_plain_chunking_is_invisible():
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    pos, box, pl = _box(13, [4], phase, rmax=4.4)
    p = torch.tensor(pos, device="cuda")
    big = _plan(phase, pl, 64, box, 1).forward_plain(p)[1].cpu().numpy()
    small_plan = _plan(phase, pl, 64, box, 1)
    from feffgpu_amd import _native as N
    per = N.lib().fm_plain_path_bytes(small_plan.edges[0].struct, 4)
    small = small_plan.forward_plain(p, workspace_bytes=7 * per)[1].cpu().numpy()   # chunks of 7 paths
    assert np.array_equal(big, small)


def test_bucket_leg_count_mismatch_gives_nan_rows():
    """A bucket that claims 4 legs for 3-leg paths (or 3 for 4): those rows are NaN, every other row is unchanged."""
    from feffgpu_amd import _native as N
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    pos, box, pl = _box(14, [6], phase, rmax=4.8)
    plan = _plan(phase, pl, 32, box, 1)
    p = torch.tensor(pos, device="cuda")
    good = plan.forward(p)[1].clone()
    b4 = [b for b in range(len(plan.buckets.nleg)) if plan.buckets.nleg[b] in (3, 4)]
    assert b4
    nlm = list(plan.buckets.nleg)
    nlm[b4[0]] = 7 - nlm[b4[0]]
    h_nlm = (N._i32 * len(nlm))(*nlm)
    ps = plan._paths_struct(p)
    N.check(N.lib().fm_chi_paths(plan.edge_array, 1, ps, plan.grid, 32, 2, plan.order_t.data_ptr(),
                                 len(plan.buckets.kind), plan._h_off, plan._h_kind, plan._h_edge, h_nlm, 0, None, None,
                                 None, plan.chi_path.data_ptr(), plan.reff.data_ptr(), None, N.stream_of(plan.device)),
            "fm_chi_paths")
    got = plan.chi_path.cpu()
    sel = plan.buckets.order[plan.buckets.offset[b4[0]]:plan.buckets.offset[b4[0] + 1]]
    bad = np.zeros(len(pl), bool)
    bad[sel] = True
    assert torch.isnan(got[bad]).all()
    assert torch.equal(got[~bad], good.cpu()[~bad])
