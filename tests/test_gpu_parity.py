"""Parity of the HIP path stage (through the C ABI) with the FP64 oracle and FEFF's own chi.dat.

Tolerances: FP64 kernels 1e-9 of max|cchi| per path (the reference's own stage-vs-port bound is 1e-8,
tests/test_pathfinder.py:229); FP32 kernels (FP64 geometry and phase) 2e-5 of max|cchi| per path and, for the
summed chi, R-factor <= 2e-5 and max|dchi|/max|chi| <= 1e-5 (the parity bar of the FP32 path).
"""
import numpy as np
import pytest
import torch

import oracle
from fixtures import CASES, group_by_nleg, load_case, r_factor_vs, rmax_of

pytestmark = pytest.mark.gpu

TOL_CCHI = {torch.complex128: 1e-9, torch.complex64: 2e-5}


def _stage(phase, dtype):
    from feffgpu_amd.stage import GPUPathStage
    return GPUPathStage(str(phase), device="cuda", dtype=dtype)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64], ids=["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES)
def test_cchi_matches_oracle(case, dtype):
    phase, paths, k, chi = load_case(case)
    orc = oracle.OraclePhase(phase)
    stage = _stage(phase, dtype)
    worst = 0.0
    for nleg, grp in sorted(group_by_nleg(paths).items()):
        print(f"[cchi] {case} {dtype} nleg={nleg} paths={len(grp)}", flush=True)
        coords = torch.tensor(np.stack([p.coords for p in grp]), device="cuda", dtype=torch.float64)
        ipots = torch.tensor(np.stack([p.ipots for p in grp]), device="cuda")
        cc, reff = stage.cchi(coords, ipots)
        cc, reff = cc.cpu().numpy(), reff.cpu().numpy()
        for i, p in enumerate(grp):
            ref, rref, _ = orc.cchi(p.coords, p.ipots)
            err = np.abs(cc[i] - ref).max() / np.abs(ref).max()
            worst = max(worst, err)
            assert err < TOL_CCHI[dtype], (case, nleg, p.index, err)
            assert abs(reff[i] - rref) < 1e-6 * rref
    print(f"[cchi] {case} {dtype} worst rel err {worst:.3e}", flush=True)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64], ids=["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES)
def test_chi_sum_matches_feff_and_oracle(case, dtype):
    """Sum over FEFF's path list of deg * chi_fine: against FEFF's chi.dat (complete lists) and the oracle."""
    phase, paths, k, chi_ref = load_case(case)
    orc = oracle.OraclePhase(phase)
    stage = _stage(phase, dtype)
    acc = np.zeros(len(k))
    acc_orc = np.zeros(len(k))
    for nleg, grp in sorted(group_by_nleg(paths).items()):
        coords = torch.tensor(np.stack([p.coords for p in grp]), device="cuda", dtype=torch.float64)
        ipots = torch.tensor(np.stack([p.ipots for p in grp]), device="cuda")
        deg = np.array([p.deg for p in grp])
        cc, reff = stage.cchi(coords, ipots)
        chi = stage.chi_fine(cc, reff, k).double().cpu().numpy()
        acc += (deg[:, None] * chi).sum(0)
        for p in grp:
            c, r, _ = orc.cchi(p.coords, p.ipots)
            acc_orc += p.deg * orc.chi_fine(c, r, k)[0]
    rmax = rmax_of(case)
    R_orc = r_factor_vs(k, acc_orc, acc, rmax)
    rel = np.abs(acc - acc_orc).max() / np.abs(acc_orc).max()
    print(f"[sum] {case} {dtype} R(vs oracle)={R_orc:.3e} maxrel={rel:.3e}", flush=True)
    if dtype == torch.complex128:
        assert R_orc < 1e-14 and rel < 1e-9
    else:
        assert R_orc < 2e-5 and rel < 1e-5
    R_feff = r_factor_vs(k, chi_ref, acc, rmax)
    print(f"[sum] {case} {dtype} R(vs FEFF chi.dat)={R_feff:.3e}", flush=True)
    assert R_feff < (1e-9 if dtype == torch.complex128 else 2e-5)
