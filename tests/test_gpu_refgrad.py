"""The drop-in stage against vectors written by the reference's own path stage (complex128 on the CPU, gradients by
its autograd through torch.utils.checkpoint as its forward models take them; tests/golden/make_ref_autograd.py):
exactly collinear and axis-aligned paths of 2-6 legs, FEFF's W path list (96 paths, 2-4 legs), and the 2-8-leg
random walks over a bcc lattice of the reference's tests/test_pathfinder.py:193-213, displaced and not.

Values (cchi, reff, chi on the fine grid) for every path. Gradients of a cchi-level and a chi-level loss with
respect to coords: FP64 within 1e-9 and FP32 within 1e-4 of the path's largest component wherever the reference's
gradient is finite. Where it is not (legs parallel to z: the reference's _trig passes the 0/0 of
d sqrt(x^2 + y^2) through torch.where, src/feffgpu/stage.py:147-157) the kernels' gradient is finite and is
checked against five-point differences of the FP64 oracle of the cchi-level loss instead."""
import numpy as np
import pytest
import torch

import oracle
from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
PHASE = GOLDEN / "W" / "phase.pad"


def _data():
    return np.load(GOLDEN / "refgrad" / "ref_autograd.npz")


def _grads(st, coords, ipots, wc, wr, q, wq):
    x = torch.tensor(coords, device="cuda", dtype=torch.float64, requires_grad=True)
    ip = torch.tensor(ipots, device="cuda")
    cc, reff = st.cchi(x, ip)
    l1 = (torch.tensor(wc, device="cuda").conj()[None] * cc.to(torch.complex128)).real.sum() + wr * reff.double().sum()
    (g1,) = torch.autograd.grad(l1, x)
    x2 = torch.tensor(coords, device="cuda", dtype=torch.float64, requires_grad=True)
    cc2, reff2 = st.cchi(x2, ip)
    chi = st.chi_fine(cc2, reff2, q)
    (g2,) = torch.autograd.grad((chi.double() * torch.tensor(wq, device="cuda")[None]).sum(), x2)
    return (cc.detach().to(torch.complex128).cpu().numpy(), reff.detach().double().cpu().numpy(),
            chi.detach().double().cpu().numpy(), g1.cpu().numpy(), g2.cpu().numpy())


@pytest.mark.parametrize("prec", [64, 32], ids=["fp64", "fp32"])
def test_stage_against_reference_autograd(prec):
    from feffgpu_amd.stage import GPUPathStage
    d = _data()
    st = GPUPathStage(str(PHASE), dtype=torch.complex128 if prec == 64 else torch.complex64)
    q, wc, wr, wq = d["q"], d["wc"], float(d["wr"]), d["wq"]
    tv, tg = (1e-9, 1e-9) if prec == 64 else (2e-5, 1e-4)
    orc = oracle.OraclePhase(PHASE)
    h = 1e-4

    def L(c, ip):
        cc, r, _ = orc.cchi(c, ip)
        return float((np.conj(wc) * cc).real.sum() + wr * r)
    worst = dict(cchi=0.0, chi=0.0, gc=0.0, gq=0.0, fd=0.0)
    nref = nfd = 0
    for g in d["groups"]:
        C, I = d[f"coords_{g}"], d[f"ipots_{g}"]
        cc, reff, chi, gc, gq = _grads(st, C, I, wc, wr, q, wq)
        rc, rchi = d[f"cchi_{g}"], d[f"chi_{g}"]
        for b in range(len(C)):
            e_c = np.abs(cc[b] - rc[b]).max() / np.abs(rc[b]).max()
            e_x = np.abs(chi[b] - rchi[b]).max() / np.abs(rchi[b]).max()
            worst["cchi"], worst["chi"] = max(worst["cchi"], e_c), max(worst["chi"], e_x)
            # FP32 cchi: 2e-5 of the path's largest value, 5e-5 for 7-8 legs (rounding over that many F products)
            tc = tv if prec == 64 or C.shape[1] <= 6 else 5e-5
            assert e_c <= tc and e_x <= tv * (1 if prec == 64 else 0.5), (g, b, e_c, e_x)
            assert abs(reff[b] - d[f"reff_{g}"][b]) <= (1e-12 if prec == 64 else 1e-6) * reff[b]
            assert np.isfinite(gc[b]).all() and np.isfinite(gq[b]).all(), (g, b)
            for name, mine, ref in (("gc", gc[b], d[f"gc_{g}"][b]), ("gq", gq[b], d[f"gq_{g}"][b])):
                if np.isfinite(ref).all():
                    e = np.abs(mine - ref).max() / np.abs(ref).max()
                    worst[name] = max(worst[name], e)
                    nref += 1
                    assert e <= tg, (g, b, name, e)
            if not np.isfinite(d[f"gc_{g}"][b]).all() and prec == 64:
                fd = np.zeros_like(C[b])
                for i in range(C.shape[1]):
                    for k in range(3):
                        e1 = np.zeros_like(C[b])
                        e1[i, k] = h
                        fd[i, k] = (-L(C[b] + 2 * e1, I[b]) + 8 * L(C[b] + e1, I[b]) - 8 * L(C[b] - e1, I[b])
                                    + L(C[b] - 2 * e1, I[b])) / (12 * h)
                e = np.abs(gc[b] - fd).max() / np.abs(fd).max()
                worst["fd"] = max(worst["fd"], e)
                nfd += 1
                assert e <= 1e-5, (g, b, e)
    print(f"[refgrad] FP{prec}: {nref} reference gradients, {nfd} against oracle differences; worst "
          + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()), flush=True)
