"""chi_fine with a dense interpolation matrix (src/feffgpu/stage.py:372-388): the "larch" convention (cubic
spline_matrix), an arbitrary S, and path lists in the larch convention, against the oracle's dense fine grid (which is
pinned on the CPU against its own linear one); gradients against finite differences of the oracle."""
import numpy as np
import pytest
import torch

import oracle
from feffgpu_amd import workloads as W
from fixtures import GOLDEN, group_by_nleg, load_case

pytestmark = pytest.mark.gpu
Q = np.arange(0.0, 18.0001, 0.05)
H = 1e-4


@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-10), (torch.complex64, 2e-5)], ids=["fp64", "fp32"])
def test_stage_chi_fine_larch_and_dense_s(dtype, tol):
    from feffgpu_amd.stage import GPUPathStage
    phase, paths, _, _ = load_case("validation/A014")
    orc = oracle.OraclePhase(phase)
    st = GPUPathStage(str(phase), dtype=dtype)
    grp = group_by_nleg(paths)[3][:12]
    coords = torch.tensor(np.stack([p.coords for p in grp]), device="cuda", dtype=torch.float64)
    ipots = torch.tensor(np.stack([p.ipots for p in grp]), device="cuda")
    cc, reff = st.cchi(coords, ipots)
    ref_c = np.stack([orc.cchi(p.coords, p.ipots)[0] for p in grp])
    ref_r = np.array([orc.cchi(p.coords, p.ipots)[1] for p in grp])
    S_rand = np.random.default_rng(1).random((len(Q), st.E)) / st.E
    for label, S_gpu, S_ref in (("larch", None, orc.spline_matrix(Q)),
                                ("dense S", torch.tensor(S_rand, device="cuda"), S_rand)):
        got = st.chi_fine(cc, reff, Q, S=S_gpu, convention="larch").double().cpu().numpy()
        ref = orc.chi_fine_dense(ref_c, ref_r, Q, S_ref)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"[larch] stage {label} {dtype}: max rel {err:.2e}", flush=True)
        assert err < tol, (label, err)
    # FEFF's linear matrix passed explicitly runs the kernels' linear interpolation: same as convention "feff"
    a = st.chi_fine(cc, reff, Q, S=st.linear_matrix(Q)).cpu()
    b = st.chi_fine(cc, reff, Q).cpu()
    assert torch.equal(a, b)


def test_stage_chi_fine_larch_gradient():
    """Autograd through cchi -> chi_fine(larch) against five-point differences of the oracle (FP64)."""
    from feffgpu_amd.stage import GPUPathStage
    phase, paths, _, _ = load_case("validation/A018")
    orc = oracle.OraclePhase(phase)
    st = GPUPathStage(str(phase), dtype=torch.complex128)
    S = orc.spline_matrix(Q)
    w = np.random.default_rng(2).normal(size=len(Q)) * Q ** 2
    rng = np.random.default_rng(3)
    for p in [x for x in paths if x.nleg == 3][:2] + [x for x in paths if x.nleg == 4][:2]:
        coords = p.coords + np.concatenate([np.zeros((1, 3)), rng.normal(0, 0.05, (p.nleg - 1, 3))])
        ct = torch.tensor(coords[None], device="cuda", requires_grad=True)
        cc, reff = st.cchi(ct, torch.tensor(p.ipots[None], device="cuda"))
        loss = (st.chi_fine(cc, reff, Q, convention="larch")[0] * torch.tensor(w, device="cuda")).sum()
        (g,) = torch.autograd.grad(loss, ct)
        g = g.cpu().numpy()[0]

        def L(c):
            c1, r1, _ = orc.cchi(c, p.ipots)
            return float(orc.chi_fine_dense(c1, r1, Q, S)[0] @ w)
        d = rng.normal(size=coords.shape)
        d /= np.linalg.norm(d)
        fd = (-L(coords + 2 * H * d) + 8 * L(coords + H * d) - 8 * L(coords - H * d) + L(coords - 2 * H * d)) / (12 * H)
        an = float((g * d).sum())
        print(f"[larch] stage gradient nleg {p.nleg}: adjoint {an:.10e}, FD {fd:.10e}", flush=True)
        assert abs(an - fd) < 1e-6 * np.abs(g).sum(), (an, fd)


def test_path_list_larch_convention():
    """A cluster path list (2-4 legs) with convention "larch": chi per segment against the oracle's per-path cchi
    through its spline fine grid; the FP64 gradient against finite differences, FP32 against FP64."""
    from feffgpu_amd.paths import PathPlan
    phase = GOLDEN / "validation" / "A018" / "phase.pad"
    orc = oracle.OraclePhase(phase)
    S = orc.spline_matrix(Q)
    cl = W.cluster("bcc", 2.87, 6.0)
    pos = W.displaced(cl, 0.05, 4, keep_first=True)
    ipot = np.full(len(pos), 2)
    ipot[::3] = 1
    pl = W.concat([W.enumerate_paths(pos, 0, ipot, 4.6, 4, segment=0, nmax=4),
                   W.enumerate_paths(pos, 1, ipot, 4.2, 4, segment=1, nmax=4)], 4)
    w = np.random.default_rng(6).normal(size=(2, len(Q))) * Q ** 2

    def seg_oracle(x):
        out = np.zeros((2, len(Q)))
        for i in range(len(pl)):
            n = int(pl.nleg[i])
            at = [int(pl.absorber[i])] + [int(a) for a in pl.atoms[i, :n - 1]]
            c, r, _ = orc.cchi(x[at] - x[at[0]], pl.ipots[i, :n])
            out[pl.segment[i]] += pl.deg[i] * orc.chi_fine_dense(c, r, Q, S)[0]
        return out
    ref = seg_oracle(pos)
    grads = {}
    for prec, tol in ((64, 1e-10), (32, 1e-5)):
        plan = PathPlan(str(phase), Q, pl.atoms, pl.nleg, pl.absorber, pl.ipots, pl.deg, segment=pl.segment, nseg=2,
                        precision=prec, device="cuda", convention="larch")
        p = torch.tensor(pos, device="cuda", requires_grad=True)
        seg, _ = plan(p)
        err = float(np.abs(seg.detach().cpu().numpy() - ref).max() / np.abs(ref).max())
        print(f"[larch] path list FP{prec}: {len(pl)} paths, max rel vs oracle {err:.2e}", flush=True)
        assert err < tol
        (g,) = torch.autograd.grad((seg * torch.tensor(w, device="cuda")).sum(), p)
        grads[prec] = g.cpu().numpy()
    d = np.random.default_rng(7).normal(size=pos.shape)
    d /= np.linalg.norm(d)
    L = lambda x: float((seg_oracle(x) * w).sum())   # noqa: E731
    fd = (-L(pos + 2 * H * d) + 8 * L(pos + H * d) - 8 * L(pos - H * d) + L(pos - 2 * H * d)) / (12 * H)
    an = float((grads[64] * d).sum())
    print(f"[larch] path list gradient: adjoint {an:.10e}, FD {fd:.10e}", flush=True)
    assert abs(an - fd) < 1e-6 * np.abs(grads[64]).sum()
    assert np.abs(grads[32] - grads[64]).max() < 1e-4 * np.abs(grads[64]).max()
