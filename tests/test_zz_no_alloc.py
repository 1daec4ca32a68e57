"""fm_cchi_backward_noalloc (its workspace a static device array, nothing allocated) against fm_cchi_backward with a
workspace from the caller: two calls in a row, the second with fewer paths, FP64 and FP32, 2-6 and 8 legs (the fused
adjoint and the prefix-product one), and the larger call's paths exceed what one chunk of the scratch holds for
nothing: both entries must give the same d loss / d coords."""
import numpy as np
import pytest
import torch

from fixtures import GOLDEN

pytestmark = pytest.mark.gpu
PHASE = GOLDEN / "W" / "phase.pad"


def _coords(rng, B, n):
    a = 3.1652
    steps = np.vstack([np.array([[i, j, l] for i in (-.5, .5) for j in (-.5, .5) for l in (-.5, .5)]),
                       np.eye(3), -np.eye(3)]) * a
    out = np.zeros((B, n, 3))
    for b in range(B):
        s = np.zeros(3)
        for j in range(1, n):
            while True:                       # a lattice step that does not land on the absorber's site
                t = s + steps[rng.integers(len(steps))]
                if np.linalg.norm(t) > 1e-6:
                    break
            s = t
            out[b, j] = s + rng.normal(0, 0.05, 3)
    return out


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64], ids=["fp64", "fp32"])
@pytest.mark.parametrize("nleg", [2, 4, 6, 8])
def test_noalloc_adjoint_equals_workspace_adjoint(dtype, nleg):
    from feffgpu_amd import _native as N
    from feffgpu_amd.stage import GPUPathStage
    st = GPUPathStage(str(PHASE), dtype=dtype)
    rng = np.random.default_rng(nleg)
    L = N.lib()
    for B in (3500, 37):                                   # the second call smaller than the first
        c = torch.tensor(_coords(rng, B, nleg), device="cuda")
        ip = torch.zeros(B, nleg, dtype=torch.int64, device="cuda")
        ip[:, 1:] = 1
        g = torch.randn(B, st.E, dtype=dtype, device="cuda")
        gr = torch.randn(B, dtype=torch.float64, device="cuda")
        ref = st._cchi_backward_raw(c, ip, g, gr)
        out = torch.full((B, nleg, 3), float("nan"), dtype=torch.float64, device="cuda")
        rc = L.fm_cchi_backward_noalloc(st._edge.struct, st.precision, 2, B, nleg, c.data_ptr(), ip.data_ptr(),
                                        g.contiguous().data_ptr(), gr.data_ptr(), out.data_ptr(), N.stream_of(st.device))
        assert rc == 0, L.fm_last_error()
        torch.cuda.synchronize()
        a, b = out.cpu().numpy(), ref.cpu().numpy()
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"[noalloc] nleg {nleg} FP{st.precision} B {B}: max rel {err:.1e}, identical {np.array_equal(a, b)}",
              flush=True)
        assert np.isfinite(a).all() and err <= (1e-12 if st.precision == 64 else 1e-6)
