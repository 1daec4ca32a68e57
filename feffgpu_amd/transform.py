"""The misfit after the path sum on the GPU: xraylarch's k window, transform to R and R-factor, with its gradient in
chi, as one HIP kernel (fm_r_factor) per call (src/feffgpu/exafs/transform.py).

    window = FitWindow(kmin=3, kmax=14, dk=1, kweight=2, rmin=1, rmax=6)
    rf = r_factor_torch(model, data, k, window)          # (B,), differentiable in model
    step = CapturedStep(plan, natom, loss_grad=RFactorLoss(data, q, window, scale=1 / nabs))

The transform is evaluated at the R points of the window only (a direct transform against a twiddle table, FP64), so
it needs no FFT and no intermediate in HBM; the kernel allocates nothing and can be captured in a hipGraph.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import asdict, dataclass

import numpy as np
import torch

from . import _native as N

ETOK = 0.2624684236075176          # eV -> 1/A^2 (src/feffgpu/exafs/transform.py:22)
KSTEP = 0.05                       # 1/A
NFFT = 2048
RSTEP = math.pi / (KSTEP * NFFT)   # A


@dataclass(frozen=True)
class FitWindow:
    """Where a model is compared with a spectrum. The field names and defaults are the reference's interface
    (src/feffgpu/exafs/transform.py:28-62), and so are the conditions that raise ValueError.

    The Hanning window rises over a sill of width dk centred on kmin and falls over one centred on kmax (1/A); chi is
    multiplied by k**kweight before the transform; the R-factor sums over rmin <= R <= rmax (A)."""
    kmin: float = 3.0
    kmax: float = 14.0
    dk: float = 1.0
    kweight: int = 2
    rmin: float = 1.0
    rmax: float = 6.0

    def __post_init__(self):
        if not (self.kmax > self.kmin >= 0.0):
            raise ValueError(f"FitWindow: kmin={self.kmin}, kmax={self.kmax}; kmin must be >= 0 and below kmax")
        if not (self.rmax > self.rmin >= 0.0):
            raise ValueError(f"FitWindow: rmin={self.rmin}, rmax={self.rmax}; rmin must be >= 0 and below rmax")
        if self.dk < 0:
            raise ValueError(f"FitWindow: sill width dk={self.dk} is negative")

    def as_dict(self):
        return asdict(self)

    @property
    def independent_points(self):
        """The number of independent points of the window, nidp = 2 dk_range dr_range / pi + 2 (Stern's estimate, as
        xraylarch counts it)."""
        return 2.0 * (self.kmax - self.kmin) * (self.rmax - self.rmin) / math.pi + 2.0


def k_grid(kmax=24.0, kstep=KSTEP):
    """The uniform k grid 0, kstep, ..., kmax (numpy)."""
    return kstep * np.arange(int(round(kmax / kstep)) + 1)


def r_grid(device=None):
    """The R grid of the transform (A), (NFFT / 2,) float32, as src/feffgpu/exafs/transform.py:125-127."""
    return RSTEP * torch.arange(NFFT // 2, device=device, dtype=torch.float32)


def r_range(window: FitWindow, kstep=KSTEP, nfft=NFFT, float32_grid=True):
    """(m0, nm): the R points m of the R-factor, rmin <= R_m <= rmax with R_m = m pi / (nfft kstep), m < nfft / 2.
    float32_grid: R_m and the bounds in float32, as r_factor_torch builds them (src/feffgpu/exafs/transform.py:140-141);
    otherwise in float64, as compare.r_factor (src/feffgpu/compare.py:24-25)."""
    m = np.arange(nfft // 2)
    if float32_grid:
        r = np.float32(math.pi / (kstep * nfft)) * m.astype(np.float32)
        keep = (r >= np.float32(window.rmin)) & (r <= np.float32(window.rmax))
    else:
        r = m * math.pi / (nfft * kstep)
        keep = (r >= window.rmin) & (r <= window.rmax)
    idx = np.nonzero(keep)[0]
    if len(idx) == 0:
        raise ValueError("the R window holds no point of the transform's grid")
    return int(idx[0]), int(idx[-1] - idx[0] + 1)


def _kstep_of(k):
    kh = np.asarray(k.detach().cpu() if torch.is_tensor(k) else k, dtype=np.float64)
    if kh.ndim != 1 or len(kh) < 2 or kh[0] != 0.0:
        raise ValueError("k must be a uniform grid from zero")
    step = float(kh[1] - kh[0])
    if np.abs(np.diff(kh) - step).max() > 1e-9 * max(step, 1.0):
        raise ValueError("k must be a uniform grid from zero")
    return step, len(kh)


def r_factor_raw(chi, data, window: FitWindow, kstep=KSTEP, nfft=NFFT, scale=1.0, grad=False, float32_grid=True):
    """fm_r_factor on device tensors: chi (S, NK) float64 on the grid k_j = j kstep, data (S, NK) or (NK,) float64.
    Returns rfac (S,) and, with grad, d rfac / d chi (S, NK)."""
    chi = chi.contiguous()
    data = data.contiguous()
    if chi.dim() != 2:
        raise ValueError("chi must be (S, NK)")
    S, nk = chi.shape
    if not ((data.dim() == 1 and data.shape[0] == nk) or (data.dim() == 2 and tuple(data.shape) == (S, nk))):
        raise ValueError("data must be (NK,) or (S, NK) like chi")
    m0, nm = r_range(window, kstep, nfft, float32_grid)
    w = N.FmWindow(float(window.kmin), float(window.kmax), float(window.dk), int(window.kweight), float(window.rmin),
                  float(window.rmax))
    rf = torch.empty(S, dtype=torch.float64, device=chi.device)
    g = torch.empty_like(chi) if grad else None
    L = N.lib()
    N.check(L.fm_r_factor(S, nk, float(kstep), int(nfft), ctypes.byref(w), m0, nm, chi.data_ptr(), float(scale),
                          data.data_ptr(), 0 if data.dim() == 1 else nk, rf.data_ptr(), N.ptr(g),
                          N.stream_of(chi.device)), "fm_r_factor")
    return rf, g


class _RFactor(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, data, window, kstep):
        rf, g = r_factor_raw(model, data, window, kstep, grad=ctx.needs_input_grad[0])
        ctx.save_for_backward(g if g is not None else torch.empty(0, device=model.device))
        return rf

    @staticmethod
    def backward(ctx, grf):
        (g,) = ctx.saved_tensors
        return grf[:, None] * g, None, None, None


def r_factor_torch(model, data, k, window: FitWindow):
    """The R-factor of every row of model (B, NK) against data (NK,), differentiable in model, on the GPU in FP64
    (src/feffgpu/exafs/transform.py:130-145): k (NK,) is the uniform grid from zero with step KSTEP, the transform's
    normalisation and R grid are KSTEP's and NFFT's, the R range is taken on the float32 R grid."""
    if not (torch.is_tensor(model) and model.is_cuda):
        raise RuntimeError("r_factor_torch runs on a ROCm device: model must be a device tensor")
    step, _ = _kstep_of(k)
    if abs(step - KSTEP) > 1e-9:   # the kernel builds k_j = j KSTEP for k^w and the window
        raise ValueError(f"k must be the uniform grid from zero with step {KSTEP} (got step {step})")
    m = model.double()
    d = torch.as_tensor(data, device=model.device).double()
    return _RFactor.apply(m, d, window, KSTEP)


class RFactorLoss:
    """The sum over segments of the R-factor of scale * chi_seg against data, as a loss_grad for CapturedStep:
    called on chi_seg (S, NK) it returns d loss / d chi_seg from one fm_r_factor launch, and keeps the R-factors of
    that call in `rfac` (a static buffer, so that it also works inside a captured graph). q: the plan's output grid
    (uniform from zero), whose step is the transform's; data: (NK,) or (S, NK). The R range is taken on the float64
    R grid, as compare.r_factor takes it."""

    def __init__(self, data, q, window: FitWindow, scale=1.0, nfft=NFFT, device="cuda"):
        self.kstep, nk = _kstep_of(q)
        self.window, self.scale, self.nfft = window, float(scale), int(nfft)
        self.data = torch.as_tensor(data, dtype=torch.float64, device=device).contiguous()
        if self.data.dim() not in (1, 2) or self.data.shape[-1] != nk:
            raise ValueError("data must be (NK,) or (S, NK), one value per point of q")
        self.m0, self.nm = r_range(window, self.kstep, self.nfft, float32_grid=False)
        self._w = N.FmWindow(float(window.kmin), float(window.kmax), float(window.dk), int(window.kweight),
                            float(window.rmin), float(window.rmax))
        self.rfac = None
        self._g = None

    def __call__(self, seg):
        S, nk = seg.shape
        if self.data.dim() == 2 and self.data.shape[0] != S:
            raise ValueError("data holds another number of segments than chi_seg")
        if self.rfac is None or self.rfac.shape[0] != S:
            self.rfac = torch.empty(S, dtype=torch.float64, device=seg.device)
            self._g = torch.empty(S, nk, dtype=torch.float64, device=seg.device)
        seg = seg.contiguous()
        N.check(N.lib().fm_r_factor(S, nk, self.kstep, self.nfft, ctypes.byref(self._w), self.m0, self.nm,
                                    seg.data_ptr(), self.scale, self.data.data_ptr(),
                                    0 if self.data.dim() == 1 else nk, self.rfac.data_ptr(), self._g.data_ptr(),
                                    N.stream_of(seg.device)), "fm_r_factor")
        return self._g
