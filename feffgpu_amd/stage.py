"""Drop-in ``GPUPathStage``: feffgpu's path-stage API (src/feffgpu/stage.py:96-388) backed by the HIP kernels.

Same constructor, attributes and methods as the reference class:

    stage = GPUPathStage(phase, device="cuda", dtype=torch.complex64, iorder=2)
    cchi, reff = stage.cchi(coords, ipots)          # (B, E) complex, (B,) A   -- coords (B, nleg, 3) A
    S = stage.linear_matrix(q)                       # (NK, E)
    chi = stage.chi_fine(cchi, reff, q, S, "feff")   # (B, NK)

``dtype=torch.complex64`` runs the FP32 kernels (FP64 leg geometry and path phase, FP32 amplitudes);
``torch.complex128`` the FP64 kernels. Everything runs on the GPU: there is no CPU path, and constructing a stage on
a device without the HIP library raises.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _native as N
from .pad import BOHR, PhaseData, load_phase, read_phase_pad
from .tables import edge_tables, interpolation, setkap
from .tables import spline_matrix as tables_spline_matrix


def _require_gpu(device: torch.device):
    if device.type != "cuda":
        raise RuntimeError("GPUPathStage runs on a ROCm GPU only (device must be 'cuda'); this package has no CPU path")
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU is available")


class _Cchi(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, ipots, stage):
        out, reff64 = stage._cchi_raw(coords, ipots)
        stage._reff64_last = reff64
        ctx.stage = stage
        ctx.save_for_backward(coords, ipots)
        return out, reff64.to(stage.rdtype)

    @staticmethod
    def backward(ctx, gcchi, greff):
        coords, ipots = ctx.saved_tensors
        g = ctx.stage._cchi_backward_raw(coords, ipots, gcchi, greff)
        return g, None, None


class _ChiFine(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cchi, reff, stage, grid, reff64):
        r64 = reff.detach().to(torch.float64) if reff64 is None else reff64
        chi = stage._chi_fine_raw(cchi, r64, grid)
        ctx.stage, ctx.grid, ctx.rdtype = stage, grid, reff.dtype
        ctx.save_for_backward(cchi, r64)
        return chi

    @staticmethod
    def backward(ctx, gchi):
        cchi, r64 = ctx.saved_tensors
        gc, gr = ctx.stage._chi_fine_backward_raw(cchi, r64, ctx.grid, gchi)
        return gc, gr.to(ctx.rdtype), None, None, None


class GPUPathStage:
    def __init__(self, phase, device="cuda", dtype=torch.complex64, iorder=2):
        P = read_phase_pad(phase) if isinstance(phase, str) else load_phase(phase)
        if P.nsp != 1:
            raise NotImplementedError("spin-dependent phase files are not ported")
        if dtype not in (torch.complex64, torch.complex128):
            raise ValueError("dtype must be torch.complex64 or torch.complex128")
        self.P, self.iorder, self.device, self.cdtype = P, int(iorder), torch.device(device), dtype
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        _require_gpu(self.device)
        self.rdtype = torch.float32 if dtype == torch.complex64 else torch.float64
        self.precision = 32 if dtype == torch.complex64 else 64
        E = self.E = P.ne1
        lmax = P.lmax[:E]
        self.L1 = int(lmax.max()) + 1
        self.lmax = torch.tensor(lmax, device=self.device)
        self.xk = torch.tensor(P.xk[:E], device=self.device, dtype=self.rdtype)
        self.ck = torch.tensor(P.ck[:E], device=self.device, dtype=dtype)
        kinit, linit = setkap(P.ihole)
        self.kinit, self.ilinit = kinit, linit + 1
        self.tables = edge_tables(P, E)
        self._edge = N.DeviceEdge(self.tables, self.device)
        self._grids = {}
        self._linear, self._spline = {}, {}
        N.lib()

    # -- the path signal on the phase grid ----------------------------------------------------------------
    def cchi(self, coords, ipots):
        """coords (B, nleg, 3) A (absorber first at the origin), ipots (B, nleg) long (absorber first). Returns cchi
        (B, E) complex (the stage's dtype) and reff (B,) A; differentiable in coords."""
        coords = torch.as_tensor(coords)
        if coords.dim() != 3 or coords.shape[-1] != 3:
            raise ValueError("coords must be (B, nleg, 3)")
        ipots = torch.as_tensor(ipots)
        if tuple(ipots.shape) != tuple(coords.shape[:2]):
            raise ValueError("ipots must be (B, nleg)")
        if ipots.device.type == "cpu" and ipots.numel() and (int(ipots.min()) < 0 or int(ipots.max()) >= self.tables.nph1):
            raise ValueError(f"potential index out of range 0..{self.tables.nph1 - 1}")
        out, reff = _Cchi.apply(coords, ipots.to(self.device), self)
        # The FP64 effective length behind the returned reff: chi_fine uses it when it is given this very tensor
        # (unchanged), so that 2 q reff (hundreds of radians) carries no FP32 rounding in complex64 mode.
        reff._fm_reff64 = (self._reff64_last, reff._version)
        self._reff64_last = None
        return out, reff

    def _reff64_of(self, reff):
        """The FP64 values behind a reff tensor returned by cchi, or None (another tensor, or changed since)."""
        tag = getattr(reff, "_fm_reff64", None) if torch.is_tensor(reff) else None
        if tag is None or tag[1] != reff._version or tag[0].shape != reff.shape:
            return None
        return tag[0]

    def _cchi_raw(self, coords, ipots):
        """Potentials outside 0..nph of a path on the device are not checked here (that would synchronise the
        device): the kernel reads potential 0's tables instead and writes NaN for that path."""
        B, nleg, _ = coords.shape
        c = coords.detach().to(device=self.device, dtype=torch.float64).contiguous()
        ip = ipots.to(device=self.device, dtype=torch.int64).contiguous()
        out = torch.empty(B, self.E, dtype=self.cdtype, device=self.device)
        reff = torch.empty(B, dtype=torch.float64, device=self.device)
        if B:
            N.check(N.lib().fm_cchi(self._edge.struct, self.precision, self.iorder, B, nleg, c.data_ptr(), ip.data_ptr(),
                                    out.data_ptr(), reff.data_ptr(), N.stream_of(self.device)), "fm_cchi")
        return out, reff

    def _cchi_backward_raw(self, coords, ipots, gcchi, greff):
        B, nleg, _ = coords.shape
        c = coords.detach().to(device=self.device, dtype=torch.float64).contiguous()
        ip = ipots.to(device=self.device, dtype=torch.int64).contiguous()
        gc = (torch.zeros(B, self.E, dtype=self.cdtype, device=self.device) if gcchi is None
              else gcchi.to(self.cdtype).contiguous())
        gr = None if greff is None else greff.to(torch.float64).contiguous()
        out = torch.zeros(B, nleg, 3, dtype=torch.float64, device=self.device)
        if B:
            ws = N.backward_workspace(self.precision, nleg, B, self.device, self.E)
            N.check(N.lib().fm_cchi_backward(self._edge.struct, self.precision, self.iorder, B, nleg, c.data_ptr(),
                                             ip.data_ptr(), gc.data_ptr(), N.ptr(gr), out.data_ptr(), ws.data_ptr(),
                                             ws.numel(), N.stream_of(self.device)), "fm_cchi_backward")
        return out.to(coords.dtype)

    # -- fine grid ----------------------------------------------------------------------------------------
    def spline_matrix(self, q):
        """Cubic-spline interpolation matrix (len(q), E) (scipy UnivariateSpline s=0, as Larch); q clamped to the
        grid. Host-built once, as in src/feffgpu/stage.py:341-351."""
        S = tables_spline_matrix(self.P.xk[: self.E] / BOHR, self._qarray(q))
        return torch.tensor(S, device=self.device, dtype=self.rdtype)

    def linear_matrix(self, q):
        """Linear interpolation matrix (len(q), E) of FEFF's final-stage convention; q clamped to the grid. The
        returned tensor is cached per grid: chi_fine recognises it and runs the kernels' own linear interpolation."""
        key = self._qkey(q)
        S = self._linear.get(key)
        if S is None:
            kf = self.P.xk[: self.E] / BOHR
            j, f = interpolation(kf, self._qarray(q))
            m = np.zeros((len(j), self.E))
            m[np.arange(len(j)), j] = 1 - f
            m[np.arange(len(j)), j + 1] = f
            S = self._linear[key] = torch.tensor(m, device=self.device, dtype=self.rdtype)
        return S

    @staticmethod
    def _qarray(q):
        return np.ascontiguousarray(np.asarray(q.detach().cpu() if torch.is_tensor(q) else q, dtype=np.float64))

    def _qkey(self, q):
        qa = self._qarray(q)
        return qa.shape[0], hash(qa.tobytes())

    @staticmethod
    def unwrap(phase):
        """Unwrap along the last axis without branches (differences brought into (-pi, pi])."""
        d = torch.diff(phase, dim=-1)
        d = d - 2 * math.pi * torch.round(d / (2 * math.pi))
        return torch.cat([phase[..., :1], phase[..., :1] + torch.cumsum(d, -1)], -1)

    def _grid(self, q, S=None):
        """The kernels' grid for q: linear interpolation (S None; cached per grid) or the dense matrix S (len(q), E),
        a grid of its own per call that holds S in FP64 for as long as the autograd graph keeps it."""
        qa = np.ascontiguousarray(np.asarray(q.detach().cpu() if torch.is_tensor(q) else q, dtype=np.float64))
        key = (qa.shape[0], hash(qa.tobytes()))
        g = self._grids.get(key)
        if g is None:
            j, f = interpolation(self.P.xk[: self.E] / BOHR, qa)
            qorder = np.argsort(qa, kind="stable").astype(np.int32)
            qstart = np.searchsorted(j[qorder], np.arange(self.E + 1), side="left").astype(np.int32)
            qt = torch.tensor(qa, device=self.device)
            jt = torch.tensor(j, device=self.device, dtype=torch.int32)
            ft = torch.tensor(f, device=self.device, dtype=torch.float64)
            ot = torch.tensor(qorder, device=self.device)
            st = torch.tensor(qstart, device=self.device)
            s = N.FmGrid()
            s.nk, s.q, s.jq, s.fq = len(qa), qt.data_ptr(), jt.data_ptr(), ft.data_ptr()
            s.qorder, s.qstart, s.qstride = ot.data_ptr(), st.data_ptr(), self.E + 1
            g = self._grids[key] = (s, qt, jt, ft, ot, st)
        if S is None:
            return g
        dense = torch.as_tensor(S).detach().to(device=self.device, dtype=torch.float64).contiguous()
        if tuple(dense.shape) != (len(qa), self.E):
            raise ValueError(f"S must be (len(q), E) = ({len(qa)}, {self.E})")
        s = N.FmGrid.from_buffer_copy(g[0])
        s.S, s.sstride = dense.data_ptr(), self.E
        return (s,) + g[1:] + (dense,)

    def chi_fine(self, cchi, reff, q, S=None, convention="feff"):
        """chi(q) = A(q) sin(Phi(q) + 2 q reff) per path, A and Phi |cchi| and its unwrapped phase interpolated from
        the phase grid by S (src/feffgpu/stage.py:372-388): linear_matrix(q) for convention "feff" (FEFF's), the
        cubic spline_matrix(q) for "larch", or the S given. Returns (B, len(q)) in the stage's real dtype;
        differentiable in cchi and reff. The matrix linear_matrix returned runs as the kernels' own linear
        interpolation; any other S as a dense interpolation matrix."""
        reff = torch.as_tensor(reff, device=self.device)
        r64 = self._reff64_of(reff)
        if S is None:
            if convention == "feff":
                return _ChiFine.apply(cchi, reff, self, self._grid(q), r64)
            if convention != "larch":
                raise ValueError("convention must be 'feff' or 'larch'")
            key = self._qkey(q)
            S = self._spline.get(key)
            if S is None:     # FP64 whatever the stage's dtype: the kernels interpolate in FP64
                kf = self.P.xk[: self.E] / BOHR
                S = self._spline[key] = torch.tensor(tables_spline_matrix(kf, self._qarray(q)), device=self.device)
        elif S is self._linear.get(self._qkey(q)):
            S = None
        return _ChiFine.apply(cchi, reff, self, self._grid(q, S), r64)

    def _chi_fine_raw(self, cchi, reff, grid):
        c = cchi.detach().to(self.cdtype).contiguous()
        r = reff.detach().to(torch.float64).contiguous()
        B = c.shape[0]
        out = torch.empty(B, grid[0].nk, dtype=self.rdtype, device=self.device)
        if B:
            N.check(N.lib().fm_chi_fine(self._edge.struct, self.precision, B, c.data_ptr(), r.data_ptr(), grid[0],
                                        out.data_ptr(), N.stream_of(self.device)), "fm_chi_fine")
        return out

    def _chi_fine_backward_raw(self, cchi, reff, grid, gchi):
        c = cchi.detach().to(self.cdtype).contiguous()
        r = reff.detach().to(torch.float64).contiguous()
        g = gchi.to(self.rdtype).contiguous()
        B = c.shape[0]
        gc = torch.zeros_like(c)
        gr = torch.zeros(B, dtype=torch.float64, device=self.device)
        if B:
            N.check(N.lib().fm_chi_fine_backward(self._edge.struct, self.precision, B, c.data_ptr(), r.data_ptr(),
                                                 grid[0], g.data_ptr(), gc.data_ptr(), gr.data_ptr(),
                                                 N.stream_of(self.device)), "fm_chi_fine_backward")
        return gc, gr
