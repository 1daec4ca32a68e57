"""Exact single scattering of whole configurations from tabulated two-leg signals (PairTables,
src/feffgpu/exafs/forward.py:178-320) on the GPU.

For every scatterer potential the two-leg chi(k) of the path stage is tabulated on the distance grid r_min, r_min + dr,
..., rmax (fm_cchi + fm_chi_fine). A configuration's single-scattering spectrum is the histogram of its
absorber-neighbor distances on that grid (linear-interpolation weights, fm_pair_histogram) times the tables; its
gradient in the positions goes through fm_pair_histogram_grad. Histograms are accumulated in fixed point, so they do not
depend on the order of the pairs.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _native as N
from .stage import GPUPathStage
from .transform import ETOK

_ONE = float(2 ** 32)        # the histogram's fixed-point unit (2^-32 of a pair)
_BIN_PAIRS = 2 ** 31          # pairs one call may add (a bin holds fewer than 2^31 pairs)


def _cell(box):
    """None, a cubic length or a 3 x 3 cell (rows the lattice vectors) -> (kind, cell[9], cellinv[9], smallest width)."""
    if box is None:
        return 0, np.zeros(9), np.zeros(9), math.inf
    if torch.is_tensor(box):
        box = box.detach().cpu().numpy()
    b = np.asarray(box, float)
    if b.ndim == 0:
        L = float(b)
        return 1, np.r_[L, np.zeros(8)], np.zeros(9), L
    c = np.diag(b) if b.ndim == 1 else b
    vol = abs(np.linalg.det(c))
    width = min(vol / np.linalg.norm(np.cross(c[(i + 1) % 3], c[(i + 2) % 3])) for i in range(3))
    return 2, c.reshape(-1), np.linalg.inv(c).reshape(-1), float(width)


class PairTables:
    """Tables of the two-leg signal per scatterer potential, and single scattering from distance histograms.

    The constructor takes the reference's arguments (src/feffgpu/exafs/forward.py:204-227). k and e0 give the grid
    q = sign(k^2 - e0 ETOK) sqrt(|k^2 - e0 ETOK|) on which the tables are evaluated; target, window, r_range, kweight
    and dk define the reference's misfit object and are kept as attributes only; ms is accepted, but the
    multiple-scattering sub-object is not built (``sub`` is None: PathPlan evaluates multiple scattering).

    Attributes: tables {potential: (nr, NK) float32}, r0, dr, nr, rmax, kt (the k grid), q, box.
    """

    def __init__(self, phase, k, target=None, e0=0.0, window=(3.0, 23.0), r_range=(1.0, 8.0), box=None, rmax=9.4,
                 r_min=1.5, dr=0.002, ms=True, nlegmax=4, critpw=0.5, chunk=2048, device="cuda",
                 dtype=torch.complex64, kweight=2, dk=1.0):
        self.device, self.rmax = torch.device(device), float(rmax)
        self.box = box
        self._kind, self._cellv, self._cellinv, width = _cell(box)
        if width < 2 * self.rmax:
            raise ValueError(f"the periodic cell (smallest width {width:.2f} A) must exceed twice the pair cutoff ({rmax} A) "
                             "for minimum-image pairs: use a larger supercell or a smaller cutoff")
        self.sub = None
        self.window, self.r_range, self.kweight, self.dk = window, r_range, int(kweight), float(dk)
        self.target = None if target is None else np.asarray(target, float)
        stage = GPUPathStage(phase, device=device, dtype=dtype)
        self.stage = stage
        kh = np.asarray(k, float)
        en = kh ** 2 - float(e0) * ETOK
        self.kt = torch.tensor(kh, device=self.device, dtype=torch.float32)
        self.q = torch.tensor(np.sign(en) * np.sqrt(np.abs(en)), device=self.device, dtype=torch.float64)
        self.S = stage.linear_matrix(self.q)
        self.r0, self.dr = float(r_min), float(dr)
        self.nr = int(round((self.rmax - self.r0) / self.dr)) + 1
        if self.nr < 2:
            raise ValueError("the distance grid needs at least two points (r_min < rmax)")
        rgrid = self.r0 + self.dr * torch.arange(self.nr, device=self.device, dtype=torch.float32)
        self.tables = {}
        with torch.no_grad():
            for ipot in range(1, stage.P.nph + 1):
                rows = []
                for s in range(0, self.nr, int(chunk)):
                    r = rgrid[s:s + int(chunk)].double()
                    coords = torch.zeros(len(r), 2, 3, device=self.device, dtype=torch.float64)
                    coords[:, 1, 2] = r
                    ipots = torch.tensor([[0, ipot]], device=self.device).expand(len(r), -1).contiguous()
                    cchi, reff = stage.cchi(coords, ipots)
                    rows.append(stage.chi_fine(cchi, reff, self.q, self.S, "feff").float())
                self.tables[ipot] = torch.cat(rows)
        if not self.tables:
            raise ValueError("the phase file has no scatterer potential")
        self.C, self.n = 0, 0

    # -- the distance grid and the cell as the kernels take them
    def _grid(self, box="self"):
        kind, cell, cellinv = (self._kind, self._cellv, self._cellinv) if isinstance(box, str) else _cell(box)[:3]
        g = N.FmPairGrid()
        g.r0, g.dr, g.rmax, g.nr, g.cell_kind = self.r0, self.dr, self.rmax, self.nr, kind
        g.cell[:] = [float(x) for x in cell]
        g.cellinv[:] = [float(x) for x in cellinv]
        return g

    @property
    def npot(self):
        return max(self.tables) + 1

    @torch.no_grad()
    def histogram(self, pos, absorbers, ipot, box=None, batch=1024):
        """The distance histogram of chosen absorbers of one configuration per scatterer potential
        (src/feffgpu/exafs/forward.py:233-280): pos (n, 3) A, absorbers (A,) indices, ipot (n,) potential of every
        atom (0: never a neighbor), box the configuration's cell (None: a cluster). Returns (hist (nph + 1, nr)
        float64, the number of pairs closer than the grid's start). batch is accepted for compatibility."""
        P = torch.as_tensor(pos, device=self.device).to(torch.float64).reshape(-1, 3).contiguous()
        n = P.shape[0]
        ab = torch.as_tensor(absorbers, device=self.device).to(torch.int32).reshape(-1).contiguous()
        ip = torch.as_tensor(ipot, device=self.device).to(torch.int32).reshape(-1).contiguous()
        if ip.shape[0] != n:
            raise ValueError("ipot needs one potential per atom")
        if ab.numel() and (int(ab.min()) < 0 or int(ab.max()) >= n):
            raise ValueError("absorber index outside the configuration")
        g = self._grid(box)
        hist = torch.zeros(self.npot, self.nr, dtype=torch.int64, device=self.device)
        short = torch.zeros(1, dtype=torch.int32, device=self.device)
        if ab.numel() == 0 or n == 0:        # no absorber: an empty histogram (a NULL list would mean every atom)
            return hist.double(), 0
        N.check(N.lib().fm_pair_histogram(1, n, P.data_ptr(), ab.numel(), ab.data_ptr(),
                                          ip.data_ptr(), self.npot, ctypes.byref(g), hist.data_ptr(), short.data_ptr(),
                                          N.stream_of(self.device)), "fm_pair_histogram")
        return hist.double() / _ONE, int(short.item())

    def table_stack(self, potentials=None, dtype=torch.float64):
        """The tables of the given potentials (default all), stacked: (P, nr, NK)."""
        potentials = sorted(self.tables) if potentials is None else list(potentials)
        return torch.stack([self.tables[p].to(dtype) for p in potentials])

    def pair_distances(self, pos):
        """All ordered pairs (absorber a, neighbor j) within the cutoff of pos (C, n, 3), minimum image: distances
        (P,) differentiable in pos, and the neighbors' potentials (those of ``self.ipot`` if set, else 1). A device
        helper in torch ops; single_scattering does not use it."""
        C, n, _ = pos.shape
        rel = pos[:, None, :, :] - pos[:, :, None, :]
        if self._kind == 1:
            L = float(self._cellv[0])
            rel = rel - L * torch.round(rel / L)
        elif self._kind == 2:
            cell = torch.tensor(self._cellv.reshape(3, 3), device=pos.device, dtype=pos.dtype)
            f = rel @ torch.linalg.inv(cell)
            rel = (f - torch.round(f)) @ cell
        r2 = (rel ** 2).sum(-1)
        mask = (r2 <= self.rmax ** 2) & (r2 > 1e-6)
        c_idx, a_idx, j_idx = torch.nonzero(mask, as_tuple=True)
        r = torch.sqrt(r2[c_idx, a_idx, j_idx])
        return r, self.ipot[j_idx] if hasattr(self, "ipot") else torch.ones_like(j_idx)

    def single_scattering(self, pos):
        """The single-scattering chi(k) (NK,) of pos (C, n, 3), every atom of every copy an absorber, averaged
        (src/feffgpu/exafs/forward.py:300-320); neighbors' potentials from ``self.ipot`` (n,) if set, else 1.
        Differentiable in pos; returned in pos's floating dtype."""
        return _SingleScattering.apply(pos, self)


class _SingleScattering(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, pt):
        if not (torch.is_tensor(pos) and pos.is_cuda):
            raise RuntimeError("single_scattering runs on a ROCm device: pos must be a device tensor")
        C, n, _ = pos.shape
        if C * n == 0:
            raise ValueError("single_scattering needs at least one atom in one configuration")
        P = pos.detach().to(torch.float64).contiguous()
        ip = (torch.as_tensor(pt.ipot, device=pos.device) if hasattr(pt, "ipot")
              else torch.ones(n, dtype=torch.int64, device=pos.device)).to(torch.int32).contiguous()
        g = pt._grid()
        # configurations in chunks of at most _BIN_PAIRS ordered pairs (the fixed-point bins' capacity), summed in FP64
        per = max(1, (_BIN_PAIRS - 1) // max(1, n * (n - 1)))
        h = torch.zeros(pt.npot, pt.nr, dtype=torch.float64, device=pos.device)
        hist = torch.empty(pt.npot, pt.nr, dtype=torch.int64, device=pos.device)
        for c0 in range(0, C, per):
            cc = min(per, C - c0)
            hist.zero_()
            N.check(N.lib().fm_pair_histogram(cc, n, P[c0:c0 + cc].data_ptr(), 0, None, ip.data_ptr(), pt.npot,
                                              ctypes.byref(g), hist.data_ptr(), None, N.stream_of(pos.device)),
                    "fm_pair_histogram")
            h += hist.double() / _ONE
        pots = sorted(pt.tables)
        T = pt.table_stack(pots)                                   # (P, nr, NK) float64
        chi = sum(torch.mv(T[i].t().contiguous(), h[p]) for i, p in enumerate(pots)) / (C * n)
        ctx.save_for_backward(P, ip, T)
        ctx.pt, ctx.pots, ctx.shape, ctx.dtype = pt, pots, (C, n), pos.dtype
        return chi.to(pos.dtype)

    @staticmethod
    def backward(ctx, gchi):
        P, ip, T = ctx.saved_tensors
        pt, (C, n) = ctx.pt, ctx.shape
        gbin = torch.zeros(pt.npot, pt.nr, dtype=torch.float64, device=P.device)
        gk = gchi.to(torch.float64).contiguous()
        for i, p in enumerate(ctx.pots):
            gbin[p] = torch.mv(T[i], gk) / (C * n)
        grad = torch.empty(C, n, 3, dtype=torch.float64, device=P.device)
        g = pt._grid()
        N.check(N.lib().fm_pair_histogram_grad(C, n, P.data_ptr(), ip.data_ptr(), pt.npot, ctypes.byref(g),
                                               gbin.contiguous().data_ptr(), grad.data_ptr(), N.stream_of(P.device)),
                "fm_pair_histogram_grad")
        return grad.to(ctx.dtype), None
