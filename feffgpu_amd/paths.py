"""The fused path stage over a path list: positions in, chi per segment and per path out, differentiable in the
positions.

This is the path-list form of feffgpu's forward model (``PathSpectra.spectra`` and
``MultipleScattering.signals/path_signals``, src/feffgpu/exafs/forward.py:143-163, 657-685, and the sum of
``chain.run``, src/feffgpu/chain.py:45-49): every path p is a sequence of atom indices after its absorber; its
coordinates are ``[0, wrap(pos[atoms[p]] - pos[absorber[p]])]`` (minimum image, src/feffgpu/cell.py:26-35); its
chi on the output grid q is the path stage's chi_fine (src/feffgpu/stage.py:372-388) times an optional Debye-Waller
factor exp(-2 sigma^2 q^2); and ``chi_seg[s] = sum_{p in s} deg[p] chi_p`` (index_add in the reference; here a
deterministic ordered sum).

A :class:`PathPlan` holds a path list on the device together with its bucketing (paths grouped by the kernel
variant their edge and number of legs need, longest first), the segment lists and the tables of every edge; its
``forward(pos)`` is one call into the C ABI. ``chi_paths`` is the one-shot functional form.
"""
from __future__ import annotations

from dataclasses import dataclass

import ctypes

import numpy as np
import torch

from . import _native as N
from .pad import load_phase
from .tables import edge_tables, energies_needed, interpolation, spline_matrix


def _cell_struct(cell):
    """(kind, cell[9], inverse[9]) of a cluster (None), a cubic box (a number) or a 3 x 3 cell (rows = vectors)."""
    if cell is None:
        return 0, np.zeros(9), np.zeros(9)
    c = np.asarray(cell.detach().cpu() if torch.is_tensor(cell) else cell, float)
    if c.ndim == 0 or c.size == 1:
        v = np.zeros(9)
        v[0] = float(c.reshape(-1)[0])
        return 1, v, np.zeros(9)
    if c.shape == (3,):
        c = np.diag(c)
    if c.shape != (3, 3):
        raise ValueError("cell must be None, a cubic length, 3 lengths or a 3 x 3 matrix")
    return 2, c.reshape(-1).copy(), np.linalg.inv(c).reshape(-1).copy()


@dataclass
class Buckets:
    order: np.ndarray          # permutation of the paths
    offset: np.ndarray         # (nb + 1,)
    kind: np.ndarray           # (nb,)
    edge: np.ndarray           # (nb,)
    nleg: np.ndarray           # (nb,) legs of every path of the bucket


def bucket_paths(nleg: np.ndarray, edge: np.ndarray, ilinit: list) -> Buckets:
    """Group the paths by kernel variant -- (K/L1 edge, 3+ legs), (L/M edge, 3+ legs), (any edge, 2 legs) --, by
    edge and by number of legs: the kernels run the legs of a launch with wave-uniform loops."""
    nleg = np.asarray(nleg)
    edge = np.asarray(edge)
    il = np.asarray(ilinit)[edge]
    kind = np.where(nleg == 2, N.FM_BUCKET_SS, np.where(il <= 1, N.FM_BUCKET_K_MS, N.FM_BUCKET_LM_MS))
    order = np.lexsort((np.arange(len(nleg)), -nleg, edge, kind)).astype(np.int32)
    key = ((kind * 16 + edge) * 16 + (15 - nleg))[order]
    off, kinds, edges, nlm = [0], [], [], []
    for kv in np.unique(key):
        sel = key == kv
        off.append(off[-1] + int(sel.sum()))
        kinds.append(int(kv) // 256)
        edges.append(int(kv) // 16 % 16)
        nlm.append(15 - int(kv) % 16)
    return Buckets(order, np.array(off, np.int32), np.array(kinds, np.int32), np.array(edges, np.int32),
                   np.array(nlm, np.int32))


class PathPlan:
    """A path list resident on the GPU, ready for repeated evaluation at new positions.

    phases: a phase file (path or PhaseData) or a list of them (one per edge; ``edge[p]`` picks one).
    q: the output grid (1/A). atoms (P, nmax-1) scatterer indices (padding beyond nleg-1 is ignored), nleg (P,),
    absorber (P,), ipots (P, nmax) (absorber first), deg (P,), segment (P,) (default 0), edge (P,) (default 0),
    sigma2 (P,) or a number (default none). cell: None, a cubic length or a 3 x 3 cell. precision 32 or 64.
    convention: "feff" (linear interpolation of amplitude and phase, FEFF's, as the reference's forward model uses)
    or "larch" (the cubic spline_matrix of every edge over all its phase-grid energies, a dense interpolation).
    keep_amplitudes: the differentiable call plan(pos) and CapturedStep keep the scattering matrices of the 3-6-leg
    multiple-scattering paths from the forward call (fm_keep_bytes per bucket: about 28-88 complex per path and
    energy, allocated at the first such call) and the adjoint at the same positions reads them instead of
    recomputing them; False: nothing is kept between forward and backward and the adjoint recomputes the amplitudes,
    as the reference's checkpointed autograd does. A plan whose buffers would exceed keep_limit_bytes keeps nothing.
    The explicit forward / backward calls keep and use them only when asked (keep=, kept=).
    """

    def __init__(self, phases, q, atoms, nleg, absorber, ipots, deg, segment=None, nseg=None, edge=None, sigma2=None,
                 cell=None, precision=32, device="cuda", iorder=2, convention="feff", keep_amplitudes=True,
                 keep_limit_bytes=16 << 30):
        self.keep_amplitudes = bool(keep_amplitudes)
        self.keep_limit_bytes = int(keep_limit_bytes)
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("PathPlan runs on a ROCm GPU only")
        if precision not in (32, 64):
            raise ValueError("precision must be 32 or 64")
        self.precision, self.iorder = int(precision), int(iorder)
        if convention not in ("feff", "larch"):
            raise ValueError("convention must be 'feff' or 'larch'")
        self.convention = convention
        self.rdtype = torch.float32 if precision == 32 else torch.float64
        phases = phases if isinstance(phases, (list, tuple)) else [phases]
        self.phases = [load_phase(p) for p in phases]
        if not 1 <= len(self.phases) <= 4:
            raise ValueError("one to four edges per plan")
        self.q = np.ascontiguousarray(np.asarray(q, float))
        nk = len(self.q)
        atoms = np.asarray(atoms, np.int64)
        nleg = np.asarray(nleg, np.int64)
        Pn = len(nleg)
        if atoms.ndim != 2 or atoms.shape[0] != Pn:
            raise ValueError("atoms must be (P, nmax-1)")
        nmax = atoms.shape[1] + 1
        if not 2 <= nmax <= N.FM_MAX_LEGS:
            raise ValueError("nmax out of range 2..9")
        if Pn and (nleg.min() < 2 or nleg.max() > nmax):
            raise ValueError(f"nleg must be within 2..{nmax}")
        ipots = np.asarray(ipots, np.int64)
        if ipots.shape != (Pn, nmax):
            raise ValueError("ipots must be (P, nmax)")
        edge = np.zeros(Pn, np.int64) if edge is None else np.asarray(edge, np.int64)
        if Pn and (edge.min() < 0 or edge.max() >= len(self.phases)):
            raise ValueError("edge index out of range")
        segment = np.zeros(Pn, np.int64) if segment is None else np.asarray(segment, np.int64)
        self.nseg = int(nseg if nseg is not None else (segment.max() + 1 if Pn else 1))
        if Pn and (segment.min() < 0 or segment.max() >= self.nseg):
            raise ValueError("segment index out of range")
        deg = np.asarray(deg, np.float64)
        for e, ph in enumerate(self.phases):
            sel = edge == e
            if sel.any():
                used = np.concatenate([ipots[sel][np.arange(nmax)[None, :] < nleg[sel][:, None]]])
                if used.min() < 0 or used.max() > ph.nph:
                    raise ValueError(f"potential index out of range 0..{ph.nph} for edge {e}")
        # per-edge tables on the energies the grid reads
        self.tables, self.edges = [], []
        jq, fq = [], []
        for ph in self.phases:
            # the linear interpolation reads the energies up to the grid's last upper point; a spline all of them
            E = energies_needed(ph, self.q) if convention == "feff" else ph.ne1
            T = edge_tables(ph, E)
            self.tables.append(T)
            self.edges.append(N.DeviceEdge(T, self.device))
            j, f = interpolation(T.kf_full, self.q)
            jq.append(j)
            fq.append(f)
        self.edge_array = (N.FmEdge * len(self.edges))(*[e.struct for e in self.edges])
        dev = self.device
        self.q_t = torch.tensor(self.q, device=dev)
        self.jq_t = torch.tensor(np.stack(jq), device=dev, dtype=torch.int32)
        self.fq_t = torch.tensor(np.stack(fq), device=dev, dtype=torch.float64)
        qorder, qstart, qstride = N.grid_runs(self.q, jq, [T.E for T in self.tables])
        self.qorder_t = torch.tensor(qorder, device=dev)
        self.qstart_t = torch.tensor(qstart, device=dev)
        self.grid = N.FmGrid()
        self.grid.nk, self.grid.q, self.grid.jq, self.grid.fq = nk, self.q_t.data_ptr(), self.jq_t.data_ptr(), self.fq_t.data_ptr()
        self.grid.qorder, self.grid.qstart, self.grid.qstride = self.qorder_t.data_ptr(), self.qstart_t.data_ptr(), qstride
        self.S_t = None
        if convention == "larch":
            emax = max(T.E for T in self.tables)
            S = np.zeros((len(self.tables), nk, emax))
            for e, T in enumerate(self.tables):
                S[e, :, :T.E] = spline_matrix(T.kf, self.q)
            self.S_t = torch.tensor(S, device=dev)
            self.grid.S, self.grid.sstride = self.S_t.data_ptr(), emax
        # path arrays
        self.npath, self.nmax, self.nk = Pn, nmax, nk
        self.nleg_h, self.edge_h, self.segment_h, self.deg_h = nleg, edge, segment, deg
        absorber = np.asarray(absorber, np.int64)
        if absorber.shape != (Pn,):
            raise ValueError("absorber must be (P,)")
        used_atoms = atoms[np.arange(nmax - 1)[None, :] < (nleg[:, None] - 1)] if Pn else np.zeros(0, np.int64)
        self.max_atom = int(max(used_atoms.max() if used_atoms.size else -1, absorber.max() if Pn else -1))
        if Pn and min(int(absorber.min()), int(used_atoms.min()) if used_atoms.size else 0) < 0:
            raise ValueError("negative atom index")
        atoms = np.where(np.arange(nmax - 1)[None, :] < (nleg[:, None] - 1), atoms, 0)
        i32 = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev, dtype=torch.int32)
        self.atoms_t = i32(atoms if atoms.size else np.zeros((0, nmax - 1)))
        self.nleg_t = i32(nleg)
        self.absorber_t = i32(np.asarray(absorber, np.int64))
        self.ipots_t = i32(ipots)
        self.deg_t = torch.tensor(deg, device=dev, dtype=torch.float64)
        self.edge_t = i32(edge)
        if sigma2 is None:
            self.sigma2_t = None
        else:
            s2 = np.broadcast_to(np.asarray(sigma2, float), (Pn,))
            self.sigma2_t = torch.tensor(np.ascontiguousarray(s2), device=dev, dtype=torch.float64)
        self.cell_kind, self.cell, self.cellinv = _cell_struct(cell)
        self.buckets = bucket_paths(nleg, edge, [T.ilinit for T in self.tables])
        self.order_t = i32(self.buckets.order)
        seg_order = np.argsort(segment, kind="stable")
        seg_ptr = np.zeros(self.nseg + 1, np.int64)
        np.add.at(seg_ptr, segment + 1, 1)
        self.seg_ptr_t = i32(np.cumsum(seg_ptr))
        self.seg_paths_t = i32(seg_order)
        self.segment_t = i32(segment)
        # per-atom references (path * nmax + slot; slot 0 = the absorber) for the gradient gather
        slots = np.arange(nmax)[None, :]
        live = slots < nleg[:, None]
        who = np.concatenate([absorber[:, None], atoms], 1) if Pn else np.zeros((0, nmax), np.int64)
        ref_atom = who[live]
        ref_id = (np.arange(Pn)[:, None] * nmax + slots)[live]
        o = np.lexsort((ref_id, ref_atom))
        self.natom_ref = int(self.max_atom + 1)
        ptr = np.zeros(self.natom_ref + 1, np.int64)
        np.add.at(ptr, ref_atom + 1, 1)
        self.atom_ptr_t = i32(np.cumsum(ptr))
        self.atom_refs_t = i32(ref_id[o])
        self.gcoord = torch.empty(Pn, nmax, 3, dtype=torch.float64, device=dev)
        big = int(np.diff(self.buckets.offset).max()) if Pn else 0
        self.workspace = N.backward_workspace(self.precision, nmax, big, dev, max(T.E for T in self.tables)) if Pn else None
        self.chi_path = torch.empty(Pn, nk, dtype=self.rdtype, device=dev)
        self.reff = torch.empty(Pn, dtype=torch.float64, device=dev)
        self._h_off = (N._i32 * len(self.buckets.offset))(*self.buckets.offset.tolist())
        self._h_kind = (N._i32 * max(1, len(self.buckets.kind)))(*self.buckets.kind.tolist())
        self._h_edge = (N._i32 * max(1, len(self.buckets.edge)))(*self.buckets.edge.tolist())
        self._h_nlm = (N._i32 * max(1, len(self.buckets.nleg)))(*self.buckets.nleg.tolist())
        if self.keep_amplitudes and self.keep_bytes() > self.keep_limit_bytes:
            self.keep_amplitudes = False
        # per-bucket launch policies of the forward (0) and adjoint (1) calls: defaults until set_launch / tune;
        # every bucket and kernel has its own zeroed counter pair for persistent launches that take groups in turn
        nb = max(1, len(self.buckets.kind))
        self._launch = [(N.FmLaunch * nb)(), (N.FmLaunch * nb)()]
        self._tickets = torch.zeros(2, nb, 2, dtype=torch.int32, device=dev)
        self._launch_set = False
        # kept scattering matrices per bucket (allocated at the first forward call that keeps them); the generation
        # counts those calls, so that an adjoint can tell whether the matrices are still those of its forward
        self._keep = None
        self.kept_generation = 0

    def set_launch(self, kernel, bucket, gpaths=0, grid=0, queue=False):
        """The launch policy of one bucket for the forward (kernel 0) or adjoint (kernel 1) calls: gpaths paths per
        workgroup (0: by the LDS budget; else 1, 2, 4 or 8), grid (0: one workgroup per group of paths; n > 0: a
        persistent grid of at most n workgroups; -1: as many as the CUs hold at once), queue (persistent grids take
        the groups from an atomic counter, else in a fixed order)."""
        if kernel not in (0, 1) or not 0 <= bucket < len(self.buckets.kind):
            raise ValueError("kernel must be 0 or 1 and bucket one of the plan's")
        if gpaths not in (0, 1, 2, 4, 8) or grid < -1:
            raise ValueError("gpaths must be 0, 1, 2, 4 or 8 and grid -1, 0 or a number of workgroups")
        L = self._launch[kernel][bucket]
        L.gpaths, L.grid = int(gpaths), int(grid)
        L.tickets = self._tickets[kernel, bucket].data_ptr() if (queue and grid != 0) else None
        self._launch_set = True

    def launch_policy(self):
        """[{kind, edge, nleg, paths, forward: (gpaths, grid, queue), adjoint: (...)}] per bucket."""
        out = []
        for b in range(len(self.buckets.kind)):
            pol = [(int(L[b].gpaths), int(L[b].grid), bool(L[b].tickets)) for L in self._launch]
            out.append(dict(kind=int(self.buckets.kind[b]), edge=int(self.buckets.edge[b]),
                            nleg=int(self.buckets.nleg[b]),
                            paths=int(self.buckets.offset[b + 1] - self.buckets.offset[b]),
                            forward=pol[0], adjoint=pol[1]))
        return out

    def _keep_sizes(self):
        L = N.lib()
        out = []
        for b in range(len(self.buckets.kind)):
            cnt = int(self.buckets.offset[b + 1] - self.buckets.offset[b])
            nb = 0
            if int(self.buckets.kind[b]) != 2 and cnt > 0:
                nb = int(L.fm_keep_bytes(self.precision, self.edges[int(self.buckets.edge[b])].struct,
                                         int(self.buckets.nleg[b]), cnt))
            out.append(nb)
        return out

    def keep_bytes(self) -> int:
        """Device bytes the kept scattering matrices of this plan take."""
        return sum(self._keep_sizes())

    def _keep_buffers(self):
        """Per bucket: (tensor, bytes) of its kept matrices, or None for buckets that keep nothing."""
        if self._keep is None:
            self._keep = [(torch.empty(nb, dtype=torch.uint8, device=self.device), nb) if nb else None
                          for nb in self._keep_sizes()]
        return self._keep

    def _policy(self, kernel, b=None, keep=False):
        """The launch options of the call: every bucket's (or bucket b's alone), with the kept matrices attached when
        `keep` (the forward writes them, the adjoint reads them) and detached otherwise."""
        bufs = self._keep_buffers() if keep else None
        rng = range(len(self.buckets.kind)) if b is None else [b]
        for i in rng:
            L = self._launch[kernel][i]
            if bufs is not None and bufs[i] is not None:
                L.keep, L.keep_bytes = bufs[i][0].data_ptr(), bufs[i][1]
            else:
                L.keep, L.keep_bytes = None, 0
        if b is not None:
            return ctypes.pointer(self._launch[kernel][b])
        return self._launch[kernel]

    def _paths_struct(self, pos64):
        s = N.FmPaths()
        s.npath, s.nmax = self.npath, self.nmax
        s.pos = pos64.data_ptr()
        s.atoms, s.nleg, s.absorber, s.ipots = (t.data_ptr() for t in (self.atoms_t, self.nleg_t, self.absorber_t, self.ipots_t))
        s.deg, s.edge = self.deg_t.data_ptr(), self.edge_t.data_ptr()
        s.sigma2 = N.ptr(self.sigma2_t)
        s.cell_kind = self.cell_kind
        for i in range(9):
            s.cell[i] = float(self.cell[i])
            s.cellinv[i] = float(self.cellinv[i])
        return s

    def forward(self, pos, want_seg=True, keep=False):
        """pos (N, 3) A. Returns (chi_seg (nseg, NK) float64, chi_path (P, NK) in the plan's precision, reff (P,) A).
        chi_path and reff are the plan's own buffers, overwritten by the next call. keep: also keep the scattering
        matrices for an adjoint at the same positions (backward(..., kept=True)); advances kept_generation."""
        pos64 = torch.as_tensor(pos).detach().to(device=self.device, dtype=torch.float64).contiguous()
        if pos64.dim() != 2 or pos64.shape[1] != 3:
            raise ValueError("pos must be (N, 3)")
        if self.max_atom >= pos64.shape[0]:
            raise ValueError(f"the path list refers to atom {self.max_atom} but pos has {pos64.shape[0]} rows")
        chi_seg = torch.empty(self.nseg, self.nk, dtype=torch.float64, device=self.device) if want_seg else None
        if self.npath:
            ps = self._paths_struct(pos64)
            N.check(N.lib().fm_chi_paths(self.edge_array, len(self.edges), ps, self.grid, self.precision, self.iorder,
                                         self.order_t.data_ptr(), len(self.buckets.kind), self._h_off, self._h_kind,
                                         self._h_edge, self._h_nlm, self.nseg, self.seg_ptr_t.data_ptr(), self.seg_paths_t.data_ptr(),
                                         N.ptr(chi_seg), self.chi_path.data_ptr(), self.reff.data_ptr(),
                                         self._policy(0, keep=keep), N.stream_of(self.device)), "fm_chi_paths")
            if keep:
                self.kept_generation += 1
        elif want_seg:
            chi_seg.zero_()
        return chi_seg, self.chi_path, self.reff

    def backward(self, pos, gseg=None, gpath=None, kept=False):
        """dL/dpos (N, 3) float64 given gseg = dL/dchi_seg (nseg, NK) and/or gpath = dL/dchi_path (P, NK), at the
        positions pos. The amplitudes are recomputed inside the kernel, or (kept) read from the matrices the last
        forward(..., keep=True) kept, which the caller guarantees ran at these positions."""
        pos64 = torch.as_tensor(pos).detach().to(device=self.device, dtype=torch.float64).contiguous()
        if self.max_atom >= pos64.shape[0]:
            raise ValueError("pos has too few rows for the path list")
        Nat = pos64.shape[0]
        grad = torch.zeros(Nat, 3, dtype=torch.float64, device=self.device)
        if not self.npath or (gseg is None and gpath is None):
            return grad
        gs = None if gseg is None else torch.as_tensor(gseg).to(device=self.device, dtype=torch.float64).contiguous()
        if gs is not None and tuple(gs.shape) != (self.nseg, self.nk):
            raise ValueError("gseg must be (nseg, NK)")
        gp = None if gpath is None else torch.as_tensor(gpath).to(device=self.device, dtype=self.rdtype).contiguous()
        if gp is not None and tuple(gp.shape) != (self.npath, self.nk):
            raise ValueError("gpath must be (P, NK)")
        ps = self._paths_struct(pos64)
        N.check(N.lib().fm_chi_paths_backward(self.edge_array, len(self.edges), ps, self.grid, self.precision,
                                              self.iorder, self.order_t.data_ptr(), len(self.buckets.kind), self._h_off,
                                              self._h_kind, self._h_edge, self._h_nlm, self.segment_t.data_ptr(), N.ptr(gs),
                                              N.ptr(gp), self.natom_ref, self.atom_ptr_t.data_ptr(),
                                              self.atom_refs_t.data_ptr(), self.gcoord.data_ptr(), grad.data_ptr(),
                                              self.workspace.data_ptr(), self.workspace.numel(),
                                              self._policy(1, keep=kept and self._keep is not None),
                                              N.stream_of(self.device)), "fm_chi_paths_backward")
        return grad

    def __call__(self, pos):
        """Differentiable evaluation: (chi_seg (nseg, NK) float64, chi_path (P, NK)) with gradients to pos."""
        return _ChiPathsFn.apply(torch.as_tensor(pos), self)

    def forward_plain(self, pos, workspace_bytes: int = 16 << 30):
        """The same list through the plain FP64 pipeline (fm_chi_paths_plain: the reference's op sequence, one kernel
        per stage, every intermediate in HBM; chunks of as many paths as `workspace_bytes` holds). Returns (chi_seg,
        chi_path, reff), float64, the plan's own buffers for this entry."""
        pos64 = torch.as_tensor(pos).detach().to(device=self.device, dtype=torch.float64).contiguous()
        if self.max_atom >= pos64.shape[0]:
            raise ValueError("pos has too few rows for the path list")
        L = N.lib()
        if getattr(self, "_plain", None) is None:
            per = max((L.fm_plain_path_bytes(self.edges[int(e)].struct, int(n))
                       for e, n in zip(self.buckets.edge, self.buckets.nleg)), default=1)
            size = int(min(max(per, workspace_bytes), per * max(self.npath, 1)))
            self._plain = (torch.empty(size, dtype=torch.uint8, device=self.device),
                           torch.empty(self.npath, self.nk, dtype=torch.float64, device=self.device),
                           torch.empty(self.npath, dtype=torch.float64, device=self.device))
        ws, chi_path, reff = self._plain
        chi_seg = torch.empty(self.nseg, self.nk, dtype=torch.float64, device=self.device)
        if not self.npath:
            return chi_seg.zero_(), chi_path, reff
        ps = self._paths_struct(pos64)
        N.check(L.fm_chi_paths_plain(self.edge_array, len(self.edges), ps, self.grid, self.iorder, self.order_t.data_ptr(),
                                     len(self.buckets.kind), self._h_off, self._h_kind, self._h_edge, self._h_nlm,
                                     self.nseg, self.seg_ptr_t.data_ptr(), self.seg_paths_t.data_ptr(), chi_seg.data_ptr(),
                                     chi_path.data_ptr(), reff.data_ptr(), ws.data_ptr(), ws.numel(),
                                     N.stream_of(self.device)), "fm_chi_paths_plain")
        return chi_seg, chi_path, reff

    def forward_bucket(self, pos, b, keep=False):
        """Only bucket b of the plan (no segment sum): one kernel launch, for timing the kernel variants."""
        pos64 = torch.as_tensor(pos).detach().to(device=self.device, dtype=torch.float64).contiguous()
        if self.max_atom >= pos64.shape[0]:
            raise ValueError("pos has too few rows for the path list")
        off = (N._i32 * 2)(int(self.buckets.offset[b]), int(self.buckets.offset[b + 1]))
        kind = (N._i32 * 1)(int(self.buckets.kind[b]))
        edge = (N._i32 * 1)(int(self.buckets.edge[b]))
        nlm = (N._i32 * 1)(int(self.buckets.nleg[b]))
        ps = self._paths_struct(pos64)
        N.check(N.lib().fm_chi_paths(self.edge_array, len(self.edges), ps, self.grid, self.precision, self.iorder,
                                     self.order_t.data_ptr(), 1, off, kind, edge, nlm, 0, None, None, None,
                                     self.chi_path.data_ptr(), self.reff.data_ptr(), self._policy(0, b, keep),
                                     N.stream_of(self.device)), "fm_chi_paths")
        if keep:
            self.kept_generation += 1
        return self.chi_path

    def backward_bucket(self, pos, gseg, b, kept=False):
        """The adjoint of bucket b only (gcoord of its paths; no per-atom sum): one kernel launch, for timing; kept:
        from the matrices forward_bucket(..., keep=True) kept at these positions."""
        pos64 = torch.as_tensor(pos).detach().to(device=self.device, dtype=torch.float64).contiguous()
        gs = torch.as_tensor(gseg).to(device=self.device, dtype=torch.float64).contiguous()
        off = (N._i32 * 2)(int(self.buckets.offset[b]), int(self.buckets.offset[b + 1]))
        kind = (N._i32 * 1)(int(self.buckets.kind[b]))
        edge = (N._i32 * 1)(int(self.buckets.edge[b]))
        nlm = (N._i32 * 1)(int(self.buckets.nleg[b]))
        ps = self._paths_struct(pos64)
        N.check(N.lib().fm_chi_paths_backward(self.edge_array, len(self.edges), ps, self.grid, self.precision,
                                              self.iorder, self.order_t.data_ptr(), 1, off, kind, edge, nlm,
                                              self.segment_t.data_ptr(), gs.data_ptr(), None, 0, None, None,
                                              self.gcoord.data_ptr(), None, self.workspace.data_ptr(), self.workspace.numel(),
                                              self._policy(1, b, kept and self._keep is not None),
                                              N.stream_of(self.device)), "fm_chi_paths_backward")
        return self.gcoord


class _ChiPathsFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, plan):
        seg, path, _ = plan.forward(pos, keep=plan.keep_amplitudes)
        ctx.plan = plan
        ctx.generation = plan.kept_generation if plan.keep_amplitudes else None
        ctx.save_for_backward(pos)
        return seg, path.clone()

    @staticmethod
    def backward(ctx, gseg, gpath):
        (pos,) = ctx.saved_tensors
        # the kept matrices are this forward's unless another keeping forward ran since
        kept = ctx.generation is not None and ctx.generation == ctx.plan.kept_generation
        g = ctx.plan.backward(pos, gseg, gpath, kept=kept)
        return g.to(pos.dtype), None


def chi_paths(phases, pos, atoms, nleg, absorber, ipots, deg, q, segment=None, nseg=None, edge=None, sigma2=None,
              cell=None, precision=32, device="cuda"):
    """One-shot fused evaluation: returns (chi_seg (nseg, NK), chi_path (P, NK), reff (P,))."""
    plan = PathPlan(phases, q, atoms, nleg, absorber, ipots, deg, segment=segment, nseg=nseg, edge=edge,
                    sigma2=sigma2, cell=cell, precision=precision, device=device)
    seg, path, reff = plan.forward(pos)
    return seg, path.clone(), reff.clone()
