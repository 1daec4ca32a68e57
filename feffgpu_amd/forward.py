"""The path-list forward models of feffgpu (src/feffgpu/exafs/forward.py:33-175, 323-730) on the fused path stage.

``PathSpectra`` and ``MultipleScattering`` keep the reference's constructors, attributes and methods for what the
path stage serves: path lists from the GPU path finder (``refresh`` / ``enumerate``), the energy shift of the grid
q = sign(k^2 - e0 ETOK) sqrt|k^2 - e0 ETOK| (forward.py:79-82), chi of explicit paths (``signals``,
``path_signals``), the spectra of copies or absorbers with their exact position gradient (``spectra``,
``chi_absorbers``, ``full_pass``, ``full_gradient``) and the R-space misfit. Where the reference loops over chunks of
paths with checkpointed autograd (forward.py:148-163, 673-685), these evaluate a held path list in one call of the
fused kernels (``PathPlan``): the degeneracy-weighted sum into copies or absorbers happens in the kernel, and the
gradient is the fused adjoint, which recomputes the amplitudes as the reference's checkpoint does.

Not here: EvAX's rule for choosing paths on a lattice (``criteria_rule="evax"``, ``set_lattice``, ``type_signals``),
a path-selection heuristic of the MD/RMC applications outside the path stage; asking for it raises
NotImplementedError.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .pathfinder import GPUPathFinder
from .paths import PathPlan
from .stage import GPUPathStage
from .transform import ETOK, KSTEP, FitWindow, r_factor_torch


def _as_cell(box):
    """None, a number (cubic) or a 3 x 3 cell (a (3,) diagonal becomes its cell), as the reference's as_cell."""
    if box is None:
        return None
    if torch.is_tensor(box):
        box = box.detach().cpu().numpy()
    b = np.asarray(box, float)
    if b.ndim == 0:
        return float(b)
    return np.diag(b) if b.ndim == 1 else b


def _min_width(cell):
    if cell is None:
        return math.inf
    if isinstance(cell, float):
        return cell
    vol = abs(np.linalg.det(cell))
    return float(min(vol / np.linalg.norm(np.cross(cell[(i + 1) % 3], cell[(i + 2) % 3])) for i in range(3)))


def _wrap(rel, cell):
    """Minimum image of displacements rel (..., 3) (a torch tensor) in the cell (src/feffgpu/cell.py:26-35)."""
    if cell is None:
        return rel
    if isinstance(cell, float):
        return rel - cell * torch.round(rel / cell)
    c = torch.as_tensor(cell, device=rel.device, dtype=rel.dtype)
    f = rel @ torch.linalg.inv(c)
    return (f - torch.round(f)) @ c


class PathSpectra:
    """Spectra of chosen absorbers through the GPU path finder and the fused path stage, with an R-space misfit to
    a target (src/feffgpu/exafs/forward.py:33-175). Parameters as the reference's: phase, k (the uniform grid from
    zero), target, e0 (eV), window (kmin, kmax), r_range, box, rmax, nlegmax, critpw, merge_degenerate, chunk (kept
    for compatibility: a held list is evaluated in one call), device, dtype (complex64: FP32 kernels), kweight, dk."""

    def __init__(self, phase, k, target, e0, window=(3.0, 23.0), r_range=(1.0, 8.0), box=None, rmax=9.4, nlegmax=4,
                 critpw=0.5, merge_degenerate=False, chunk=2048, device="cuda", dtype=torch.complex64, kweight=2,
                 dk=1.0):
        self.device, self.chunk, self.box = torch.device(device), int(chunk), _as_cell(box)
        self.kweight, self.dk = int(kweight), float(dk)
        if self.box is not None and _min_width(self.box) < 2 * float(rmax):
            raise ValueError(f"the cell's smallest width ({_min_width(self.box):.2f} A) is less than twice the path "
                             f"cutoff ({rmax} A): minimum-image paths need a larger cell or a smaller cutoff")
        self.phase = phase
        self.finder = GPUPathFinder(phase, device=device, rmax=rmax, nlegmax=nlegmax, critpw=critpw,
                                    merge_degenerate=merge_degenerate)
        self.stage = GPUPathStage(phase, device=device, dtype=dtype)
        self.rdtype = self.stage.rdtype
        self.precision = self.stage.precision
        k = np.asarray(k, float)
        self.k = k
        en = k ** 2 - float(e0) * ETOK
        self.q_np = np.sign(en) * np.sqrt(np.abs(en))
        self.q = torch.tensor(self.q_np, device=self.device, dtype=self.rdtype)
        self.S = self.stage.linear_matrix(self.q_np)
        self.kt = torch.tensor(k, device=self.device, dtype=torch.float32)
        self.window = (float(window[0]), float(window[1]))
        self._target = np.asarray(target, float)
        self.set_r_range(*r_range)
        self.lists, self.C, self.n, self.A = None, 0, 0, 0
        self._plans = {}

    # -- the misfit -------------------------------------------------------------------------------------------
    def set_r_range(self, r1, r2):
        """The R window of the misfit (A)."""
        self.r_range = (float(r1), float(r2))
        self.fit_window = FitWindow(self.window[0], self.window[1], self.dk, self.kweight, self.r_range[0],
                                    self.r_range[1])
        self.set_target(self._target)

    def set_target(self, target):
        """The target spectrum (NK,) on the k grid."""
        t = target.detach().cpu().numpy() if torch.is_tensor(target) else target
        self._target = np.asarray(t, float)
        self.target = torch.as_tensor(self._target, device=self.device, dtype=torch.float64)

    def misfit(self, chi_bar):
        """R-space misfit of one spectrum (NK,) to the target: sum |chi_R - target_R|^2 / sum |target_R|^2 over the
        R window (the R-factor kernel; src/feffgpu/exafs/forward.py:166-169)."""
        if abs(float(self.k[1] - self.k[0]) - KSTEP) > 1e-9 or self.k[0] != 0.0:
            raise ValueError(f"the misfit needs the uniform k grid from zero with step {KSTEP}")
        return r_factor_torch(chi_bar[None].to(torch.float64), self.target, self.k, self.fit_window)[0]

    # -- path lists -------------------------------------------------------------------------------------------
    @staticmethod
    def sample_absorbers(C, n, A, rng):
        """(C, A) random distinct absorber indices per copy."""
        return np.stack([rng.choice(n, size=min(A, n), replace=False) for _ in range(C)])

    def refresh(self, positions, absorbers, ipot=None):
        """Find the paths of the chosen absorbers (C, A) of the copies at positions (C, n, 3) and hold them: per
        number of legs, global atom indices (copy c, atom i -> c n + i), absorber, potentials, degeneracy, copy.
        ipot (n,) is every atom's scatterer potential (default 1). Returns the number of paths per number of legs."""
        pos = torch.as_tensor(positions, device=self.device, dtype=torch.float32)
        C, n, _ = pos.shape
        absorbers = np.asarray(absorbers).reshape(C, -1)
        self.C, self.n, self.A = C, n, absorbers.shape[1]
        ip = torch.ones(n, dtype=torch.long, device=self.device) if ipot is None else \
            torch.as_tensor(ipot, device=self.device)
        # every copy's absorbers in one pass of the finder, numbered over the copies (c n + i)
        ab = (torch.as_tensor(absorbers, device=self.device, dtype=torch.long)
              + n * torch.arange(C, device=self.device)[:, None]).reshape(-1)
        kept, _ = self.finder.enumerate_many(pos, self.box, ab, ip)
        self.lists = {}
        for nleg, v in sorted(kept.items()):
            if v["deg"].numel() == 0:
                continue
            a = ab[v["slot"].long()]
            self.lists[nleg] = dict(atoms=v["atoms"].long(), absorber=a, ipots=v["ipots"], deg=v["deg"],
                                    copy=a // n)
        self.counts = {nleg: int(v["atoms"].shape[0]) for nleg, v in self.lists.items()}
        self._plans = {}
        return self.counts

    def _plan(self, lists, segment_key, nseg, key):
        """A PathPlan of the given lists (one per number of legs), segments from segment_key(L) (cached by key)."""
        plan = self._plans.get(key)
        if plan is not None:
            return plan
        rows = [L for L in lists.values() if L["atoms"].shape[0]]
        if not rows:
            return None
        nmax = max(L["atoms"].shape[1] + 1 for L in rows)
        pad = lambda a, w: torch.nn.functional.pad(a, (0, w - a.shape[1]))
        atoms = torch.cat([pad(L["atoms"], nmax - 1) for L in rows]).cpu().numpy()
        ipots = torch.cat([pad(L["ipots"], nmax) for L in rows]).cpu().numpy()
        nleg = np.concatenate([np.full(L["atoms"].shape[0], L["atoms"].shape[1] + 1) for L in rows])
        absorber = torch.cat([L["absorber"] for L in rows]).cpu().numpy()
        deg = torch.cat([L["deg"] for L in rows]).double().cpu().numpy()
        seg = torch.cat([segment_key(L) for L in rows]).cpu().numpy()
        plan = PathPlan(self.phase if isinstance(self.phase, str) else self.stage.P, self.q_np, atoms, nleg, absorber,
                        ipots, deg, segment=seg, nseg=nseg, cell=self.box, precision=self.precision,
                        device=self.device)
        self._plans[key] = plan
        return plan

    def spectra(self, pos_flat, use_checkpoint=True, nlegs=None):
        """chi(k) per copy, (C, NK), averaged over that copy's absorbers; pos_flat (C n, 3) may require a gradient.
        nlegs restricts the sum to these numbers of legs (default all). use_checkpoint is accepted for
        compatibility: the adjoint recomputes the amplitudes in any case."""
        if self.lists is None:
            raise RuntimeError("no path lists: call refresh(positions, absorbers) first")
        sel = {k: v for k, v in self.lists.items() if nlegs is None or k in nlegs}
        plan = self._plan(sel, lambda L: L["copy"], self.C, ("spectra", None if nlegs is None else tuple(sorted(nlegs))))
        if plan is None:
            return torch.zeros(self.C, len(self.k), device=self.device, dtype=self.rdtype)
        seg, _ = plan(pos_flat)
        return (seg / self.A).to(self.rdtype)

    @torch.no_grad()
    def spectrum_only(self, positions):
        """The mean spectrum (NK,) of the held lists at positions (C, n, 3), its misfit, and the spectra per copy."""
        pos = torch.as_tensor(positions, device=self.device, dtype=self.rdtype).reshape(-1, 3)
        chi = self.spectra(pos, use_checkpoint=False)
        return chi.mean(0), self.misfit(chi.mean(0)), chi


class MultipleScattering:
    """The multiple-scattering paths (three legs up to nlegmax) of every absorber of every copy, from the GPU path
    finder, evaluated by the fused path stage: ``rows`` holds one spectrum per absorber (``full_pass``);
    ``chi_absorbers`` evaluates a subset with a gradient; ``full_gradient`` gives the exact gradient of a loss of the
    mean spectrum (src/feffgpu/exafs/forward.py:323-730). Only criteria_rule "feff" (FEFF's importance filter on the
    structure as it is)."""

    def __init__(self, phase, k, e0, box, rmax=7.0, critpw=2.0, chunk=1024, device="cuda", nlegmax=4,
                 criteria_rule="feff"):
        if criteria_rule not in ("feff", "evax"):
            raise ValueError(f"criteria_rule must be 'feff' or 'evax' (got {criteria_rule!r})")
        if criteria_rule == "evax":
            raise NotImplementedError("criteria_rule 'evax' (paths chosen on a lattice by EvAX's rule) is not part of "
                                      "this path stage")
        self.full = PathSpectra(phase, k, np.zeros(len(k)), e0, box=box, rmax=rmax, nlegmax=nlegmax, critpw=critpw,
                                chunk=chunk, device=device)
        self.device, self.chunk, self.box = torch.device(device), int(chunk), _as_cell(box)
        self.NK = len(k)
        self.k = np.asarray(k, float)
        self.critpw, self.criteria_rule = float(critpw), criteria_rule
        self.lists, self.rows, self.n_abs = None, None, 0
        self._plans = {}

    def enumerate(self, pos, ipot=None, absorbers=None):
        """Find the paths of three and more legs of every absorber of every copy of pos (C, n, 3). Atom indices are
        global (copy c, atom i -> c n + i); row c A + j is absorber j of copy c (``row_of``: global atom -> row, -1
        for a non-absorber; ``abs_atoms``: row -> global atom). Returns the number of paths per number of legs."""
        pos = torch.as_tensor(pos, device=self.device, dtype=torch.float32)
        C, n, _ = pos.shape
        ipot = torch.ones(n, dtype=torch.long, device=self.device) if ipot is None else \
            torch.as_tensor(ipot, device=self.device, dtype=torch.long)
        absorbers = torch.arange(n, device=self.device) if absorbers is None else \
            torch.as_tensor(absorbers, device=self.device, dtype=torch.long)
        A = int(absorbers.shape[0])
        # every copy's absorbers in one pass of the finder, numbered over the copies (c n + i)
        ab = (absorbers[None, :] + n * torch.arange(C, device=self.device)[:, None]).reshape(-1)
        kept, _ = self.full.finder.enumerate_many(pos, self.box, ab, ipot)
        self.lists = {nleg: dict(atoms=v["atoms"].long(), absorber=ab[v["slot"].long()], ipots=v["ipots"],
                                 deg=v["deg"]) for nleg, v in sorted(kept.items()) if nleg >= 3 and v["deg"].numel()}
        self.C, self.n, self.A, self.n_abs = C, n, A, C * A
        self.abs_atoms = (absorbers[None, :] + n * torch.arange(C, device=self.device)[:, None]).reshape(-1)
        self.row_of = torch.full((C * n,), -1, device=self.device, dtype=torch.long)
        self.row_of[self.abs_atoms] = torch.arange(C * A, device=self.device)
        self.counts = {nleg: int(v["atoms"].shape[0]) for nleg, v in self.lists.items()}
        self._plans = {}
        return self.counts

    def set_cell(self, box):
        """Use another periodic cell from now on (a cubic length, a 3 x 3 cell, or None); it must be at least twice as
        wide as the path cutoff."""
        cell = _as_cell(box)
        rmax = float(self.full.finder.rmax)
        if cell is not None and _min_width(cell) < 2 * rmax - 1e-6:
            raise ValueError(f"the cell's smallest width ({_min_width(cell):.2f} A) is less than twice the path "
                             f"cutoff ({rmax} A): minimum-image paths need a larger cell or a smaller cutoff")
        self.box = cell
        self.full.box = cell
        self._plans = {}
        self.full._plans = {}

    def signals(self, coords, ipots):
        """The signals of explicit paths: coords (B, nleg, 3) A with the absorber first at the origin, ipots (B, nleg)
        with 0 for the absorber. Returns chi (B, NK) on this object's (energy-shifted) grid without the degeneracy,
        and the effective half path lengths (B,) A."""
        st = self.full.stage
        cchi, reff = st.cchi(coords, ipots)
        return st.chi_fine(cchi, reff, self.full.q_np, self.full.S, "feff"), reff

    def path_signals(self, pos_flat, L, start=0, stop=None):
        """The signals of the paths start:stop of one list L (an entry of ``self.lists``) at the positions pos_flat
        (C n, 3): chi (B, NK) without the degeneracy, and reff (B,)."""
        atoms, absr = L["atoms"][start:stop], L["absorber"][start:stop]
        rel = _wrap(pos_flat[atoms] - pos_flat[absr][:, None, :], self.box)
        coords = torch.cat([torch.zeros(rel.shape[0], 1, 3, device=rel.device, dtype=rel.dtype), rel], 1)
        return self.signals(coords, L["ipots"][start:stop])

    def _rows_plan(self, absorbers=None):
        """A plan of the paths of the given absorbers (global indices; all absorbers: None), one segment per
        absorber (its position in the given list, or its row)."""
        if absorbers is None:
            local, nseg, key = self.row_of, self.n_abs, ("rows",)
            lists = self.lists
        else:
            absorbers = torch.as_tensor(absorbers, device=self.device, dtype=torch.long)
            local = torch.full((self.C * self.n,), -1, device=self.device, dtype=torch.long)
            local[absorbers] = torch.arange(absorbers.shape[0], device=self.device)
            nseg, key = int(absorbers.shape[0]), ("subset", tuple(absorbers.tolist()))
            lists = {}
            for nleg, L in self.lists.items():
                sel = local[L["absorber"]] >= 0
                lists[nleg] = {k: v[sel] for k, v in L.items()}
        plan = self._plans.get(key)
        if plan is None:
            plan = self.full._plan(lists, lambda L: local[L["absorber"]], nseg, key)
            self._plans[key] = plan
        return plan, nseg

    def chi_absorbers(self, pos_flat, absorbers):
        """(A, NK): the multiple-scattering chi of the given absorbers (global atom indices), differentiable in
        pos_flat (C n, 3)."""
        plan, nseg = self._rows_plan(absorbers)
        if plan is None:
            return torch.zeros(nseg, self.NK, device=self.device, dtype=self.full.rdtype)
        seg, _ = plan(pos_flat)
        return seg.to(self.full.rdtype)

    @torch.no_grad()
    def full_pass(self, pos_flat):
        """Every absorber's multiple-scattering chi at pos_flat (C n, 3) -> ``self.rows`` (n_abs, NK)."""
        plan, nseg = self._rows_plan()
        if plan is None:
            self.rows = torch.zeros(nseg, self.NK, device=self.device, dtype=self.full.rdtype)
        else:
            self.rows = plan.forward(pos_flat)[0].to(self.full.rdtype)
        return self.rows

    def full_gradient(self, pos_flat, loss_fn, absorbers_per_pass=512):
        """The exact gradient with respect to pos_flat of loss_fn(mean spectrum), the mean over every absorber of
        ``self.rows`` (filled by full_pass at these positions). One adjoint launch for all absorbers:
        absorbers_per_pass is accepted for compatibility."""
        rows = self.rows.detach().double().requires_grad_(True)
        with torch.enable_grad():
            (grows,) = torch.autograd.grad(loss_fn(rows.mean(0).to(self.full.rdtype)), rows)
        plan, _ = self._rows_plan()
        if plan is None:
            return torch.zeros_like(pos_flat)
        return plan.backward(pos_flat, gseg=grows).to(pos_flat.dtype)

    @property
    def mean(self):
        """The mean of the held rows over every absorber, (NK,)."""
        return self.rows.mean(0)
