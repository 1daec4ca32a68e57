"""Drop-in ``GPUPathFinder``: FEFF's path finder (src/feffgpu/pathfinder.py:29-279) backed by HIP kernels.

    finder = GPUPathFinder(phase, device="cuda", rmax=9.4, nlegmax=4, critpw=0.1, merge_degenerate=True)
    kept, ref = finder.enumerate(pos, box, absorber, ipot)
    # kept: {nleg: (coords (B, nleg, 3), ipots (B, nleg), deg (B,), xport (B,), atoms (B, nleg - 1))}

Same constructor, methods and results as the reference class, in float32 as the reference computes them. Every path
of two to nlegmax legs whose total length is at most 2 rmax (the first and last atoms not the absorber, no atom
following itself, the absorber allowed in between; a path and its reverse one path of degeneracy 2, 1 if it is its
own reverse), its plane-wave importance, and the filter: a path is kept when its degeneracy-weighted importance is at
least critpw percent of the shortest single-scattering path's. With merge_degenerate the degeneracy that the filter
weighs is the total over symmetry-equivalent multiple-scattering paths (every member stays in the list), and
single-scattering paths at coincident distances become one path (FEFF's accidental-degeneracy merge, which the
reference also performs on the host).

``enumerate_many`` finds the paths of many absorbers of one configuration in one pass of kernels (what the forward
models' ``refresh`` / ``enumerate`` need); ``enumerate`` is its one-absorber form. The kernels are in
csrc/fm_finder.hip: candidates, two-pass enumeration (counts, exclusive scan, write), the partial-path extension of
five and more legs, importance, degeneracy grouping and the filter.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N
from .pad import BOHR, load_phase, read_phase_pad

NBETA = 40
NLEG_MAX = 9
# absorbers enumerated together: the four-leg counts take nc (nc + 1) int32 per absorber (nc candidates within rmax)
ABSORBERS_PER_PASS = 4096
PARTIAL_LIMIT = 30_000_000     # partial paths held per absorber (on average over a call) while 5+-leg paths grow
_SEQ = N.FM_MAX_LEGS - 1


def _legendre(lmax: int, x: np.ndarray) -> np.ndarray:
    """P_l(x) for l = 0..lmax by the three-term recurrence (FP64), (lmax + 1, len(x))."""
    P = np.zeros((lmax + 1, len(x)))
    P[0] = 1.0
    if lmax >= 1:
        P[1] = x
    for l in range(1, lmax):
        P[l + 1] = ((2 * l + 1) * x * P[l] - l * P[l - 1]) / (l + 1)
    return P


def _cell32(box):
    """(kind, cell[9], cellinv[9]) in float32 from None, a cubic length, 3 lengths or a 3 x 3 cell (rows the lattice
    vectors), as the reference's as_cell makes it (src/feffgpu/cell.py:11-23)."""
    if box is None:
        return 0, np.zeros(9, np.float32), np.zeros(9, np.float32)
    b = box.detach().cpu().numpy() if torch.is_tensor(box) else np.asarray(box, float)
    if np.ndim(b) == 0:
        c = np.zeros(9, np.float32)
        c[0] = np.float32(float(b))
        return 1, c, np.zeros(9, np.float32)
    c = np.diag(b) if np.ndim(b) == 1 else np.asarray(b)
    if c.shape != (3, 3):
        raise ValueError("box must be None, a number, 3 lengths or a 3 x 3 cell")
    c32 = c.astype(np.float32)
    return 2, c32.reshape(-1), np.linalg.inv(c32.astype(np.float64)).astype(np.float32).reshape(-1)


class GPUPathFinder:
    def __init__(self, phase, device="cuda", rmax=9.4, nlegmax=4, critpw=0.1, merge_degenerate=True):
        P = read_phase_pad(phase) if isinstance(phase, str) else load_phase(phase)
        self.P, self.device = P, torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("GPUPathFinder runs on a ROCm GPU only (device must be 'cuda')")
        self.rmax, self.nlegmax, self.critpw = float(rmax), int(nlegmax), float(critpw)
        if not 2 <= self.nlegmax <= NLEG_MAX:
            raise ValueError(f"nlegmax must be 2..{NLEG_MAX}, the most FEFF's path finder allows (got {nlegmax})")
        self.merge_degenerate = bool(merge_degenerate)
        E = P.ne1
        lmax = P.lmax[:E]
        cosb = 0.025 * np.arange(-NBETA, NBETA + 1)
        cosb[0], cosb[-1] = -1.0, 1.0
        Pl = _legendre(int(lmax.max()), cosb)
        fb = np.zeros((E, P.nph + 1, 2 * NBETA + 1))
        for iph in range(P.nph + 1):
            for ie in range(E):
                L = int(lmax[ie, iph]) + 1
                tl = np.array([(np.exp(2j * P.phase(ie, -l, iph)) - 1) / 2j for l in range(L)])
                fb[ie, iph] = np.abs(((tl * (2 * np.arange(L) + 1))[:, None] * Pl[:L]).sum(0))
        ck = P.ck[:E]
        dev = self.device
        self.fbeta = torch.tensor(fb, device=dev, dtype=torch.float32).contiguous()
        self.cksp = torch.tensor(ck.real / BOHR, device=dev, dtype=torch.float32)
        im = np.where(np.abs(ck.imag) > 1e-16, ck.imag, 1.0)
        self.xlam = torch.tensor(np.where(np.abs(ck.imag) > 1e-16, 1 / im, 1e10) * BOHR, device=dev,
                                 dtype=torch.float32)
        self.ik0 = P.ik0 - 1
        x = ck.real[self.ik0:] / BOHR
        w = np.zeros(len(x))
        w[0] = x[1] - x[0]
        w[1:-1] = x[2:] - x[:-2]
        w[-1] = x[-1] - x[-2]
        self.trap = torch.tensor(w / 2, device=dev, dtype=torch.float32)
        t = N.FmFindTables()
        t.E, t.nph1, t.ik0 = E, P.nph + 1, self.ik0
        t.fbeta, t.cksp, t.xlam, t.trap = (v.data_ptr() for v in (self.fbeta, self.cksp, self.xlam, self.trap))
        self._tables = t
        N.lib()

    # -- importance and degeneracy ----------------------------------------------------------------------------
    def importance(self, pts, ipots):
        """pts (B, m+1, 3): the path's atoms with the absorber at index 0; ipots (B, m+1). Returns xport (B,)."""
        return self._importance(pts, ipots)[0]

    def _importance(self, pts, ipots):
        c = torch.as_tensor(pts, device=self.device).to(torch.float32).contiguous()
        ip = torch.as_tensor(ipots, device=self.device).to(torch.int64).contiguous()
        B, n = c.shape[0], c.shape[1]
        xp = torch.empty(B, dtype=torch.float32, device=self.device)
        rt = torch.empty(B, dtype=torch.float32, device=self.device)
        if B:
            N.check(N.lib().fm_find_importance(ctypes.byref(self._tables), B, n, c.data_ptr(), ip.data_ptr(),
                                               xp.data_ptr(), rt.data_ptr(), N.stream_of(self.device)),
                    "fm_find_importance")
        return xp, rt

    @staticmethod
    def degenerate_total(coords, deg, slot=None):
        """The total degeneracy of the group of every path (paths of equal leg lengths and angles to 1e-3, in either
        direction; src/feffgpu/pathfinder.py:135-158), (B,); slot (B,) keeps the groups of different absorbers
        apart."""
        c = torch.as_tensor(coords).to(torch.float32).contiguous()
        d = torch.as_tensor(deg, device=c.device).to(torch.float32).contiguous()
        B, n = c.shape[0], c.shape[1]
        s = (torch.zeros(B, dtype=torch.int32, device=c.device) if slot is None
             else torch.as_tensor(slot, device=c.device).to(torch.int32).contiguous())
        out = torch.empty(B, dtype=torch.float32, device=c.device)
        if B:
            L = N.lib()
            ws = torch.empty(int(L.fm_find_group_workspace_bytes(B)), dtype=torch.uint8, device=c.device)
            N.check(L.fm_find_degenerate_total(B, n, c.data_ptr(), d.data_ptr(), s.data_ptr(), out.data_ptr(),
                                               ws.data_ptr(), ws.numel(), N.stream_of(c.device)),
                    "fm_find_degenerate_total")
        return out

    # -- enumeration ------------------------------------------------------------------------------------------
    def _scan(self, cnt):
        L = N.lib()
        n = cnt.numel()
        pos = torch.empty(n, dtype=torch.int64, device=self.device)
        total = torch.zeros(1, dtype=torch.int64, device=self.device)
        ws = torch.empty(int(L.fm_scan_workspace_bytes(n)), dtype=torch.uint8, device=self.device)
        N.check(L.fm_exclusive_scan(cnt.data_ptr(), n, pos.data_ptr(), total.data_ptr(), ws.data_ptr(), ws.numel(),
                                    N.stream_of(self.device)), "fm_exclusive_scan")
        return pos, int(total.item())

    def _alloc(self, P, nleg):
        dev = self.device
        t = dict(atoms=torch.empty(P, nleg - 1, dtype=torch.int64, device=dev),
                 coords=torch.empty(P, nleg, 3, dtype=torch.float32, device=dev),
                 ipots=torch.empty(P, nleg, dtype=torch.int64, device=dev),
                 deg=torch.empty(P, dtype=torch.float32, device=dev),
                 slot=torch.empty(P, dtype=torch.int32, device=dev))
        s = N.FmFindOut()
        s.atoms, s.coords, s.ipots, s.deg, s.slot = (t[k].data_ptr() for k in ("atoms", "coords", "ipots", "deg", "slot"))
        return t, s

    def _partials(self, P):
        dev = self.device
        t = dict(seq=torch.empty(P, _SEQ, dtype=torch.int16, device=dev),
                 len=torch.empty(P, dtype=torch.float32, device=dev),
                 slot=torch.empty(P, dtype=torch.int32, device=dev))
        s = N.FmFindPartial()
        s.seq, s.len, s.slot = t["seq"].data_ptr(), t["len"].data_ptr(), t["slot"].data_ptr()
        return t, s

    def enumerate_many(self, pos, box, absorbers, ipot):
        """The paths of every absorber in ``absorbers`` (A,) in one pass: pos (n, 3) one configuration, or (C, n, 3)
        configurations of the same atoms (frames of a trajectory), whose atoms count c n + i; every absorber's paths
        run over the atoms of its own configuration. ipot (n,) is the potential of every atom of a configuration.

        Returns ({nleg: dict(coords, ipots, deg, xport, atoms, slot)}, ref (A,) float32): slot is the position of a
        path's absorber in ``absorbers``; the order is by absorber, then as ``enumerate`` orders the paths of one.
        Absorbers are taken ABSORBERS_PER_PASS at a time (bounding the work arrays), the results joined."""
        ab = torch.as_tensor(absorbers).reshape(-1)
        if ab.numel() <= ABSORBERS_PER_PASS:
            return self._enumerate_pass(pos, box, ab, ipot)
        parts, refs = [], []
        for s0 in range(0, ab.numel(), ABSORBERS_PER_PASS):
            kept, ref = self._enumerate_pass(pos, box, ab[s0:s0 + ABSORBERS_PER_PASS], ipot)
            for v in kept.values():
                v["slot"] = v["slot"] + s0
            parts.append(kept)
            refs.append(ref)
        out = {}
        for nleg in sorted({k for kept in parts for k in kept}):
            rows = [kept[nleg] for kept in parts if nleg in kept]
            out[nleg] = {k: torch.cat([r[k] for r in rows]) for k in rows[0]}
        return out, np.concatenate(refs)

    def _enumerate_pass(self, pos, box, absorbers, ipot):
        L = N.lib()
        dev = self.device
        st = N.stream_of(dev)
        if np.size(absorbers) == 0:
            return {}, np.zeros(0, np.float32)
        p32 = torch.as_tensor(pos, device=dev).to(torch.float32)
        n = p32.shape[-2] if p32.dim() == 3 else p32.reshape(-1, 3).shape[0]
        p32 = p32.reshape(-1, 3).contiguous()
        nconf = p32.shape[0] // max(n, 1)
        ab = torch.as_tensor(absorbers, device=dev).to(torch.int32).reshape(-1).contiguous()
        A = ab.numel()
        ip = torch.as_tensor(ipot, device=dev).to(torch.int32).reshape(-1).contiguous()
        if ip.numel() != n:
            raise ValueError("ipot needs one potential per atom")
        if A and (int(ab.min()) < 0 or int(ab.max()) >= n * nconf):
            raise ValueError("absorber index outside the configurations")
        if nconf > 1:
            ip = ip.repeat(nconf)
        sysd = N.FmFindSys()
        sysd.natom, sysd.pos = n, p32.data_ptr()
        kind, cell, cinv = _cell32(box)
        sysd.cell_kind = kind
        for i in range(9):
            sysd.cell[i], sysd.cellinv[i] = float(cell[i]), float(cinv[i])
        sysd.rmax = float(np.float32(self.rmax))
        maxc = max(1, min(n, 512))
        while True:
            cidx = torch.empty(A, maxc, dtype=torch.int32, device=dev)
            crel = torch.empty(A, maxc, 3, dtype=torch.float32, device=dev)
            cd0 = torch.empty(A, maxc, dtype=torch.float32, device=dev)
            ncand = torch.zeros(max(A, 1), dtype=torch.int32, device=dev)
            N.check(L.fm_find_candidates(ctypes.byref(sysd), A, ab.data_ptr(), maxc, cidx.data_ptr(), crel.data_ptr(),
                                         cd0.data_ptr(), ncand.data_ptr(), st), "fm_find_candidates")
            nc = ncand[:A].cpu().numpy().astype(np.int64)
            if A == 0 or nc.max() <= maxc:
                break
            maxc = int(nc.max())
        rmax32 = float(np.float32(self.rmax))
        # work items: two-leg paths and seeds (nc per absorber), three legs (nc), four legs (nc (nc + 1)); absorbers
        # with a single candidate have single scattering only, as in the reference (the kernels skip their 3+-leg
        # items, and they have none here)
        multi = nc > 1
        off2 = np.concatenate([[0], np.cumsum(nc)])
        n3 = np.where(multi, nc, 0)
        n4 = np.where(multi, nc * (nc + 1), 0)
        off3 = np.concatenate([[0], np.cumsum(n3)])
        off4 = np.concatenate([[0], np.cumsum(n4)])
        t64 = lambda a: torch.tensor(np.ascontiguousarray(a, np.int64), device=dev)
        off2_t, off3_t, off4_t = t64(off2), t64(off3), t64(off4)
        ncmax = int(nc.max()) if A else 0
        nlm = self.nlegmax
        cnt3 = torch.zeros(max(int(off3[-1]), 1), dtype=torch.int32, device=dev)
        cnt4 = torch.zeros(max(int(off4[-1]), 1), dtype=torch.int32, device=dev)
        if nlm >= 3 and ncmax > 1:
            N.check(L.fm_find_count(A, off3_t.data_ptr(), off4_t.data_ptr(), maxc, crel.data_ptr(), cd0.data_ptr(),
                                    ncand.data_ptr(), ncmax, rmax32, nlm, cnt3.data_ptr(), cnt4.data_ptr(), st),
                    "fm_find_count")
        res = {}
        o2, s2 = self._alloc(int(off2[-1]), 2)
        if nlm >= 3:
            pos3, tot3 = self._scan(cnt3[:max(int(off3[-1]), 0)].contiguous()) if off3[-1] else (cnt3.long(), 0)
            o3, s3 = self._alloc(tot3, 3)
        if nlm >= 4:
            pos4, tot4 = self._scan(cnt4[:max(int(off4[-1]), 0)].contiguous()) if off4[-1] else (cnt4.long(), 0)
            o4, s4 = self._alloc(tot4, 4)
        if ncmax > 0:
            N.check(L.fm_find_write(A, ab.data_ptr(), off3_t.data_ptr(), off4_t.data_ptr(), maxc, cidx.data_ptr(),
                                    crel.data_ptr(), cd0.data_ptr(), ncand.data_ptr(), ncmax, ip.data_ptr(), rmax32,
                                    min(nlm, 4) if multi.any() else 2, off2_t.data_ptr(), ctypes.byref(s2),
                                    pos3.data_ptr() if nlm >= 3 else None, ctypes.byref(s3) if nlm >= 3 else None,
                                    pos4.data_ptr() if nlm >= 4 else None, ctypes.byref(s4) if nlm >= 4 else None,
                                    st), "fm_find_write")
        res[2] = o2
        if nlm >= 3:
            res[3] = o3
        if nlm >= 4:
            res[4] = o4
        if nlm >= 5 and multi.any():
            # seeds: the candidates of absorbers with more than one (the first atom of a path is a scatterer)
            seeds, sd = self._partials(int(off2[-1]))
            N.check(L.fm_find_seed(A, off2_t.data_ptr(), maxc, cd0.data_ptr(), ncand.data_ptr(), ncmax,
                                   ctypes.byref(sd), st), "fm_find_seed")
            keep_seed = torch.tensor(np.repeat(multi, nc), device=dev)
            cur = {k: v[keep_seed] for k, v in seeds.items()}
            for m in range(2, nlm):
                npart = cur["len"].numel()
                if npart == 0:
                    break
                ci = N.FmFindPartial()
                ci.seq, ci.len, ci.slot = cur["seq"].data_ptr(), cur["len"].data_ptr(), cur["slot"].data_ptr()
                cnt = torch.empty(npart, dtype=torch.int32, device=dev)
                N.check(L.fm_find_extend(npart, m - 1, ctypes.byref(ci), maxc, crel.data_ptr(), cd0.data_ptr(),
                                         ncand.data_ptr(), rmax32, cnt.data_ptr(), None, None, st), "fm_find_extend")
                posx, tot = self._scan(cnt)
                if tot > PARTIAL_LIMIT * max(int(multi.sum()), 1) or tot * 24 > (8 << 30):
                    raise RuntimeError(f"more than {PARTIAL_LIMIT} partial paths of {m + 1} legs per absorber within "
                                       f"{self.rmax} A: lower the path cutoff or the largest number of legs")
                nxt, so = self._partials(tot)
                N.check(L.fm_find_extend(npart, m - 1, ctypes.byref(ci), maxc, crel.data_ptr(), cd0.data_ptr(),
                                         ncand.data_ptr(), rmax32, None, posx.data_ptr(), ctypes.byref(so), st),
                        "fm_find_extend")
                cur = nxt
                if m + 1 < 5 or tot == 0:
                    continue
                cn = N.FmFindPartial()
                cn.seq, cn.len, cn.slot = cur["seq"].data_ptr(), cur["len"].data_ptr(), cur["slot"].data_ptr()
                cnt = torch.empty(tot, dtype=torch.int32, device=dev)
                N.check(L.fm_find_emit(tot, m, ctypes.byref(cn), ab.data_ptr(), maxc, cidx.data_ptr(), crel.data_ptr(),
                                       cd0.data_ptr(), ncand.data_ptr(), ip.data_ptr(), cnt.data_ptr(), None, None, st),
                        "fm_find_emit")
                pose, tote = self._scan(cnt)
                if tote == 0:
                    continue
                oe, se = self._alloc(tote, m + 1)
                N.check(L.fm_find_emit(tot, m, ctypes.byref(cn), ab.data_ptr(), maxc, cidx.data_ptr(), crel.data_ptr(),
                                       cd0.data_ptr(), ncand.data_ptr(), ip.data_ptr(), None, pose.data_ptr(),
                                       ctypes.byref(se), st), "fm_find_emit")
                res[m + 1] = oe
        for nleg, o in res.items():
            o["xport"], o["rtot"] = self._importance(o["coords"], o["ipots"])
        # the reference importance per absorber: the shortest single-scattering path (after the merge, times its
        # merged count)
        ref = np.ones(A, np.float32)
        if self.merge_degenerate and multi.any():
            res[2] = self._merge_single_scattering(res, A, multi)
        xp2 = res[2]["xport"].cpu().numpy()
        slot2 = res[2]["slot"].cpu().numpy()
        d2 = torch.linalg.vector_norm(res[2]["coords"][:, 1], dim=-1).cpu().numpy() if len(slot2) else np.zeros(0)
        deg2 = res[2]["deg"].cpu().numpy()
        start = np.searchsorted(slot2, np.arange(A + 1))
        for b in range(A):
            s0, s1 = start[b], start[b + 1]
            if s1 > s0:
                k = s0 + int(np.argmin(d2[s0:s1]))
                ref[b] = np.float32(xp2[k] * deg2[k]) if (self.merge_degenerate and multi[b]) else xp2[k]
        ref_t = torch.tensor(ref, device=dev)
        kept = {}
        for nleg, o in res.items():
            P = o["deg"].numel()
            if nleg > 2 and self.merge_degenerate and P:
                dtot = self.degenerate_total(o["coords"], o["deg"], o["slot"])
            else:
                dtot = o["deg"]
            keep = torch.empty(P, dtype=torch.uint8, device=dev)
            if P:
                N.check(L.fm_find_filter(P, dtot.data_ptr(), o["xport"].data_ptr(), o["slot"].data_ptr(),
                                         ref_t.data_ptr(), float(np.float32(self.critpw)), keep.data_ptr(), st),
                        "fm_find_filter")
            sel = keep.bool()
            kept[nleg] = {k: o[k][sel] for k in ("coords", "ipots", "deg", "xport", "atoms", "slot")}
        return kept, ref

    def _merge_single_scattering(self, res, A, multi):
        """FEFF's accidental-degeneracy merge of two-leg paths (src/feffgpu/pathfinder.py:96-133), per absorber with
        more than one candidate, on the host as in the reference: every path of the absorber sorted by total length
        and cut into greedy windows of 1e-3 A; two-leg paths whose distance rounds to the same 1e-3 A, of the same
        potential and window, become one path (the shortest) whose degeneracy is their number."""
        eps3 = np.float32(1.0e-3)
        o2 = res[2]
        slot_all, rtot_all = [], []
        for nleg in sorted(res):
            slot_all.append(res[nleg]["slot"].cpu().numpy())
            rtot_all.append(res[nleg]["rtot"].cpu().numpy())
        slot_all = np.concatenate(slot_all)
        rtot_all = np.concatenate(rtot_all).astype(np.float32)
        n2 = o2["deg"].numel()
        d = torch.linalg.vector_norm(o2["coords"][:, 1], dim=-1).cpu().numpy().astype(np.float32)
        ip2 = o2["ipots"][:, 1].cpu().numpy()
        slot2 = slot_all[:n2]
        reps, counts = [], []
        for b in range(A):
            sel2 = np.nonzero(slot2 == b)[0]
            if not multi[b]:
                reps.extend(sel2.tolist())
                counts.extend([1.0] * len(sel2))
                continue
            # the absorber's paths in the reference's order: two legs first, then three, four, ...
            mine = np.nonzero(slot_all == b)[0]
            allr = rtot_all[mine]
            order = np.argsort(allr, kind="stable")
            sr = allr[order]
            window = np.zeros(len(sr), np.int64)
            i, w = 0, 0
            while i < len(sr):
                j = max(int(np.searchsorted(sr, np.float32(sr[i] + eps3), side="left")), i + 1)
                window[i:j] = w
                w += 1
                i = j
            win_of = np.empty(len(sr), np.int64)
            win_of[order] = window
            win2 = win_of[:len(sel2)]           # the absorber's two-leg paths come first in its list
            dd = d[sel2]
            key = np.round(dd * np.float32(1000)).astype(np.int64)
            groups = {}
            for k in np.argsort(dd, kind="stable"):
                groups.setdefault((int(key[k]), int(ip2[sel2[k]]), int(win2[k])), []).append(int(sel2[k]))
            for g in groups.values():
                reps.append(g[0])
                counts.append(float(len(g)))
        idx = torch.tensor(np.asarray(reps, np.int64), device=self.device)
        out = {k: v[idx] for k, v in o2.items()}
        out["deg"] = torch.tensor(np.asarray(counts, np.float32), device=self.device)
        return out

    def enumerate(self, pos, box, absorber, ipot, chunk=256):
        """pos (n, 3) A, the periodic cell (a cubic length, a 3 x 3 cell or None), the absorber's index, ipot (n,) the
        potential of every atom. Returns ({nleg: (coords (B, nleg, 3) relative to the absorber, ipots (B, nleg),
        deg (B,), xport (B,), atoms (B, nleg - 1))} after the importance filter, and the reference importance).
        chunk is accepted for compatibility."""
        kept, ref = self.enumerate_many(pos, box, [int(absorber)], ipot)
        return ({nleg: (v["coords"], v["ipots"], v["deg"], v["xport"], v["atoms"]) for nleg, v in kept.items()},
                float(ref[0]))
