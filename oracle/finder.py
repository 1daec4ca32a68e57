"""CPU restatement of feffgpu's path finder in float32 numpy: TEST INFRASTRUCTURE ONLY (the checker of the HIP path
finder; the product never imports it).

It follows src/feffgpu/pathfinder.py:
  * ``fbeta_tables``: |f(beta)| on FEFF's cos(beta) grid of 0.025 from the j = l - 1/2 phase shifts (:41-59);
  * ``enumerate``: every closed path of 2..nlegmax legs within 2 rmax (:211-279, the 5+-leg extension :160-209);
  * ``importance``: the plane-wave importance of a path (:72-94) with the bin rule of ``_indbet`` (:66-71);
  * ``degenerate_total``: symmetry-degenerate multiple-scattering paths counted together (:135-158);
  * ``merge_single_scattering``: FEFF's accidental-degeneracy merge of two-leg paths (:96-133);
  * the importance filter relative to the shortest single-scattering path (:266-279).
Everything is float32 as in the reference (positions, lengths, angles, importance). Distances are computed directly
(the reference's torch.cdist may take its matrix-product form), so a decision can differ only for a path whose
length or importance ties its cutoff to float32 rounding.

Pinned in tests/test_finder_oracle.py against FEFF 8.5L's own path lists (paths.dat of the W set and the six
validation cases: the degeneracy-weighted number of paths per number of legs) and against a plain enumeration.
"""
from __future__ import annotations

import numpy as np
from scipy.special import eval_legendre

BOHR = 0.52917721067
NBETA = 40
F32 = np.float32


def fbeta_tables(oph):
    """From an oracle.OraclePhase: fbeta (E, nph+1, 81) float32, cksp (E,) 1/A, xlam (E,) A, ik0 (0-based), trap
    (E - ik0,) float32 trapezoid weights over cksp[ik0:]."""
    lm, xk, ck = oph.tables()
    sh, ik0 = oph.phase_shifts(-1)                   # ph(ie, -l, iph)
    E = oph.ne1
    cosb = 0.025 * np.arange(-NBETA, NBETA + 1)
    cosb[0], cosb[-1] = -1.0, 1.0
    P = np.stack([eval_legendre(l, cosb) for l in range(sh.shape[1])])          # (ltot+1, 81)
    fb = np.zeros((E, oph.nph + 1, 2 * NBETA + 1))
    for iph in range(oph.nph + 1):
        for ie in range(E):
            L = int(lm[ie, iph]) + 1
            tl = (np.exp(2j * sh[ie, :L, iph]) - 1) / 2j
            fb[ie, iph] = np.abs(((tl * (2 * np.arange(L) + 1))[:, None] * P[:L]).sum(0))
    cksp = (ck.real / BOHR).astype(F32)
    xlam = (np.where(np.abs(ck.imag) > 1e-16, 1 / np.where(ck.imag == 0, 1, ck.imag), 1e10) * BOHR).astype(F32)
    k0 = ik0 - 1
    x = ck.real[k0:] / BOHR
    w = np.zeros(len(x))
    w[0] = x[1] - x[0]
    w[1:-1] = x[2:] - x[:-2]
    w[-1] = x[-1] - x[-2]
    return fb.astype(F32), cksp, xlam, k0, (w / 2).astype(F32)


def _wrap(rel, box):
    if box is None:
        return rel
    b = np.asarray(box, np.float64)
    if b.ndim == 0:
        L = F32(b)
        return (rel - L * np.round(rel / L)).astype(F32)
    c = (np.diag(b) if b.ndim == 1 else b).astype(F32)
    f = (rel @ np.linalg.inv(c.astype(np.float64)).astype(F32)).astype(F32)
    return ((f - np.round(f)) @ c).astype(F32)


def _norm(v):
    return np.sqrt((v * v).sum(-1, dtype=F32)).astype(F32)


def _bins(cosb):
    t = np.abs(cosb).astype(F32)
    n = np.floor(t / F32(0.025)).astype(F32)
    n = np.where(t - n * F32(0.025) > F32(0.0125), n + 1, n)
    n = np.where(cosb < 0, -n, n)
    return np.clip(n.astype(np.int64) + NBETA, 0, 2 * NBETA)


class OracleFinder:
    def __init__(self, oph, rmax, nlegmax=4, critpw=0.1, merge_degenerate=True):
        self.fb, self.cksp, self.xlam, self.ik0, self.trap = fbeta_tables(oph)
        self.rmax, self.nlegmax, self.critpw, self.merge = float(rmax), int(nlegmax), float(critpw), merge_degenerate

    def importance(self, pts, ipots):
        """pts (B, m+1, 3) float32, absorber at index 0; ipots (B, m+1). Returns (xport (B,), rtot (B,))."""
        pts = np.asarray(pts, F32)
        B, m1, _ = pts.shape
        m = m1 - 1
        closed = np.concatenate([pts, pts[:, :1]], 1)
        legs = (closed[:, 1:] - closed[:, :-1]).astype(F32)
        ri = _norm(legs)
        rtot = ri.sum(1, dtype=F32)
        u = (legs / np.maximum(ri, F32(1e-12))[..., None]).astype(F32)
        cos_sc = (u[:, :-1] * u[:, 1:]).sum(-1, dtype=F32)
        cos_abs = (u[:, -1] * u[:, 0]).sum(-1, dtype=F32)
        ies = np.arange(self.ik0, self.fb.shape[0])
        ck = self.cksp[ies]
        crit = (np.maximum(np.abs(cos_abs), F32(0.3))[:, None] / (ri[:, -1][:, None] * ck[None])).astype(F32)
        ib = _bins(cos_sc)
        for j in range(m):
            f = self.fb[ies][:, ipots[:, j + 1], ib[:, j]].T
            crit = ((crit * f) / (ri[:, j][:, None] * ck[None])).astype(F32)
        crit = (crit * np.exp(-rtot[:, None] / self.xlam[ies][None]).astype(F32)).astype(F32)
        return (np.abs(crit) * self.trap[None]).sum(1, dtype=F32), rtot

    @staticmethod
    def degenerate_total(coords, deg):
        """Group totals of the degeneracy over paths of equal leg lengths and angles (to 1e-3, either direction)."""
        c = np.asarray(coords, F32)
        closed = np.concatenate([c, c[:, :1]], 1)
        legs = (closed[:, 1:] - closed[:, :-1]).astype(F32)
        ri = _norm(legs)
        u = (legs / np.maximum(ri, F32(1e-12))[..., None]).astype(F32)
        cosb = (u * np.roll(u, -1, 1)).sum(-1, dtype=F32)
        n = ri.shape[1]
        fwd = np.concatenate([ri, cosb], 1)
        rev = np.concatenate([ri[:, ::-1], cosb[:, :n - 1][:, ::-1], cosb[:, n - 1:]], 1)
        kf = np.round(fwd * F32(1000)).astype(np.int64)
        kr = np.round(rev * F32(1000)).astype(np.int64)
        diff = kf != kr
        first = np.argmax(diff, 1)
        rows = np.arange(len(kf))
        less = diff.any(1) & (kr[rows, first] < kf[rows, first])
        key = np.where(less[:, None], kr, kf)
        _, inv = np.unique(key, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        tot = np.zeros(inv.max() + 1 if len(inv) else 0, F32)
        np.add.at(tot, inv, np.asarray(deg, F32))
        return tot[inv]

    def _longer(self, cl, inter, rel, two_r):
        nc = len(cl)
        P = rel[inter]
        D = _norm(P[:, None] - P[None])
        back = _norm(P)
        seq = np.arange(nc)[:, None]
        length = back[:nc].copy()
        for m in range(2, self.nlegmax):
            tot = (length[:, None] + D[seq[:, -1]]).astype(F32)
            ok = ((tot + back[None]).astype(F32) <= two_r) & (np.arange(len(inter))[None] != seq[:, -1:])
            i, j = np.nonzero(ok)
            seq, length = np.concatenate([seq[i], j[:, None]], 1), tot[i, j]
            if len(seq) == 0:
                return
            if m + 1 < 5:
                continue
            done = seq[seq[:, -1] != nc]
            if len(done) == 0:
                continue
            rev = done[:, ::-1]
            differ = done != rev
            first = np.argmax(differ, 1)
            rows = np.arange(len(done))
            pal = ~differ.any(1)
            keep = pal | (done[rows, first] < rev[rows, first])
            yield m + 1, inter[done[keep]], np.where(pal[keep], F32(1), F32(2))

    def enumerate(self, pos, box, absorber, ipot):
        """{nleg: (coords (B, nleg, 3) float32, ipots (B, nleg), deg (B,) float32, xport (B,), atoms (B, nleg-1))}
        after the filter, and the reference importance."""
        pos = np.asarray(pos, F32)
        ipot = np.asarray(ipot)
        rel = _wrap((pos - pos[absorber]).astype(F32), box)
        d0 = _norm(rel)
        two_r = F32(2 * self.rmax)
        cl = np.nonzero((d0 <= F32(self.rmax)) & (np.arange(len(pos)) != absorber))[0]
        inter = np.concatenate([cl, [absorber]])
        rc, ri_ = rel[cl], rel[inter]
        dc0 = d0[cl]
        ipot_of = lambda a: np.where(a == absorber, 0, ipot[a])
        out = {}

        def finish(atoms, deg):
            coords = np.concatenate([np.zeros((len(atoms), 1, 3), F32), rel[atoms]], 1)
            ipots = np.concatenate([np.zeros((len(atoms), 1), np.int64), ipot_of(atoms)], 1)
            xp = self.importance(coords, ipots)[0] if len(atoms) else np.zeros(0, F32)
            return coords, ipots, np.asarray(deg, F32), xp, atoms

        out[2] = finish(cl[:, None], np.ones(len(cl), F32))
        ref = out[2][3][np.argmin(dc0)] if len(cl) else F32(1)
        if self.nlegmax >= 3 and len(cl) > 1:
            dij = _norm(rc[:, None] - rc[None])
            tot = ((dc0[:, None] + dij).astype(F32) + dc0[None]).astype(F32)
            ii, jj = np.nonzero((tot <= two_r) & (np.arange(len(cl))[:, None] < np.arange(len(cl))[None]))
            out[3] = finish(np.stack([cl[ii], cl[jj]], 1), np.full(len(ii), 2, F32))
        if self.nlegmax >= 4 and len(cl) > 1:
            dci = _norm(rc[:, None] - ri_[None])                                   # (nc, ni)
            tot = (((dc0[:, None, None] + dci[:, :, None]).astype(F32) + dci.T[None]).astype(F32)
                   + dc0[None, None]).astype(F32)
            ok = tot <= two_r
            ok &= np.arange(len(cl))[:, None, None] <= np.arange(len(cl))[None, None, :]
            ok &= (cl[:, None, None] != inter[None, :, None]) & (inter[None, :, None] != cl[None, None, :])
            a, b, c = np.nonzero(ok)
            atoms4 = np.stack([cl[a], inter[b], cl[c]], 1)
            out[4] = finish(atoms4, np.where(atoms4[:, 0] == atoms4[:, 2], F32(1), F32(2)))
        if self.nlegmax >= 5 and len(cl) > 1:
            for nleg, atoms, deg in self._longer(cl, inter, rel, two_r):
                out[nleg] = finish(atoms, deg)
        if self.merge and len(cl) > 1:
            out = self.merge_single_scattering(out)
            d = _norm(out[2][0][:, 1])
            k = int(np.argmin(d))
            ref = F32(out[2][3][k] * out[2][2][k])
        ref = F32(ref)
        kept = {}
        for nleg, (coords, ipots, deg, xp, atoms) in out.items():
            dtot = self.degenerate_total(coords, deg) if (nleg > 2 and self.merge and len(deg)) else deg
            frac = ((F32(100) * dtot).astype(F32) * xp).astype(F32) / ref
            keep = frac.astype(F32) >= F32(self.critpw)
            kept[nleg] = (coords[keep], ipots[keep], deg[keep], xp[keep], atoms[keep])
        return kept, float(ref)

    @staticmethod
    def merge_single_scattering(out):
        eps3 = 1.0e-3
        rtot = []
        for nleg, (coords, *_rest) in out.items():
            closed = np.concatenate([coords, coords[:, :1]], 1)
            rtot.append(_norm((closed[:, 1:] - closed[:, :-1]).astype(F32)).sum(1, dtype=F32))
        allr = np.concatenate(rtot).astype(F32)
        order = np.argsort(allr, kind="stable")
        sr = allr[order]
        window = np.zeros(len(sr), np.int64)
        i, w = 0, 0
        while i < len(sr):
            j = int(np.searchsorted(sr, F32(sr[i] + F32(eps3)), side="left"))
            j = max(j, i + 1)
            window[i:j] = w
            w += 1
            i = j
        win_of = np.empty(len(sr), np.int64)
        win_of[order] = window
        coords, ipots, deg, xp, atoms = out[2]
        n2 = len(coords)
        d = _norm(coords[:, 1])
        key = np.round(d * F32(1000)).astype(np.int64)
        groups = {}
        for k in np.argsort(d, kind="stable"):
            groups.setdefault((int(key[k]), int(ipots[k, 1]), int(win_of[:n2][k])), []).append(int(k))
        rep = np.array([g[0] for g in groups.values()], np.int64)
        cnt = np.array([len(g) for g in groups.values()], F32)
        out = dict(out)
        out[2] = (coords[rep], ipots[rep], cnt, xp[rep], atoms[rep])
        return out
