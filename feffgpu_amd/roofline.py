"""Algorithmic work of the path stage, counted once here for every roofline figure.

Per path, summed over the phase-grid energies e the kernel evaluates, in complex multiply-adds (cMAC, 8 flops
each; transcendentals and the fine-grid epilogue are not counted):

    cMAC(e) = sum_legs 2 mu L_j(e)                 spherical-wave recursion of every leg
            + sum_blocks (r_b + c_b) L_b(e)        gam / gamtl construction of every scattering matrix
            + sum_blocks r_b c_b L_b(e)            the contraction over partial waves
            + sum_blocks r_b c_b                   the row phases exp(-i eta m)
            + sum_chain r c lambda0                the matrix chain
            + n_ch lambda0^2 + lambda0^2           termination and trace

L_j(e) = max(lmax(e) at the leg's two atoms) + 1; L_b(e) = lmax(e, scatterer) + 1; mu = mnp1 (4, or 3 for
two-leg K paths); blocks (lambda x lambda0) for two legs, else (lambda x lambda0), (n-3) (lambda x lambda),
(lambda0 x lambda); lambda = 6, lambda0 = 4 (K) or 6; two-leg paths use the single-scattering set (6 x 6 for K, 15 x 15 for L/M).
The HBM traffic of a launch is small (path indices and positions in, NK floats per
path out): the kernels are bound by the FP32 (or FP64) vector pipes, not by HBM.
"""
from __future__ import annotations

import numpy as np

FLOPS_PER_CMAC = 8
MI355X_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}      # published dense peaks (MFMA); FP32 vector needs packed math
MI355X_HBM_TBPS = 8.0


def path_cmacs(tables, nleg: np.ndarray, ipots: np.ndarray) -> np.ndarray:
    """cMACs per path for paths of one edge: nleg (P,), ipots (P, nmax) (absorber first). tables: EdgeTables."""
    nleg = np.asarray(nleg)
    ipots = np.asarray(ipots)
    E = tables.E
    Lp = tables.lmax.astype(np.int64) + 1                 # (nph1, E): lmax + 1 per potential and energy
    nch = len(tables.chan_il)
    out = np.zeros(len(nleg))
    for n in np.unique(nleg):
        sel = nleg == n
        ip = ipots[sel][:, :n]                              # atoms 0..n-1 (0 = absorber)
        atom = np.concatenate([ip, ip[:, :1]], 1)            # atoms 0..n, atom n = absorber
        Lat = Lp[atom]                                      # (B, n+1, E)
        Lleg = np.maximum(Lat[:, 1:], Lat[:, :-1])          # legs 1..n
        if n == 2:                                          # single scattering: the nsc = 1 lambda set
            lam = lam0 = 6 if tables.ilinit <= 1 else 15
            mu = {1: 3, 2: 5, 3: 6}[tables.ilinit]
            blocks, chain = [(lam, lam0, 1)], []
        else:
            lam, lam0, mu = 6, (4 if tables.ilinit <= 1 else 6), 4
            blocks = [(lam, lam0, 1)] + [(lam, lam, s) for s in range(2, n - 1)] + [(lam0, lam, n - 1)]
            chain = [(lam, lam)] * (n - 3) + [(lam0, lam)]
        lam0_eff = lam0
        c = (2 * mu * Lleg).sum((1, 2)).astype(float)
        for r, cc, s in blocks:
            Ls = Lat[:, s].sum(1)                           # scatterer s, summed over energies
            c += (r + cc) * Ls + r * cc * Ls + r * cc * E
        for r, cc in chain:
            c += r * cc * lam0_eff * E
        c += (nch * lam0_eff ** 2 + lam0_eff ** 2) * E
        out[sel] = c
    return out


def plan_flops(plan) -> float:
    """Algorithmic flops of one forward launch of a PathPlan (all edges)."""
    total = 0.0
    for e, T in enumerate(plan.tables):
        sel = plan.edge_h == e
        if sel.any():
            ip = plan.ipots_t.detach().cpu().numpy()[sel] if hasattr(plan, "ipots_t") else None
            total += path_cmacs(T, plan.nleg_h[sel], ip).sum() * FLOPS_PER_CMAC
    return float(total)
