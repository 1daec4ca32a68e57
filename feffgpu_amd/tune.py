"""Launch-policy autotuner of the path-stage kernels.

The kernels put G paths in a workgroup (1, 2, 4 or 8; bands of energies per wavefront, see tables.energy_bands).
More paths per workgroup balance the wavefronts better but take more LDS, which lowers the workgroups a CU holds and,
for the adjoint, decides between the register budgets of 1 and 2 wavefronts per SIMD. A bucket's launch is either one
workgroup per group of G paths, or a persistent grid of as many workgroups as the CUs hold that loops over the groups
(in a fixed order, or taking them from an atomic counter). Which is faster depends on the edge, the legs, the bucket's
size and the GPU, so it is measured:
* ``tune(plan, pos)`` times the plan's forward and adjoint under each candidate LDS budget and keeps the fastest for
  each (fm_set_lds_budget; process-wide, per kernel and precision);
* ``tune_buckets(plan, pos)`` times every bucket's forward and adjoint alone under every (G, grid) candidate and sets
  the fastest on that bucket of the plan (PathPlan.set_launch).
"""
from __future__ import annotations

import torch

from . import _native as N

BUDGETS_KIB = (40, 80, 160)


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        fn()
    ev[1].record()
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


def set_budget(kernel: int, kib: int, precision: int = 32) -> int:
    """kernel 0 (forward) or 1 (adjoint) of the given precision."""
    prev = N.lib().fm_set_lds_budget(int(kernel) + (2 if precision == 64 else 0), int(kib))
    if prev < 0:
        raise ValueError("kernel must be 0 (forward) or 1 (adjoint) and the budget 16..160 KiB")
    return prev


def tune(plan, pos, budgets=BUDGETS_KIB, reps: int = 3) -> dict:
    """Times plan.forward and plan.backward (dL/dchi_seg = 1) under each LDS budget and sets the fastest for each.
    Returns {"forward": {kib: ms}, "backward": {kib: ms}, "chosen": {"forward": kib, "backward": kib}}."""
    pos = torch.as_tensor(pos).to(device=plan.device, dtype=torch.float64)
    gseg = torch.ones(plan.nseg, plan.nk, dtype=torch.float64, device=plan.device)
    out = {"forward": {}, "backward": {}, "chosen": {}}
    for kernel, name, fn in ((0, "forward", lambda: plan.forward(pos)), (1, "backward", lambda: plan.backward(pos, gseg))):
        for kib in budgets:
            set_budget(kernel, kib, plan.precision)
            out[name][kib] = _time(fn, reps)
        best = min(out[name], key=out[name].get)
        set_budget(kernel, best, plan.precision)
        out["chosen"][name] = best
    return out


# (grid, queue): one workgroup per group; a persistent grid in a fixed order; a persistent grid with a counter
GRIDS = ((0, False), (-1, False), (-1, True))
GPATHS = (1, 2, 4, 8)


def _policy_name(gpaths: int, grid: int, queue: bool) -> str:
    g = "budget" if gpaths == 0 else f"G{gpaths}"
    return f"{g}/" + ("per-group" if grid == 0 else ("persistent-queue" if queue else "persistent"))


def tune_buckets(plan, pos, reps: int = 3, gpaths=GPATHS, grids=GRIDS) -> list:
    """For every bucket of the plan and each of its kernels (forward, adjoint): the bucket's launch alone timed under
    the default policy and under every candidate of gpaths x grids that the bucket's edge and LDS allow (the others
    are refused by the library and skipped); the fastest is set on the plan. Returns, per bucket, {kind, edge, nleg,
    paths, forward: {chosen, ms: {policy: ms}}, adjoint: {...}} (policies named G<n>/per-group, .../persistent,
    .../persistent-queue; "budget/per-group" is the default)."""
    pos = torch.as_tensor(pos).to(device=plan.device, dtype=torch.float64)
    gseg = torch.ones(plan.nseg, plan.nk, dtype=torch.float64, device=plan.device)
    keep = plan.keep_amplitudes     # as the captured step runs them: the forward keeps, the adjoint reads
    out = []
    for b in range(len(plan.buckets.kind)):
        row = dict(kind=int(plan.buckets.kind[b]), edge=int(plan.buckets.edge[b]), nleg=int(plan.buckets.nleg[b]),
                   paths=int(plan.buckets.offset[b + 1] - plan.buckets.offset[b]))
        for kernel, name, fn in ((0, "forward", lambda: plan.forward_bucket(pos, b, keep=keep)),
                                 (1, "adjoint", lambda: plan.backward_bucket(pos, gseg, b, kept=keep))):
            times, choice = {}, {}
            for g, grid, queue in [(0, 0, False)] + [(g, gr, qu) for g in gpaths for gr, qu in grids]:
                plan.set_launch(kernel, b, g, grid, queue)
                try:
                    t = _time(fn, reps)
                except ValueError:          # refused: no band plan for this G, or its paths do not fit one CU
                    continue
                key = _policy_name(g, grid, queue)
                times[key], choice[key] = t, (g, grid, queue)
            best = min(times, key=times.get)
            plan.set_launch(kernel, b, *choice[best])
            row[name] = dict(chosen=best, policy=choice[best], ms=times)
        out.append(row)
    return out


def apply_buckets(plan, rows) -> None:
    """Sets the policies tune_buckets chose (its returned rows) on the plan again."""
    for b, row in enumerate(rows):
        for kernel, name in ((0, "forward"), (1, "adjoint")):
            plan.set_launch(kernel, b, *row[name]["policy"])


def set_all(plan, grid: int, queue: bool) -> None:
    """Every bucket of the plan, both kernels: paths per workgroup by the LDS budget, the given grid."""
    for b in range(len(plan.buckets.kind)):
        for kernel in (0, 1):
            plan.set_launch(kernel, b, 0, grid, queue)
