"""Launch-policy autotuner of the path-stage kernels.

The kernels put G paths in a workgroup (1, 2, 4 or 8; bands of energies per wavefront, see tables.energy_bands).
More paths per workgroup balance the wavefronts better but take more LDS, which lowers the workgroups a CU holds and,
for the adjoint, decides between the register budgets of 1 and 2 wavefronts per SIMD. Which is faster depends on the
edge, the legs and the GPU, so it is measured: ``tune(plan, pos)`` times the plan's forward and adjoint under each
candidate LDS budget and keeps the fastest for each (fm_set_lds_budget; process-wide, per kernel and precision).
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
