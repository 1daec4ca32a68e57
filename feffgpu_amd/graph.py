"""One step of the path stage captured as a hipGraph: the fused forward over a plan's whole path list and, if asked,
its adjoint, replayed on static buffers.

Every launch of a step is stream-ordered and allocation-free once the plan's buffers exist (fm_chi_paths,
fm_chi_paths_backward, torch's elementwise kernels; no memset nodes, no host synchronisation), so the sequence
captures as one graph: bucket kernels, segment sum, the loss gradient, adjoint kernels, per-atom gradient sum. A
replay re-reads `pos` (and `gseg`) from the same device addresses; copy new values into them before replaying.

    step = CapturedStep(plan, natom, loss_grad=lambda seg: 2 * w * (seg - target))
    step.pos.copy_(pos)            # or step(pos)
    chi_seg, grad = step()
"""
from __future__ import annotations

from typing import Callable, Optional

import torch


class CapturedStep:
    """plan: a PathPlan. natom: rows of the position array. adjoint: capture the backward too. loss_grad: maps
    chi_seg (nseg, NK) float64 to dL/dchi_seg inside the graph (torch ops, or feffgpu_amd.transform.RFactorLoss:
    launches on the current stream that allocate nothing once warmed up); without it the adjoint reads the
    static buffer `gseg`. warmup: eager runs before capture (they also set the kernels' LDS attributes)."""

    def __init__(self, plan, natom: int, adjoint: bool = True, loss_grad: Optional[Callable] = None, warmup: int = 2):
        if plan.device.type != "cuda":
            raise RuntimeError("CapturedStep needs a ROCm device")
        self.plan, self.adjoint, self.loss_grad = plan, bool(adjoint), loss_grad
        dev = plan.device
        self.pos = torch.zeros(int(natom), 3, dtype=torch.float64, device=dev)
        self.gseg = torch.zeros(plan.nseg, plan.nk, dtype=torch.float64, device=dev)
        side = torch.cuda.Stream(device=dev)          # default priority
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for _ in range(max(1, int(warmup))):
                self._body()
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = self._body()

    def _body(self):
        keep = self.adjoint and self.plan.keep_amplitudes
        seg, _, _ = self.plan.forward(self.pos, keep=keep)
        if not self.adjoint:
            return seg, None
        g = self.loss_grad(seg) if self.loss_grad is not None else self.gseg
        return seg, self.plan.backward(self.pos, g, kept=keep)

    def __call__(self, pos=None, gseg=None):
        """Replays the step (after copying pos / gseg into the static buffers when given). Returns (chi_seg,
        grad_pos or None): the graph's own output tensors, overwritten by the next replay."""
        if pos is not None:
            self.pos.copy_(torch.as_tensor(pos, device=self.pos.device, dtype=torch.float64))
        if gseg is not None:
            self.gseg.copy_(torch.as_tensor(gseg, device=self.gseg.device, dtype=torch.float64))
        self.graph.replay()
        return self.out
