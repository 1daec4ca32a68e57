"""feffgpu_amd: feffgpu's path stage (the batched, position-differentiable Rehr-Albers path amplitudes and their sum
to chi(k)) on MI355X, as hand-written HIP kernels behind a C ABI (include/feffgpu_amd.h).

Public API (drop-in names of the reference):
    GPUPathStage           -- src/feffgpu/stage.py:96 (cchi, linear_matrix, spline_matrix, chi_fine, unwrap)
    read_phase_pad, PhaseData, load_phase, BOHR -- src/feffgpu/pad.py
    r_factor, hanning      -- src/feffgpu/compare.py
    PathPlan, chi_paths    -- the fused path-list entry (positions in, chi per segment and per path out)
"""
from .pad import BOHR, HART, PhaseData, load_phase, read_phase_pad
from .compare import hanning, r_factor

__version__ = "0.1.0"


def __getattr__(name):
    # the GPU classes import torch; keep the light names importable without it
    if name == "GPUPathStage":
        from .stage import GPUPathStage
        return GPUPathStage
    if name in ("PathPlan", "chi_paths"):
        from . import paths
        return getattr(paths, name)
    raise AttributeError(name)


__all__ = ["BOHR", "HART", "PhaseData", "load_phase", "read_phase_pad", "hanning", "r_factor", "GPUPathStage",
           "PathPlan", "chi_paths"]
