"""Build the HIP library libfeffgpu_amd.so in-tree for gfx950 (hipcc; cross-compiles without a GPU).

    python -m feffgpu_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ["csrc/fm_forward.hip", "csrc/fm_backward.hip", "csrc/fm_plain.hip", "csrc/fm_loss.hip", "csrc/fm_pairs.hip", "csrc/fm_finder.hip"]
OUT = HERE / "libfeffgpu_amd.so"
ARCH = os.environ.get("FEFFGPU_AMD_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
         # packed FP32 instructions are disabled for every kernel of this library
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise FileNotFoundError("hipcc not found (ROCm)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")
    if verbose and err.strip():
        print(err)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compiles every source to an object in parallel (one hipcc per file), then links the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    srcs = [HERE / s for s in SOURCES if (HERE / s).exists()]
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (s.stem + ".o") for s in srcs]
    cflags = [f for f in FLAGS if f != "-shared"]
    jobs = jobs or min(len(srcs), os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([hipcc(), *cflags, "-c", "-o", str(so[1]), str(so[0])], verbose), zip(srcs, objs)))
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)], verbose)
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
