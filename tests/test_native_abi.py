"""The C ABI library: builds, loads without a GPU, exports every entry point include/feffgpu_amd.h declares, and its
structs have the same layout in C (gcc) and in the ctypes binding. No compute calls here (no GPU)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "feffgpu_amd.h"


def declared_functions():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*(fm_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from feffgpu_amd import _native as N
    if not N.LIB_PATH.exists():
        from feffgpu_amd.build import build
        build()
    lib = ctypes.CDLL(str(N.LIB_PATH))
    names = declared_functions()
    assert "fm_chi_paths" in names and "fm_cchi" in names and "fm_backward_workspace_bytes" in names
    assert len(names) >= 9
    for name in names:
        assert hasattr(lib, name), name
    lib.fm_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.fm_version()


def test_struct_layout_matches_ctypes(tmp_path):
    from feffgpu_amd import _native as N
    src = tmp_path / "probe.c"
    src.write_text(f"""#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
#define F(S, m) printf("%s.%s %zu\\n", #S, #m, offsetof(S, m))
int main(void) {{
  printf("fm_edge %zu\\nfm_paths %zu\\nfm_grid %zu\\nfm_window %zu\\n", sizeof(fm_edge), sizeof(fm_paths), sizeof(fm_grid),
         sizeof(fm_window));
  F(fm_edge, lmax1); F(fm_edge, lmax); F(fm_edge, rotn); F(fm_paths, pos); F(fm_paths, cell_kind); F(fm_paths, cellinv);
  F(fm_grid, q); F(fm_grid, fq); F(fm_grid, S); F(fm_grid, sstride); F(fm_window, kweight); F(fm_window, rmin);
  return 0;
}}
""")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    assert int(out["fm_edge"]) == ctypes.sizeof(N.FmEdge)
    assert int(out["fm_paths"]) == ctypes.sizeof(N.FmPaths)
    assert int(out["fm_grid"]) == ctypes.sizeof(N.FmGrid)
    assert int(out["fm_window"]) == ctypes.sizeof(N.FmWindow)
    for key, (cls, field) in {"fm_edge.lmax1": (N.FmEdge, "lmax1"), "fm_edge.lmax": (N.FmEdge, "lmax"),
                              "fm_edge.rotn": (N.FmEdge, "rotn"), "fm_paths.pos": (N.FmPaths, "pos"),
                              "fm_paths.cell_kind": (N.FmPaths, "cell_kind"), "fm_paths.cellinv": (N.FmPaths, "cellinv"),
                              "fm_grid.q": (N.FmGrid, "q"), "fm_grid.fq": (N.FmGrid, "fq"), "fm_grid.S": (N.FmGrid, "S"),
                              "fm_grid.sstride": (N.FmGrid, "sstride"), "fm_window.kweight": (N.FmWindow, "kweight"),
                              "fm_window.rmin": (N.FmWindow, "rmin")}.items():
        assert int(out[key]) == getattr(cls, field).offset, key


def test_gpu_classes_refuse_cpu_device():
    import torch
    from feffgpu_amd.stage import GPUPathStage
    with pytest.raises(RuntimeError):
        GPUPathStage(str(ROOT / "tests/golden/W/phase.pad"), device="cpu")
