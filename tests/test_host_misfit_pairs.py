"""Host logic of the misfit and the pair tables, without a GPU: FitWindow's checks (src/feffgpu/exafs/transform.py:55-62),
the R ranges of the two grid conventions, the uniform-grid check, and the cell handling of PairTables."""
import math

import numpy as np
import pytest


def test_fit_window_checks_and_independent_points():
    from feffgpu_amd.transform import FitWindow
    w = FitWindow()
    assert (w.kmin, w.kmax, w.dk, w.kweight, w.rmin, w.rmax) == (3.0, 14.0, 1.0, 2, 1.0, 6.0)
    assert abs(w.independent_points - (2 * 11 * 5 / math.pi + 2)) < 1e-12
    for bad in (dict(kmin=5.0, kmax=4.0), dict(kmin=-1.0), dict(rmin=3.0, rmax=2.0), dict(dk=-0.5)):
        with pytest.raises(ValueError):
            FitWindow(**bad)


def test_r_ranges():
    from feffgpu_amd.transform import KSTEP, NFFT, RSTEP, FitWindow, r_range
    w = FitWindow(rmin=1.0, rmax=6.0)
    m0, nm = r_range(w)
    r = RSTEP * np.arange(NFFT // 2)
    sel = np.nonzero((r >= 1.0) & (r <= 6.0))[0]
    assert (m0, nm) == (sel[0], len(sel))
    assert r_range(FitWindow(rmin=0.0, rmax=1e3)) == (0, NFFT // 2)
    # a bound exactly on a grid point is inside in both conventions
    w2 = FitWindow(rmin=float(np.float32(RSTEP) * np.float32(40)), rmax=3.0)
    assert r_range(w2)[0] == 40
    with pytest.raises(ValueError):
        r_range(FitWindow(rmin=1.001, rmax=1.002))
    assert r_range(w, KSTEP, NFFT, float32_grid=False) == (m0, nm)


def test_uniform_grid_check():
    from feffgpu_amd.transform import _kstep_of, k_grid
    step, n = _kstep_of(k_grid(18.0))
    assert abs(step - 0.05) < 1e-15 and n == 361
    for bad in (k_grid(18.0) + 0.1, np.r_[0.0, 0.05, 0.11, 0.15], np.array([0.0])):
        with pytest.raises(ValueError):
            _kstep_of(bad)


def test_pair_cells():
    from feffgpu_amd.pairs import _cell
    kind, cell, inv, width = _cell(None)
    assert kind == 0 and width == math.inf
    kind, cell, inv, width = _cell(14.35)
    assert kind == 1 and cell[0] == 14.35 and width == 14.35
    c = np.array([[10.0, 0, 0], [5.0, 8.660254037844386, 0], [0, 0, 12.0]])   # 60 degree cell
    kind, cell, inv, width = _cell(c)
    assert kind == 2 and np.allclose(cell.reshape(3, 3) @ inv.reshape(3, 3), np.eye(3))
    assert abs(width - 8.660254037844386) < 1e-12
    kind, cell, _, width = _cell(np.array([9.0, 10.0, 11.0]))
    assert kind == 2 and np.allclose(cell.reshape(3, 3), np.diag([9.0, 10.0, 11.0])) and abs(width - 9.0) < 1e-12
