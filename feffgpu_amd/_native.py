"""ctypes binding of the HIP library libfeffgpu_amd.so (the C ABI of include/feffgpu_amd.h).

The library is built in-tree by ``feffgpu_amd.build`` (hipcc, gfx950). Loading fails loudly when it is missing:
there is no CPU fallback of the path stage in this package.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import torch

from . import tables as _tables

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfeffgpu_amd.so"

FM_MAX_LEGS, FM_MAX_CHAN, FM_MAX_POT, FM_NREP = 9, 4, 16, 20
FM_BUCKET_K_MS, FM_BUCKET_LM_MS, FM_BUCKET_SS = 0, 1, 2
FM_OK, FM_EINVAL, FM_EUNSUPPORTED, FM_EHIP = 0, -1, -2, -3

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32


class FmBands(ctypes.Structure):
    _fields_ = [("nband", _i32), ("start", _i32 * 17), ("wave", _i32 * 16)]


class FmEdge(ctypes.Structure):
    _fields_ = [("E", _i32), ("L1", _i32), ("nph1", _i32), ("ilinit", _i32), ("nchan", _i32),
                ("chan_il", _i32 * FM_MAX_CHAN), ("lmax1", _i32 * FM_MAX_POT),
                ("lmax", _vp), ("ck", _vp), ("ckmxk", _vp), ("kf", _vp), ("tl32", _vp), ("tl64", _vp),
                ("chw32", _vp), ("chw64", _vp), ("rotc", _vp), ("rotn", _vp), ("bands", FmBands * 4)]


class FmPaths(ctypes.Structure):
    _fields_ = [("npath", _i32), ("nmax", _i32), ("pos", _vp), ("atoms", _vp), ("nleg", _vp), ("absorber", _vp),
                ("ipots", _vp), ("deg", _vp), ("edge", _vp), ("sigma2", _vp), ("cell_kind", _i32),
                ("cell", ctypes.c_double * 9), ("cellinv", ctypes.c_double * 9)]


class FmGrid(ctypes.Structure):
    _fields_ = [("nk", _i32), ("q", _vp), ("jq", _vp), ("fq", _vp), ("qorder", _vp), ("qstart", _vp), ("qstride", _i32),
                ("S", _vp), ("sstride", _i32)]


class FmWindow(ctypes.Structure):
    _fields_ = [("kmin", ctypes.c_double), ("kmax", ctypes.c_double), ("dk", ctypes.c_double),
                ("kweight", _i32), ("rmin", ctypes.c_double), ("rmax", ctypes.c_double)]


class FmPairGrid(ctypes.Structure):
    _fields_ = [("r0", ctypes.c_double), ("dr", ctypes.c_double), ("rmax", ctypes.c_double), ("nr", _i32),
                ("cell_kind", _i32), ("cell", ctypes.c_double * 9), ("cellinv", ctypes.c_double * 9)]


class FmFindSys(ctypes.Structure):
    _fields_ = [("natom", _i32), ("pos", _vp), ("cell_kind", _i32), ("cell", ctypes.c_float * 9),
                ("cellinv", ctypes.c_float * 9), ("rmax", ctypes.c_float)]


class FmFindTables(ctypes.Structure):
    _fields_ = [("E", _i32), ("nph1", _i32), ("ik0", _i32), ("fbeta", _vp), ("cksp", _vp), ("xlam", _vp), ("trap", _vp)]


class FmFindOut(ctypes.Structure):
    _fields_ = [("atoms", _vp), ("coords", _vp), ("ipots", _vp), ("deg", _vp), ("slot", _vp)]


class FmFindPartial(ctypes.Structure):
    _fields_ = [("seq", _vp), ("len", _vp), ("slot", _vp)]


class FmLaunch(ctypes.Structure):
    """The launch policy of one bucket: gpaths (0 by the LDS budget, else 1, 2, 4, 8), grid (0 one workgroup per
    group, n > 0 a persistent grid of n, -1 as many as the CUs hold), tickets (a zeroed device uint32 pair: groups
    taken from a counter; None: fixed assignment), keep / keep_bytes (the bucket's kept scattering matrices: written
    by the forward call, read by the adjoint call; None: kept nothing, recomputed)."""
    _fields_ = [("gpaths", _i32), ("grid", _i32), ("tickets", _vp), ("keep", _vp), ("keep_bytes", ctypes.c_size_t)]


def grid_runs(q, jq_rows, E_list):
    """qorder (q indices in ascending q) and qstart [nedge][stride]: for every edge, the first position in qorder
    whose lower interpolation index is >= e, for e = 0..E (the runs the adjoint of the fine grid gathers)."""
    jq_rows = [np.asarray(j) for j in jq_rows]
    stride = max(E_list) + 1
    order = np.argsort(np.asarray(q), kind="stable")
    qstart = np.zeros((len(jq_rows), stride), np.int32)
    for r, (j, E) in enumerate(zip(jq_rows, E_list)):
        js = j[order]
        if np.any(np.diff(js) < 0):
            order = None
            break
        qstart[r] = np.searchsorted(js, np.arange(stride), side="left")
    if order is None:
        raise ValueError("the interpolation indices of the grid are not monotone in a common order")
    return order.astype(np.int32), qstart, stride


class NativeError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeError(f"{LIB_PATH} is missing: build it with `python -m feffgpu_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        L.fm_last_error.restype = ctypes.c_char_p
        L.fm_version.restype = ctypes.c_char_p
        L.fm_cchi.argtypes = [ctypes.POINTER(FmEdge), _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp]
        L.fm_cchi_backward.argtypes = [ctypes.POINTER(FmEdge), _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp,
                                       ctypes.c_size_t, _vp]
        L.fm_cchi_backward_noalloc.argtypes = [ctypes.POINTER(FmEdge), _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp,
                                               _vp, _vp]
        _i64, _f32, _sz = ctypes.c_int64, ctypes.c_float, ctypes.c_size_t
        _P = ctypes.POINTER
        L.fm_find_candidates.argtypes = [_P(FmFindSys), _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp]
        L.fm_find_count.argtypes = [_i32, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _f32, _i32, _vp, _vp, _vp]
        L.fm_scan_workspace_bytes.restype = _sz
        L.fm_scan_workspace_bytes.argtypes = [_i64]
        L.fm_exclusive_scan.argtypes = [_vp, _i64, _vp, _vp, _vp, _sz, _vp]
        L.fm_find_write.argtypes = [_i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _i32, _vp, _f32, _i32, _vp,
                                    _P(FmFindOut), _vp, _P(FmFindOut), _vp, _P(FmFindOut), _vp]
        L.fm_find_seed.argtypes = [_i32, _vp, _i32, _vp, _vp, _i32, _P(FmFindPartial), _vp]
        L.fm_find_extend.argtypes = [_i64, _i32, _P(FmFindPartial), _i32, _vp, _vp, _vp, _f32, _vp, _vp,
                                     _P(FmFindPartial), _vp]
        L.fm_find_emit.argtypes = [_i64, _i32, _P(FmFindPartial), _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _P(FmFindOut), _vp]
        L.fm_find_importance.argtypes = [_P(FmFindTables), _i64, _i32, _vp, _vp, _vp, _vp, _vp]
        L.fm_find_group_workspace_bytes.restype = _sz
        L.fm_find_group_workspace_bytes.argtypes = [_i64]
        L.fm_find_degenerate_total.argtypes = [_i64, _i32, _vp, _vp, _vp, _vp, _vp, _sz, _vp]
        L.fm_find_filter.argtypes = [_i64, _vp, _vp, _vp, _vp, _f32, _vp, _vp]
        L.fm_backward_workspace_bytes.restype = ctypes.c_size_t
        L.fm_backward_workspace_bytes.argtypes = [_i32, _i32, _i32]
        L.fm_chi_fine.argtypes = [ctypes.POINTER(FmEdge), _i32, _i32, _vp, _vp, ctypes.POINTER(FmGrid), _vp, _vp]
        L.fm_chi_fine_backward.argtypes = [ctypes.POINTER(FmEdge), _i32, _i32, _vp, _vp, ctypes.POINTER(FmGrid), _vp,
                                           _vp, _vp, _vp]
        L.fm_chi_paths.argtypes = [ctypes.POINTER(FmEdge), _i32, ctypes.POINTER(FmPaths), ctypes.POINTER(FmGrid), _i32,
                                   _i32, _vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                   ctypes.POINTER(_i32), _i32, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(FmLaunch), _vp]
        L.fm_chi_paths_backward.argtypes = [ctypes.POINTER(FmEdge), _i32, ctypes.POINTER(FmPaths),
                                            ctypes.POINTER(FmGrid), _i32, _i32, _vp, _i32, ctypes.POINTER(_i32),
                                            ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32), _vp, _vp,
                                            _vp, _i32, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t,
                                            ctypes.POINTER(FmLaunch), _vp]
        L.fm_plain_path_bytes.restype = ctypes.c_size_t
        L.fm_keep_bytes.restype = ctypes.c_size_t
        L.fm_keep_bytes.argtypes = [_i32, ctypes.POINTER(FmEdge), _i32, _i32]
        L.fm_plain_path_bytes.argtypes = [ctypes.POINTER(FmEdge), _i32]
        L.fm_chi_paths_plain.argtypes = [ctypes.POINTER(FmEdge), _i32, ctypes.POINTER(FmPaths), ctypes.POINTER(FmGrid),
                                         _i32, _vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                         ctypes.POINTER(_i32), ctypes.POINTER(_i32), _i32, _vp, _vp, _vp, _vp, _vp, _vp,
                                         ctypes.c_size_t, _vp]
        L.fm_r_factor.argtypes = [_i32, _i32, ctypes.c_double, _i32, ctypes.POINTER(FmWindow), _i32, _i32, _vp,
                                  ctypes.c_double, _vp, _i32, _vp, _vp, _vp]
        L.fm_pair_histogram.argtypes = [_i32, _i32, _vp, _i32, _vp, _vp, _i32, ctypes.POINTER(FmPairGrid), _vp, _vp,
                                        _vp]
        L.fm_pair_histogram_grad.argtypes = [_i32, _i32, _vp, _vp, _i32, ctypes.POINTER(FmPairGrid), _vp, _vp, _vp]
        L.fm_set_lds_budget.restype = _i32
        L.fm_set_lds_budget.argtypes = [_i32, _i32]
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != FM_OK:
        msg = lib().fm_last_error().decode()
        if rc == FM_EUNSUPPORTED:
            raise NotImplementedError(f"{what}: {msg}")
        if rc == FM_EINVAL:
            raise ValueError(f"{what}: {msg}")
        raise NativeError(f"{what}: {msg}")


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def backward_workspace(precision: int, nlegmax: int, npath: int, device, E: int, max_slots: int = 4096,
                       max_region_bytes: int = 1 << 30):
    """A device buffer for the adjoints: band slots for the workgroups in flight (paths of 7-9 legs: a workgroup
    takes one per band of its plan, at most 16), at most max_slots and no more than npath workgroups need; and, for
    the fused adjoint (2-6 legs), a region of (4 nleg + 1) E complex per path of a launch plus one spare, for up to
    npath paths and at most max_region_bytes (larger buckets run in chunks)."""
    slots = max(16, min(int(max_slots), 16 * max(int(npath), 1)))
    nbytes = lib().fm_backward_workspace_bytes(int(precision), int(nlegmax), slots)
    per = (4 * min(int(nlegmax), 6) + 1) * int(E) * (8 if precision == 32 else 16)
    nbytes = max(nbytes, min((int(npath) + 1) * per, int(max_region_bytes)))
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)


def stream_of(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class DeviceEdge:
    """The tables of one edge uploaded to a device, with the fm_edge struct that points at them."""

    def __init__(self, T: _tables.EdgeTables, device):
        dev = torch.device(device)
        self.host = T

        def put(a, dtype):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev).contiguous()

        def cplx(a):
            a = np.asarray(a)
            return np.stack([a.real, a.imag], -1)

        self.lmax = put(T.lmax, torch.int32)
        self.ck = put(cplx(T.ck), torch.float64)
        self.ckmxk = put(cplx(T.ckmxk), torch.float64)
        self.kf = put(T.kf, torch.float64)
        tl = cplx(np.transpose(T.tl, (0, 1, 2)))          # (nph1, L1+1, E, 2)
        self.tl64 = put(tl, torch.float64)
        self.tl32 = put(tl, torch.float32)
        chw = cplx(T.chw)
        self.chw64 = put(chw, torch.float64)
        self.chw32 = put(chw, torch.float32)
        self.rotc = put(T.rotc, torch.float64)
        self.rotn = put(T.rotn, torch.float64)
        s = FmEdge()
        s.E, s.L1, s.nph1, s.ilinit, s.nchan = T.E, T.L1, T.nph1, T.ilinit, len(T.chan_il)
        if T.nph1 > FM_MAX_POT:
            raise NotImplementedError(f"more than {FM_MAX_POT} potentials")
        for i, v in enumerate(T.chan_il):
            s.chan_il[i] = v
        for i, v in enumerate(T.lmax1):
            s.lmax1[i] = int(v)
        s.lmax, s.ck, s.ckmxk, s.kf = (t.data_ptr() for t in (self.lmax, self.ck, self.ckmxk, self.kf))
        s.tl32, s.tl64, s.chw32, s.chw64 = (t.data_ptr() for t in (self.tl32, self.tl64, self.chw32, self.chw64))
        s.rotc, s.rotn = self.rotc.data_ptr(), self.rotn.data_ptr()
        for g, (gp, starts, waves) in enumerate(T.bands):
            bp = s.bands[g]
            bp.nband = len(starts) - 1 if starts else 0
            for i, v in enumerate(starts):
                bp.start[i] = int(v)
            for i, v in enumerate(waves):
                bp.wave[i] = int(v)
        self.struct = s
