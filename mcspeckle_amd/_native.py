"""ctypes binding of libscmc.so (include/scmc.h).

The library is built in-tree by ``mcspeckle_amd.build`` (hipcc, gfx950).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libscmc.so")

SCMC_OK = 0
SCMC_E_ARG = -1
SCMC_E_DEVICE = -2
SCMC_E_NOMEM = -3
SCMC_E_UNSUPPORTED = -4
SCMC_E_CANCELLED = -5
SCMC_E_BUSY = -6
SCMC_E_NOTFOUND = -7

FP64 = 0
MIXED = 1
PRECISIONS = {"fp64": FP64, "mixed": MIXED}

EXPORTS = [
    "scmc_mccov", "scmc_mcfield", "scmc_mcfield_noalloc", "scmc_mcfield_fourier_quad", "scmc_quad_grid",
    "scmc_mcfield_wave", "scmc_mcfield_fourier_quad_wave", "scmc_quad_grid_wave", "scmc_mccov_hetro", "scmc_mcfield_hetro",
    "scmc_trace", "scmc_mccov_replay", "scmc_mcfield_replay",
    "scmc_session_create", "scmc_session_run", "scmc_session_result", "scmc_session_stats",
    "scmc_session_destroy",
    "scmc_server_start", "scmc_server_stop", "scmc_submit_mccov", "scmc_submit_mcfield", "scmc_poll",
    "scmc_wait", "scmc_cancel", "scmc_request_footprint",
    "scmc_last_error", "scmc_version", "scmc_device_count",
]

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class Phase(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("g", C.c_double),
        ("n_amp", C.c_int32), ("amp_re", _dp), ("amp_im", _dp),
        ("n_icdf", C.c_int32), ("icdf", _dp),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("sigt", C.c_double), ("albedo", C.c_double), ("lam", C.c_double),
        ("box_min", C.c_double * 3), ("box_max", C.c_double * 3),
        ("nl", C.c_int32), ("nv", C.c_int32), ("l", _dp), ("v", _dp),
        ("is_ff_l", C.c_int32), ("is_ff_v", C.c_int32), ("do_cbs", C.c_int32), ("smp_flg", C.c_int32),
        ("amp", Phase), ("has_amp0", C.c_int32), ("amp0", Phase),
        ("has_lmean0", C.c_int32), ("lmean0", C.c_double * 3),
        ("seed", C.c_uint64), ("path_offset", C.c_uint64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("paths", C.c_int64), ("vertices", C.c_int64), ("chunks", C.c_int64),
        ("ms_trace", C.c_double), ("ms_nee", C.c_double), ("ms_accum", C.c_double),
        ("ms_total", C.c_double), ("fp16_scale_log2", C.c_double),
        ("ms_herk", C.c_double), ("herk_launches", C.c_int64), ("herk_flops", C.c_double),
        ("ms_expand", C.c_double), ("nonfinite", C.c_int64), ("ms_queued", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Medium(C.Structure):
    _fields_ = [("nb", C.c_int32 * 3), ("box_bin", C.c_double * 3), ("sigt", _dp), ("albedo", _dp),
                ("has_lrad", C.c_int32), ("lrad", C.c_double)]


class PathSet(C.Structure):
    _fields_ = [
        ("n_paths", C.c_int64), ("n_vertices", C.c_int64), ("offset", _i64p),
        ("x1", _dp), ("px", _dp), ("w1", _dp), ("w0p_re", _dp), ("w0p_im", _dp),
        ("x", _dp), ("ow", _dp), ("weight", _dp), ("xi", _dp),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("max_active", C.c_int32),
        ("hbm_budget_gb", C.c_double), ("paths_per_slice", C.c_int64),
        ("shared_launches", C.c_int32), ("resident_loop", C.c_int32), ("resident_workgroups", C.c_int32),
    ]


class ScmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    """Load libscmc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m mcspeckle_amd.build` "
            "(hipcc, gfx950). The speckle engines have no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    P, S = C.POINTER(Problem), C.POINTER(Stats)
    sig = {
        "scmc_mccov": ([P, C.c_int64, C.c_int32, _dp, _dp, _dp, _dp, S], C.c_int),
        "scmc_mcfield": ([P, C.c_int64, C.c_int32, _dp, _dp, S], C.c_int),
        "scmc_mcfield_noalloc": ([P, C.c_int64, C.c_int32, _dp, _dp, S], C.c_int),
        "scmc_mcfield_fourier_quad": ([P, C.c_double, C.c_double, C.c_int64, C.c_double, C.c_int32, _dp, _dp, S], C.c_int),
        "scmc_quad_grid": ([_dp, _dp, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32),
                            C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "scmc_mcfield_wave": ([P, _dp, _dp, C.c_int64, C.c_int32, _dp, _dp, S], C.c_int),
        "scmc_mcfield_fourier_quad_wave": ([P, _dp, _dp, C.c_double, C.c_double, C.c_int64, C.c_double, C.c_int32,
                                            _dp, _dp, S], C.c_int),
        "scmc_quad_grid_wave": ([_dp, _dp, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "scmc_mccov_hetro": ([P, C.POINTER(Medium), C.c_int64, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, S], C.c_int),
        "scmc_mcfield_hetro": ([P, C.POINTER(Medium), C.c_int64, C.c_int32, _dp, _dp, S], C.c_int),
        "scmc_trace": ([P, C.c_int64, _i64p, _i64p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "scmc_mccov_replay": ([P, C.POINTER(PathSet), C.c_int64, C.c_int32, _dp, _dp, _dp, _dp, S], C.c_int),
        "scmc_mcfield_replay": ([P, C.POINTER(PathSet), C.c_int64, C.c_int32, _dp, _dp, S], C.c_int),
        "scmc_session_create": ([P, C.c_int32, C.c_int32, _i64p], C.c_int),
        "scmc_session_run": ([C.c_int64, C.c_int64, C.c_int64], C.c_int),
        "scmc_session_result": ([C.c_int64, C.c_int64, _dp, _dp, _dp, _dp], C.c_int),
        "scmc_session_stats": ([C.c_int64, S], C.c_int),
        "scmc_session_destroy": ([C.c_int64], C.c_int),
        "scmc_server_start": ([C.POINTER(ServerConfig)], C.c_int),
        "scmc_server_stop": ([], C.c_int),
        "scmc_submit_mccov": ([P, C.c_int64, C.c_int32, _i64p], C.c_int),
        "scmc_submit_mcfield": ([P, C.c_int64, C.c_int32, _i64p], C.c_int),
        "scmc_poll": ([C.c_int64, _dp, _dp, _dp, _dp, S], C.c_int),
        "scmc_wait": ([C.c_int64, _dp, _dp, _dp, _dp, S], C.c_int),
        "scmc_cancel": ([C.c_int64], C.c_int),
        "scmc_request_footprint": ([C.c_int32, C.c_int32, C.c_int32], C.c_int64),
        "scmc_last_error": ([], C.c_char_p),
        "scmc_version": ([], C.c_char_p),
        "scmc_device_count": ([], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc):
    if rc != SCMC_OK:
        msg = lib().scmc_last_error().decode(errors="replace")
        raise ScmcError(rc, msg)
    return rc


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_i64p)
