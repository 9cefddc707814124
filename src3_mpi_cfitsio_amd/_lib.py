"""ctypes binding of libmcrt.so (include/mcrt.h).

The library is loaded from the package directory; there is no fallback:
if it is missing or cannot find a GPU, calls raise.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcrt.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p

MCRT_FP64 = 0
MCRT_MIXED = 1
MCRT_T_REFERENCE = 0
MCRT_T_PERCELL = 1

SIGNATURES = {
    "mcrt_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "mcrt_destroy": (C.c_int, [_vp]),
    "mcrt_last_error": (C.c_char_p, [_vp]),
    "mcrt_set_stream": (C.c_int, [_vp, _vp]),
    "mcrt_synchronize": (C.c_int, [_vp]),
    "mcrt_set_grid": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32] + [_dp] * 9),
    "mcrt_set_medium": (C.c_int, [_vp, _dp, _u8p, _ip]),
    "mcrt_set_dust": (C.c_int, [_vp, C.c_int32, C.c_int32] + [_dp] * 6 + [C.c_int32, C.c_int32] + [_dp] * 5),
    "mcrt_setemiss": (C.c_int, [_vp, C.c_int32, C.c_int32, _ip, _dp, _dp, _dp, C.c_int32, _dp, C.c_int32, _dp, _ip,
                                _dp, _dp, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _fp]),
    "mcrt_setemiss_noalloc": (C.c_int, [_vp, C.c_int32, C.c_int32, _ip, _vp, _vp, _vp, C.c_int32, _vp, C.c_int32,
                                        _vp, _ip, _vp, _vp, _vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mcrt_build_models": (C.c_int, [_vp, C.c_int32, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _u8p, _dp,
                                    C.POINTER(C.c_int64), _fp]),
    "mcrt_set_isrf": (C.c_int, [_vp, C.c_int32, _dp, _dp, _dp, C.c_double]),
    "mcrt_set_spots": (C.c_int, [_vp, C.c_int32] + [C.c_double] * 5),
    "mcrt_set_accretion": (C.c_int, [_vp, _dp, _dp, C.c_double]),
    "mcrt_set_diffusion": (C.c_int, [_vp, C.c_int32, _dp, _dp, C.c_double, C.c_int32]),
    "mcrt_setdiffus": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int64)]),
    "mcrt_get_diffusion": (C.c_int, [_vp, _ip, C.POINTER(C.c_uint32)]),
    "mcrt_set_source": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double]),
    "mcrt_set_observer": (C.c_int, [_vp, C.c_int32, C.c_int32, _dp, _dp, C.c_int32, C.c_double,
                                    C.c_int32, C.c_int32, _ip, _dp, _dp,
                                    C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp, _dp,
                                    C.c_double, C.c_double]),
    "mcrt_set_tdust": (C.c_int, [_vp, _dp]),
    "mcrt_zero_tallies": (C.c_int, [_vp]),
    "mcrt_run_packets": (C.c_int, [_vp, C.c_int64, C.c_int64, C.c_uint64, C.c_int32]),
    "mcrt_set_launch": (C.c_int, [_vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "mcrt_set_peel_gather": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "mcrt_get_tallies": (C.c_int, [_vp, _dp, _ip, _ip, _ip, _fp, _fp, _fp, _dp]),
    "mcrt_dusttemp": (C.c_int, [_vp, C.c_double, C.c_int32, C.c_int32, _dp]),
    "mcrt_dusttemp_sigma": (C.c_int, [_vp, _dp, _dp, C.POINTER(C.c_int64), _ip]),
    "mcrt_trace_rays": (C.c_int, [_vp, C.c_int32, C.c_int32, _dp, _dp, _ip, _u8p, _dp, C.c_int32,
                                  _ip, _ip, _dp, _ip, _dp, _ip, _ip]),
    "mcrt_mueller": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp]),
    "mcrt_enable_trace": (C.c_int, [_vp, C.c_int64]),
    "mcrt_get_trace": (C.c_int, [_vp, _dp, C.c_int64]),
    "mcrt_last_run_ms": (C.c_int, [_vp, _fp]),
    "mcrt_run_stats": (C.c_int, [_vp, _dp]),
    "mcrt_adjoint_image": (C.c_int, [_vp, C.c_int64, C.c_int64, C.c_uint64, C.c_int32, _fp, _dp, _dp]),
    "mcrt_set_batch": (C.c_int, [_vp, C.c_int32, _ip, _ip] + [_dp] * 10 + [_u8p, _ip]),
    "mcrt_set_batch_tdust": (C.c_int, [_vp, _dp]),
    "mcrt_set_batch_launch": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "mcrt_run_batch": (C.c_int, [_vp, C.POINTER(C.c_int64), C.c_uint64, C.c_int32]),
    "mcrt_get_batch_tallies": (C.c_int, [_vp, _dp, _ip, _ip, _ip, _fp, _dp, _dp]),
    "mcrt_batch_dusttemp": (C.c_int, [_vp, C.c_int32, C.c_int32, _dp]),
    "mcrt_batch_autotune": (C.c_int, [_vp, C.c_int64, C.c_uint64, C.c_int32, _ip, _dp, C.c_int32, _ip]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libmcrt.so and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m src3_mpi_cfitsio_amd.build` "
                           "(the engine has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None
