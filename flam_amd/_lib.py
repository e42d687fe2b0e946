"""ctypes binding of librskx.so (the C ABI declared in include/rskx.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback for the factorization or the sweeps.
"""
import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_int, c_int64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librskx.so")

RSKX_OK = 0
RSKX_ERR_ARG = -1
RSKX_ERR_HIP = -2
RSKX_ERR_NUM = -3
RSKX_ERR_UNSUP = -4
RSKX_ERR_MEM = -5
RSKX_PREC_FP64 = 0
RSKX_PREC_FP32 = 1
RSKX_F64 = 0
RSKX_F32 = 1
RSKX_Z64 = 2
RSKX_Z32 = 3
RSKX_KERNEL_LAPLACE2D_DL = 1
RSKX_KERNEL_LAPLACE3D_SL = 2
RSKX_KERNEL_HELMHOLTZ2D_CFIE = 3
RSKX_KERNEL_CALLBACK = 4
RSKX_PROXY_NONE = 0
RSKX_PROXY_RING_SL = 1
RSKX_PROXY_RING_DL = 2
RSKX_PROXY_POINTS_SL = 3
RSKX_PROXY_RING_CFIE = 4
RSKX_PROXY_CALLBACK = 5

# host callbacks of RSKX_KERNEL_CALLBACK / RSKX_PROXY_CALLBACK
ENTRIES_FN = ctypes.CFUNCTYPE(c_int, c_void_p, POINTER(c_int64), c_int64, POINTER(c_int64), c_int64,
                              POINTER(c_double))
PROXY_FN = ctypes.CFUNCTYPE(c_int, c_void_p, POINTER(c_int64), c_int64, POINTER(c_int64),
                            POINTER(c_int64), POINTER(c_double), POINTER(c_double),
                            POINTER(c_double), c_int64, POINTER(c_int64))


class Kernel(ctypes.Structure):
    _fields_ = [("kind", c_int), ("nu", c_void_p), ("w", c_void_p), ("diag", c_void_p),
                ("scale", c_double), ("wavenumber", c_double), ("eta", c_double),
                ("entries", ENTRIES_FN), ("ctx", c_void_p)]


class Proxy(ctypes.Structure):
    _fields_ = [("kind", c_int), ("p", c_int), ("radius", c_double), ("weight", c_double),
                ("pts", c_void_p), ("fn", PROXY_FN), ("ctx", c_void_p)]


class Opts(ctypes.Structure):
    _fields_ = [("lvlmax", c_int), ("ext", POINTER(c_double)), ("Tmax", c_double),
                ("rrqr_iter", c_double), ("symm", c_char), ("verb", c_int),
                ("precision", c_int), ("profile", c_int), ("stop", c_int), ("id_gram", c_int),
                ("rank_model", c_void_p), ("rank_scale", c_double), ("record_boxes", c_int)]


class Info(ctypes.Structure):
    _fields_ = [("N", c_int64), ("d", c_int), ("nlvl", c_int), ("nlvl_done", c_int),
                ("nfactor", c_int64), ("bytes", c_int64), ("precision", c_int),
                ("symm", c_char), ("id_scratch", c_int64), ("rank_fallbacks", c_int64)]


class FactorView(ctypes.Structure):
    _fields_ = [("k", c_int64), ("r", c_int64),
                ("sk", POINTER(c_int64)), ("rd", POINTER(c_int64)),
                ("T", POINTER(c_double)), ("L", POINTER(c_double)), ("U", POINTER(c_double)),
                ("p", POINTER(c_int64)), ("E", POINTER(c_double)), ("F", POINTER(c_double))]


# every entry point of include/rskx.h with its ctypes signature
SIGNATURES = {
    "rskx_opts_default": (c_int, [POINTER(Opts)]),
    "rskx_rskelf": (c_int, [c_void_p, c_int, c_int64, c_int, c_double, POINTER(Kernel),
                            POINTER(Proxy), POINTER(Opts), POINTER(c_void_p)]),
    "rskx_rskelf_batch": (c_int, [c_int, POINTER(c_int64), c_void_p, c_int, c_int, c_double,
                                  POINTER(Kernel), POINTER(Proxy), POINTER(Opts),
                                  POINTER(c_void_p)]),
    "rskx_rskelf_sv": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int]),
    "rskx_rskelf_mv": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int]),
    "rskx_rskelf_sv_noalloc": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int]),
    "rskx_rskelf_logdet": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    "rskx_rskelf_cholmv": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int]),
    "rskx_rskelf_cholsv": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int]),
    "rskx_sweep_sv": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int, c_int]),
    "rskx_sweep_mv": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int, c_int]),
    "rskx_partial_info": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64),
                                  POINTER(c_int64), POINTER(c_int64), POINTER(c_double)]),
    "rskx_free": (None, [c_void_p]),
    "rskx_release_cache": (c_int, []),
    "rskx_last_error": (c_char_p, []),
    "rskx_set_stream": (c_int, [c_void_p]),
    "rskx_factor_info": (c_int, [c_void_p, POINTER(Info)]),
    "rskx_level_table": (c_int, [c_void_p, POINTER(c_int64), c_int]),
    "rskx_factor_lvp": (c_int, [c_void_p, POINTER(c_int64), c_int]),
    "rskx_export_factor": (c_int, [c_void_p, c_int64, POINTER(FactorView)]),
    "rskx_kernel_times": (c_int, [c_void_p, POINTER(c_char_p), POINTER(c_double),
                                  POINTER(c_int64), POINTER(c_int64), POINTER(c_double), c_int]),
    "rskx_kernel_matvec": (c_int, [c_void_p, c_int, c_int64, POINTER(Kernel), c_void_p, c_int64,
                                   c_void_p, c_int64, c_int]),
    "rskx_sv_graph": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_char, c_int, POINTER(c_void_p)]),
    "rskx_graph_launch": (c_int, [c_void_p]),
    "rskx_graph_free": (None, [c_void_p]),
    "rskx_set_tuning": (c_int, [c_char_p, c_double]),
    "rskx_get_tuning": (c_int, [c_char_p, POINTER(c_double)]),
    "rskx_import": (c_int, [c_int64, c_char, c_int, c_int, POINTER(c_int64), c_void_p,
                            POINTER(c_void_p)]),
    "rskx_rank_model_create": (c_int, [ctypes.c_uint64, POINTER(c_void_p)]),
    "rskx_rank_model_free": (None, [c_void_p]),
    "rskx_rank_model_nparams": (c_int, []),
    "rskx_rank_model_params": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rskx_rank_model_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                     POINTER(c_double)]),
    "rskx_rank_model_train": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int64,
                                      c_double, POINTER(c_double)]),
    "rskx_rank_model_predict": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "rskx_box_features": (c_int, [c_void_p, POINTER(c_int64), c_void_p, c_void_p]),
    "rskx_cgs2": (c_int, [c_void_p, c_int64, c_int64, c_int, c_void_p, c_int64, c_int64, c_int,
                          c_void_p]),
    "rskx_kernel_matvec_batch": (c_int, [c_int, POINTER(c_int64), c_void_p, c_int,
                                         POINTER(Kernel), c_void_p, c_int64, c_void_p, c_int64,
                                         c_int]),
}

_lib = None


def lib():
    """The loaded library; raises if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -m flam_amd.build or "
                "__graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    msg = lib().rskx_last_error()
    return msg.decode() if msg else ""
