"""ctypes binding of libnadvibs.so (include/nadvibs.h).

The library is loaded from the package directory; a missing library is an
error (there is no CPU fallback for the product path).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnadvibs.so")

NV_OK = 0
NV_WARN_NONORTHOGONAL = 6
NV_QUIRK_Q3ELEM = 1
NV_QUIRK_LOWERDIAG = 2
NV_QUIRKS_REF = 3
NV_QUIRKS_EXACT = 0
NV_PREC_F64 = 0
NV_PREC_F32 = 1
NV_PREC_BF16 = 2
PREC = {"f64": NV_PREC_F64, "f32": NV_PREC_F32, "bf16": NV_PREC_BF16}

_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)


class NvProblem(ctypes.Structure):
    _fields_ = [
        ("nstates", ctypes.c_int64), ("nmodes", ctypes.c_int64), ("ordr", ctypes.c_int64),
        ("nfunc", _i64p), ("aomega", _dp), ("concoef", _dp), ("potterms", _dp),
        ("ztoler", ctypes.c_double), ("quirks", ctypes.c_int32), ("neworigin", ctypes.c_int32),
        ("statew", _dp), ("bomega", _dp), ("dvec", _dp), ("tmat", _dp),
        ("nirreps", ctypes.c_int64), ("npirrep", _i64p), ("istate", _i64p),
    ]


class NvLanczosOpts(ctypes.Structure):
    _fields_ = [
        ("niter", ctypes.c_int64), ("reorthog", ctypes.c_int32), ("prec", ctypes.c_int32),
        ("chkorthog", ctypes.c_int64), ("seed", ctypes.c_uint64), ("noise_scale", ctypes.c_double),
        ("pro_eps", ctypes.c_double), ("savevecs", ctypes.c_int32), ("checkpoint", ctypes.c_int64),
        ("pro_tau", ctypes.c_double), ("pro_eta", ctypes.c_double),
    ]


# (name, restype, argtypes) for every entry point of include/nadvibs.h
SIGNATURES = [
    ("nv_create", ctypes.c_int32, [ctypes.POINTER(NvProblem), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    ("nv_destroy", ctypes.c_int32, [ctypes.c_void_p]),
    ("nv_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("nv_dims", ctypes.c_int32, [ctypes.c_void_p, _i64p, _i64p]),
    ("nv_term_summary", ctypes.c_int32, [ctypes.c_void_p, _i64p]),
    ("nv_term_counts", ctypes.c_int32, [ctypes.c_void_p, _i64p]),
    ("nv_hv", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, _dp, _dp]),
    ("nv_hv_device", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    ("nv_start_vector", ctypes.c_int32, [ctypes.c_void_p, _dp]),
    ("nv_lanczos_init", ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(NvLanczosOpts)]),
    ("nv_update_lanczos", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64]),
    ("nv_hv_iter", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64]),
    ("nv_lanczos_step", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64]),
    ("nv_set_omega", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64]),
    ("nv_partial_orthog", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64]),
    ("nv_lanczos_run", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]),
    ("nv_get_coefficients", ctypes.c_int32, [ctypes.c_void_p, _dp, _dp, _dp]),
    ("nv_pro_stats", ctypes.c_int32, [ctypes.c_void_p, _i64p, _i64p]),
    ("nv_get_lanczos_vector", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp]),
    ("nv_ritz", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _i32p]),
    ("nv_lanczos_adjoint", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, ctypes.c_int32, _dp, _dp,
                                            _dp]),
    ("nv_spectrum", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, ctypes.c_int64, _dp, ctypes.c_double,
                                     ctypes.c_int32, ctypes.c_double, _dp]),
    ("nv_spectrum_adjoint", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, ctypes.c_int64, _dp,
                                             ctypes.c_double, ctypes.c_int32, ctypes.c_double, _dp, _dp, _dp, _dp]),
    ("nv_krylov_gradient", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, _dp, _dp]),
    ("nv_select_hv_kernel", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    ("nv_profile_hv", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp]),
    ("nv_hv_flops", ctypes.c_int32, [ctypes.c_void_p, _i64p]),
    ("nv_batch_create", ctypes.c_int32, [ctypes.POINTER(NvProblem), ctypes.c_int64, ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    ("nv_batch_destroy", ctypes.c_int32, [ctypes.c_void_p]),
    ("nv_batch_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("nv_batch_init", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32]),
    ("nv_batch_tune", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    ("nv_batch_run", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]),
    ("nv_batch_coefficients", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp]),
    ("nv_batch_context", ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int64]),
    ("nv_identify_roots", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, ctypes.c_int64, ctypes.c_int64,
                                           _dp, _dp, _i32p, _i64p, _i32p, _i32p, _dp, _dp, _dp, _dp]),
    ("nv_convolve", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, ctypes.c_int64, _dp,
                                     ctypes.c_double, _dp]),
    ("nv_convolve_noalloc", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int64, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    ("nv_save_vectors", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p]),
    ("nv_load_vectors", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64, _dp, _dp, _dp]),
    ("nv_save_archive", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]),
    ("nv_load_archive", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64]),
    ("nv_profile_enable", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    ("nv_profile_read", ctypes.c_int32, [ctypes.c_void_p, _dp, _i64p]),
    ("nv_hv_kernel_info", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64]),
]

_lib = None


def load():
    """Loads the library (once) and declares every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m nadvibs_amd.build` (no CPU fallback exists)")
    # PyTorch-ROCm bundles its own HIP runtime under the same soname: loaded first, it serves
    # both (device pointers and streams then cross the binding); loaded second, it would meet
    # the system runtime already in the process and find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nadvibs error {code}: {msg}")
        self.code = code


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_i64p) if a is not None else None
