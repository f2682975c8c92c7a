"""ctypes binding of libisle_hfa.so (the C ABI declared in include/isle_hfa.h).

The library is built in-tree by ``__graft_entry__.build()``.  Nothing here falls
back to the CPU: a missing library or a missing GPU raises immediately.
"""
import ctypes
import pathlib

import torch

LIB_PATH = pathlib.Path(__file__).resolve().with_name("libisle_hfa.so")

_c_int, _c_long, _c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); one entry per declaration in include/isle_hfa.h
SIGNATURES = {
    "isle_last_error": (ctypes.c_char_p, []),
    "isle_hfm_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _dp, _c_double, _c_int]),
    "isle_hfm_destroy": (_c_int, [_vp]),
    "isle_hfm_set_stream": (_c_int, [_vp, _vp]),
    "isle_hfm_set_recompute": (_c_int, [_vp, _c_int]),
    "isle_hfm_set_tile": (_c_int, [_vp, _c_int, _c_int]),
    "isle_hfm_sync": (_c_int, [_vp]),
    "isle_hfm_slice_matrix": (_c_int, [_vp, _c_int, _c_int, _dp]),
    "isle_hfm_F": (_c_int, [_vp, _c_int, _vp, _c_int, _c_int, _c_int, _vp]),
    "isle_logdetM": (_c_int, [_vp, _vp, _c_int, _c_int, _dp]),
    "isle_logdetM_batch": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _dp]),
    "isle_force_parts": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp]),
    "isle_force_parts_multi": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _c_int, _vp, _c_int, _c_int, _vp]),
    "isle_solveM": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp]),
    "isle_solveM_ex": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int]),
    "isle_applyM": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_long, _vp]),
    "isle_qlu": (_c_int, [_vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "isle_solveq": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "isle_combine_force": (_c_int, [_vp, _vp, _vp, _c_long, _c_int, _c_int, _vp]),
    "isle_combine_force_batch": (_c_int, [_vp, _vp, _vp, _c_long, _c_int, _c_long, _c_long, _c_int, _c_int, _vp]),
    "isle_gauge_force": (_c_int, [_vp, _vp, _c_long, _c_double, _c_int, _vp]),
    "isle_field_sums": (_c_int, [_vp, _c_long, _c_long, _c_int, _dp, _vp]),
    "isle_leapfrog_kick": (_c_int, [_vp, _vp, _vp, _c_long, _c_double, _c_double, _vp]),
    "isle_leapfrog_drift": (_c_int, [_vp, _vp, _c_long, _c_double, _vp]),
    "isle_cscale": (_c_int, [_vp, _vp, _c_long, _c_double, _c_double, _vp]),
    "isle_logdet": (_c_int, [_vp, _c_int, _dp, _vp]),
    "isle_logdet_noalloc": (_c_int, [_vp, _c_int, _dp, _vp]),
    "isle_profile_kernels": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _dp,
                                      _c_int]),
    "isle_invert": (_c_int, [_vp, _c_int, _vp, _vp]),
    "isle_invert_logdet": (_c_int, [_vp, _c_int, _vp, _dp, _vp]),
}

_ERRORS = {1: ValueError, 2: IndexError, 3: RuntimeError, 4: RuntimeError}
_lib = None


def load():
    """Load the shared library (no GPU needed for loading)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH.name} is not built; run __graft_entry__.build() first")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code):
    if code != 0:
        msg = load().isle_last_error().decode(errors="replace")
        raise _ERRORS.get(code, RuntimeError)(msg)


def device():
    """The HIP device all computations run on; raises if there is none."""
    if not torch.cuda.is_available():
        raise RuntimeError("isle_amd needs an AMD GPU (gfx950); torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def dptr(a):
    """Pointer to a contiguous float64 numpy array."""
    return a.ctypes.data_as(_dp)
