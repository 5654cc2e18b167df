"""ctypes binding of liblm.so (declared in include/lm.h).

torch is imported first: torch's bundled libamdhip64.so carries the SONAME
libamdhip64.so.7 that liblm.so links against, so loading torch first makes
liblm share torch's HIP runtime (one runtime, one set of streams).

Every wrapper checks shapes, dtypes and devices on the host before a launch
and raises RuntimeError with lm_last_error() on a non-zero status.  There is no
CPU fallback: without the built library the calls raise.
"""
import ctypes
import os

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblm.so")

c_i32, c_i64, c_u32, c_u64, c_dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double
c_vp = ctypes.c_void_p


class LmGrid(ctypes.Structure):
    _fields_ = [("nx", c_i32), ("ny", c_i32), ("x0", c_dbl), ("y0", c_dbl), ("inv_dx", c_dbl), ("inv_dy", c_dbl)]


# name -> (restype, argtypes); mirrors include/lm.h
_SIGNATURES = {
    "lm_abi_version": (ctypes.c_int, []),
    "lm_last_error": (ctypes.c_char_p, []),
    "lm_field_blend_f64": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "lm_advect_rk4_f64": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp, c_i64, ctypes.POINTER(LmGrid), c_i32, c_dbl,
                                         c_vp, c_vp, c_i64, c_dbl, c_u64, c_u64, c_u64, c_vp, c_vp]),
    "lm_field_blend_f32": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "lm_advect_rk4_f32": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp, c_i64, ctypes.POINTER(LmGrid), c_i32, c_dbl,
                                         c_vp, c_vp, c_i64, c_dbl, c_u64, c_u64, c_u64, c_vp, c_vp]),
    "lm_pairs_ws_create": (ctypes.c_int, [c_i64, c_i32, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl,
                                          ctypes.POINTER(c_vp)]),
    "lm_pairs_ws_create_ex": (ctypes.c_int, [c_i64, c_i32, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl,
                                             ctypes.POINTER(c_vp)]),
    "lm_pairs_ws_destroy": (None, [c_vp]),
    "lm_pairs_ws_bytes": (c_i64, [c_vp]),
    "lm_pairs_ws_buckets": (c_i64, [c_vp]),
    "lm_query_pairs": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_dbl, c_dbl, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "lm_query_pairs_batch": (ctypes.c_int, [c_vp, c_i32, c_i64, c_vp, c_vp, c_i64, c_dbl, c_dbl, c_vp, c_vp,
                                            c_i64, c_vp, c_vp]),
    "lm_query_pairs_runs": (ctypes.c_int, [c_vp, c_i32, c_i64, c_vp, c_vp, c_i64, c_vp, c_dbl, c_dbl, c_vp, c_vp,
                                           c_i64, c_vp, c_vp]),
    "lm_rps_ws_create": (ctypes.c_int, [c_i64, c_i32, ctypes.POINTER(c_vp)]),
    "lm_rps_ws_destroy": (None, [c_vp]),
    "lm_rps_ws_bytes": (c_i64, [c_vp]),
    "lm_rps_update": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_dbl, c_dbl, c_dbl,
                                     c_u64, c_u64, c_u64, c_vp]),
    "lm_rps_prepare": (ctypes.c_int, [c_vp, c_i32, c_i64, c_vp, c_vp, c_i64, c_vp, c_dbl, c_dbl, c_dbl,
                                      c_u64, c_u64, c_u64, c_vp]),
    "lm_rps_prepare_runs": (ctypes.c_int, [c_vp, c_i32, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_i32, c_vp, c_vp,
                                           c_u64, c_vp]),
    "lm_rps_play": (ctypes.c_int, [c_vp, c_i32, c_i64, c_vp, c_vp, c_vp]),
    "lm_rps_ws_set_play_blocks": (ctypes.c_int, [c_vp, c_i32, c_i32]),
    "lm_rps_ws_reserve": (ctypes.c_int, [c_vp, c_i64, c_vp]),
    "lm_rps_component_stats": (ctypes.c_int, [c_vp, c_i32, c_vp, c_vp]),
    "lm_philox_uniforms": (ctypes.c_int, [c_i64, c_u64, c_u64, c_u64, c_u64, c_u64, c_u64, c_i32, c_vp, c_vp]),
    "lm_species_counts": (ctypes.c_int, [c_vp, c_i64, c_i32, c_i64, c_vp, c_vp]),
    "lm_distance_histogram": (ctypes.c_int, [c_i64, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "lm_adjoint_forward": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp, c_i32, ctypes.POINTER(LmGrid), c_vp, c_vp, c_vp,
                                          c_i32, c_dbl, c_i32, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "lm_adjoint_forward_noalloc": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp, c_i32, ctypes.POINTER(LmGrid), c_vp,
                                                  c_vp, c_vp, c_i32, c_dbl, c_i32, c_vp, c_vp, c_vp]),
    "lm_adjoint_scratch_bytes": (c_i64, [c_i64, c_i32, ctypes.POINTER(LmGrid), c_i32]),
    "lm_adjoint_backward": (ctypes.c_int, [c_i64, c_vp, c_vp, c_i32, ctypes.POINTER(LmGrid), c_vp, c_vp, c_vp,
                                           c_i32, c_dbl, c_i32, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "lm_copy_bytes": (ctypes.c_int, [c_vp, c_vp, c_i64, c_vp]),
    "lm_sort_pairs_u32": (ctypes.c_int, [c_i32, c_vp, c_i64, c_i64, c_i64, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "lm_timing_enable": (ctypes.c_int, [ctypes.c_int]),
    "lm_timing_summary": (ctypes.c_int, [c_vp, c_vp, c_i32]),
    "lm_timing_kernel_name": (ctypes.c_char_p, [c_i32]),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load(path=LIB_PATH):
    """Load liblm.so and declare its signatures (no GPU needed)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(
                f"liblm.so not found at {path}: build it with `python -m lagrangian_microbes_amd.build` "
                "(the HIP path has no CPU fallback)")
        L = ctypes.CDLL(path)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.lm_abi_version() != 6:
            raise RuntimeError("liblm.so ABI version mismatch; rebuild it")
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().lm_last_error()
        raise RuntimeError(f"liblm error {rc}: {msg.decode() if msg else '?'}")


def stream_ptr(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return c_vp(s.cuda_stream)


def ptr(t):
    return c_vp(t.data_ptr()) if t is not None else c_vp(0)


def require(t, dtype, numel=None, name="tensor", min_numel=None):
    """Host-side validation before a pointer reaches a kernel."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor (got {t.device})")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements (got {t.numel()})")
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError(f"{name} must have at least {min_numel} elements (got {t.numel()})")
    return t
