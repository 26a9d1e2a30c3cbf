"""ctypes binding of libmtip.so (the C ABI declared in include/mtip.h).

The library is loaded from the package directory; if it is missing the
import of the GPU path fails loudly -- there is no CPU fallback.
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmtip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "mtip.h")

MTIP_OK = 0
MTIP_ERR_ARG = -1
MTIP_ERR_DOMAIN = -2
MTIP_ERR_HIP = -3
MTIP_ERR_NOMEM = -4
MTIP_ERR_STATE = -5
MTIP_ERR_UNSUPPORTED = -6

MTIP_FP32, MTIP_BF16, MTIP_FP8 = 0, 1, 2
PRECISIONS = {"fp32": MTIP_FP32, "bf16": MTIP_BF16, "fp8": MTIP_FP8}

_c = ctypes
_vp = _c.c_void_p
_i64 = _c.c_int64
_i32 = _c.c_int
_f64 = _c.c_double

# name -> (restype, argtypes)
_SIGS = {
    "mtip_abi_version": (_i32, []),
    "mtip_last_error": (_c.c_char_p, []),
    "mtip_mem_info": (_i32, [_i32, _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)]),
    "mtip_plan_create": (_i32, [_i32, _f64, _i32, _c.POINTER(_vp)]),
    "mtip_plan_destroy": (_i32, [_vp]),
    "mtip_plan_info": (_i32, [_vp, _c.POINTER(_i32), _c.POINTER(_i32), _c.POINTER(_f64),
                              _c.POINTER(_i32)]),
    "mtip_plan_set_model": (_i32, [_vp, _vp, _vp]),
    "mtip_plan_execute": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "mtip_plan_slices": (_i32, [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _i32, _vp, _f64, _vp, _i64,
                                _vp, _vp, _vp, _vp, _vp]),
    "mtip_rotate_positions": (_i32, [_vp, _i64, _vp, _i64, _f64, _vp, _vp, _vp, _vp]),
    "mtip_plan_spread_begin": (_i32, [_vp, _vp]),
    "mtip_plan_spread_slices": (_i32, [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _i32, _f64, _vp,
                                       _i64, _vp, _vp]),
    "mtip_plan_spread_points": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "mtip_plan_spread_finish": (_i32, [_vp, _vp, _f64, _vp, _vp]),
    "mtip_phase_amplitudes": (_i32, [_vp, _vp, _i64, _vp]),
    "mtip_phase_update": (_i32, [_vp, _vp, _vp, _i64, _i32, _f64, _f64, _vp]),
    "mtip_clip_bin": (_i32, [_vp, _i32, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp]),
    "mtip_merge_pad": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _vp]),
    "mtip_merge_combine": (_i32, [_vp, _vp, _vp, _vp, _i32, _i32, _f64, _f64, _vp, _vp]),
    "mtip_dot_workspace_size": (_c.c_size_t, []),
    "mtip_dot": (_i32, [_vp, _vp, _i64, _vp, _c.c_size_t, _vp, _vp]),
    "mtip_cg_update": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "mtip_cg_direction": (_i32, [_vp, _vp, _vp, _vp, _i64, _i32, _vp]),
    "mtip_row_stats": (_i32, [_vp, _i64, _i64, _i64, _vp, _vp, _vp]),
    "mtip_scaling_ratio": (_i32, [_vp, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp]),
    "mtip_segment_scaling_ratio": (_i32, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp]),
    "mtip_quantize_rows": (_i32, [_vp, _i64, _i64, _i64, _i32, _vp, _i64, _vp, _vp]),
    "mtip_residual_rows": (_i32, [_vp, _i64, _i64, _i64, _vp, _i32, _vp, _i32, _vp, _i64, _vp,
                                  _vp, _vp]),
    "mtip_col_mean_workspace_size": (_c.c_size_t, [_i64]),
    "mtip_col_mean": (_i32, [_vp, _i64, _i64, _i64, _vp, _c.c_size_t, _vp, _vp]),
    "mtip_match_workspace_size": (_c.c_size_t, [_i64, _i64, _i32, _i32]),
    "mtip_match_topk": (_i32, [_i32, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _i64,
                               _vp, _i32, _vp, _c.c_size_t, _vp, _vp, _vp]),
    "mtip_group_workspace_size": (_c.c_size_t, [_i64, _i64]),
    "mtip_group_by_winner": (_i32, [_vp, _i64, _i32, _i64, _i64, _vp, _c.c_size_t, _vp]),
    "mtip_group_topk": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _i64,
                               _i32, _vp, _vp, _i32, _vp]),
    "mtip_group_topk_prec": (_i32, [_i32, _vp, _i32, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp,
                                    _vp, _i64, _vp, _i64, _i32, _vp, _vp, _i32, _vp]),
    "mtip_group_bad_count": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp]),
    "mtip_group_perm_offset": (_c.c_size_t, [_i64, _i64]),
    "mtip_merge_distinct": (_i32, [_vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp]),
    "mtip_unique_rows": (_i32, [_vp, _i64, _i32, _vp]),
    "mtip_set_gemm_tile": (_i32, [_i32, _i32]),
    "mtip_gemm_tile": (_i32, [_i32, _i64, _i64]),
    "mtip_match_partial": (_i32, [_i32, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _i64,
                                  _vp, _i32, _i32, _vp, _c.c_size_t, _vp]),
    "mtip_match_partial_ex": (_i32, [_i32, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64,
                                     _i64, _vp, _i32, _i32, _i32, _vp, _c.c_size_t, _vp]),
    "mtip_topk_merge": (_i32, [_i32, _vp, _i64, _i64, _i32, _i32, _vp, _vp, _vp]),
    "mtip_match_topk_static": (_i32, [_i32, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64,
                                      _i64, _vp, _i32, _vp, _vp, _vp]),
    "mtip_rescore": (_i32, [_vp, _i64, _i64, _vp, _i64, _i64, _i64, _vp, _i32, _vp, _i32, _vp,
                            _vp, _vp, _vp, _vp, _vp]),
    "mtip_euclidean_distance": (_i32, [_vp, _vp, _i64, _i64, _i64, _vp, _vp]),
    "mtip_heap_sort": (_i32, [_vp, _i64, _i64, _i32, _vp, _vp]),
}

_lib = None


class MtipError(RuntimeError):
    """A libmtip call failed (HIP error, allocation, call order)."""


def header_functions(path=HEADER_PATH):
    """Names of every function declared in include/mtip.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mtip_[a-z0-9_]+)\s*\(", text)))


def load(path=LIB_PATH):
    """Load libmtip.so and declare every entry point (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"libmtip.so not found at {path}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().mtip_last_error().decode(errors="replace")


def check(rc, what=""):
    """Raise the Python exception that matches a libmtip return code."""
    if rc == MTIP_OK:
        return
    msg = f"{what}: {last_error()} (code {rc})"
    if rc == MTIP_ERR_ARG:
        raise ValueError(msg)
    if rc == MTIP_ERR_DOMAIN:
        raise AssertionError(msg)
    if rc == MTIP_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == MTIP_ERR_NOMEM:
        raise MemoryError(msg)
    raise MtipError(msg)
