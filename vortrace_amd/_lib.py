"""Loader for libvortrace_hip.so (the C-ABI of include/vortrace_hip.h).

The library is built in-tree (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C vortrace_amd/csrc``).  There is no fallback: if the
library is missing, importing vortrace_amd raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os
import warnings

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvortrace_hip.so")

VT_OK, VT_EINVAL, VT_ERUNTIME = 0, 1, 2


class TreeInfo(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("leaves", C.c_int64),
                ("max_depth", C.c_int64), ("build_seconds", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("walk_batch", C.c_int), ("leaf_batch", C.c_int), ("ray_parts", C.c_int),
                ("help_events", C.c_int)]


class View(C.Structure):
    _fields_ = [("extent", C.c_double * 4), ("bounds", C.c_double * 2),
                ("center", C.c_double * 3), ("rot", C.c_double * 9),
                ("nx", C.c_int64), ("ny", C.c_int64), ("rotate", C.c_int32),
                ("pad_", C.c_int32)]


class CallStats(C.Structure):
    _fields_ = [("segments", C.c_int64), ("queries", C.c_int64),
                ("max_walk_depth", C.c_int64), ("max_knn_depth", C.c_int64),
                ("slow_queries", C.c_int64), ("deferred_rays", C.c_int64),
                ("kernel_ms", C.c_float), ("intervals", C.c_int64),
                ("intervals_back", C.c_int64)]


WARN_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_void_p)

# name -> (restype, argtypes); every symbol include/vortrace_hip.h declares.
P, S, I, D = C.c_void_p, C.c_size_t, C.c_int, C.c_double
PP = C.POINTER(C.c_void_p)
SIGNATURES = {
    "vt_last_error": (C.c_char_p, []),
    "vt_set_warning_callback": (None, [WARN_FN, P]),
    "vt_set_verbose": (None, [I]),
    "vt_get_verbose": (I, []),
    "vt_cloud_create": (I, [P, S, P, S, P, P, S, I, PP]),
    "vt_cloud_build_tree": (I, [P]),
    "vt_cloud_build_host_tree": (I, [P]),
    "vt_cloud_upload": (I, [P]),
    "vt_cloud_tree_export": (I, [P, P, P, P, P, P, P, P]),
    "vt_cloud_destroy": (None, [P]),
    "vt_cloud_npart": (S, [P]),
    "vt_cloud_nfields": (S, [P]),
    "vt_cloud_periodic": (I, [P]),
    "vt_cloud_valid_rgba": (I, [P]),
    "vt_cloud_tree_built": (I, [P]),
    "vt_cloud_pad": (D, [P]),
    "vt_cloud_subbox": (None, [P, P]),
    "vt_cloud_orig_ids": (None, [P, P]),
    "vt_cloud_tree_info": (I, [P, C.POINTER(TreeInfo)]),
    "vt_query_knn": (I, [P, P, S, I, P, P, P]),
    "vt_project": (I, [P, P, P, S, I, P, P]),
    "vt_project_noalloc": (I, [P, P, P, S, I, P, P]),
    "vt_project_count": (I, [P, P, P, S, P, P]),
    "vt_project_segments": (I, [P, P, P, S, I, P, P, P, P, P, P]),
    "vt_cloud_set_fields": (I, [P, P, P]),
    "vt_cloud_set_ray_rows": (I, [P, C.c_int64]),
    "vt_cloud_set_ray_order": (I, [P, P, C.c_int64, P]),
    "vt_cloud_order_rays": (I, [P, P, P, P, C.c_int64, C.c_int64, C.c_int64, P, P, P]),
    "vt_cloud_save_tree": (I, [P, P, S, C.POINTER(S)]),
    "vt_cloud_load_tree": (I, [P, P, S, I]),
    "vt_slice": (I, [P, P, S, S, C.c_double, P, P]),
    "vt_adjoint": (I, [P, P, P, S, I, P, P, P, P]),
    "vt_project_counted": (I, [P, P, P, S, I, P, P, P]),
    "vt_grid_rays": (I, [C.POINTER(View), S, P, P, P]),
    "vt_project_recorded": (I, [P, P, P, S, I, P, P, P, S, C.POINTER(S), P]),
    "vt_adjoint_recorded": (I, [P, S, I, P, P, S, P, P, P]),
    "vt_set_tuning": (I, [C.POINTER(Tuning)]),
    "vt_get_tuning": (I, [C.POINTER(Tuning)]),
    "vt_last_call_stats": (I, [C.POINTER(CallStats)]),
    "vt_cloud_lanes": (I, [P, I, I, C.POINTER(C.c_int64)]),
}

_lib = None


def _on_warning(msg, _user):
    warnings.warn(msg.decode(errors="replace"), UserWarning, stacklevel=3)


_warn_cb = WARN_FN(_on_warning)


def lib():
    """The loaded library (raises ImportError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C "
                f"{os.path.join(_HERE, 'csrc')}` (hipcc, gfx950)")
        # torch first: its HIP runtime (same soname) then serves this library
        # too, so torch's device pointers and streams are valid here.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.vt_set_warning_callback(_warn_cb, None)
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Raise the reference's exception type for a status code."""
    if rc == VT_OK:
        return
    msg = lib().vt_last_error().decode(errors="replace")
    if rc == VT_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def last_call_stats() -> CallStats:
    st = CallStats()
    check(lib().vt_last_call_stats(C.byref(st)))
    return st
