"""ctypes binding of libd2mi.so (include/d2mi.h).

The library is built in-tree by ``d2_kmeans_amd.build``; loading fails loudly
when it is missing -- there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("D2MI_LIB") or os.path.join(PKG, "libd2mi.so")

D2MI_ENGINE = 0
D2MI_PLAIN = 1

_ERRNAMES = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EEMPTY", -5: "ENUMERIC", -6: "EUNSUPPORTED"}


class D2MIError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"d2mi error {code} ({_ERRNAMES.get(code, '?')}): {msg}")
        self.code = code


class BadmmStats(C.Structure):
    _fields_ = [("rho", C.c_double), ("obj", C.c_double), ("n_trace", C.c_int),
                ("trace", (C.c_double * 4) * 64), ("ms_total", C.c_double),
                ("ms_plan", C.c_double), ("n_plan_launches", C.c_int),
                ("plan_elems", C.c_double), ("graph", C.c_int), ("plan_forms", C.c_char * 200)]


class LabelStats(C.Structure):
    _fields_ = [("changes", C.c_ulonglong), ("dist_count", C.c_ulonglong), ("cost", C.c_double),
                ("ms", C.c_double)]


class ClusterStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("changes", C.c_longlong * 256),
                ("dist_count", C.c_longlong * 256), ("final_cost", C.c_double),
                ("ms_labeling", C.c_double), ("ms_badmm", C.c_double)]


class ServeOpts(C.Structure):
    _fields_ = [("max_batch", C.c_int), ("streams", C.c_int), ("hbm_bytes", C.c_size_t),
                ("max_m", C.c_int), ("use_triangle", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in ("submitted", "completed", "groups", "lps",
                                             "memory_waits", "max_inflight_groups", "max_group")]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_sp = C.POINTER(C.c_size_t)
_cp = C.c_char_p
_vp = C.c_void_p

# every function the header declares, with its argument types
SIGNATURES = {
    "d2mi_last_error": (C.c_char_p, []),
    "d2mi_device_count": (C.c_int, [_ip]),
    "d2mi_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "d2mi_destroy": (C.c_int, [_vp]),
    "d2mi_synchronize": (C.c_int, [_vp]),
    "d2mi_set_phase": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_size_t, _ip, _sp,
                                 _dp, _dp, _ip, C.c_int, _dp, _dp]),
    "d2mi_set_labels": (C.c_int, [_vp, _ip, _vp]),
    "d2mi_get_labels": (C.c_int, [_vp, _ip, _vp]),
    "d2mi_set_centroids": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "d2mi_get_centroids": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "d2mi_badmm": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                             C.POINTER(BadmmStats)]),
    "d2mi_badmm_set_graph": (C.c_int, [_vp, C.c_int]),
    "d2mi_badmm_tune": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_size_t]),
    "d2mi_badmm_tune_clear": (C.c_int, [_vp]),
    "d2mi_get_plan": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "d2mi_stream_probe": (C.c_int, [_vp, C.c_size_t, C.c_int, _dp, C.c_char_p, C.c_int]),
    "d2mi_set_plan": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "d2mi_pdist2": (C.c_int, [_vp, C.c_int, C.c_size_t, C.c_size_t, _dp, _dp, _dp]),
    "d2mi_transport_batch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _ip]),
    "d2mi_transport_profile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _ip, _dp]),
    "d2mi_transport_batch_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                               _ip]),
    "d2mi_pdist2_mfma": (C.c_int, [_vp, C.c_int, C.c_size_t, C.c_size_t, _dp, _dp, _dp, C.c_int,
                                   C.POINTER(C.c_float)]),
    "d2mi_distances": (C.c_int, [_vp, C.c_int, C.c_size_t, _ip, _ip, _dp]),
    "d2mi_elkan_reset": (C.c_int, [_vp, C.c_int]),
    "d2mi_labeling": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(LabelStats)]),
    "d2mi_save_centroids": (C.c_int, [_vp, C.c_int]),
    "d2mi_elkan_post": (C.c_int, [_vp, C.c_int]),
    "d2mi_clustering": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                  C.c_int, C.c_int, C.POINTER(ClusterStats)]),
    "d2mi_serve_open": (C.c_int, [_vp, C.c_int, C.POINTER(ServeOpts), C.POINTER(_vp)]),
    "d2mi_serve_submit": (C.c_int, [_vp, C.c_int, _dp, _dp, C.POINTER(C.c_longlong)]),
    "d2mi_serve_progress": (C.c_int, [_vp, _ip]),
    "d2mi_serve_result": (C.c_int, [_vp, C.c_longlong, _ip, _dp, _dp]),
    "d2mi_serve_drain": (C.c_int, [_vp]),
    "d2mi_serve_stats_get": (C.c_int, [_vp, C.POINTER(ServeStats)]),
    "d2mi_serve_close": (C.c_int, [_vp]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m d2_kmeans_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise D2MIError(rc, lib().d2mi_last_error().decode(errors="replace"))
    return rc


def ptr(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def vptr(a):
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)
