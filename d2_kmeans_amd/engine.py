"""Python host side of libd2mi.so, mirroring the reference's entry points.

:class:`Engine` owns one device context (one HIP stream on one GPU).  Its
methods follow the reference's interface for the hot path:

* :meth:`Engine.centroid_sphBregman` -- ``d2_centroid_sphBregman``
  (include/d2/clustering.h:226-230, centroid_Bregman.c:51-325);
* :meth:`Engine.pdist2` -- ``_dpdist2`` (blas_like64.c:199-208).

Errors raise :class:`~d2_kmeans_amd._lib.D2MIError`; the reference asserts
or exits in the same places (empty cluster, non-positive normaliser).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import BadmmStats, check, lib, ptr, vptr
from .data import BADMMOptions, Centroids, D2Data, Phase

VARIANTS = {"engine": _lib.D2MI_ENGINE, "plain": _lib.D2MI_PLAIN}


def device_count() -> int:
    n = C.c_int(0)
    check(lib().d2mi_device_count(C.byref(n)))
    return n.value


class Engine:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().d2mi_create(C.byref(h), device))
        self._h = h
        self.data: Optional[D2Data] = None
        self.cents = {}

    def close(self):
        if self._h:
            lib().d2mi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        check(lib().d2mi_synchronize(self._h))

    # -------------------------------------------------------------- data
    def set_phase(self, idx: int, ph: Phase):
        check(lib().d2mi_set_phase(
            self._h, idx, ph.metric, ph.dim, ph.size, ph.col, ptr(ph.p_str, _lib._ip),
            ph.p_str_cum.astype(np.uint64).ctypes.data_as(_lib._sp), ptr(ph.p_w),
            ptr(ph.p_supp), ptr(ph.p_supp_sym, _lib._ip), ph.vocab_size, ptr(ph.vocab_vec),
            ptr(ph.dist_mat)))

    def set_data(self, data: D2Data):
        for k, ph in enumerate(data.phases):
            self.set_phase(k, ph)
        self.data = data

    def set_labels(self, label, label_switch=None):
        lab = np.ascontiguousarray(label, np.int32)
        lsw = None if label_switch is None else np.ascontiguousarray(label_switch, np.int8)
        check(lib().d2mi_set_labels(self._h, ptr(lab, _lib._ip), vptr(lsw)))

    def get_labels(self):
        n = self.data.size
        lab = np.empty(n, np.int32)
        lsw = np.empty(n, np.int8)
        check(lib().d2mi_get_labels(self._h, ptr(lab, _lib._ip), vptr(lsw)))
        return lab, lsw

    def set_centroids(self, idx: int, c: Centroids):
        check(lib().d2mi_set_centroids(self._h, idx, c.K, c.s, ptr(np.ascontiguousarray(c.p_w)),
                                       ptr(None if c.p_supp is None else np.ascontiguousarray(c.p_supp))))
        self.cents[idx] = (c.K, c.s, c.dim, c.p_supp is not None)

    def get_centroids(self, idx: int) -> Centroids:
        K, s, dim, has_supp = self.cents[idx]
        c = Centroids.zeros(K, s, dim, with_supp=has_supp)
        check(lib().d2mi_get_centroids(self._h, idx, ptr(c.p_w), ptr(c.p_supp)))
        return c

    def get_plan(self, idx: int, precision: int = 32, which: str = "Z") -> np.ndarray:
        K, s, _, _ = self.cents[idx]
        out = np.empty(s * self.data.phases[idx].col)
        check(lib().d2mi_get_plan(self._h, idx, precision, 0 if which == "Z" else 1, ptr(out)))
        return out

    def set_plan(self, idx: int, Z, precision: int = 32):
        Z = np.ascontiguousarray(Z, np.float64)
        check(lib().d2mi_set_plan(self._h, idx, precision, ptr(Z)))

    def stream_probe(self, n: int, reps: int = 5):
        """(best GB/s, report) of two-array in-place FP32 streaming over n
        elements per array: the plan kernels' practical HBM ceiling."""
        g = C.c_double()
        buf = C.create_string_buffer(1024)
        check(lib().d2mi_stream_probe(self._h, int(n), int(reps), C.byref(g), buf, len(buf)))
        return g.value, buf.value.decode()

    # ------------------------------------------------------------ B-ADMM
    def badmm(self, idx: int, opts: BADMMOptions, precision: int = 32,
              variant: str = "engine") -> dict:
        st = BadmmStats()
        check(lib().d2mi_badmm(self._h, idx, opts.maxIters, opts.rhoCoeff, opts.updatePerLoops,
                               precision, VARIANTS[variant], C.byref(st)))
        tr = np.array([[st.trace[i][c] for c in range(4)] for i in range(st.n_trace)])
        return dict(rho=st.rho, obj=st.obj, trace=tr.reshape(-1, 4), ms_total=st.ms_total,
                    ms_plan=st.ms_plan, n_plan_launches=st.n_plan_launches,
                    plan_elems=st.plan_elems, graph=bool(st.graph),
                    plan_forms=st.plan_forms.decode(errors="replace"))

    def set_graph(self, on: bool):
        """Inner iterations of an update as one replayed HIP graph (default)
        or launched one by one; results are identical."""
        check(lib().d2mi_badmm_set_graph(self._h, 1 if on else 0))

    def tune(self, idx: int, precision: int = 32, reps: int = 3) -> str:
        """Times every feasible plan-kernel form and workgroups-per-CU
        setting per object-size class on the resident plans (restored
        afterwards) and keeps the fastest for later updates; returns the
        report, one line per class."""
        buf = C.create_string_buffer(8192)
        check(lib().d2mi_badmm_tune(self._h, idx, precision, reps, buf, len(buf)))
        return buf.value.decode(errors="replace")

    def tune_clear(self):
        check(lib().d2mi_badmm_tune_clear(self._h))

    def centroid_sphBregman(self, idx_ph: int, opts: BADMMOptions = None, precision: int = 32):
        """``d2_centroid_sphBregman(p_data, var_work, idx_ph, c0, c)`` with the
        warm start c0 == c (clustering.c:493): labels, label switches and
        centroids must already be on the device; the update runs there and
        the centroids are returned."""
        opts = opts or BADMMOptions()
        st = self.badmm(idx_ph, opts, precision)
        return self.get_centroids(idx_ph), st

    # ------------------------------------------------------------ pdist2
    def pdist2(self, A, B) -> np.ndarray:
        """_dpdist2: returns C with C[i*n + j] = ||A_j - B_i||^2 (A: n x d,
        B: m x d), computed on the device in FP64."""
        A = np.ascontiguousarray(A, np.float64)
        B = np.ascontiguousarray(B, np.float64)
        n, d = A.shape
        m = B.shape[0]
        out = np.empty(n * m)
        check(lib().d2mi_pdist2(self._h, d, n, m, ptr(A), ptr(B), ptr(out)))
        return out

    def pdist2_mfma(self, A, B, bf16: bool = False, timing: bool = False):
        """The same table as :meth:`pdist2` from FP32 inputs on the matrix
        cores (FP32 or BF16 products, FP32 norms and accumulation)."""
        A = np.ascontiguousarray(A, np.float64)
        B = np.ascontiguousarray(B, np.float64)
        n, d = A.shape
        m = B.shape[0]
        out = np.empty(n * m)
        ms = C.c_float(0)
        check(lib().d2mi_pdist2_mfma(self._h, d, n, m, ptr(A), ptr(B), ptr(out), int(bf16),
                                     C.byref(ms)))
        return (out, ms.value) if timing else out

    # ------------------------------------------------------- transport LP
    def transport_batch(self, C, wX, wY, pivots=False):
        """Exact LP values of a batch of same-shape transport problems
        (d2_match_by_distmat's LP, solver_mosek.cc:61-143).  C: (B, n*m) with
        element (i, j) at [i + j*n]; wX: (B, n); wY: (B, m)."""
        wX = np.ascontiguousarray(np.atleast_2d(wX), np.float64)
        wY = np.ascontiguousarray(np.atleast_2d(wY), np.float64)
        nb, n = wX.shape
        m = wY.shape[1]
        C = np.ascontiguousarray(C, np.float64).reshape(nb, n * m)
        out = np.empty(nb)
        piv = np.empty(nb, np.int32)
        check(lib().d2mi_transport_batch(self._h, n, m, nb, ptr(C), ptr(wX), ptr(wY), ptr(out),
                                         ptr(piv, _lib._ip)))
        return (out, piv) if pivots else out

    LP_PHASES = ("table", "duals", "start", "root", "jump", "price", "pivot", "value",
                 "start_argmax", "start_take", "start_scan", "start_walk")

    def transport_profile(self, C, wX, wY):
        """transport_batch plus each LP's time per solver phase:
        (values, pivots, phase_us of shape (B, 12) in the order LP_PHASES;
        the last four split the "start" phase)."""
        wX = np.ascontiguousarray(np.atleast_2d(wX), np.float64)
        wY = np.ascontiguousarray(np.atleast_2d(wY), np.float64)
        nb, n = wX.shape
        m = wY.shape[1]
        C = np.ascontiguousarray(C, np.float64).reshape(nb, n * m)
        out = np.empty(nb)
        piv = np.empty(nb, np.int32)
        ph = np.empty((nb, len(self.LP_PHASES)))
        check(lib().d2mi_transport_profile(self._h, n, m, nb, ptr(C), ptr(wX), ptr(wY), ptr(out),
                                           ptr(piv, _lib._ip), ptr(ph)))
        return out, piv, ph

    # -------------------------------------------------------- assignment
    def distances(self, obj, cent, selected_phase: int = -1) -> np.ndarray:
        """Batched d2_compute_distance (clustering.c:33-128): exact
        sqrt(sum over phases of the transport-LP value) per (object, centroid)."""
        obj = np.ascontiguousarray(obj, np.int32)
        cent = np.ascontiguousarray(cent, np.int32)
        out = np.empty(len(obj))
        check(lib().d2mi_distances(self._h, selected_phase, len(obj), ptr(obj, _lib._ip),
                                   ptr(cent, _lib._ip), ptr(out)))
        return out

    def elkan_reset(self, selected_phase: int = -1):
        check(lib().d2mi_elkan_reset(self._h, selected_phase))

    def labeling(self, use_elkan: bool = False, selected_phase: int = -1) -> dict:
        """d2_labeling (brute force) or d2_labeling_prep (Elkan) on the device."""
        st = _lib.LabelStats()
        check(lib().d2mi_labeling(self._h, selected_phase, int(use_elkan), C.byref(st)))
        return dict(changes=st.changes, dist_count=st.dist_count, cost=st.cost, ms=st.ms)

    def save_centroids(self, selected_phase: int = -1):
        check(lib().d2mi_save_centroids(self._h, selected_phase))

    def elkan_post(self, selected_phase: int = -1):
        check(lib().d2mi_elkan_post(self._h, selected_phase))

    def clustering(self, K: int, max_iter: int, opts: BADMMOptions = None, use_triangle=True,
                   selected_phase: int = -1, precision: int = 32) -> dict:
        """The d2_clustering outer loop (clustering.c:409-525) on the device,
        from the resident centroids; labels are read back at the end."""
        opts = opts or BADMMOptions()
        st = _lib.ClusterStats()
        check(lib().d2mi_clustering(self._h, K, max_iter, selected_phase, int(use_triangle),
                                    opts.maxIters, opts.rhoCoeff, opts.updatePerLoops, precision,
                                    C.byref(st)))
        it = st.iterations
        return dict(iterations=it, changes=list(st.changes[:it]),
                    dist_count=list(st.dist_count[:it]), final_cost=st.final_cost,
                    ms_labeling=st.ms_labeling, ms_badmm=st.ms_badmm)
