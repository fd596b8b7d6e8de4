"""ctypes bindings for the oracle (test infrastructure only).

Two libraries, both CPU, FP64:

* ``oracle/liboracle.so`` -- this project's restatement (d2_oracle.c);
* ``oracle/_ref/libad2c_ref.so`` -- the reference's own sources built by
  ``oracle/Makefile.ref`` where ``/root/reference`` exists.  Loaded lazily
  (RTLD_LAZY): its Mosek solver symbols are undefined and never reached.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module.  The product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libad2c_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_sp = C.POINTER(C.c_size_t)
_cp = C.POINTER(C.c_char)


def build(force=False):
    root = os.path.dirname(HERE)
    if force or not os.path.exists(LIB_ORACLE):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-f", "oracle/Makefile"], cwd=root)


def build_ref(force=False):
    """Compile the reference sources (only where /root/reference exists)."""
    root = os.path.dirname(HERE)
    if not os.path.isdir("/root/reference"):
        return False
    args = ["make", "-s", "-j8", "-f", "oracle/Makefile.ref"]
    if force:
        args.insert(2, "-B")
    subprocess.check_call(args, cwd=root)
    # the reference's own program on the product library, when that is built
    if os.path.exists(os.path.join(root, "d2_kmeans_amd", "libd2mi.so")):
        subprocess.check_call(args + ["main"], cwd=root)
    return True


REF_MAIN = os.path.join(HERE, "_ref", "d2_main_on_d2mi")


def ref_available():
    return os.path.exists(LIB_REF)


def ptr(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


# ------------------------------------------------------------ the restatement

class OrPhase(C.Structure):
    _fields_ = [("metric", C.c_int), ("dim", C.c_int), ("size", C.c_size_t), ("col", C.c_size_t),
                ("p_str", _ip), ("p_str_cum", _sp), ("p_w", _dp), ("p_supp", _dp),
                ("p_supp_sym", _ip), ("vocab_size", C.c_int), ("vocab_vec", _dp),
                ("dist_mat", _dp)]


class OrCentroids(C.Structure):
    _fields_ = [("K", C.c_int), ("s", C.c_int), ("dim", C.c_int), ("p_w", _dp), ("p_supp", _dp)]


class OrOpts(C.Structure):
    _fields_ = [("maxIters", C.c_int), ("rhoCoeff", C.c_double), ("updatePerLoops", C.c_int)]


class OrElkan(C.Structure):
    _fields_ = [("l", _dp), ("u", _dp), ("s", _dp), ("c", _dp), ("r", _cp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_ORACLE)
        L.or_pdist2.argtypes = [C.c_int, C.c_size_t, C.c_size_t, _dp, _dp, _dp]
        L.or_badmm.argtypes = [C.POINTER(OrPhase), _ip, _cp, C.POINTER(OrCentroids), _dp, _dp, _dp,
                               _dp, C.POINTER(OrOpts), _dp, _dp, C.c_int, _ip, _dp]
        L.or_transport.restype = C.c_double
        L.or_transport.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_transport_ex.restype = C.c_double
        L.or_transport_ex.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _ip]
        L.or_distance_obj_cent.restype = C.c_double
        L.or_distance_obj_cent.argtypes = [C.c_int, C.POINTER(OrPhase), C.c_int, C.c_size_t,
                                           C.POINTER(OrCentroids), C.c_int]
        L.or_distance_cent_cent.restype = C.c_double
        L.or_distance_cent_cent.argtypes = [C.c_int, C.POINTER(OrPhase), C.c_int,
                                            C.POINTER(OrCentroids), C.c_int,
                                            C.POINTER(OrCentroids), C.c_int]
        L.or_labeling.restype = C.c_size_t
        L.or_labeling.argtypes = [C.c_int, C.POINTER(OrPhase), C.c_int, C.POINTER(OrCentroids),
                                  C.c_int, _ip, _cp, _dp]
        L.or_labeling_prep.restype = C.c_size_t
        L.or_labeling_prep.argtypes = [C.c_int, C.POINTER(OrPhase), C.c_int,
                                       C.POINTER(OrCentroids), C.c_int, _ip, _cp,
                                       C.POINTER(OrElkan), _sp]
        L.or_labeling_post.argtypes = [C.c_int, C.POINTER(OrPhase), C.c_int,
                                       C.POINTER(OrCentroids), C.POINTER(OrCentroids), C.c_int,
                                       _ip, C.POINTER(OrElkan)]
        L.or_calculate_distmat.argtypes = [C.POINTER(OrPhase), _ip, C.POINTER(OrCentroids), _dp]
        L.or_shuffle.argtypes = [_sp, C.c_size_t, C.c_uint]
        L.or_merge.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int]
        L.or_centroid_rands.argtypes = [C.POINTER(OrPhase), C.c_int, C.c_int,
                                        C.POINTER(OrCentroids), C.c_uint]
        L.or_clustering.argtypes = [C.c_int, C.POINTER(OrPhase), C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(OrOpts), C.POINTER(OrCentroids), _ip, _sp, C.c_int,
                                    _dp]
        _lib = L
    return _lib


class _Keep:
    """Holds numpy buffers alive while ctypes structs point into them."""

    def __init__(self):
        self.bufs = []

    def add(self, a):
        self.bufs.append(a)
        return a


def or_phase(ph, keep):
    s = OrPhase()
    s.metric, s.dim, s.size, s.col = ph.metric, ph.dim, ph.size, ph.col
    s.p_str = ptr(keep.add(ph.p_str), _ip)
    s.p_str_cum = ptr(keep.add(ph.p_str_cum.astype(np.uint64)), _sp)
    s.p_w = ptr(keep.add(ph.p_w))
    s.p_supp = ptr(ph.p_supp)
    s.p_supp_sym = ptr(ph.p_supp_sym, _ip)
    s.vocab_size = ph.vocab_size
    s.vocab_vec = ptr(ph.vocab_vec)
    s.dist_mat = ptr(ph.dist_mat)
    return s


def or_cent(c):
    s = OrCentroids()
    s.K, s.s, s.dim = c.K, c.s, c.dim
    s.p_w = ptr(c.p_w)
    s.p_supp = ptr(c.p_supp)
    return s


def pdist2(A, B):
    """C (m x n, C[i*n+j]) for A: n x d, B: m x d."""
    A = np.ascontiguousarray(A, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    n, d = A.shape
    m = B.shape[0]
    out = np.empty(n * m)
    lib().or_pdist2(d, n, m, ptr(A), ptr(B), ptr(out))
    return out


def transport(Cmat, wX, wY, pivots=False):
    """Exact LP value; Cmat is n x m with element (i, j) at [i + j*n]."""
    Cf = np.ascontiguousarray(Cmat, np.float64).reshape(-1)
    wX = np.ascontiguousarray(wX, np.float64)
    wY = np.ascontiguousarray(wY, np.float64)
    piv = C.c_int(0)
    v = lib().or_transport_ex(len(wX), len(wY), ptr(Cf), ptr(wX), ptr(wY), C.byref(piv))
    return (v, piv.value) if pivots else v


def badmm(phase, label, label_switch, cent, Z, opts, max_trace=64):
    """One B-ADMM centroid update (centroid_Bregman.c:51-325) in place on
    ``cent`` and ``Z``.  Returns dict(rho, obj, trace, X, Y, C, rc)."""
    keep = _Keep()
    s = cent.s
    sc = s * phase.col
    Cb = np.empty(sc)
    X = np.empty(sc)
    Y = np.empty(sc)
    lab = keep.add(np.ascontiguousarray(label, np.int32))
    lsw = keep.add(np.ascontiguousarray(label_switch, np.int8))
    ph = or_phase(phase, keep)
    cs = or_cent(cent)
    o = OrOpts(opts.maxIters, opts.rhoCoeff, opts.updatePerLoops)
    rho = C.c_double(0)
    obj = C.c_double(0)
    trace = np.zeros(4 * max_trace)
    ntr = C.c_int(0)
    rc = lib().or_badmm(C.byref(ph), ptr(lab, _ip), lsw.ctypes.data_as(_cp), C.byref(cs), ptr(Cb),
                        ptr(X), ptr(Z), ptr(Y), C.byref(o), C.byref(rho), ptr(trace), max_trace,
                        C.byref(ntr), C.byref(obj))
    return dict(rc=rc, rho=rho.value, obj=obj.value,
                trace=trace[:4 * min(ntr.value, max_trace)].reshape(-1, 4), X=X, Y=Y, C=Cb)


def hist_cost(phase, s):
    """The restatement's cost blocks of a histogram phase (or_calculate_distmat
    with centroids of s = V bins; labels do not enter)."""
    from d2_kmeans_amd.data import Centroids
    keep = _Keep()
    out = np.zeros(s * phase.col)
    lab = keep.add(np.zeros(phase.size, np.int32))
    c = Centroids(1, s, 0, keep.add(np.full(s, 1.0 / s)))
    cs = or_cent(c)
    lib().or_calculate_distmat(C.byref(or_phase(phase, keep)), ptr(lab, _ip), C.byref(cs),
                               ptr(out))
    return out


def _phases(data, keep):
    arr = (OrPhase * data.s_ph)()
    for k, p in enumerate(data.phases):
        arr[k] = or_phase(p, keep)
    return arr


def _cents(cents):
    arr = (OrCentroids * len(cents))()
    for k, c in enumerate(cents):
        if c is not None:
            arr[k] = or_cent(c)
    return arr


def distance(data, i, cents, j, selected_phase=-1):
    keep = _Keep()
    return lib().or_distance_obj_cent(data.s_ph, _phases(data, keep), selected_phase, i,
                                      _cents(cents), j)


def distance_cc(data, ca, i, cb, j, selected_phase=-1):
    keep = _Keep()
    return lib().or_distance_cent_cent(data.s_ph, _phases(data, keep), selected_phase,
                                       _cents(ca), i, _cents(cb), j)


def labeling(data, cents, label=None, selected_phase=-1):
    """Brute-force labeling; returns (label, label_switch, changes, cost)."""
    keep = _Keep()
    N = data.size
    lab = np.array(data.label if label is None else label, dtype=np.int32)
    lsw = np.zeros(N, np.int8)
    cost = C.c_double(0)
    K = cents[max(selected_phase, 0)].K
    ch = lib().or_labeling(data.s_ph, _phases(data, keep), selected_phase, _cents(cents), K,
                           ptr(lab, _ip), lsw.ctypes.data_as(_cp), C.byref(cost))
    return lab, lsw, ch, cost.value


class Elkan:
    def __init__(self, N, K):
        self.l = np.zeros(N * K)
        self.u = np.full(N, np.finfo(np.float64).max)
        self.s = np.zeros(K)
        self.c = np.zeros(K * K)
        self.r = np.ones(N, np.int8)

    def struct(self):
        e = OrElkan()
        e.l, e.u, e.s, e.c = ptr(self.l), ptr(self.u), ptr(self.s), ptr(self.c)
        e.r = self.r.ctypes.data_as(_cp)
        return e


def labeling_prep(data, cents, label, elk, selected_phase=-1):
    keep = _Keep()
    N = data.size
    lsw = np.zeros(N, np.int8)
    dc = C.c_size_t(0)
    K = cents[max(selected_phase, 0)].K
    es = elk.struct()
    ch = lib().or_labeling_prep(data.s_ph, _phases(data, keep), selected_phase, _cents(cents), K,
                                ptr(label, _ip), lsw.ctypes.data_as(_cp), C.byref(es),
                                C.byref(dc))
    return lsw, ch, dc.value


def labeling_post(data, c_old, c_new, label, elk, selected_phase=-1):
    keep = _Keep()
    K = c_new[max(selected_phase, 0)].K
    es = elk.struct()
    lib().or_labeling_post(data.s_ph, _phases(data, keep), selected_phase, _cents(c_old),
                           _cents(c_new), K, ptr(label, _ip), C.byref(es))


def shuffle(n, seed=0):
    a = np.arange(n, dtype=np.uint64)
    lib().or_shuffle(a.ctypes.data_as(_sp), n, seed)
    return a


def merge(supp, w, n):
    supp = np.ascontiguousarray(supp, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    m, d = supp.shape
    cs = np.zeros(n * d)
    cw = np.zeros(n)
    lib().or_merge(d, ptr(supp), ptr(w), m, ptr(cs), ptr(cw), n)
    return cs.reshape(n, d), cw


def centroid_rands(phase, K, s, seed=0):
    from d2_kmeans_amd.data import Centroids
    keep = _Keep()
    c = Centroids.zeros(K, s, phase.dim, with_supp=phase.dim > 0)
    cs = or_cent(c)
    rc = lib().or_centroid_rands(C.byref(or_phase(phase, keep)), K, s, C.byref(cs), seed)
    if rc:
        raise RuntimeError(f"too few objects with >= {s} supports for {K} centroids")
    return c


def clustering(data, cents, K, max_iter, opts, use_triangle=True, selected_phase=-1):
    """Full outer loop; cents (list per phase) updated in place.
    Returns (label, iterations, changes_trace, final_cost)."""
    keep = _Keep()
    N = data.size
    lab = np.full(N, -1, np.int32)
    trace = np.zeros(max_iter, np.uint64)
    cost = C.c_double(0)
    o = OrOpts(opts.maxIters, opts.rhoCoeff, opts.updatePerLoops)
    it = lib().or_clustering(data.s_ph, _phases(data, keep), selected_phase, K, max_iter,
                             int(use_triangle), C.byref(o), _cents(cents), ptr(lab, _ip),
                             trace.ctypes.data_as(_sp), max_iter, C.byref(cost))
    return lab, it, trace[:it].astype(np.int64), cost.value


# -------------------------------------------------------- the reference build

class RefSph(C.Structure):
    """``sph`` with SCALAR=double (include/d2/clustering.h:14-95)."""
    _fields_ = [("dim", C.c_int), ("str", C.c_int), ("max_str", C.c_int), ("col", C.c_size_t),
                ("max_col", C.c_size_t), ("p_str", _ip), ("p_str_cum", _sp), ("p_w", _dp),
                ("metric_type", C.c_int), ("p_supp", _dp), ("p_supp_sym", _ip),
                ("dist_mat", _dp), ("vocab_size", C.c_int), ("vocab_vec", _dp),
                ("is_meta_allocated", C.c_char)]


class RefMph(C.Structure):
    """``mph`` (clustering.h:117-124)."""
    _fields_ = [("s_ph", C.c_int), ("size", C.c_size_t), ("global_size", C.c_size_t),
                ("label", _ip), ("num_of_labels", C.c_int), ("ph", C.POINTER(RefSph))]


class RefVarSph(C.Structure):
    _fields_ = [("C", _dp), ("X", _dp), ("L", _dp)]


class RefVarBregman(C.Structure):
    _fields_ = [("X", _dp), ("Z", _dp), ("Y", _dp), ("Xc", _dp), ("Zr", _dp)]


class RefTrieq(C.Structure):
    _fields_ = [("l", _dp), ("u", _dp), ("s", _dp), ("c", _dp), ("r", _cp)]


class RefVarMph(C.Structure):
    _fields_ = [("s_ph", C.c_int), ("g_var", C.POINTER(RefVarSph)),
                ("l_var_sphBregman", C.POINTER(RefVarBregman)), ("label_switch", _cp),
                ("tr", RefTrieq)]


class RefOpts(C.Structure):
    _fields_ = [("maxIters", C.c_int), ("rhoCoeff", C.c_double), ("updatePerLoops", C.c_int)]


_ref = None


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libad2c_ref.so is not built")
        L = C.CDLL(LIB_REF, mode=os.RTLD_LAZY | os.RTLD_GLOBAL)
        L.d2_centroid_sphBregman.argtypes = [C.POINTER(RefMph), C.POINTER(RefVarMph), C.c_int,
                                             C.POINTER(RefSph), C.POINTER(RefSph)]
        L._dpdist2.argtypes = [C.c_int, C.c_size_t, C.c_size_t, _dp, _dp, _dp]
        L._dpdist2_submat.argtypes = [C.c_size_t, _ip, _dp, C.c_int, _dp]
        L.merge.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int]
        L.shuffle.argtypes = [_sp, C.c_size_t]
        L.d2_centroid_rands.argtypes = [C.POINTER(RefMph), C.c_int, C.POINTER(RefSph)]
        _ref = L
    return _ref


def _ref_sph(ph, keep, s_override=None):
    s = RefSph()
    s.dim = ph.dim
    s.str = s_override if s_override is not None else ph.str
    s.max_str = int(ph.p_str.max()) if ph.size else 0
    s.col = ph.col
    s.max_col = ph.col
    s.p_str = ptr(keep.add(ph.p_str.copy()), _ip)
    s.p_str_cum = ptr(keep.add(ph.p_str_cum.astype(np.uint64)), _sp)
    s.p_w = ptr(keep.add(ph.p_w.copy()))
    s.metric_type = ph.metric
    s.p_supp = ptr(keep.add(ph.p_supp.copy())) if ph.p_supp is not None else None
    s.p_supp_sym = ptr(keep.add(ph.p_supp_sym.copy()), _ip) if ph.p_supp_sym is not None else None
    s.dist_mat = ptr(ph.dist_mat)
    s.vocab_size = ph.vocab_size
    s.vocab_vec = ptr(ph.vocab_vec)
    s.is_meta_allocated = b"\0"
    return s


def _ref_cent_sph(c, metric, keep):
    s = RefSph()
    s.dim, s.str, s.max_str = c.dim, c.s, c.s
    s.col = s.max_col = c.K * c.s
    s.p_str = ptr(keep.add(np.full(c.K, c.s, np.int32)), _ip)
    s.p_str_cum = ptr(keep.add((np.arange(c.K) * c.s).astype(np.uint64)), _sp)
    s.p_w = ptr(c.p_w)
    s.metric_type = metric
    s.p_supp = ptr(c.p_supp)
    return s


def ref_set_badmm_options(opts):
    """Points the reference's ``p_badmm_options`` at ``badmm_clu_options``
    and overwrites its fields (the way main.cc:232-236 repoints it)."""
    L = ref()
    clu = RefOpts.in_dll(L, "badmm_clu_options")
    clu.maxIters, clu.rhoCoeff, clu.updatePerLoops = opts.maxIters, opts.rhoCoeff, opts.updatePerLoops
    pp = C.c_void_p.in_dll(L, "p_badmm_options")
    pp.value = C.addressof(clu)
    C.c_double.in_dll(L, "time_budget").value = 1e30  # never break on wall clock


def ref_prefill_hist_cost(phase, s, Cb):
    """The ground-cost blocks the reference precomputes once for histogram
    phases, because its calculate_distmat skips them (centroid_util.c:26-28):
    clustering_util.c:183-199.  Dense histograms copy the phase's V x V
    ``dist_mat`` into every object's block (the reference indexes ``ph[k]``
    by object there, a bug; the phase's own matrix is what it means); sparse
    histograms take the rows of their bins through the reference's own
    ``_dpdist2_submat`` (blas_like64.c:222-230)."""
    L = ref()
    cum = phase.p_str_cum
    if phase.metric == 5:  # D2_HISTOGRAM
        for k in range(phase.size):
            n = s * int(phase.p_str[k])
            Cb[s * int(cum[k]):s * int(cum[k]) + n] = phase.dist_mat[:n]
    elif phase.metric == 12:  # D2_SPARSE_HISTOGRAM
        dm = np.ascontiguousarray(phase.dist_mat, np.float64)
        sym = np.ascontiguousarray(phase.p_supp_sym, np.int32)
        for k in range(phase.size):
            off = int(cum[k])
            L._dpdist2_submat(int(phase.p_str[k]), sym[off:].ctypes.data_as(_ip),
                              Cb[s * off:].ctypes.data_as(_dp), int(phase.vocab_size), ptr(dm))


def ref_work(phase, K, s):
    """The work arrays d2_allocate_work gives one phase (clustering_util.c:
    152-230, centroid_Bregman.c:26-35): C, X, Y, Xc, Zr, all touched."""
    N = phase.size
    return dict(Cb=np.zeros(s * (phase.col + K * s)), X=np.zeros(s * phase.col),
                Y=np.zeros(s * phase.col), Xc=np.zeros(phase.col), Zr=np.zeros(s * max(N, K)))


def ref_badmm(phase, label, label_switch, cent, Z, opts, work=None):
    """Runs the reference's d2_centroid_sphBregman (centroid_Bregman.c:51)
    on one phase, in place on ``cent`` and ``Z``; returns (X, Y, C).
    Histogram phases get their cost blocks pre-filled first, as the
    reference's work allocation does (``ref_prefill_hist_cost``); a fresh
    block per call, so every call starts from the raw ground cost.  ``work``
    (from ``ref_work``) supplies the work arrays instead of fresh ones."""
    L = ref()
    ref_set_badmm_options(opts)
    keep = _Keep()
    N = phase.size
    sph = _ref_sph(phase, keep, s_override=cent.s)
    sph_arr = (RefSph * 1)(sph)
    lab = keep.add(np.ascontiguousarray(label, np.int32).copy())
    m = RefMph(1, N, N, ptr(lab, _ip), cent.K, sph_arr)
    s = cent.s
    sc = s * phase.col
    w = work if work is not None else ref_work(phase, cent.K, s)
    Cb, X, Y, Xc, Zr = w["Cb"], w["X"], w["Y"], w["Xc"], w["Zr"]
    if phase.metric in (5, 12):
        Cb[:] = 0.0
        ref_prefill_hist_cost(phase, s, Cb)
    gv = (RefVarSph * 1)(RefVarSph(ptr(Cb), None, None))
    bv = (RefVarBregman * 1)(RefVarBregman(ptr(X), ptr(Z), ptr(Y), ptr(Xc), ptr(Zr)))
    lsw = np.ascontiguousarray(label_switch, np.int8).copy()
    vm = RefVarMph(1, gv, bv, lsw.ctypes.data_as(_cp), RefTrieq())
    csph = _ref_cent_sph(cent, phase.metric, keep)
    rc = L.d2_centroid_sphBregman(C.byref(m), C.byref(vm), 0, C.byref(csph), C.byref(csph))
    assert rc == 0
    return X, Y, Cb[:sc]


def ref_pdist2(A, B):
    A = np.ascontiguousarray(A, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    n, d = A.shape
    m = B.shape[0]
    out = np.empty(n * m)
    ref()._dpdist2(d, n, m, ptr(A), ptr(B), ptr(out))
    return out


def ref_merge(supp, w, n):
    supp = np.ascontiguousarray(supp, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    m, d = supp.shape
    cs = np.zeros(n * d)
    cw = np.zeros(n)
    ref().merge(d, ptr(supp), ptr(w), m, ptr(cs), ptr(cw), n)
    return cs.reshape(n, d), cw


def ref_centroid_rands(phase, K, s):
    """The reference's d2_centroid_rands (centroid_rand.c:161), serial."""
    from d2_kmeans_amd.data import Centroids
    L = ref()
    keep = _Keep()
    sph = _ref_sph(phase, keep, s_override=s)
    sph_arr = (RefSph * 1)(sph)
    lab = keep.add(np.full(phase.size, -1, np.int32))
    m = RefMph(1, phase.size, phase.size, ptr(lab, _ip), K, sph_arr)
    c = Centroids.zeros(K, s, phase.dim, with_supp=phase.dim > 0)
    csph = _ref_cent_sph(c, phase.metric, keep)
    L.d2_centroid_rands(C.byref(m), 0, C.byref(csph))
    return c
