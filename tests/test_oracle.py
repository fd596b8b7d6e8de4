"""Pins the FP64 oracle (oracle/d2_oracle.c) before anything is checked
against it: golden vectors of the reference's own build, the live reference
build where it exists, an independent LP solver, closed forms and invariants.
CPU only."""
import itertools
import os

import numpy as np
import pytest

from d2_kmeans_amd import synth
from d2_kmeans_amd.data import (BADMMOptions, Centroids, D2Data, D2_EUCLIDEAN_L2, Phase,
                                read_d2)
from cases import hist_golden_case, hist_phase, hist_state
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _badmm_case(g):
    ph = Phase(metric=D2_EUCLIDEAN_L2, dim=int(g["dim"]), p_str=g["p_str"], p_w=g["p_w"],
               p_supp=g["p_supp"])
    K, s = int(g["K"]), int(g["s"])
    c0 = Centroids(K, s, ph.dim, g["c0_w"].copy(), g["c0_supp"].copy())
    opts = BADMMOptions(int(g["maxIters"]), float(g["rhoCoeff"]), int(g["updatePerLoops"]))
    return ph, c0, opts


# ------------------------------------------------------------ B-ADMM vs golden

def test_badmm_matches_reference_golden():
    """centroid_Bregman.c:51-325 through the reference's own build: a cold
    call then a warm call with a third of the labels switched."""
    g = gold("ref_badmm.npz")
    ph, c, opts = _badmm_case(g)
    Z = np.zeros(c.s * ph.col)
    r1 = O.badmm(ph, g["label"], g["lsw"], c, Z, opts)
    assert r1["rc"] == 0
    np.testing.assert_allclose(c.p_w, g["c1_w"], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(c.p_supp, g["c1_supp"], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(Z, g["Z1"], rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(r1["Y"], g["Y1"], rtol=1e-9, atol=1e-14)
    r2 = O.badmm(ph, g["label2"], g["lsw2"], c, Z, opts)
    np.testing.assert_allclose(c.p_w, g["c2_w"], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(c.p_supp, g["c2_supp"], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(Z, g["Z2"], rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(r2["X"], g["X2"], rtol=1e-10, atol=1e-15)


def test_pdist2_matches_reference_golden():
    g = gold("ref_pdist2.npz")
    np.testing.assert_array_equal(O.pdist2(g["A"], g["B"]), g["C"])


def test_centroid_rands_matches_reference_golden():
    g = gold("ref_rands.npz")
    data, _ = synth.gaussian_mixture()
    c = O.centroid_rands(data.phases[0], 8, 6, seed=0)
    np.testing.assert_array_equal(c.p_w, g["p_w"])
    np.testing.assert_array_equal(c.p_supp, g["p_supp"])


# ------------------------------------------------------- live reference build

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")


@needs_ref
@pytest.mark.parametrize("seed,K,s,iters,upl", [(1, 8, 6, 10, 10), (2, 3, 4, 25, 7),
                                                (3, 1, 6, 45, 10)])
def test_badmm_matches_reference_live(seed, K, s, iters, upl):
    data, _ = synth.gaussian_mixture(n=300, seed=seed)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, K, s, seed=0)
    lab, lsw, _, _ = O.labeling(data, [c0])
    if len(np.unique(lab)) < K:
        pytest.skip("empty cluster")
    opts = BADMMOptions(iters, 2.0 if K > 1 else 1.0, upl)
    ca, cb = c0.copy(), c0.copy()
    Za, Zb = np.zeros(s * ph.col), np.zeros(s * ph.col)
    O.badmm(ph, lab, lsw, ca, Za, opts)
    O.ref_badmm(ph, lab, lsw, cb, Zb, opts)
    np.testing.assert_allclose(ca.p_w, cb.p_w, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(ca.p_supp, cb.p_supp, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(Za, Zb, rtol=1e-9, atol=1e-15)


@needs_ref
@pytest.mark.parametrize("kind", ["word_embed", "euclid_d16"])
def test_badmm_stored_cost_phases_match_reference_live(kind):
    """Word embeddings (centroid_Bregman.c:229-263, _dpdist2_sym) and a
    Euclidean phase with d = 16: the restatement against the reference build."""
    if kind == "word_embed":
        data, _ = synth.word_embed_docs(n=200, V=300, d=10, m_lo=5, m_hi=30, seed=7, s=6)
    else:
        data, _ = synth.serving_requests(n=200, d=16, m_lo=6, m_hi=40, seed=8)
    ph = data.phases[0]
    K, s = 4, 6
    c0 = O.centroid_rands(ph, K, s, seed=0)
    lab, lsw, _, _ = O.labeling(data, [c0])
    if len(np.unique(lab)) < K:
        pytest.skip("empty cluster")
    opts = BADMMOptions(12, 2.0, 5)
    ca, cb = c0.copy(), c0.copy()
    Za, Zb = np.zeros(s * ph.col), np.zeros(s * ph.col)
    O.badmm(ph, lab, lsw, ca, Za, opts)
    O.ref_badmm(ph, lab, lsw, cb, Zb, opts)
    np.testing.assert_allclose(ca.p_w, cb.p_w, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(ca.p_supp, cb.p_supp, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(Za, Zb, rtol=1e-9, atol=1e-15)


@needs_ref
@pytest.mark.parametrize("kind,n,V,K", [("dense", 100, 64, 4), ("sparse", 60, 1024, 3)])
def test_badmm_histogram_phases_match_reference_live(kind, n, V, K):
    """Histogram B-ADMM (no support update, centroid_Bregman.c:265-267) of
    the restatement against the reference build, whose cost blocks are
    pre-filled as clustering_util.c:183-199 does: a cold call, then a warm
    call with a third of the labels switched.  The same restatement run on
    the transposed ground cost must NOT agree, which shows the check sees
    the cost layout."""
    ph = hist_phase(kind, n, V, seed=V + n)
    c0, label = hist_state(ph, K, seed=5)
    opts = BADMMOptions(10, 2.0, 10)
    lsw = np.ones(n, np.int8)
    ca, cb = c0.copy(), c0.copy()
    Za, Zb = np.zeros(V * ph.col), np.zeros(V * ph.col)
    ra = O.badmm(ph, label, lsw, ca, Za, opts)
    Xb, _, _ = O.ref_badmm(ph, label, lsw, cb, Zb, opts)
    assert ra["rc"] == 0
    np.testing.assert_allclose(ca.p_w, cb.p_w, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(Za, Zb, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(ra["X"], Xb, rtol=1e-9, atol=1e-15)
    rng = np.random.default_rng(9)
    label2 = label.copy()
    mv = rng.random(n) < 1 / 3
    label2[mv] = (label2[mv] + 1) % K
    O.badmm(ph, label2, mv.astype(np.int8), ca, Za, opts)
    O.ref_badmm(ph, label2, mv.astype(np.int8), cb, Zb, opts)
    np.testing.assert_allclose(ca.p_w, cb.p_w, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(Za, Zb, rtol=1e-9, atol=1e-15)
    # the transposed cost gives other centroids
    pt = Phase(metric=ph.metric, dim=0, p_str=ph.p_str, p_w=ph.p_w, p_supp_sym=ph.p_supp_sym,
               vocab_size=V, dist_mat=ph.dist_mat.reshape(V, V).T.copy().reshape(-1), str=V)
    ct = c0.copy()
    O.badmm(pt, label, lsw, ct, np.zeros(V * ph.col), opts)
    c1 = c0.copy()
    O.ref_badmm(ph, label, lsw, c1, np.zeros(V * ph.col), opts)
    assert np.max(np.abs(ct.p_w - c1.p_w)) > 1e-4 * np.max(c1.p_w)


@needs_ref
def test_pdist2_submat_matches_reference_live():
    """or_pdist2_submat (the sparse-histogram cost gather) against the
    reference's _dpdist2_submat, bit for bit."""
    ph = hist_phase("sparse", 20, 50, seed=3)
    s = 50
    Ca = np.zeros(s * ph.col)
    O.ref_prefill_hist_cost(ph, s, Ca)
    Cb = O.hist_cost(ph, s)
    np.testing.assert_array_equal(Ca, Cb)


def test_badmm_histograms_match_reference_golden():
    """The reference build's histogram B-ADMM (cost blocks pre-filled as
    clustering_util.c:183-199 intends), kept as golden vectors: a dense
    16-bin and a sparse 64-bin phase with an asymmetric ground cost, a cold
    call then a warm call with a third of the labels switched."""
    g = gold("ref_badmm_hist.npz")
    for kind in ("dense", "sparse"):
        ph, c, label, lsw, label2, lsw2, opts = hist_golden_case(g, kind)
        Z = np.zeros(c.s * ph.col)
        r1 = O.badmm(ph, label, lsw, c, Z, opts)
        assert r1["rc"] == 0
        np.testing.assert_allclose(c.p_w, g[kind + "_c1_w"], rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(Z, g[kind + "_Z1"], rtol=1e-9, atol=1e-15)
        r2 = O.badmm(ph, label2, lsw2, c, Z, opts)
        np.testing.assert_allclose(c.p_w, g[kind + "_c2_w"], rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(Z, g[kind + "_Z2"], rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(r2["X"], g[kind + "_X2"], rtol=1e-9, atol=1e-15)


@needs_ref
def test_merge_matches_reference_live():
    rng = np.random.default_rng(0)
    for m, n in [(12, 6), (9, 2), (40, 39), (80, 40)]:
        supp = rng.standard_normal((m, 3))
        w = rng.gamma(1.0, size=m)
        w /= w.sum()
        a = O.merge(supp, w, n)
        b = O.ref_merge(supp, w, n)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


# -------------------------------------------------------- exact transport LP

def test_lp_matches_highs():
    """The oracle's transportation simplex against scipy HiGHS on the LP the
    reference poses (solver_mosek.cc:61-143)."""
    g = gold("lp_highs.npz")
    oc = ox = oy = 0
    for n, m, val in zip(g["n"], g["m"], g["val"]):
        n, m = int(n), int(m)
        Cm = g["C"][oc:oc + n * m]
        wX = g["wX"][ox:ox + n]
        wY = g["wY"][oy:oy + m]
        oc += n * m
        ox += n
        oy += m
        v = O.transport(Cm, wX, wY)
        assert abs(v - val) <= 1e-9 * max(1.0, abs(val)), (n, m, v, val)


def test_lp_matches_highs_serving_shapes():
    """The same at the serving shapes, 48 x 200 ... 48 x 500 and 64 x 500."""
    from cases import lp_large_cases
    for n, m, Cm, wX, wY, val in lp_large_cases(gold("lp_highs_large.npz")):
        v = O.transport(Cm, wX, wY)
        assert abs(v - val) <= 1e-9 * max(1.0, abs(val)), (n, m, v, val)


def test_lp_two_by_two_closed_form():
    """2x2: x11 = t on [max(0, b1 - a2), min(a1, b1)], objective linear in
    t, so the optimum is at an end point."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = rng.dirichlet([1, 1])
        b = rng.dirichlet([1, 1])
        Cm = rng.uniform(0, 5, 4)  # [i + j*2]
        def obj(t):
            x = [t, b[0] - t, a[0] - t, a[1] - b[0] + t]  # x00, x10, x01, x11
            return Cm[0] * x[0] + Cm[1] * x[1] + Cm[2] * x[2] + Cm[3] * x[3]
        lo, hi = max(0.0, b[0] - a[1]), min(a[0], b[0])
        expect = min(obj(lo), obj(hi))
        assert abs(O.transport(Cm, a, b) - expect) < 1e-12


def _w1_1d(x, wx, y, wy):
    """W1 on the line = integral of |F - G| (closed form)."""
    pts = np.union1d(x, y)
    F = np.array([wx[x <= p].sum() for p in pts])
    G = np.array([wy[y <= p].sum() for p in pts])
    return float(np.sum(np.abs(F - G)[:-1] * np.diff(pts)))


def _w2sq_1d(x, wx, y, wy):
    """W2^2 on the line: quantile coupling of the sorted supports."""
    ox, oy = np.argsort(x), np.argsort(y)
    x, wx, y, wy = x[ox], wx[ox], y[oy], wy[oy]
    cx, cy = np.cumsum(wx), np.cumsum(wy)
    qs = np.union1d(cx, cy)
    qs = qs[qs <= min(cx[-1], cy[-1]) + 1e-15]
    prev, tot = 0.0, 0.0
    for q in qs:
        mid = 0.5 * (prev + q)
        i = min(np.searchsorted(cx, mid), len(x) - 1)
        j = min(np.searchsorted(cy, mid), len(y) - 1)
        tot += (q - prev) * (x[i] - y[j]) ** 2
        prev = q
    return tot


@pytest.mark.parametrize("n,m", [(3, 5), (8, 8), (20, 13), (64, 40)])
def test_lp_one_dimensional_closed_forms(n, m):
    rng = np.random.default_rng(n * 100 + m)
    for _ in range(5):
        x = rng.uniform(0, 10, n)
        y = rng.uniform(0, 10, m)
        wx = rng.dirichlet(np.ones(n))
        wy = rng.dirichlet(np.ones(m))
        C1 = np.abs(x[:, None] - y[None, :]).T.reshape(-1)
        C2 = ((x[:, None] - y[None, :]) ** 2).T.reshape(-1)
        assert abs(O.transport(C1, wx, wy) - _w1_1d(x, wx, y, wy)) < 1e-9
        assert abs(O.transport(C2, wx, wy) - _w2sq_1d(x, wx, y, wy)) < 1e-8


def test_lp_invariances():
    rng = np.random.default_rng(4)
    n, m = 7, 9
    Cm = rng.uniform(0, 3, n * m)
    a = rng.dirichlet(np.ones(n))
    b = rng.dirichlet(np.ones(m))
    v = O.transport(Cm, a, b)
    assert abs(O.transport(2.5 * Cm, a, b) - 2.5 * v) < 1e-12
    # transposed problem has the same value (roles of rows/columns swap)
    Ct = Cm.reshape(m, n).T.reshape(-1)
    assert abs(O.transport(Ct, b, a) - v) < 1e-12
    # identical distributions under a metric cost: zero
    pts = rng.standard_normal((n, 2))
    D = ((pts[:, None] - pts[None]) ** 2).sum(-1).reshape(-1)
    assert abs(O.transport(D, a, a)) < 1e-14
    # a product plan is feasible: the LP value is below its cost
    assert v <= float(np.sum(Cm.reshape(m, n) * np.outer(b, a))) + 1e-14


# ----------------------------------------------------------- B-ADMM closed forms

def test_badmm_single_support_is_weighted_mean():
    """s = 1 forces X[0, j] = w_j, so the support update (:206-221) is the
    weight-averaged support of the cluster."""
    data, _ = synth.gaussian_mixture(n=120, seed=9)
    ph = data.phases[0]
    K = 3
    lab = (np.arange(ph.size) % K).astype(np.int32)
    c = Centroids(K, 1, 3, np.ones(K), np.zeros(K * 3))
    O.badmm(ph, lab, np.ones(ph.size, np.int8), c, np.zeros(ph.col), BADMMOptions(0, 2.0, 10))
    supp = ph.p_supp.reshape(-1, 3)
    obj = np.repeat(np.arange(ph.size), ph.p_str)
    for k in range(K):
        sel = lab[obj] == k
        mean = (ph.p_w[sel, None] * supp[sel]).sum(0) / ph.p_w[sel].sum()
        np.testing.assert_allclose(c.p_supp[3 * k:3 * k + 3], mean, rtol=1e-12)
    np.testing.assert_allclose(c.p_w, 1.0)


def test_badmm_invariants():
    """After an update: every centroid's weights sum to 1 (:190), every
    object's X columns sum to its weights (:158-159) and Z row sums equal the
    centroid weights used in that iteration (:169-170)."""
    data, _ = synth.gaussian_mixture(n=200, seed=5)
    ph = data.phases[0]
    K, s = 4, 5
    c = O.centroid_rands(ph, K, s)
    lab, lsw, _, _ = O.labeling(data, [c])
    Z = np.zeros(s * ph.col)
    r = O.badmm(ph, lab, lsw, c, Z, BADMMOptions(0, 2.0, 10))
    np.testing.assert_allclose(c.p_w.reshape(K, s).sum(1), 1.0, rtol=1e-13)
    X = r["X"].reshape(-1, s)  # column j of the global s x col matrix
    np.testing.assert_allclose(X.sum(1), ph.p_w, rtol=1e-12)


# ------------------------------------------------------ labeling / Elkan / loop

def test_elkan_labels_equal_brute_force():
    data, _ = synth.gaussian_mixture(n=300, seed=12)
    ph = data.phases[0]
    c = O.centroid_rands(ph, 6, 5)
    lab_b, _, _, _ = O.labeling(data, [c])
    lab_e = np.full(ph.size, -1, np.int32)
    elk = O.Elkan(ph.size, 6)
    _, ch, dc = O.labeling_prep(data, [c], lab_e, elk)
    np.testing.assert_array_equal(lab_e, lab_b)
    assert ch == ph.size and dc < ph.size * 6 + 15  # the bounds prune
    # bounds hold: l <= d <= u for every pair
    D = np.array([[O.distance(data, i, [c], k) for k in range(6)] for i in range(ph.size)])
    assert np.all(elk.l.reshape(-1, 6) <= D + 1e-12)
    assert np.all(elk.u >= D[np.arange(ph.size), lab_e] - 1e-12)


def test_clustering_triangle_equals_plain():
    """Elkan bounds are exact, so the closed loop with and without them
    visits the same labels and centroids."""
    data, _ = synth.gaussian_mixture(n=250, seed=21)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 5, 6)
    opts = BADMMOptions(10, 2.0, 10)
    ca, cb = c0.copy(), c0.copy()
    la, ita, tra, costa = O.clustering(data, [ca], 5, 8, opts, use_triangle=True)
    lb, itb, trb, costb = O.clustering(data, [cb], 5, 8, opts, use_triangle=False)
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(tra, trb)
    np.testing.assert_allclose(ca.p_supp, cb.p_supp, rtol=1e-12)
    assert abs(costa - costb) <= 1e-12 * costa


@needs_ref
def test_clustering_loop_with_reference_badmm():
    """The oracle's outer loop (clustering.c:455-503) against the same loop
    written here around the reference's own B-ADMM, plain labeling."""
    data, _ = synth.gaussian_mixture(n=200, seed=31)
    ph = data.phases[0]
    K, s = 4, 6
    c0 = O.centroid_rands(ph, K, s)
    opts = BADMMOptions(10, 2.0, 10)
    ca = c0.copy()
    la, it, _, _ = O.clustering(data, [ca], K, 5, opts, use_triangle=False)
    cb = c0.copy()
    lab = np.full(ph.size, -1, np.int32)
    Z = np.zeros(s * ph.col)
    for _ in range(it):
        lab, lsw, ch, _ = O.labeling(data, [cb], label=lab)
        if ch < 0.001 * ph.size:
            break
        O.ref_badmm(ph, lab, lsw, cb, Z, opts)
    np.testing.assert_array_equal(lab, la)
    np.testing.assert_allclose(cb.p_supp, ca.p_supp, rtol=1e-9)


def test_reads_reference_data_file():
    """clustering_io.c:14-202 on the first objects of data/mountaindat.d2."""
    d = read_d2(os.path.join(GOLD, "mountaindat_head.d2"), 300, [3, 3], strides=[6, 8])
    assert d.size == 300 and d.s_ph == 2
    for ph in d.phases:
        starts = ph.p_str_cum.astype(np.int64)
        np.testing.assert_allclose(np.add.reduceat(ph.p_w, starts), 1.0, rtol=1e-12)
        assert 2 <= ph.p_str.min() and ph.p_str.max() <= 16
