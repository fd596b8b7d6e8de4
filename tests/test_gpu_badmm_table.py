"""B-ADMM with stored ground costs (Euclidean d > 8, word embeddings) and
exact distances for word-embedding phases, against the FP64 oracle (which
the CPU suite pins to the reference build for both kinds of phase).

Tolerances as for the fused path: FP64 to 1e-9 relative, the FP32 objective
within 1e-5 of FP64."""
import os

import numpy as np
import pytest

from d2_kmeans_amd import synth
from d2_kmeans_amd.data import BADMMOptions
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def device_update(engine, data, lab, lsw, c0, opts, precision):
    engine.set_data(data)
    engine.set_centroids(0, c0)
    engine.set_labels(lab, lsw)
    st = engine.badmm(0, opts, precision=precision)
    return st, engine.get_centroids(0)


def case(kind):
    if kind == "word_embed":
        data, _ = synth.word_embed_docs(n=600, V=500, d=24, m_lo=5, m_hi=40, seed=17, s=8)
    else:
        data, _ = synth.serving_requests(n=600, d=64, m_lo=6, m_hi=90, seed=18)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 5, 8, seed=0)
    lab, lsw, _, _ = O.labeling(data, [c0])
    assert len(np.unique(lab)) == 5
    return data, ph, c0, lab, lsw


@pytest.mark.parametrize("kind", ["word_embed", "euclid_d64"])
def test_stored_cost_update_matches_oracle(engine, kind):
    say(f"stored-cost B-ADMM vs oracle: {kind}")
    data, ph, c0, lab, lsw = case(kind)
    opts = BADMMOptions(12, 2.0, 5)
    st64, c64 = device_update(engine, data, lab, lsw, c0, opts, 64)
    co = c0.copy()
    ro = O.badmm(ph, lab, np.ones(ph.size, np.int8), co, np.zeros(c0.s * ph.col), opts)
    assert rel(c64.p_w, co.p_w) < 1e-9
    assert rel(c64.p_supp, co.p_supp) < 1e-9
    assert abs(st64["obj"] - ro["obj"]) <= 1e-9 * abs(ro["obj"])
    st32, c32 = device_update(engine, data, lab, lsw, c0, opts, 32)
    assert abs(st32["obj"] - st64["obj"]) <= 1e-5 * abs(st64["obj"])
    assert rel(c32.p_supp, c64.p_supp) < 1e-4


@pytest.mark.parametrize("stages", ["3", "2"])
@pytest.mark.parametrize("shape", ["word_d24_s8", "euclid_d64_s8", "word_d300_s64", "word_d100_s32"])
def test_packed_support_numerators_equal_per_object(engine, shape, stages):
    """Step 5 numerators (centroid_Bregman.c:199-212) over the label-sorted
    columns packed into blocks, with their rows copied straight into LDS,
    against the per-object blocks of the earlier kernel
    (D2MI_SUPP_PER_OBJECT=1): both FP32 matrix-core sums of the same
    products, grouped differently, so they agree to FP32 rounding.  Shapes
    give 1, 2 and 5 output tiles per wave (d = 24 and 64, 100, 300); the
    blocks go through three LDS stages, or two (D2MI_SUPP_STAGES=2)."""
    say(f"packed vs per-object support numerators: {shape}, {stages} stages")
    if shape == "euclid_d64_s8":
        data, ph, c0, lab, lsw = case("euclid_d64")
    else:
        d, s = {"word_d24_s8": (24, 8), "word_d300_s64": (300, 64), "word_d100_s32": (100, 32)}[shape]
        data, _ = synth.word_embed_docs(n=500, V=1500, d=d, m_lo=5, m_hi=120, seed=31, s=s)
        ph = data.phases[0]
        c0 = O.centroid_rands(ph, 4, s, seed=1)
        lab = (np.arange(ph.size) % 4).astype(np.int32)
        lsw = np.ones(ph.size, np.int8)
    opts = BADMMOptions(10, 2.0, 3)
    os.environ["D2MI_SUPP_STAGES"] = stages
    try:
        sa, ca = device_update(engine, data, lab, lsw, c0, opts, 32)
    finally:
        os.environ.pop("D2MI_SUPP_STAGES", None)
    os.environ["D2MI_SUPP_PER_OBJECT"] = "1"
    try:
        sb, cb = device_update(engine, data, lab, lsw, c0, opts, 32)
    finally:
        os.environ.pop("D2MI_SUPP_PER_OBJECT", None)
    assert rel(ca.p_supp, cb.p_supp) < 1e-5
    assert rel(ca.p_w, cb.p_w) < 1e-5
    assert abs(sa["obj"] - sb["obj"]) <= 1e-6 * abs(sb["obj"])


@pytest.mark.parametrize("precision", [64, 32])
def test_rho_from_flat_table_sum(engine, precision):
    """rho (centroid_Bregman.c:109) of a word-embedding phase from one flat
    sum over the cost table (every object's block back to back) against the
    object-by-object sum of the setup kernel (D2MI_RHO_DIRECT=1): the same
    terms in another order, so 1e-13 relative; FP64 also against the oracle.
    (Euclidean phases take rho from object moments instead.)"""
    say(f"rho from the flat table sum: FP{precision}")
    data, ph, c0, lab, lsw = case("word_embed")
    opts = BADMMOptions(2, 2.0, 5)
    sa, _ = device_update(engine, data, lab, lsw, c0, opts, precision)
    os.environ["D2MI_RHO_DIRECT"] = "1"
    try:
        sb, _ = device_update(engine, data, lab, lsw, c0, opts, precision)
    finally:
        os.environ.pop("D2MI_RHO_DIRECT", None)
    assert abs(sa["rho"] - sb["rho"]) <= 1e-13 * sb["rho"], (sa["rho"], sb["rho"])
    if precision == 64:
        co = c0.copy()
        ro = O.badmm(ph, lab, np.ones(ph.size, np.int8), co, np.zeros(c0.s * ph.col), opts)
        assert abs(sa["rho"] - ro["rho"]) <= 1e-12 * ro["rho"]


@pytest.mark.parametrize("s", [8, 12])
def test_word_gather_forms_agree(engine, s):
    """Cost blocks gathered from the pdist2 table 16 bytes per thread (s a
    multiple of 4) are the same copy as the element-wise gather
    (D2MI_GATHER_SCALAR=1): an update from the same state ends bit for bit
    the same."""
    say(f"word gather forms, s={s}")
    data, _ = synth.word_embed_docs(n=400, V=800, d=24, m_lo=5, m_hi=90, seed=37, s=s)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 4, s, seed=2)
    lab = (np.arange(ph.size) % 4).astype(np.int32)
    lsw = np.ones(ph.size, np.int8)
    opts = BADMMOptions(8, 2.0, 3)
    sa, ca = device_update(engine, data, lab, lsw, c0, opts, 32)
    os.environ["D2MI_GATHER_SCALAR"] = "1"
    try:
        sb, cb = device_update(engine, data, lab, lsw, c0, opts, 32)
    finally:
        os.environ.pop("D2MI_GATHER_SCALAR", None)
    np.testing.assert_array_equal(ca.p_supp, cb.p_supp)
    np.testing.assert_array_equal(ca.p_w, cb.p_w)
    assert sa["obj"] == sb["obj"]


def test_stored_and_fused_costs_agree(engine):
    """The same Euclidean d = 3 update with costs recomputed in the plan
    kernel and with the stored table (D2MI_BADMM_TABLE=1)."""
    say("stored vs fused costs, d=3")
    data, _ = synth.gaussian_mixture(n=800, seed=23)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 6, 6, seed=0)
    lab, lsw, _, _ = O.labeling(data, [c0])
    opts = BADMMOptions(15, 2.0, 5)
    sa, ca = device_update(engine, data, lab, lsw, c0, opts, 64)
    os.environ["D2MI_BADMM_TABLE"] = "1"
    try:
        sb, cb = device_update(engine, data, lab, lsw, c0, opts, 64)
    finally:
        os.environ.pop("D2MI_BADMM_TABLE", None)
    assert rel(cb.p_w, ca.p_w) < 1e-11
    assert rel(cb.p_supp, ca.p_supp) < 1e-11
    assert abs(sb["obj"] - sa["obj"]) <= 1e-11 * abs(sa["obj"])


def test_word_embed_distances_and_labels(engine):
    say("word-embedding exact distances and labeling")
    data, ph, c0, lab, lsw = case("word_embed")
    engine.set_data(data)
    engine.set_centroids(0, c0)
    engine.set_labels(np.zeros(ph.size, np.int32), np.zeros(ph.size, np.int8))
    obj = np.repeat(np.arange(40), c0.K).astype(np.int32)
    cen = np.tile(np.arange(c0.K), 40).astype(np.int32)
    got = engine.distances(obj, cen)
    ref = np.array([O.distance(data, int(i), [c0], int(j)) for i, j in zip(obj, cen)])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13)
    st = engine.labeling(use_elkan=False)
    np.testing.assert_array_equal(engine.get_labels()[0], lab)


def test_word_embed_clustering_matches_oracle(engine):
    say("word-embedding clustering loop vs oracle")
    data, ph, c0, lab, lsw = case("word_embed")
    K = c0.K
    opts = BADMMOptions(10, 2.0, 10)
    engine.set_data(data)
    engine.set_centroids(0, c0)
    res = engine.clustering(K, max_iter=6, opts=opts, use_triangle=True, precision=64)
    co = [c0.copy()]
    lab_o, it, _, cost = O.clustering(data, co, K, 6, opts, use_triangle=True)
    np.testing.assert_array_equal(engine.get_labels()[0], lab_o)
    assert rel(engine.get_centroids(0).p_supp, co[0].p_supp) < 1e-8


@pytest.mark.parametrize("precision", [64, 32])
def test_hbm_scratch_plan_matches_lds_plan(engine, precision):
    """The plan kernel with its scratch in HBM (used when an object's plan
    does not fit LDS), forced on a case the LDS form handles."""
    say(f"HBM-scratch plan vs LDS plan, fp{precision}")
    data, _ = synth.gaussian_mixture(n=800, seed=29)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 6, 6, seed=0)
    lab, lsw, _, _ = O.labeling(data, [c0])
    opts = BADMMOptions(20, 2.0, 10)  # includes a residual pass (iteration 19)
    sa, ca = device_update(engine, data, lab, lsw, c0, opts, precision)
    os.environ["D2MI_BADMM_HBM_SCRATCH"] = "1"
    try:
        sb, cb = device_update(engine, data, lab, lsw, c0, opts, precision)
    finally:
        os.environ.pop("D2MI_BADMM_HBM_SCRATCH", None)
    tol = 1e-12 if precision == 64 else 1e-5
    assert rel(cb.p_w, ca.p_w) < tol
    assert rel(cb.p_supp, ca.p_supp) < tol
    assert abs(sb["obj"] - sa["obj"]) <= tol * abs(sa["obj"])
    assert len(sb["trace"]) == 1 and len(sa["trace"]) == 1
    np.testing.assert_allclose(sb["trace"], sa["trace"], rtol=1e-6)


def test_large_objects_d64_match_oracle(engine):
    """Objects of up to 500 points at d = 64 with s = 48: stored costs and the
    HBM-scratch plan (the serving workload's training shape)."""
    say("m up to 500, d=64, s=48 vs oracle")
    data, _ = synth.serving_requests(n=160, d=64, m_lo=40, m_hi=500, seed=31)
    ph = data.phases[0]
    c0 = synth.seed_centroids(ph, 4, 48)
    lab, lsw, _, _ = O.labeling(data, [c0])
    assert len(np.unique(lab)) == 4
    opts = BADMMOptions(10, 2.0, 10)
    st64, c64 = device_update(engine, data, lab, lsw, c0, opts, 64)
    co = c0.copy()
    ro = O.badmm(ph, lab, np.ones(ph.size, np.int8), co, np.zeros(c0.s * ph.col), opts)
    assert rel(c64.p_w, co.p_w) < 1e-9
    assert rel(c64.p_supp, co.p_supp) < 1e-9
    assert abs(st64["obj"] - ro["obj"]) <= 1e-9 * abs(ro["obj"])


def test_word_embed_update_at_configured_shape(engine):
    """Word-embedding B-ADMM at the configured vocabulary and centroid
    shape: V = 50,000 words of d = 300 (N(0,1) rows, L2-normalised), Zipf
    document lengths in [10, 200], s = 64 support points per centroid,
    K = 6, on 300 documents (centroid_Bregman.c:229-263 with _dpdist2_sym,
    blas_like64.c:210-220).  FP64 agrees with the oracle to 1e-9; the FP32
    update (cost table from the MFMA pdist2 with FP32 products, several
    k-slices of the 300-long rows) and the same with BF16 products keep the
    objective within 1e-5 of the oracle's (the metric's gate); the centroid
    supports stay within 1e-4 with FP32 products and within 1e-3 with BF16
    products (8-bit significands: each product carries up to 2^-8 relative
    error before the FP32 accumulation)."""
    say("word embeddings at V=50k, d=300, s=64")
    data, _ = synth.word_embed_docs(n=300, V=50_000, d=300, seed=29, s=64)
    ph = data.phases[0]
    K = 6
    c0 = synth.seed_centroids(ph, K, 64)
    rng = np.random.default_rng(3)
    lab = (np.arange(ph.size) % K).astype(np.int32)
    rng.shuffle(lab)
    lsw = np.ones(ph.size, np.int8)
    opts = BADMMOptions(10, 2.0, 10)
    co = c0.copy()
    ro = O.badmm(ph, lab, lsw, co, np.zeros(64 * ph.col), opts)
    st64, c64 = device_update(engine, data, lab, lsw, c0, opts, 64)
    say(f"fp64 obj {st64['obj']:.12g} oracle {ro['obj']:.12g}")
    assert rel(c64.p_w, co.p_w) < 1e-9
    assert rel(c64.p_supp, co.p_supp) < 1e-9
    assert abs(st64["obj"] - ro["obj"]) <= 1e-9 * abs(ro["obj"])
    for bf16 in (False, True):
        if bf16:
            os.environ["D2MI_PDIST2_BF16"] = "1"
        try:
            st32, c32 = device_update(engine, data, lab, lsw, c0, opts, 32)
        finally:
            os.environ.pop("D2MI_PDIST2_BF16", None)
        err = abs(st32["obj"] - ro["obj"]) / abs(ro["obj"])
        say(f"fp32 {'bf16' if bf16 else 'fp32'} products: obj rel err {err:.2e}, "
            f"supp rel err {rel(c32.p_supp, co.p_supp):.2e}")
        assert err <= 1e-5
        assert rel(c32.p_supp, co.p_supp) < (1e-3 if bf16 else 1e-4)
        assert rel(c32.p_w, co.p_w) < (1e-3 if bf16 else 1e-4)
