"""The plan kernel's forms on one workload: objects launched by size class
(register form with 8 or 16 column slots, LDS form, HBM-scratch form) against
the FP64 oracle, the forms against each other in FP32, and bit-identical
repeats.  The image-like shape (s = 40, m in [20, 80]) spans two size classes.

Tolerances: FP64 within 1e-9 relative of the oracle; FP32 forms within 1e-5
of each other on the objective and 1e-4 on supports (their sums run in
different orders)."""
import os

import numpy as np
import pytest

from d2_kmeans_amd import synth
from d2_kmeans_amd.data import BADMMOptions, Centroids, D2Data
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def strokes(n, seed=0):
    data, meta = synth.grid_strokes(n=n)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 10, 40, seed=seed)
    rng = np.random.default_rng(seed)
    lab = rng.integers(0, 10, size=ph.size).astype(np.int32)
    return data, ph, c0, lab


def update(engine, data, lab, c0, opts, precision, env=None):
    saved = {}
    for k, v in (env or {}).items():
        saved[k] = os.environ.get(k)
        os.environ[k] = v
    try:
        engine.set_data(data)
        engine.set_centroids(0, c0)
        engine.set_labels(lab, np.ones(len(lab), np.int8))
        st = engine.badmm(0, opts, precision=precision)
        c = engine.get_centroids(0)
        return st, c, engine.get_plan(0, precision, "Z")
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_size_classes_fp64_match_oracle(engine):
    """FP64, image-like shape: both size classes (register form for m <= 48,
    LDS form above) against the oracle, including a residual pass."""
    say("grid strokes N=600 s=40 fp64, size classes vs oracle")
    data, ph, c0, lab = strokes(600)
    opts = BADMMOptions(20, 2.0, 10)  # iteration 19 is a residual pass
    st, cg, _ = update(engine, data, lab, c0, opts, 64)
    co = c0.copy()
    ro = O.badmm(ph, lab, np.ones(ph.size, np.int8), co, np.zeros(c0.s * ph.col), opts)
    assert rel(cg.p_w, co.p_w) < 1e-9
    assert rel(cg.p_supp, co.p_supp) < 1e-9
    assert abs(st["obj"] - ro["obj"]) <= 1e-9 * abs(ro["obj"])
    np.testing.assert_allclose(st["trace"], ro["trace"], rtol=1e-8)


@pytest.mark.parametrize("env", [
    {"D2MI_PLAN_NO_BUCKETS": "1"},                       # one launch, 16-slot register form
    {"D2MI_PLAN_NOREG": "1"},                            # LDS form
    {"D2MI_PLAN_LEGACY": "1"},                           # general form
    {"D2MI_PLAN_LEGACY": "1", "D2MI_BADMM_HBM_SCRATCH": "1"},  # plan scratch in HBM
    {"D2MI_PLAN_NO_REVERSE": "1"},                       # every pass walks its chunk forwards
], ids=["one_launch", "lds", "general", "hbm_scratch", "forward_only"])
def test_fp32_forms_agree(engine, env):
    say(f"grid strokes N=4000 fp32: size classes vs {env}")
    data, ph, c0, lab = strokes(4000, seed=1)
    opts = BADMMOptions(10, 2.0, 10)
    sa, ca, _ = update(engine, data, lab, c0, opts, 32)
    sb, cb, _ = update(engine, data, lab, c0, opts, 32, env)
    assert abs(sa["obj"] - sb["obj"]) <= 1e-5 * abs(sb["obj"])
    assert rel(ca.p_w, cb.p_w) < 1e-5
    assert rel(ca.p_supp, cb.p_supp) < 1e-4


def test_size_classes_bit_reproducible(engine):
    say("grid strokes N=4000 fp32: repeat")
    data, ph, c0, lab = strokes(4000, seed=2)
    opts = BADMMOptions(10, 2.0, 10)
    outs = [update(engine, data, lab, c0, opts, 32) for _ in range(2)]
    np.testing.assert_array_equal(outs[0][1].p_w, outs[1][1].p_w)
    np.testing.assert_array_equal(outs[0][1].p_supp, outs[1][1].p_supp)
    np.testing.assert_array_equal(outs[0][2], outs[1][2])
    assert outs[0][0]["obj"] == outs[1][0]["obj"]


@pytest.mark.parametrize("precision", [32, 64])
def test_size_classes_large_objects(engine, precision):
    """Objects of 40 to 500 points at d = 64 (stored costs), s = 48: four
    size classes (register, LDS, LDS, HBM scratch) against one launch in the
    HBM-scratch form, and FP32 against FP64."""
    say(f"m up to 500, d=64, s=48: size classes, fp{precision}")
    data, _ = synth.serving_requests(n=160, d=64, m_lo=40, m_hi=500, seed=37)
    ph = data.phases[0]
    c0 = synth.seed_centroids(ph, 4, 48)
    lab, _, _, _ = O.labeling(data, [c0])
    opts = BADMMOptions(10, 2.0, 10)
    sa, ca, _ = update(engine, data, lab, c0, opts, precision)
    sb, cb, _ = update(engine, data, lab, c0, opts, precision, {"D2MI_PLAN_NO_BUCKETS": "1"})
    tol = 1e-11 if precision == 64 else 1e-5
    assert abs(sa["obj"] - sb["obj"]) <= tol * abs(sb["obj"])
    assert rel(ca.p_w, cb.p_w) < tol
    assert rel(ca.p_supp, cb.p_supp) < (tol if precision == 64 else 1e-4)
    if precision == 32:
        s64, c64, _ = update(engine, data, lab, c0, opts, 64)
        assert abs(sa["obj"] - s64["obj"]) <= 1e-5 * abs(s64["obj"])
        assert rel(ca.p_supp, c64.p_supp) < 1e-4


@pytest.mark.parametrize("d", [2, 3, 16])
def test_rho_from_object_moments_matches_direct_sum(engine, d):
    """rho (centroid_Bregman.c:109) from each object's mean support point and
    spread against the element-by-element ground-cost sum, FP64: the identity
    sum_k sum_j |c_k - x_j|^2 = m sum_k |c_k - mu|^2 + s sum_j |x_j - mu|^2 is
    exact, so the two agree to rounding (1e-13 relative), and both agree with
    the oracle's rho to 1e-12."""
    say(f"rho from object moments, d={d}")
    data, _ = synth.gaussian_mixture(n=1500, d=d, m_lo=4, m_hi=30, seed=5 + d)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 5, 7)
    lab, lsw, _, _ = O.labeling(data, [c0])
    opts = BADMMOptions(2, 2.0, 10)
    sa, _, _ = update(engine, data, lab, c0, opts, 64)
    sb, _, _ = update(engine, data, lab, c0, opts, 64, {"D2MI_RHO_DIRECT": "1"})
    assert abs(sa["rho"] - sb["rho"]) <= 1e-13 * sb["rho"], (sa["rho"], sb["rho"])
    co = c0.copy()
    ro = O.badmm(ph, lab, np.ones(ph.size, np.int8), co, np.zeros(c0.s * ph.col), opts)
    assert abs(sa["rho"] - ro["rho"]) <= 1e-12 * ro["rho"]


@pytest.mark.parametrize("V", [200, 1024])
def test_tall_form_sparse_histograms(engine, V):
    """Sparse histograms with s = V bins and 5-40 nonzeros: the tall register
    form (one thread per bin row, size classes of 8 columns) against the
    HBM-scratch form in FP32 (objective within 1e-5, weights within 1e-4:
    different summation orders) and against the FP64 oracle (1e-9).  V = 200
    leaves the last wave partly idle."""
    say(f"tall form: sparse histograms, V={V}")
    data, meta = synth.multiphase_histograms(n=300, V2=V, seed=23 + V)
    ph = data.phases[2]
    one = D2Data([ph])
    K = 3
    rng = np.random.default_rng(V)
    c0 = Centroids(K, V, 0, rng.dirichlet(np.ones(V), size=K).reshape(-1))
    lab = (meta["classes"] % K).astype(np.int32)
    opts = BADMMOptions(10, 2.0, 10)
    st_t, c_t, _ = update(engine, one, lab, c0, opts, 32)
    assert "tall" in st_t["plan_forms"], st_t["plan_forms"]
    st_h, c_h, _ = update(engine, one, lab, c0, opts, 32, {"D2MI_PLAN_NO_TALL": "1"})
    assert "tall" not in st_h["plan_forms"]
    assert abs(st_t["obj"] - st_h["obj"]) <= 1e-5 * abs(st_h["obj"])
    assert np.max(np.abs(c_t.p_w - c_h.p_w)) <= 1e-4 * np.max(c_h.p_w)
    st64, c64, _ = update(engine, one, lab, c0, opts, 64)
    assert "tall" in st64["plan_forms"], st64["plan_forms"]
    co = c0.copy()
    ro = O.badmm(ph, lab, np.ones(ph.size, np.int8), co, np.zeros(c0.s * ph.col), opts)
    assert rel(c64.p_w, co.p_w) < 1e-9
    assert abs(st64["obj"] - ro["obj"]) <= 1e-9 * abs(ro["obj"])
    assert abs(st_t["obj"] - ro["obj"]) <= 1e-5 * abs(ro["obj"])


@pytest.mark.parametrize("which", [1, 2])
def test_rho_from_line_sums_histograms(engine, which):
    """rho of a dense (1) or sparse (2) histogram phase from the ground-cost
    table's line sums against the element-by-element sum, FP64: the same
    terms in another order, so they agree to 1e-13 relative; and the update
    still matches the oracle's rho to 1e-12."""
    say(f"rho from line sums, histogram phase {which}")
    data, meta = synth.multiphase_histograms(n=250, V2=300, seed=41 + which)
    ph = data.phases[which]
    one = D2Data([ph])
    V, K = ph.vocab_size, 3
    rng = np.random.default_rng(which)
    c0 = Centroids(K, V, 0, rng.dirichlet(np.ones(V), size=K).reshape(-1))
    lab = (meta["classes"] % K).astype(np.int32)
    opts = BADMMOptions(1, 2.0, 10)
    sa, _, _ = update(engine, one, lab, c0, opts, 64)
    sb, _, _ = update(engine, one, lab, c0, opts, 64, {"D2MI_RHO_DIRECT": "1"})
    assert abs(sa["rho"] - sb["rho"]) <= 1e-13 * sb["rho"], (sa["rho"], sb["rho"])
    co = c0.copy()
    ro = O.badmm(ph, lab, np.ones(ph.size, np.int8), co, np.zeros(c0.s * ph.col), opts)
    assert abs(sa["rho"] - ro["rho"]) <= 1e-12 * ro["rho"]


@pytest.mark.parametrize("n,seed", [(4000, 3), (600, 9)])
def test_prefetch_form_bit_identical(engine, n, seed):
    """The register form's prefetching variant (the next object's plan copied
    into LDS while the current one computes) does the same arithmetic in the
    same order: FP32 updates agree bit for bit with the plain register form,
    across size classes (reg8p for m <= 48, reg16p above)."""
    say(f"prefetch form, grid strokes N={n}")
    data, ph, c0, lab = strokes(n, seed=seed)
    opts = BADMMOptions(10, 2.0, 10)
    sa, ca, za = update(engine, data, lab, c0, opts, 32, {"D2MI_PLAN_NO_BUCKETS": "0"})
    sb, cb, zb = update(engine, data, lab, c0, opts, 32, {"D2MI_PLAN_PREFETCH": "1"})
    assert "p" in sb["plan_forms"].split(":")[1].split("/")[0], sb["plan_forms"]
    np.testing.assert_array_equal(ca.p_w, cb.p_w)
    np.testing.assert_array_equal(ca.p_supp, cb.p_supp)
    np.testing.assert_array_equal(za, zb)
    assert sa["obj"] == sb["obj"]


def test_stream_probe(engine):
    """The two-array stream that bench.py reports as the plan kernels'
    practical ceiling: six variants and a positive rate (at this size the
    arrays sit in the last-level cache, so no HBM bound applies), and bad
    arguments refused."""
    from d2_kmeans_amd._lib import D2MIError
    say("stream probe 1<<22 elements")
    g, rep = engine.stream_probe(1 << 22, 2)
    assert 0 < g < 1e6
    assert len(rep.strip().splitlines()) == 6
    with pytest.raises(D2MIError):
        engine.stream_probe(16, 2)
