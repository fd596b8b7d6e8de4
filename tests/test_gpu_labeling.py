"""Assignment on the MI355X: exact distances, brute-force and Elkan labeling,
Elkan bound relaxation and the whole clustering loop, against the FP64
oracle.  Labels must be identical (the inputs are continuous, so tie-free);
distances and costs agree to 1e-10 relative (both sides solve the LP
exactly in FP64)."""
import os

import numpy as np
import pytest

from d2_kmeans_amd import synth
from d2_kmeans_amd.data import BADMMOptions, Centroids, D2Data, read_d2
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)


def setup(engine, data, cents, label=None):
    engine.set_data(data)
    for k, c in enumerate(cents):
        engine.set_centroids(k, c)
    lab = np.full(data.size, -1, np.int32) if label is None else label
    engine.set_labels(lab, np.zeros(data.size, np.int8))


def c1_case(n=1000, K=8, s=6, seed=0):
    data, _ = synth.gaussian_mixture(n=n, seed=synth.SEED_BASE + seed)
    c = O.centroid_rands(data.phases[0], K, s)
    return data, [c]


def test_distances_vs_oracle(engine):
    say("distances")
    data, cents = c1_case(n=300)
    setup(engine, data, cents)
    rng = np.random.default_rng(0)
    obj = rng.integers(0, data.size, 200)
    cent = rng.integers(0, 8, 200)
    got = engine.distances(obj, cent)
    ref = np.array([O.distance(data, int(i), cents, int(k)) for i, k in zip(obj, cent)])
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)


def test_distances_two_phases_reference_data(engine):
    """The reference's data/mountaindat.d2 (first 300 objects, 2 phases,
    d=3), both phases and each phase alone."""
    say("distances mountaindat 2 phases")
    data = read_d2(os.path.join(GOLD, "mountaindat_head.d2"), 300, [3, 3], strides=[6, 8])
    cents = [O.centroid_rands(data.phases[0], 5, 6), O.centroid_rands(data.phases[1], 5, 8)]
    setup(engine, data, cents)
    rng = np.random.default_rng(1)
    obj = rng.integers(0, data.size, 100)
    cent = rng.integers(0, 5, 100)
    for sel in (-1, 0, 1):
        got = engine.distances(obj, cent, selected_phase=sel)
        ref = np.array([O.distance(data, int(i), cents, int(k), selected_phase=sel)
                        for i, k in zip(obj, cent)])
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("use_elkan", [False, True])
def test_labeling_vs_oracle(engine, use_elkan):
    say(f"labeling elkan={use_elkan}")
    data, cents = c1_case()
    setup(engine, data, cents)
    if use_elkan:
        engine.elkan_reset()
    st = engine.labeling(use_elkan=use_elkan)
    lab, lsw = engine.get_labels()
    ref_lab, ref_lsw, ref_ch, ref_cost = O.labeling(data, cents)
    np.testing.assert_array_equal(lab, ref_lab)
    np.testing.assert_array_equal(lsw, ref_lsw)
    assert st["changes"] == ref_ch == data.size
    if not use_elkan:
        assert abs(st["cost"] - ref_cost) <= 1e-10 * ref_cost
    else:
        elk = O.Elkan(data.size, 8)
        lab_o = np.full(data.size, -1, np.int32)
        _, ch_o, dc_o = O.labeling_prep(data, cents, lab_o, elk)
        np.testing.assert_array_equal(lab_o, lab)
        assert st["dist_count"] < data.size * 8  # the bounds prune


def test_elkan_post_and_second_round(engine):
    """Move the centroids, relax the bounds (d2_labeling_post) and relabel:
    labels equal brute force, bounds equal the oracle's post-update bounds."""
    say("elkan post + second round")
    data, cents = c1_case(n=600, seed=3)
    setup(engine, data, cents)
    engine.elkan_reset()
    engine.labeling(use_elkan=True)
    lab, _ = engine.get_labels()
    opts = BADMMOptions(10, 2.0, 10)
    engine.save_centroids()
    engine.badmm(0, opts, precision=64)
    engine.elkan_post()
    c_new = engine.get_centroids(0)
    st = engine.labeling(use_elkan=True)
    lab2, _ = engine.get_labels()
    ref_lab, _, _, _ = O.labeling(data, [c_new], label=lab)
    np.testing.assert_array_equal(lab2, ref_lab)
    assert st["dist_count"] < data.size * 8


@pytest.mark.parametrize("use_triangle", [True, False])
def test_clustering_loop_vs_oracle(engine, use_triangle):
    """d2_clustering's outer loop end to end (3-d Gaussian mixture, N=1000,
    K=8, s=6, 10-iteration B-ADMM): FP64 device run against the FP64 oracle
    -- identical label trajectory; FP32 run -- identical final labels and the
    final mean cost within 1e-5."""
    say(f"clustering loop triangle={use_triangle}")
    data, cents = c1_case()
    opts = BADMMOptions(10, 2.0, 10)
    co = cents[0].copy()
    lab_o, it_o, tr_o, cost_o = O.clustering(data, [co], 8, 50, opts, use_triangle=use_triangle)
    for precision in (64, 32):
        setup(engine, data, cents)
        st = engine.clustering(8, 50, opts, use_triangle=use_triangle, precision=precision)
        lab, _ = engine.get_labels()
        np.testing.assert_array_equal(lab, lab_o)
        if precision == 64:
            assert st["iterations"] == it_o
            assert st["changes"] == list(tr_o)
            cg = engine.get_centroids(0)
            np.testing.assert_allclose(cg.p_supp, co.p_supp, rtol=1e-8, atol=1e-9)
            assert abs(st["final_cost"] - cost_o) <= 1e-9 * cost_o
        else:
            assert abs(st["final_cost"] - cost_o) <= 1e-5 * cost_o


def test_image_like_clustering_loop_vs_oracle_golden(engine):
    """The closed clustering loop (clustering.c:409-525: Elkan labeling with
    exact LPs, B-ADMM, label switching, bound relaxation) on the image-like
    workload at N = 3000 (16 x 16 grid strokes, m in [20, 80], K = 10,
    s = 40, 5 outer iterations) in FP64: the label changes of every outer
    iteration, the final labels, centroids and cost equal the FP64 oracle's
    (tests/golden/oracle_c2_loop.npz; the oracle needs minutes for this
    loop, so its trajectory is a fixture).  The FP32 run keeps the final
    labels and the cost within 1e-5."""
    from cases import c2_loop_case
    g = np.load(os.path.join(GOLD, "oracle_c2_loop.npz"), allow_pickle=False)
    data, c0, digest = c2_loop_case()
    assert digest == str(g["digest"]), "the generator no longer makes the fixture's input"
    say("image-like clustering loop, N=3000, K=10, s=40")
    opts = BADMMOptions(10, 2.0, 10)
    for precision in (64, 32):
        setup(engine, data, [c0])
        st = engine.clustering(10, 5, opts, use_triangle=True, precision=precision)
        lab, _ = engine.get_labels()
        say(f"fp{precision}: changes {st['changes']} oracle {list(g['changes'])}, "
            f"labeling {st['ms_labeling']:.1f} ms, B-ADMM {st['ms_badmm']:.1f} ms")
        np.testing.assert_array_equal(lab, g["label"].astype(np.int32))
        if precision == 64:
            assert st["iterations"] == int(g["iterations"])
            assert st["changes"] == [int(x) for x in g["changes"]]
            cg = engine.get_centroids(0)
            np.testing.assert_allclose(cg.p_supp, g["c_supp"], rtol=1e-8, atol=1e-9)
            np.testing.assert_allclose(cg.p_w, g["c_w"], rtol=1e-8, atol=1e-12)
            assert abs(st["final_cost"] - float(g["cost"])) <= 1e-9 * float(g["cost"])
        else:
            assert abs(st["final_cost"] - float(g["cost"])) <= 1e-5 * float(g["cost"])
