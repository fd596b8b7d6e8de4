"""The B-ADMM update's inner iterations as a replayed HIP graph, and the
per-size-class plan-kernel tuner.

Graph: the same updates launched one by one and replayed from the captured
graph must agree bit for bit.  Each graph case is replayed three times, with
the program's own eager work (uploads, an FP64 pdist2, synchronisations)
between the replays, and every replay is checked against the eager result.
The plans are re-uploaded before every replay, so the graph is refreshed with
new buffer addresses each time.

Tuner: it must leave the resident state untouched (plans restored bit for
bit), report one line per size class, and the update it then selects must
agree with the untuned one to the FP32 tolerance (objective within 1e-5
relative, supports within 1e-4 relative: the forms sum in different orders).
"""
import numpy as np
import pytest

from d2_kmeans_amd import synth
from d2_kmeans_amd.data import BADMMOptions
from d2_kmeans_amd.engine import Engine
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPTS = BADMMOptions(10, 2.0, 10)


def say(*a):
    print(*a, flush=True)


def two_updates(eng, data, lab, c0, precision, K):
    """A cold update, then a warm one with every 7th object switched."""
    eng.set_data(data)
    eng.set_centroids(0, c0)
    eng.set_labels(lab, np.ones(len(lab), np.int8))
    st1 = eng.badmm(0, OPTS, precision=precision)
    lab2 = lab.copy()
    lab2[::7] = (lab2[::7] + 1) % K
    eng.set_labels(lab2, (lab2 != lab).astype(np.int8))
    st2 = eng.badmm(0, OPTS, precision=precision)
    c = eng.get_centroids(0)
    return (st1, st2), (c.p_w.copy(), None if c.p_supp is None else c.p_supp.copy(),
                        eng.get_plan(0, precision, "Z"), eng.get_plan(0, precision, "Y"))


def eager_noise(eng):
    a = np.random.default_rng(1).standard_normal((50, 3))
    eng.pdist2(a, a[:17])
    eng.synchronize()


def strokes_case():
    data, meta = synth.grid_strokes(n=3000, seed=synth.SEED_BASE + 71)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 10, 40)
    lab = np.random.default_rng(3).integers(0, 10, size=ph.size).astype(np.int32)
    return data, c0, lab, 10


def word_case():
    data, _ = synth.word_embed_docs(n=600, V=2000, d=32, m_lo=10, m_hi=120, s=24)
    ph = data.phases[0]
    c0 = synth.seed_centroids(ph, 6, 24)
    lab = (np.arange(ph.size) % 6).astype(np.int32)
    return data, c0, lab, 6


@pytest.mark.parametrize("case,precision", [("strokes", 32), ("strokes", 64), ("word", 32),
                                            ("word", 64)])
def test_graph_replays_match_eager(case, precision):
    data, c0, lab, K = strokes_case() if case == "strokes" else word_case()
    with Engine(0) as eng:
        say(f"{case} fp{precision}: eager")
        eng.set_graph(False)
        sts, ref = two_updates(eng, data, lab, c0, precision, K)
        assert not sts[0]["graph"] and not sts[1]["graph"]
        eng.set_graph(True)
        for rep in range(3):
            say(f"{case} fp{precision}: graph replay {rep}")
            eager_noise(eng)
            sts_g, got = two_updates(eng, data, lab, c0, precision, K)
            assert sts_g[0]["graph"] and sts_g[1]["graph"], "the update did not run as a graph"
            for a, b in zip(ref, got):
                if a is not None:
                    np.testing.assert_array_equal(a, b)
            assert sts_g[1]["obj"] == sts[1]["obj"]
            assert sts_g[1]["n_plan_launches"] == OPTS.maxIters + 1
            # graph nodes cannot be timed with HIP events: no per-pass time is
            # reported for a graphed update, every pass is timed eagerly
            for st_g, st_e in zip(sts_g, sts):
                assert st_g["ms_plan"] == -1.0
                assert 0.0 < st_e["ms_plan"] <= st_e["ms_total"]


def test_graph_plain_variant_untouched():
    """The one-thread-per-object form is never captured."""
    data, c0, lab, K = strokes_case()
    with Engine(0) as eng:
        eng.set_data(data)
        eng.set_centroids(0, c0)
        eng.set_labels(lab, np.ones(len(lab), np.int8))
        st = eng.badmm(0, OPTS, precision=32, variant="plain")
        assert not st["graph"] and st["plan_forms"] == "plain"


@pytest.mark.parametrize("precision", [32, 64])
def test_tuner_restores_state_and_agrees(precision):
    # m in [20, 200] at s = 40: three size classes in FP32, four in FP64
    data, _ = synth.grid_strokes(n=3000, m_lo=20, m_hi=200, seed=synth.SEED_BASE + 72)
    ph = data.phases[0]
    K = 10
    c0 = O.centroid_rands(ph, K, 40)
    lab = np.random.default_rng(4).integers(0, K, size=ph.size).astype(np.int32)
    with Engine(0) as eng:
        eng.set_data(data)
        eng.set_centroids(0, c0)
        eng.set_labels(lab, np.ones(ph.size, np.int8))
        with pytest.raises(Exception):
            eng.tune(0, precision)  # no update has run: nothing to time on
        eng.set_data(data)
        eng.set_centroids(0, c0)
        eng.set_labels(lab, np.ones(ph.size, np.int8))
        eng.badmm(0, OPTS, precision=precision)
        c1 = eng.get_centroids(0)
        Z1 = eng.get_plan(0, precision, "Z")
        zero = np.zeros(ph.size, np.int8)

        def update_from_state():
            eng.set_centroids(0, c1)
            eng.set_plan(0, Z1, precision)
            eng.set_labels(lab, zero)
            st = eng.badmm(0, OPTS, precision=precision)
            return st, eng.get_centroids(0)

        st_u, cu = update_from_state()
        say(f"fp{precision} untuned forms: {st_u['plan_forms']}")
        eng.set_centroids(0, c1)
        eng.set_plan(0, Z1, precision)
        Yb = eng.get_plan(0, precision, "Y")
        rep = eng.tune(0, precision, reps=2)
        say(rep)
        lines = [l for l in rep.splitlines() if l.strip()]
        assert len(lines) == (3 if precision == 32 else 4), rep
        assert all("->" in l for l in lines)
        np.testing.assert_array_equal(eng.get_plan(0, precision, "Z"), Z1)
        # Y is the last update's; the tuner's passes must not have changed it either
        np.testing.assert_array_equal(eng.get_plan(0, precision, "Y"), Yb)
        st_t, ct = update_from_state()
        say(f"fp{precision} tuned forms: {st_t['plan_forms']}")
        tol_obj, tol_supp = (1e-12, 1e-10) if precision == 64 else (1e-5, 1e-4)
        assert abs(st_t["obj"] - st_u["obj"]) <= tol_obj * abs(st_u["obj"])
        assert np.max(np.abs(ct.p_supp - cu.p_supp)) <= tol_supp * np.max(np.abs(cu.p_supp))
        assert np.max(np.abs(ct.p_w - cu.p_w)) <= tol_supp
        # a tuned update repeats bit for bit
        st_t2, ct2 = update_from_state()
        np.testing.assert_array_equal(ct2.p_supp, ct.p_supp)
        assert st_t2["obj"] == st_t["obj"]
        eng.tune_clear()
        st_c, cc = update_from_state()
        assert st_c["plan_forms"] == st_u["plan_forms"]
        np.testing.assert_array_equal(cc.p_supp, cu.p_supp)
