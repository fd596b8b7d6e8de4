"""Assignment server (csrc/serve.hip): exact nearest-centroid queries with
continuous batching.  Labels must equal the FP64 oracle's brute-force
labeling (d2_labeling, clustering.c:324-376: first minimum on ties) and
distances its exact LP values; each result must be the same served alone
(one query per launch) and in the mix."""
import numpy as np
import pytest

from d2_kmeans_amd import synth
from d2_kmeans_amd.data import D2Data
from d2_kmeans_amd.serve import Server
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def queries_of(ph):
    return [ph.object(i) for i in range(ph.size)]


def oracle_answers(qph, cent):
    """Brute force with the oracle: labels and the exact distance to the label."""
    lab, _, _, _ = O.labeling(D2Data([qph]), [cent])
    K, s, d = cent.K, cent.s, qph.dim
    CS = cent.p_supp.reshape(K, s, d)
    CW = cent.p_w.reshape(K, s)
    dist = np.empty(qph.size)
    for i in range(qph.size):
        w, X = qph.object(i)
        k = lab[i]
        Cm = ((CS[k][:, None, :] - X[None, :, :]) ** 2).sum(-1)  # s x m, element (a, j)
        dist[i] = np.sqrt(max(O.transport(Cm.T.reshape(-1).copy(), CW[k], w), 0.0))
    return lab, dist


def oracle_answers_threads(qph, cent, workers=8):
    """The same brute force, one query per thread: every centroid's exact LP
    (the oracle's simplex releases the interpreter lock while it solves),
    first minimum on ties (clustering.c:344)."""
    from concurrent.futures import ThreadPoolExecutor
    O.lib()
    K, s, d = cent.K, cent.s, qph.dim
    CS = cent.p_supp.reshape(K, s, d)
    CW = cent.p_w.reshape(K, s)

    def one(i):
        w, X = qph.object(i)
        best, lab = None, -1
        for k in range(K):
            Cm = O.pdist2(CS[k], X)  # element (a, j) at [a + j*s]
            v = np.sqrt(max(O.transport(Cm, CW[k], w), 0.0))
            if best is None or v < best:
                best, lab = v, k
        return lab, best

    with ThreadPoolExecutor(workers) as ex:
        res = list(ex.map(one, range(qph.size)))
    return np.array([r[0] for r in res]), np.array([r[1] for r in res])


@pytest.fixture(scope="module")
def setup(engine):
    say("serve setup: d=64 training phase, K=12 centroids of 16 points")
    train, _ = synth.serving_requests(n=400, seed=101, m_lo=5, m_hi=300)
    cent = synth.seed_centroids(train.phases[0], 12, 16)
    engine.set_data(train)
    engine.set_centroids(0, cent)
    qdata, _ = synth.serving_requests(n=60, seed=102, m_lo=5, m_hi=300)
    qph = qdata.phases[0]
    lab, dist = oracle_answers(qph, cent)
    return engine, cent, qph, lab, dist, train


def _resident(setup):
    """Puts the fixture's training set and centroids back on the device
    (other tests upload their own)."""
    eng, cent, qph, lab, dist, train = setup
    eng.set_data(train)
    eng.set_centroids(0, cent)
    return eng, cent, qph, lab, dist


def test_server_matches_oracle(setup):
    eng, cent, qph, lab, dist = _resident(setup)
    say("server vs oracle brute force")
    srv = Server(eng, 0, max_batch=16, streams=4)
    got_l, got_d = srv.assign(queries_of(qph))
    st = srv.stats()
    srv.close()
    np.testing.assert_array_equal(got_l, lab)
    np.testing.assert_allclose(got_d, dist, rtol=1e-9, atol=1e-12)
    assert st["completed"] == qph.size
    # pruning: far fewer LPs than brute force
    assert st["lps"] < qph.size * cent.K
    say("stats", st)


def test_alone_and_in_the_mix_agree(setup):
    eng, cent, qph, lab, dist = _resident(setup)
    say("alone vs mix")
    plain = Server(eng, 0, plain=True)
    alone_l, alone_d = [], []
    for q in queries_of(qph):
        l, d = plain.assign([q])
        alone_l.append(l[0])
        alone_d.append(d[0])
    plain.close()
    rng = np.random.default_rng(5)
    arrivals = np.sort(rng.exponential(0.002, qph.size).cumsum())
    mix = Server(eng, 0, max_batch=8, streams=4)
    mix_l, mix_d, lat, _ = mix.replay(queries_of(qph), arrivals)
    st = mix.stats()
    mix.close()
    # the LP's pivot sequence and final sum do not depend on the group it
    # runs in: the same query gives the same bits alone and in the mix
    np.testing.assert_array_equal(np.array(alone_l), mix_l)
    np.testing.assert_array_equal(np.array(alone_d), mix_d)
    np.testing.assert_array_equal(mix_l, lab)
    assert (lat > 0).all()
    say("mix stats", st)


def test_admission_waits_for_memory(setup):
    eng, cent, qph, lab, dist = _resident(setup)
    say("small HBM budget")
    big = int(qph.p_str.max())
    per_query = big * (qph.dim + 1) * 8 + 4096 * 8
    srv = Server(eng, 0, max_batch=32, streams=4, hbm_bytes=3 * per_query)
    got_l, got_d = srv.assign(queries_of(qph))
    st = srv.stats()
    srv.close()
    np.testing.assert_array_equal(got_l, lab)
    np.testing.assert_allclose(got_d, dist, rtol=1e-9, atol=1e-12)
    assert st["memory_waits"] > 0 or st["max_group"] < 32


def test_ties_go_to_the_lowest_index(engine):
    say("ties")
    data, _ = synth.serving_requests(n=50, seed=103, d=8, m_lo=5, m_hi=40)
    cent = synth.seed_centroids(data.phases[0], 6, 5)
    K, s, d = 6, 5, 8
    # centroids 1 and 4 identical: every query is equidistant from both
    cent.p_w[4 * s:5 * s] = cent.p_w[1 * s:2 * s]
    cent.p_supp[4 * s * d:5 * s * d] = cent.p_supp[1 * s * d:2 * s * d]
    engine.set_data(data)
    engine.set_centroids(0, cent)
    ph = data.phases[0]
    lab, dist = oracle_answers(ph, cent)
    srv = Server(engine, 0, max_batch=16)
    got_l, got_d = srv.assign(queries_of(ph))
    srv.close()
    np.testing.assert_array_equal(got_l, lab)
    assert not (got_l == 4).any()
    np.testing.assert_allclose(got_d, dist, rtol=1e-9, atol=1e-12)


def test_oversized_query_is_rejected(setup):
    eng, cent, qph, lab, dist = _resident(setup)
    say("oversized")
    from d2_kmeans_amd._lib import D2MIError
    srv = Server(eng, 0, max_m=100)
    rng = np.random.default_rng(0)
    with pytest.raises(D2MIError):
        srv.submit(np.full(101, 1 / 101), rng.standard_normal((101, srv.dim)))
    srv.close()


def test_unnormalised_weights_are_renormalised(setup):
    """Weights that do not sum to 1 are renormalised on submission, as
    d2_read renormalises every object (clustering_io.c:143-151): labels and
    distances equal the oracle's on the normalised query.  (The server's
    lower bounds assume unit mass; without this a bound could exceed the
    LP value and drop the true nearest centroid.)  A weight <= 1e-9 is
    refused."""
    eng, cent, qph, lab, dist = _resident(setup)
    say("unnormalised weights")
    from d2_kmeans_amd._lib import D2MIError
    rng = np.random.default_rng(8)
    srv = Server(eng, 0, max_batch=16, streams=4)
    qs = [(w * rng.choice([0.01, 0.3, 2.5, 40.0]), X) for w, X in queries_of(qph)]
    got_l, got_d = srv.assign(qs)
    np.testing.assert_array_equal(got_l, lab)
    np.testing.assert_allclose(got_d, dist, rtol=1e-9, atol=1e-12)
    w, X = qph.object(0)
    bad = w.copy()
    bad[0] = 0.0
    with pytest.raises(D2MIError):
        srv.submit(bad, X)
    srv.close()


def test_server_at_serving_shape(engine):
    """The serving workload's shape: K = 64 centroids of 48 points in d = 64,
    queries of lognormal(ln 50, 1) size clipped to [5, 500] plus four of the
    maximum 500 points, against the oracle's brute force over all 64
    centroids (exact LPs up to 48 x 500)."""
    say("serving shape: K=64 x 48, d=64, m <= 500")
    rng = np.random.default_rng(77)
    means = rng.standard_normal((64, 64)) * 2.0
    train, _ = synth.serving_requests(n=600, seed=201, means=means)
    cent = synth.seed_centroids(train.phases[0], 64, 48)
    engine.set_data(train)
    engine.set_centroids(0, cent)
    q1, _ = synth.serving_requests(n=72, seed=202, means=means)
    q2, _ = synth.serving_requests(n=4, seed=203, m_lo=500, m_hi=500, mu=np.log(500.0),
                                   sigma=0.01, means=means)
    qs = queries_of(q1.phases[0]) + queries_of(q2.phases[0])
    assert max(len(w) for w, _ in qs) == 500
    from d2_kmeans_amd.data import Phase
    qph = Phase(metric=0, dim=64, p_str=np.array([len(w) for w, _ in qs], np.int32),
                p_w=np.concatenate([w for w, _ in qs]),
                p_supp=np.concatenate([X.reshape(-1) for _, X in qs]), str=48)
    lab, dist = oracle_answers_threads(qph, cent)
    srv = Server(engine, 0, max_batch=16, streams=6)
    got_l, got_d = srv.assign(qs)
    st = srv.stats()
    srv.close()
    np.testing.assert_array_equal(got_l, lab)
    np.testing.assert_allclose(got_d, dist, rtol=1e-9, atol=1e-12)
    assert st["lps"] < len(qs) * 64
    say("stats", st)
