"""Cases shared by the CPU and GPU tests and by tests/golden/make_golden.py:
histogram phases with an asymmetric ground cost and the loader of
ref_badmm_hist.npz; the serving-shape LPs of lp_highs_large.npz; the
image-like clustering-loop case of oracle_c2_loop.npz."""
import hashlib

import numpy as np

from d2_kmeans_amd.data import (BADMMOptions, Centroids, D2_HISTOGRAM, D2_SPARSE_HISTOGRAM,
                                Phase)


def hist_phase(kind, n, V, seed, m_lo=3, m_hi=10):
    """A dense (every bin) or sparse (m_lo..m_hi bins) histogram phase over V
    bins whose ground cost is NOT symmetric, D[a, b] = (a-b)^2/V +
    0.5*max(a-b, 0) for object bin a and centroid bin b, so that a
    transposed cost layout changes every result."""
    rng = np.random.default_rng(seed)
    i = np.arange(V, dtype=np.float64)
    D = (i[:, None] - i[None, :]) ** 2 / V + 0.5 * np.maximum(i[:, None] - i[None, :], 0)
    if kind == "dense":
        h = rng.dirichlet(np.ones(V), size=n)
        return Phase(metric=D2_HISTOGRAM, dim=0, p_str=np.full(n, V, np.int32),
                     p_w=h.reshape(-1), vocab_size=V, dist_mat=D.reshape(-1), str=V)
    m = rng.integers(m_lo, m_hi + 1, size=n).astype(np.int32)
    sym = np.concatenate([np.sort(rng.choice(V, k, replace=False)) for k in m]).astype(np.int32)
    w = np.concatenate([rng.dirichlet(np.ones(k)) for k in m])
    return Phase(metric=D2_SPARSE_HISTOGRAM, dim=0, p_str=m, p_w=w, p_supp_sym=sym,
                 vocab_size=V, dist_mat=D.reshape(-1), str=V)


def hist_state(ph, K, seed):
    rng = np.random.default_rng(seed)
    c0 = Centroids(K, ph.vocab_size, 0, rng.dirichlet(np.ones(ph.vocab_size), size=K).reshape(-1))
    label = (np.arange(ph.size) % K).astype(np.int32)
    rng.shuffle(label)
    return c0, label


def hist_golden_case(g, kind):
    V = int(g[kind + "_V"])
    kw = dict(dim=0, p_str=g[kind + "_p_str"], p_w=g[kind + "_p_w"], vocab_size=V,
              dist_mat=g[kind + "_dist_mat"], str=V)
    if kind == "dense":
        ph = Phase(metric=D2_HISTOGRAM, **kw)
    else:
        ph = Phase(metric=D2_SPARSE_HISTOGRAM, p_supp_sym=g[kind + "_sym"], **kw)
    K = int(g[kind + "_K"])
    c = Centroids(K, V, 0, g[kind + "_c0_w"].copy())
    opts = BADMMOptions(int(g["maxIters"]), float(g["rhoCoeff"]), int(g["updatePerLoops"]))
    return (ph, c, g[kind + "_label"], g[kind + "_lsw"], g[kind + "_label2"],
            g[kind + "_lsw2"], opts)


def lp_large_cases(g):
    """(n, m, C, wX, wY, HiGHS value) of lp_highs_large.npz, the cost being
    the oracle's pdist2 of the stored support points (element (i, j) at
    [i + j*n])."""
    from oracle import oracle as O
    oa = ob = ox = oy = 0
    for n, m, val in zip(g["n"], g["m"], g["val"]):
        n, m = int(n), int(m)
        A = g["A"][oa:oa + n * 8].reshape(n, 8)
        B = g["B"][ob:ob + m * 8].reshape(m, 8)
        wX = g["wX"][ox:ox + n]
        wY = g["wY"][oy:oy + m]
        oa += n * 8
        ob += m * 8
        ox += n
        oy += m
        yield n, m, O.pdist2(A, B), wX, wY, float(val)


def c2_loop_case():
    """The image-like workload at N = 3000 (16 x 16 grid strokes, m in
    [20, 80], K = 10, s = 40) and its starting centroids (one object per
    stroke class, synth.seed_centroids), with a hash of both."""
    from d2_kmeans_amd import synth
    data, meta = synth.grid_strokes(n=3000, seed=20261205)
    ph = data.phases[0]
    c0 = synth.seed_centroids(ph, 10, 40, labels=meta["classes"])
    h = hashlib.sha256()
    for a in (ph.p_str, ph.p_w, ph.p_supp, c0.p_w, c0.p_supp):
        h.update(np.ascontiguousarray(a).tobytes())
    return data, c0, h.hexdigest()
