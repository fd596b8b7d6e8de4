"""Seeded synthetic D2 workloads (the reference ships no data for these shapes).

Each generator returns a :class:`~d2_kmeans_amd.data.D2Data` plus a dict of
the parameters it used (seed included), so a run can be re-created exactly.
Seeds default to 20261203 + workload index.

* :func:`gaussian_mixture` -- 1 Euclidean phase, d=3, N=1000, support sizes
  U{4..12}, points from a mixture of 8 isotropic Gaussians (means U[-5,5]^3,
  sigma 0.5; each object draws a dominant component and takes 80% of its
  points from it), Dirichlet(1) weights.  K=8, s=6.
* :func:`grid_strokes` -- 1 Euclidean phase, d=2, supports on the 16x16
  integer grid, N=200k, support sizes U{20..80} distinct cells, 10 stroke
  templates; 90% of an object's cells come from its template, 10% are
  jittered off it; intensities U(0,1] normalised.  K=10, s=40.
* :func:`word_embed_docs` -- word-embedding phase: V embeddings in d dims
  (N(0,1) rows, L2-normalised), documents with Zipf(1.1) lengths in
  [10,200], Zipf word ids, lognormal(0,1) weights.  K=20, s=64.
* :func:`serving_requests` -- request distributions for assignment: sizes
  lognormal(ln 50, 1) clipped to [5,500], Dirichlet(1) weights, points from
  a d-dim Gaussian mixture.
* :func:`multiphase_histograms` -- three phases per object: Euclidean d=3
  with 8-32 Gaussian-mixture points; a dense 64-bin histogram with ground
  cost (i-j)^2; a sparse 1024-bin histogram with 5-40 nonzeros, ground cost
  |i-j| and Dirichlet(0.5) weights.  K=32.
"""
from __future__ import annotations

import numpy as np

from .data import (D2_EUCLIDEAN_L2, D2_HISTOGRAM, D2_SPARSE_HISTOGRAM, D2_WORD_EMBED, D2Data,
                   Phase)

SEED_BASE = 20261203


def _dirichlet_ragged(rng, p_str):
    w = rng.gamma(1.0, 1.0, size=int(p_str.sum()))
    w = np.maximum(w, 1e-6)
    starts = np.concatenate([[0], np.cumsum(p_str)[:-1]])
    sums = np.add.reduceat(w, starts)
    return w / np.repeat(sums, p_str)


def gaussian_mixture(n=1000, d=3, n_comp=8, m_lo=4, m_hi=12, sigma=0.5, seed=SEED_BASE + 0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-5, 5, size=(n_comp, d))
    p_str = rng.integers(m_lo, m_hi + 1, size=n).astype(np.int32)
    col = int(p_str.sum())
    dom = np.repeat(rng.integers(0, n_comp, size=n), p_str)
    other = rng.integers(0, n_comp, size=col)
    comp = np.where(rng.random(col) < 0.8, dom, other)
    pts = means[comp] + sigma * rng.standard_normal((col, d))
    w = _dirichlet_ragged(rng, p_str)
    ph = Phase(metric=D2_EUCLIDEAN_L2, dim=d, p_str=p_str, p_w=w, p_supp=pts.reshape(-1), str=6)
    return D2Data([ph]), dict(kind="gaussian_mixture", n=n, d=d, n_comp=n_comp, m=(m_lo, m_hi),
                              sigma=sigma, seed=seed)


def _stroke_templates(rng, n_templates, grid):
    temps = []
    for _ in range(n_templates):
        mask = np.zeros((grid, grid), dtype=bool)
        while mask.sum() < 96:
            p0 = rng.uniform(0, grid - 1, 2)
            p1 = rng.uniform(0, grid - 1, 2)
            t = np.linspace(0, 1, 4 * grid)[:, None]
            pts = p0 + t * (p1 - p0)
            for off in ((0, 0), (1, 0), (0, 1)):
                idx = np.clip(np.round(pts + off).astype(int), 0, grid - 1)
                mask[idx[:, 1], idx[:, 0]] = True
        temps.append(np.flatnonzero(mask.reshape(-1)))
    return temps


def grid_strokes(n=200_000, grid=16, n_templates=10, m_lo=20, m_hi=80, jitter=0.10,
                 seed=SEED_BASE + 1, chunk=50_000):
    rng = np.random.default_rng(seed)
    temps = _stroke_templates(rng, n_templates, grid)
    G = grid * grid
    in_t = np.zeros((n_templates, G), dtype=bool)
    for t, cells in enumerate(temps):
        in_t[t, cells] = True
    p_str = rng.integers(m_lo, m_hi + 1, size=n).astype(np.int32)
    cls = rng.integers(0, n_templates, size=n)
    cells_all = []
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = p_str[a:b]
        n_in = np.minimum(np.round((1 - jitter) * m).astype(np.int64), in_t[cls[a:b]].sum(1))
        score = rng.random((b - a, G), dtype=np.float32)
        member = in_t[cls[a:b]]
        # template cells rank first (score in [2,3)), the rest after ([0,1))
        s_in = np.where(member, score + 2.0, -1.0)
        s_out = np.where(member, -1.0, score)
        o_in = np.argsort(-s_in, axis=1)
        o_out = np.argsort(-s_out, axis=1)
        rank = np.arange(G)[None, :]
        take_in = rank < n_in[:, None]
        take_out = rank < (m - n_in)[:, None]
        for r in range(b - a):
            cells_all.append(np.concatenate([o_in[r, take_in[r]], o_out[r, take_out[r]]]))
    cells = np.concatenate(cells_all)
    pts = np.stack([cells % grid, cells // grid], axis=1).astype(np.float64)
    w = rng.uniform(0, 1, size=len(cells))
    w = 1.0 - w  # U(0,1]
    starts = np.concatenate([[0], np.cumsum(p_str)[:-1]])
    w = w / np.repeat(np.add.reduceat(w, starts), p_str)
    ph = Phase(metric=D2_EUCLIDEAN_L2, dim=2, p_str=p_str, p_w=w, p_supp=pts.reshape(-1), str=40)
    return D2Data([ph]), dict(kind="grid_strokes", n=n, grid=grid, n_templates=n_templates,
                              m=(m_lo, m_hi), jitter=jitter, seed=seed, classes=cls.astype(np.int32))


def _zipf_trunc(rng, a, lo, hi, size):
    k = np.arange(lo, hi + 1, dtype=np.float64)
    p = k ** (-a)
    p /= p.sum()
    return rng.choice(k.astype(np.int64), size=size, p=p)


def word_embed_docs(n=20_000, V=50_000, d=300, m_lo=10, m_hi=200, seed=SEED_BASE + 2, s=64):
    rng = np.random.default_rng(seed)
    vocab = rng.standard_normal((V, d))
    vocab /= np.linalg.norm(vocab, axis=1, keepdims=True)
    p_str = _zipf_trunc(rng, 1.1, m_lo, m_hi, n).astype(np.int32)
    rankp = 1.0 / np.arange(1, V + 1)
    cdf = np.cumsum(rankp / rankp.sum())
    cdf[-1] = 1.0
    ids = []
    for m in p_str:  # distinct word ids per document, Zipf over the vocabulary
        ids.append(np.unique(np.searchsorted(cdf, rng.random(int(m) * 2), side="right"))[:m])
    # np.unique may return fewer than m ids: top up with uniform draws
    for t, m in enumerate(p_str):
        while len(ids[t]) < m:
            ids[t] = np.unique(np.concatenate([ids[t], rng.integers(0, V, m - len(ids[t]))]))[:m]
    sym = np.concatenate(ids).astype(np.int32)
    w = rng.lognormal(0.0, 1.0, size=len(sym))
    starts = np.concatenate([[0], np.cumsum(p_str)[:-1]])
    w = w / np.repeat(np.add.reduceat(w, starts), p_str)
    ph = Phase(metric=D2_WORD_EMBED, dim=d, p_str=p_str, p_w=w, p_supp_sym=sym, vocab_size=V,
               vocab_vec=vocab.reshape(-1), str=s)
    return D2Data([ph]), dict(kind="word_embed_docs", n=n, V=V, d=d, m=(m_lo, m_hi), seed=seed)


def serving_requests(n=1000, d=64, n_comp=64, m_lo=5, m_hi=500, mu=np.log(50.0), sigma=1.0,
                     seed=SEED_BASE + 4, means=None):
    rng = np.random.default_rng(seed)
    if means is None:
        means = rng.standard_normal((n_comp, d)) * 2.0
    p_str = np.clip(np.round(rng.lognormal(mu, sigma, size=n)), m_lo, m_hi).astype(np.int32)
    col = int(p_str.sum())
    dom = np.repeat(rng.integers(0, len(means), size=n), p_str)
    other = rng.integers(0, len(means), size=col)
    comp = np.where(rng.random(col) < 0.8, dom, other)
    pts = means[comp] + 0.5 * rng.standard_normal((col, d))
    w = _dirichlet_ragged(rng, p_str)
    ph = Phase(metric=D2_EUCLIDEAN_L2, dim=d, p_str=p_str, p_w=w, p_supp=pts.reshape(-1), str=48)
    # mixture component most of each query's points come from
    dominant = dom[np.concatenate([[0], np.cumsum(p_str)[:-1]])] if n else dom
    return D2Data([ph]), dict(kind="serving_requests", n=n, d=d, m=(m_lo, m_hi), mu=float(mu),
                              sigma=sigma, seed=seed, dominant=dominant)


def seed_centroids(phase: Phase, K: int, s: int, labels=None, seed=SEED_BASE):
    """Starting centroids for benchmarks: for each cluster k, the first
    object (in a seeded order, restricted to label k when labels are given)
    with at least s supports, cut to its s heaviest supports, weights
    renormalised.  (The clustering driver's own initialisation restates
    d2_centroid_rands; this helper only needs a valid, seeded start.)"""
    from .data import Centroids
    rng = np.random.default_rng(seed)
    order = rng.permutation(phase.size)
    c = Centroids.zeros(K, s, phase.dim, with_supp=phase.dim > 0)
    for k in range(K):
        cand = order if labels is None else order[np.asarray(labels)[order] == k]
        cand = cand[phase.p_str[cand] >= s]
        if labels is None:
            cand = cand[k:]
        if len(cand) == 0:
            raise ValueError(f"no object with >= {s} supports for cluster {k}")
        w, sup = phase.object(int(cand[0]))
        if phase.metric == D2_WORD_EMBED:  # supports are the words' embeddings
            sup = phase.vocab_vec.reshape(phase.vocab_size, phase.dim)[sup]
        top = np.sort(np.argsort(-w, kind="stable")[:s])
        c.p_w[k * s:(k + 1) * s] = w[top] / w[top].sum()
        if c.p_supp is not None:
            c.p_supp[k * s * phase.dim:(k + 1) * s * phase.dim] = sup[top].reshape(-1)
    return c


def multiphase_histograms(n=100_000, V1=64, V2=1024, n_comp=32, seed=SEED_BASE + 3):
    """Co-clustering workload with three phases.  Each object draws a
    component; its points, its dense histogram (a discretised Gaussian bump
    around a component centre) and its sparse histogram's bins (around
    another centre) follow that component."""
    rng = np.random.default_rng(seed)
    comp = rng.integers(0, n_comp, size=n)
    # phase 0: Euclidean d=3
    means = rng.uniform(-5, 5, size=(n_comp, 3))
    m0 = rng.integers(8, 33, size=n).astype(np.int32)
    c0 = np.repeat(comp, m0)
    pts = means[c0] + 0.5 * rng.standard_normal((len(c0), 3))
    w0 = _dirichlet_ragged(rng, m0)
    ph0 = Phase(metric=D2_EUCLIDEAN_L2, dim=3, p_str=m0, p_w=w0, p_supp=pts.reshape(-1), str=16)
    # phase 1: dense V1-bin histograms, cost (i-j)^2
    ctr1 = rng.uniform(8, V1 - 8, size=n_comp)
    bins = np.arange(V1)
    centre = ctr1[comp] + rng.normal(0, 2.0, size=n)
    width = rng.uniform(3.0, 8.0, size=n)
    h = np.exp(-0.5 * ((bins[None, :] - centre[:, None]) / width[:, None]) ** 2) + 1e-3
    h /= h.sum(1, keepdims=True)
    d1 = (bins[:, None] - bins[None, :]).astype(np.float64) ** 2
    ph1 = Phase(metric=D2_HISTOGRAM, dim=0, p_str=np.full(n, V1, np.int32), p_w=h.reshape(-1),
                vocab_size=V1, dist_mat=d1.reshape(-1), str=V1)
    # phase 2: sparse V2-bin histograms, 5-40 nonzeros, cost |i-j|, Dirichlet(0.5)
    ctr2 = rng.uniform(64, V2 - 64, size=n_comp)
    m2 = rng.integers(5, 41, size=n).astype(np.int32)
    sym = []
    for i in range(n):
        cand = np.unique(np.clip(np.round(rng.normal(ctr2[comp[i]], 40.0, size=3 * m2[i])), 0,
                                 V2 - 1).astype(np.int64))
        while len(cand) < m2[i]:
            cand = np.unique(np.concatenate([cand, rng.integers(0, V2, m2[i])]))
        sym.append(np.sort(rng.choice(cand, size=m2[i], replace=False)))
    sym = np.concatenate(sym).astype(np.int32)
    w2 = np.maximum(rng.gamma(0.5, 1.0, size=len(sym)), 1e-6)
    starts = np.concatenate([[0], np.cumsum(m2)[:-1]])
    w2 = w2 / np.repeat(np.add.reduceat(w2, starts), m2)
    b2 = np.arange(V2)
    d2 = np.abs(b2[:, None] - b2[None, :]).astype(np.float64)
    ph2 = Phase(metric=D2_SPARSE_HISTOGRAM, dim=0, p_str=m2, p_w=w2, p_supp_sym=sym,
                vocab_size=V2, dist_mat=d2.reshape(-1), str=V2)
    return D2Data([ph0, ph1, ph2]), dict(kind="multiphase_histograms", n=n, V1=V1, V2=V2,
                                         n_comp=n_comp, seed=seed, classes=comp.astype(np.int32))
