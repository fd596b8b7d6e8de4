"""D2 data model: discrete distributions with CSR-style ragged supports.

Mirrors the reference's ``sph`` / ``mph`` structs (include/d2/clustering.h:14-124):
one :class:`Phase` per phase holds ``p_str[N]`` support counts, ``p_str_cum[N]``
column offsets, ``p_w[col]`` weights and, by metric, ``p_supp[col*dim]``
(Euclidean, row-major per support), ``p_supp_sym[col]`` (0-based word ids /
histogram bins), ``vocab_vec[V*dim]`` or ``dist_mat[V*V]``.  :class:`Centroids`
is the centroid side: K objects of exactly ``s`` supports each
(centroid_rand.c:177-183).

Text I/O follows the reference's ``.d2`` formats (clustering_io.c:14-250,
data/README.md): weights are renormalised to sum 1 on read (:149-151); ids
are 1-based on disk and id 0 maps to V-1 (:164-166); writes print ``%lf``.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

# metric ids, include/d2/param.h:23-28
D2_EUCLIDEAN_L2 = 0
D2_CITYBLOCK_L1 = 1
D2_HISTOGRAM = 5
D2_N_GRAM = 6
D2_WORD_EMBED = 7
D2_SPARSE_HISTOGRAM = 12

# centroid methods, param.h:17-19
D2_CENTROID_BADMM = 0
D2_CENTROID_ADMM = 1
D2_CENTROID_GRADDEC = 2


@dataclasses.dataclass
class BADMMOptions:
    """``BADMM_options`` (param.h:4-8); defaults of ``badmm_clu_options``
    (centroid_Bregman.c:17).  ``badmm_cen_options`` is (2000, 1.0, 10)."""

    maxIters: int = 100
    rhoCoeff: float = 2.0
    updatePerLoops: int = 10


BADMM_CLU_OPTIONS = BADMMOptions(100, 2.0, 10)
BADMM_CEN_OPTIONS = BADMMOptions(2000, 1.0, 10)


@dataclasses.dataclass
class Phase:
    metric: int
    dim: int
    p_str: np.ndarray  # int32 [N]
    p_w: np.ndarray  # float64 [col]
    p_supp: Optional[np.ndarray] = None  # float64 [col*dim]
    p_supp_sym: Optional[np.ndarray] = None  # int32 [col]
    vocab_size: int = 0
    vocab_vec: Optional[np.ndarray] = None  # float64 [V*dim]
    dist_mat: Optional[np.ndarray] = None  # float64 [V*V]
    str: int = 0  # requested centroid support count (-s)

    def __post_init__(self):
        self.p_str = np.ascontiguousarray(self.p_str, dtype=np.int32)
        self.p_w = np.ascontiguousarray(self.p_w, dtype=np.float64)
        if self.p_supp is not None:
            self.p_supp = np.ascontiguousarray(self.p_supp, dtype=np.float64).reshape(-1)
        if self.p_supp_sym is not None:
            self.p_supp_sym = np.ascontiguousarray(self.p_supp_sym, dtype=np.int32)
        if self.vocab_vec is not None:
            self.vocab_vec = np.ascontiguousarray(self.vocab_vec, dtype=np.float64).reshape(-1)
        if self.dist_mat is not None:
            self.dist_mat = np.ascontiguousarray(self.dist_mat, dtype=np.float64).reshape(-1)
        self.p_str_cum = np.zeros(len(self.p_str), dtype=np.uint64)
        if len(self.p_str) > 1:
            self.p_str_cum[1:] = np.cumsum(self.p_str[:-1], dtype=np.uint64)

    @property
    def size(self) -> int:
        return int(len(self.p_str))

    @property
    def col(self) -> int:
        return int(self.p_w.shape[0])

    def object(self, i: int):
        """(weights, supports-or-ids) of object i."""
        a = int(self.p_str_cum[i])
        b = a + int(self.p_str[i])
        if self.metric == D2_EUCLIDEAN_L2:
            return self.p_w[a:b], self.p_supp[a * self.dim:b * self.dim].reshape(-1, self.dim)
        if self.p_supp_sym is not None:
            return self.p_w[a:b], self.p_supp_sym[a:b]
        return self.p_w[a:b], None

    def subset(self, idx) -> "Phase":
        """A new phase holding objects ``idx`` (in that order)."""
        idx = np.asarray(idx, dtype=np.int64)
        starts = self.p_str_cum[idx].astype(np.int64)
        lens = self.p_str[idx].astype(np.int64)
        cols = np.concatenate([np.arange(s, s + l) for s, l in zip(starts, lens)]) if len(idx) else np.zeros(0, np.int64)
        kw = dict(metric=self.metric, dim=self.dim, p_str=self.p_str[idx], p_w=self.p_w[cols],
                  vocab_size=self.vocab_size, vocab_vec=self.vocab_vec, dist_mat=self.dist_mat,
                  str=self.str)
        if self.p_supp is not None:
            kw["p_supp"] = self.p_supp.reshape(-1, self.dim)[cols].reshape(-1) if self.dim else None
        if self.p_supp_sym is not None:
            kw["p_supp_sym"] = self.p_supp_sym[cols]
        return Phase(**kw)


@dataclasses.dataclass
class D2Data:
    """``mph``: objects with one or more phases and a shared label vector."""

    phases: List[Phase]
    label: Optional[np.ndarray] = None

    def __post_init__(self):
        n = self.phases[0].size
        for p in self.phases:
            assert p.size == n, "every phase must hold the same objects"
        if self.label is None:
            self.label = np.full(n, -1, dtype=np.int32)  # clustering_util.c:117-118

    @property
    def size(self) -> int:
        return self.phases[0].size

    @property
    def s_ph(self) -> int:
        return len(self.phases)

    def subset(self, idx) -> "D2Data":
        return D2Data([p.subset(idx) for p in self.phases], np.asarray(self.label)[idx].copy())


@dataclasses.dataclass
class Centroids:
    """Centroids of one phase: K x s supports (``sph`` with p_str[k] = s)."""

    K: int
    s: int
    dim: int
    p_w: np.ndarray  # float64 [K*s]
    p_supp: Optional[np.ndarray] = None  # float64 [K*s*dim]

    @staticmethod
    def zeros(K: int, s: int, dim: int, with_supp: bool = True) -> "Centroids":
        return Centroids(K, s, dim, np.zeros(K * s), np.zeros(K * s * dim) if with_supp and dim else None)

    def copy(self) -> "Centroids":
        return Centroids(self.K, self.s, self.dim, self.p_w.copy(),
                         None if self.p_supp is None else self.p_supp.copy())


# ------------------------------------------------------------------ text I/O

def _tokens(path):
    with open(path) as f:
        for line in f:
            for tok in line.split():
                yield tok


def read_d2(path: str, n: int, dims, types=None, meta: Optional[str] = None,
            strides=None) -> D2Data:
    """Read up to ``n`` objects of ``len(dims)`` phases (d2_read,
    clustering_io.c:14-202).  Stops early, like the reference, when the file
    ends.  ``meta`` overrides the ``.hist<n>`` / ``.vocab<n>`` side file for a
    one-phase read (:60-63, :77-80)."""
    s_ph = len(dims)
    types = list(types) if types is not None else [D2_EUCLIDEAN_L2] * s_ph
    strides = list(strides) if strides is not None else [0] * s_ph
    metas = [dict() for _ in range(s_ph)]
    for k in range(s_ph):
        if types[k] in (D2_HISTOGRAM, D2_SPARSE_HISTOGRAM):
            fn = meta if (meta and s_ph == 1) else f"{path}.hist{k}"
            toks = _tokens(fn)
            V = int(next(toks))
            metas[k]["dist_mat"] = np.array([float(next(toks)) for _ in range(V * V)])
            metas[k]["vocab_size"] = V
        elif types[k] == D2_WORD_EMBED:
            fn = meta if (meta and s_ph == 1) else f"{path}.vocab{k}"
            toks = _tokens(fn)
            d = int(next(toks))
            assert d == dims[k]
            V = int(next(toks))
            metas[k]["vocab_vec"] = np.array([float(next(toks)) for _ in range(V * d)])
            metas[k]["vocab_size"] = V
    toks = _tokens(path)
    p_str = [[] for _ in range(s_ph)]
    p_w = [[] for _ in range(s_ph)]
    p_supp = [[] for _ in range(s_ph)]
    p_sym = [[] for _ in range(s_ph)]
    count = 0
    done = False
    while count < n and not done:
        for k in range(s_ph):
            try:
                d = int(next(toks))
            except StopIteration:
                done = True
                break
            assert d == dims[k]
            m = int(next(toks))
            if m == 0:
                continue
            w = np.array([float(next(toks)) for _ in range(m)])
            if types[k] != D2_HISTOGRAM:
                assert np.all(w > 1e-9), "support weights must be positive (clustering_io.c:145)"
            p_str[k].append(m)
            p_w[k].append(w / w.sum())
            if types[k] == D2_EUCLIDEAN_L2:
                p_supp[k].append(np.array([float(next(toks)) for _ in range(m * d)]))
            elif types[k] in (D2_WORD_EMBED, D2_SPARSE_HISTOGRAM):
                ids = np.array([int(next(toks)) for _ in range(m)], dtype=np.int64) - 1
                ids[ids < 0] = metas[k]["vocab_size"] - 1
                p_sym[k].append(ids.astype(np.int32))
        else:
            count += 1
    phases = []
    for k in range(s_ph):
        kw = dict(metric=types[k], dim=dims[k], p_str=np.array(p_str[k], np.int32),
                  p_w=np.concatenate(p_w[k]) if p_w[k] else np.zeros(0), str=strides[k], **metas[k])
        if types[k] == D2_EUCLIDEAN_L2:
            kw["p_supp"] = np.concatenate(p_supp[k])
        elif types[k] in (D2_WORD_EMBED, D2_SPARSE_HISTOGRAM):
            kw["p_supp_sym"] = np.concatenate(p_sym[k])
        if types[k] == D2_SPARSE_HISTOGRAM:
            kw["str"] = metas[k]["vocab_size"]  # clustering_io.c:182-185
        phases.append(Phase(**kw))
    return D2Data(phases)


def write_d2(path: str, phases: List[Phase]) -> None:
    """d2_write (clustering_io.c:204-250): ``%lf`` weights and supports."""
    n = phases[0].size
    with open(path, "w") as f:
        for i in range(n):
            for ph in phases:
                w, sup = ph.object(i)
                f.write(f"{ph.dim}\n{len(w)}\n")
                f.write("".join(f"{x:f} " for x in w) + "\n")
                if ph.metric == D2_EUCLIDEAN_L2:
                    for row in sup:
                        f.write("".join(f"{x:f} " for x in row) + "\n")
                elif ph.metric in (D2_WORD_EMBED, D2_SPARSE_HISTOGRAM):
                    f.write("".join(f"{int(x) + 1} " for x in sup) + "\n")


def centroids_as_phase(c: Centroids, metric: int = D2_EUCLIDEAN_L2) -> Phase:
    return Phase(metric=metric, dim=c.dim, p_str=np.full(c.K, c.s, np.int32), p_w=c.p_w,
                 p_supp=c.p_supp, str=c.s)


def write_labels(path: str, label) -> None:
    """d2_write_labels (clustering_io.c:252-275): one label per line."""
    with open(path, "w") as f:
        for l in np.asarray(label):
            f.write(f"{int(l)}\n")
