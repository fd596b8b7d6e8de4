"""One shard of the CPU baseline (test infrastructure: only bench.py's
``cpu_baseline`` leg starts it).

    python -m oracle.cpu_shard SHARD.npz

Loads one shard of a one-phase Euclidean workload (p_str, p_w, p_supp,
labels, starting centroids, B-ADMM options), prints ``ready``, waits for a
line on stdin, runs ONE B-ADMM centroid update on it with the reference's
own d2_centroid_sphBregman (oracle/_ref, serial FP64, MKL sequential: one
thread) -- or with the FP64 restatement where the reference build is absent
-- and prints a JSON line with the seconds it took.  Several of these side
by side over disjoint shards are the reference's MPI split
(clustering_io.c:329-480 shards, one rank per core) without its per-pass
MPI_Allreduce of the centroid sums (centroid_Bregman.c:188, 215-217), which
only makes the CPU figure faster.  The process never touches a GPU.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


class _Quiet:
    """Silences C-level stdout (the reference build prints its tables)."""

    def __enter__(self):
        sys.stdout.flush()
        self.fd = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *a):
        import ctypes
        ctypes.CDLL(None).fflush(None)  # C stdio buffers what the library printed
        os.dup2(self.fd, 1)
        os.close(self.fd)
        os.close(self.null)


def main(path):
    from d2_kmeans_amd.data import BADMMOptions, Centroids, D2_EUCLIDEAN_L2, Phase
    from oracle import oracle as O
    z = np.load(path, allow_pickle=False)
    ph = Phase(metric=D2_EUCLIDEAN_L2, dim=int(z["dim"]), p_str=z["p_str"], p_w=z["p_w"],
               p_supp=z["p_supp"])
    K, s = int(z["K"]), int(z["s"])
    c = Centroids(K, s, ph.dim, z["c_w"].copy(), z["c_supp"].copy())
    opts = BADMMOptions(int(z["maxIters"]), float(z["rhoCoeff"]), int(z["updatePerLoops"]))
    lab = z["label"].astype(np.int32)
    lsw = np.ones(ph.size, np.int8)
    Z = np.zeros(s * ph.col)
    kind = "reference" if O.ref_available() else "port"
    if kind == "reference":
        O.ref()  # load the library before the clock starts
        # the work arrays d2_allocate_work makes once per run, touched, as in
        # the reference's own loop from its second outer iteration on
        work = O.ref_work(ph, K, s)
        for a in work.values():
            a.fill(1.0)
            a.fill(0.0)
    else:
        O.lib()
    Z.fill(1.0)
    Z.fill(0.0)
    print("ready", flush=True)
    sys.stdin.readline()
    t0 = time.perf_counter()
    if kind == "reference":
        with _Quiet():
            O.ref_badmm(ph, lab, lsw, c, Z, opts, work=work)
    else:
        O.badmm(ph, lab, lsw, c, Z, opts)
    dt = time.perf_counter() - t0
    print("RESULT " + json.dumps({"seconds": dt, "objects": int(ph.size), "kind": kind}),
          flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
