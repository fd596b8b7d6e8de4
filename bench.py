"""B-ADMM plan-update throughput on one MI355X (the image-like D2 workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: one Euclidean phase, d=2, supports on the 16x16 integer grid,
N=200k objects with U{20..80} distinct cells (10 stroke templates, 10%
jitter), intensities U(0,1] normalised; K=10 centroids of s=40 supports;
B-ADMM with maxIters=10 (11 inner passes, iterations 0..10), rho coefficient
2.0, supports refreshed every 10 passes.  A step is one whole centroid update
(setup + 11 passes), the way d2_clustering calls it once per outer
iteration; inputs are resident in HBM before timing starts.

Metric: plan updates/s = N * 11 / seconds per step (one plan update = one
inner pass over one object's s x m_i plan).  It counts only when the FP32
objective stays within 1e-5 of an FP64 run from the same state
(``gate_obj_rel_err``).

The ``clustering_loop`` object runs the whole clustering loop (Elkan
labeling with exact LPs, B-ADMM, label switching) for a few outer
iterations at the same N in FP64 and FP32, with the final labels of a
random sample checked against the FP64 oracle's brute force.

The ``assign`` object is the serving half of the metric: exact
nearest-centroid queries (the reference's --eval assignment) against K=64
resident centroids of 48 points in d=64; query sizes lognormal(ln 50, 1)
clipped to [5, 500], Dirichlet(1) weights, Gaussian-mixture supports;
Poisson arrivals at a fixed rate, served with continuous batching, with
p50/p99 latency and queries/s beside the one-query-per-launch form on the
same arrivals.  The ``word_embed`` object covers the word-embedding
workload (V=50k, d=300): FP32 B-ADMM with MFMA cost tables gated against
FP64, and the pdist2 GEMM against the FP32 matrix peak.  The ``multiphase``
object covers 3-phase co-clustering (N=100k; Euclidean, dense and sparse
histogram phases, K=32) phase by phase, each gated against FP64.  Prints ONE
JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The assignment server keeps one HIP stream per launch group in flight, and
# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default):
# groups whose streams share a queue run one after the other.  16 gives each
# of the server's streams a queue of its own.  Read when HIP initialises, so
# it is set before anything touches the GPU (--hw-queues 0 keeps the
# environment's value).
HW_QUEUES = 16
if "--hw-queues" in sys.argv:
    HW_QUEUES = int(sys.argv[sys.argv.index("--hw-queues") + 1])
if HW_QUEUES > 0:
    os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

METRIC = "B-ADMM plan updates/s at obj rel err ≤1e-5 vs FP64; assign p50/p99 ms"
HBM_PEAK_GBPS = 8000.0
PDIST2_REPS = 10  # pdist2 GEMM timings: best of this many calls


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic():
    """HBM bytes per plan-kernel launch measured by rocprofv3 PMC counters
    (FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel) and committed under
    profiles/; None when the summary is absent."""
    fn = os.path.join(ROOT, "profiles", "k_badmm_plan_grid_strokes_pmc_traffic.txt")
    try:
        for line in open(fn):
            if line.startswith("TRAFFIC_BYTES_PER_LAUNCH"):
                return float(line.split()[1])
    except OSError:
        pass
    return None


def _cpu_info():
    """(CPU model, nproc, cores this process may run on, cores used): the
    cores used are the affinity set, capped by OMP_NUM_THREADS where the
    environment sets one (a GPU box's host share)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    use = min(aff, int(share)) if share and share.isdigit() and int(share) > 0 else aff
    return model, os.cpu_count(), aff, max(1, use)


def _cpu_shards(ph, labels, c0, opts, idx_list, tmp):
    """Runs one oracle.cpu_shard process per shard, side by side, and
    returns (wall seconds from the common start to the last finish, per-shard
    reports).  The shards load their data first; the clock starts when all
    are ready."""
    import subprocess
    procs = []
    for r, idx in enumerate(idx_list):
        sub = ph.subset(idx)
        fn = os.path.join(tmp, f"shard{r}.npz")
        np.savez(fn, dim=sub.dim, p_str=sub.p_str, p_w=sub.p_w, p_supp=sub.p_supp,
                 label=np.asarray(labels)[idx].astype(np.int32), K=c0.K, s=c0.s, c_w=c0.p_w,
                 c_supp=c0.p_supp, maxIters=opts.maxIters, rhoCoeff=opts.rhoCoeff,
                 updatePerLoops=opts.updatePerLoops)
        env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
        procs.append(subprocess.Popen([sys.executable, "-m", "oracle.cpu_shard", fn], cwd=ROOT,
                                      stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True,
                                      env=env))
    for p in procs:
        line = p.stdout.readline()
        if line.strip() != "ready":
            raise RuntimeError(f"cpu shard did not start: {line!r}")
    t0 = time.perf_counter()
    for p in procs:
        p.stdin.write("go\n")
        p.stdin.flush()
    reps = []
    for p in procs:
        out, _ = p.communicate(timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"cpu shard failed with {p.returncode}")
        res = [l for l in out.splitlines() if l.startswith("RESULT ")]
        if not res:
            raise RuntimeError(f"cpu shard printed no result: {out[-500:]!r}")
        reps.append(json.loads(res[-1][len("RESULT "):]))
    return time.perf_counter() - t0, reps


def cpu_baseline(data, labels, c0, opts, n_one, n_all):
    """The reference's own B-ADMM (oracle/_ref, built from its sources,
    serial FP64, MKL sequential) timed on this host: on one core over the
    first n_one objects, and on every core of the host share at once, one
    process per core over disjoint shards of the first n_all objects (the
    reference's MPI split without its per-pass allreduce).  Falls back to
    this project's FP64 restatement where the reference build is absent."""
    import shutil
    import tempfile
    ph = data.phases[0]
    model, nproc, aff, P = _cpu_info()
    passes = opts.maxIters + 1
    tmp = tempfile.mkdtemp(prefix="d2mi_cpu_")
    try:
        n_one = min(n_one, ph.size)
        w1, r1 = _cpu_shards(ph, labels, c0, opts, [np.arange(n_one)], tmp)
        n_all = min(n_all, ph.size)
        bounds = np.linspace(0, n_all, P + 1).astype(np.int64)
        wP, rP = _cpu_shards(ph, labels, c0, opts,
                             [np.arange(bounds[r], bounds[r + 1]) for r in range(P)], tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    kind = r1[0]["kind"]
    return {"value": n_all * passes / wP, "unit": "plan updates/s", "cores": P, "kind": kind,
            "seconds": wP, "cpu_model": model, "nproc": nproc, "affinity_cores": aff,
            "sample": f"all {P} cores: the first {n_all} objects of the same workload (s=40, "
                      f"K=10) in {P} disjoint shards, one serial FP64 process each running one "
                      f"{passes}-pass B-ADMM update (the MPI split without its allreduce); "
                      f"the clock runs from the common start to the last shard's end",
            "shard_seconds_max": max(r["seconds"] for r in rP),
            "one_core": {"value": n_one * passes / w1, "cores": 1, "seconds": w1,
                         "sample": f"first {n_one} objects, one {passes}-pass update"}}


def oracle_assign(qph, cent, idx):
    """FP64 oracle brute force (d2_labeling) on the queries idx: labels and
    exact distances to the label."""
    from d2_kmeans_amd.data import D2Data
    from oracle import oracle as O
    sub = D2Data([qph]).subset(np.asarray(idx))
    lab, _, _, _ = O.labeling(sub, [cent])
    K, s, d = cent.K, cent.s, qph.dim
    CS = cent.p_supp.reshape(K, s, d)
    CW = cent.p_w.reshape(K, s)
    dist = np.empty(len(idx))
    for t, i in enumerate(idx):
        w, X = qph.object(int(i))
        Cm = ((CS[lab[t]][:, None, :] - X[None, :, :]) ** 2).sum(-1)
        dist[t] = np.sqrt(max(O.transport(Cm.T.reshape(-1).copy(), CW[lab[t]], w), 0.0))
    return lab, dist


def assign_leg(eng, args):
    """Serving: Poisson arrivals at a fixed rate against resident centroids,
    continuous batching vs one query per launch."""
    from d2_kmeans_amd import synth
    from d2_kmeans_amd.serve import Server
    K, s, d = 64, 48, 64
    log(f"[assign] training K={K} centroids of {s} points in d={d} on {args.assign_train} "
        "queries with the clustering loop (FP32 B-ADMM, Elkan labeling)")
    from d2_kmeans_amd.data import BADMMOptions
    train, tmeta = synth.serving_requests(n=args.assign_train, seed=synth.SEED_BASE + 40)
    eng.set_data(train)
    # one seed per mixture component (an object of that component with >= s points)
    eng.set_centroids(0, synth.seed_centroids(train.phases[0], K, s, labels=tmeta["dominant"]))
    t0 = time.perf_counter()
    tr = eng.clustering(K, max_iter=args.assign_train_iters, opts=BADMMOptions(10, 2.0, 10),
                        use_triangle=True, precision=32)
    train_s = time.perf_counter() - t0
    cent = eng.get_centroids(0)
    log(f"[assign] trained in {train_s:.1f} s: {tr['iterations']} iterations, "
        f"changes {tr['changes']}")
    qd, qmeta = synth.serving_requests(n=args.assign_n, seed=synth.SEED_BASE + 41)
    qph = qd.phases[0]
    Q = [qph.object(i) for i in range(qph.size)]
    rng = np.random.default_rng(synth.SEED_BASE + 42)
    arrivals = np.cumsum(rng.exponential(1.0 / args.assign_qps, qph.size))
    # warm up both forms (kernels loaded, streams and staging created)
    srv = Server(eng, 0, max_batch=args.assign_batch, streams=args.assign_streams)
    srv.assign(Q[:64])
    log(f"[assign] continuous batching ({args.assign_batch} per group, {args.assign_streams} "
        f"streams), {qph.size} queries at {args.assign_qps} q/s")
    lab, dist, lat, el = srv.replay(Q, arrivals)
    st = srv.stats()
    t0 = time.perf_counter()
    srv.assign(Q)
    burst = qph.size / (time.perf_counter() - t0)
    srv.close()
    n_plain = min(qph.size, args.assign_plain_n)
    plain = Server(eng, 0, plain=True)
    plain.assign(Q[:8])
    log(f"[assign] one query per launch on the first {n_plain} arrivals")
    lab_p, dist_p, lat_p, el_p = plain.replay(Q[:n_plain], arrivals[:n_plain])
    t0 = time.perf_counter()
    plain.assign(Q[:n_plain])
    burst_p = n_plain / (time.perf_counter() - t0)
    plain.close()
    same = bool((lab_p == lab[:n_plain]).all()) and float(
        np.max(np.abs(dist_p - dist[:n_plain]) / np.maximum(dist[:n_plain], 1e-300))) <= 1e-12
    nchk = min(qph.size, args.assign_check)
    log(f"[assign] FP64 oracle brute force on {nchk} queries")
    lo, do = oracle_assign(qph, cent, np.arange(nchk))
    res = {
        "workload": f"K={K} centroids x {s} points, d={d}; queries m~lognormal(ln 50, 1) in "
                    f"[5,500] (mean {qph.p_str.mean():.1f}), Dirichlet(1) weights, 64-component "
                    f"Gaussian mixture; centroids trained on {args.assign_train} queries of the "
                    f"same law by the clustering loop ({tr['iterations']} iterations, FP32 "
                    "B-ADMM, Elkan labeling)",
        "train_seconds": train_s,
        "queries": int(qph.size), "qps_offered": args.assign_qps,
        "max_batch": args.assign_batch, "streams": args.assign_streams,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "qps_achieved": qph.size / el, "burst_qps": burst,
        "lps_per_query": st["lps"] / max(st["completed"], 1), "groups": st["groups"],
        "max_inflight_groups": st["max_inflight_groups"],
        "plain": {"queries": n_plain, "p50_ms": float(np.percentile(lat_p, 50)),
                  "p99_ms": float(np.percentile(lat_p, 99)), "qps_achieved": n_plain / el_p,
                  "burst_qps": burst_p},
        "same_alone_and_in_mix": same,
        "oracle_check": {"queries": nchk, "labels_equal": bool((lo == lab[:nchk]).all()),
                         "dist_max_rel_err": float(np.max(np.abs(do - dist[:nchk]) /
                                                          np.maximum(do, 1e-300)))},
    }
    # latency by query size, and where one exact LP of each size spends its
    # time (d2mi_transport_profile: one LP per launch, a query against the
    # centroid nearest by mean, as the server's first round would pick)
    m_of = qph.p_str
    res["by_size"] = {}
    for lo_m, hi_m in ((5, 64), (65, 192), (193, 500)):
        sel = (m_of >= lo_m) & (m_of <= hi_m)
        if sel.any():
            res["by_size"][f"m{lo_m}-{hi_m}"] = {
                "queries": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                "p99_ms": float(np.percentile(lat[sel], 99))}
    cw = cent.p_w.reshape(K, s)
    cs = cent.p_supp.reshape(K, s, d)
    mu_c = cs.mean(1)
    res["lp_phases"] = {}
    for lo_m, hi_m in ((45, 55), (190, 210), (450, 500)):
        idx = [i for i in range(qph.size) if lo_m <= m_of[i] <= hi_m][:8]
        if not idx:
            continue
        rows = []
        for i in idx:
            w, x = Q[i]
            k = int(np.argmin(((mu_c - x.mean(0)) ** 2).sum(1)))
            C = ((cs[k][:, None, :] - x[None, :, :]) ** 2).sum(-1)  # (i, j) at i + j*s
            _, piv, ph_us = eng.transport_profile(C.T.reshape(1, -1), cw[k][None], w[None])
            rows.append((int(piv[0]), ph_us[0]))
        P = np.array([r[1] for r in rows])
        pv = np.array([r[0] for r in rows])
        res["lp_phases"][f"{s}x{lo_m}-{hi_m}"] = {
            "lps": len(rows), "median_us": float(np.median(P[:, :8].sum(1))),
            "median_pivots": float(np.median(pv)),
            "phase_median_us": {nm: float(x) for nm, x in zip(eng.LP_PHASES, np.median(P, 0))}}
    log(f"[assign] {res}")
    return res


def word_embed_leg(eng):
    """Word-embedding D2 (V=50k, d=300, N=20k documents, K=20, s=64): one
    B-ADMM update in FP32 with the cost table from the matrix cores, gated
    against an FP64 update from the same state; pdist2 GEMM throughput in
    FP32 and BF16 against the FP32 matrix peak."""
    from d2_kmeans_amd import synth
    from d2_kmeans_amd.data import BADMMOptions
    data, meta = synth.word_embed_docs()
    ph = data.phases[0]
    K, s = 20, 64
    c0 = synth.seed_centroids(ph, K, s)
    N = ph.size
    labels = (np.arange(N) % K).astype(np.int32)  # balanced, deterministic
    eng.set_data(data)
    opts = BADMMOptions(10, 2.0, 10)
    out = {"workload": f"word-embedding D2: V={ph.vocab_size}, d={ph.dim}, N={N} documents "
                       f"(Zipf lengths, mean {ph.p_str.mean():.1f}), K={K}, s={s}, labels i mod K",
           "plan_elements": int(s * ph.col)}
    objs = {}
    for name, prec, env in (("fp64", 64, {}), ("fp32_mfma", 32, {}),
                            ("fp32_bf16", 32, {"D2MI_PDIST2_BF16": "1"})):
        os.environ.update(env)
        eng.set_centroids(0, c0)
        eng.set_labels(labels, np.ones(N, np.int8))
        eng.badmm(0, opts, precision=prec)  # warm (plans allocated)
        eng.set_centroids(0, c0)
        eng.set_labels(labels, np.zeros(N, np.int8))
        t0 = time.perf_counter()
        st = eng.badmm(0, opts, precision=prec)
        dt = time.perf_counter() - t0
        for k in env:
            os.environ.pop(k)
        objs[name] = st["obj"]
        out[name] = {"plan_updates_per_s": N * (opts.maxIters + 1) / dt, "ms_per_update": 1e3 * dt,
                     "plan_kernel_ms_total": st["ms_plan"] if st["ms_plan"] >= 0 else None,
                     "inner_loop_graph": st["graph"], "obj": st["obj"]}
    for name in ("fp32_mfma", "fp32_bf16"):
        out[name]["obj_rel_err_vs_fp64"] = abs(objs[name] - objs["fp64"]) / abs(objs["fp64"])
    V, d = ph.vocab_size, ph.dim
    vocab = ph.vocab_vec.reshape(V, d)
    cs = c0.p_supp.reshape(K * s, d)
    flops = 2.0 * V * K * s * d
    dp = (d + 31) // 32 * 32
    for bf in (False, True):
        eng.pdist2_mfma(cs[:8], vocab[:64], bf16=bf)
        ms = min(eng.pdist2_mfma(cs, vocab, bf16=bf, timing=True)[1] for _ in range(PDIST2_REPS))
        peak = 2500.0 if bf else 157.3
        rec = {"shape": f"{V} x {K * s} x {d}", "ms": ms, "tflops": flops / ms / 1e9,
               "peak_tflops": peak, "frac": flops / ms / 1e9 / peak,
               "timing": f"HIP events around the GEMM kernel alone, best of {PDIST2_REPS}"}
        if bf:
            # 16x faster products make the FP32 result table the bound: HBM
            # bytes = the table written + the BF16 operands read once
            hbm = 4.0 * V * K * s + 2.0 * (V + K * s) * dp
            rec.update({"bound": "hbm", "hbm_bytes": hbm, "achieved_GBps": hbm / ms / 1e6,
                        "hbm_frac": hbm / ms / 1e6 / HBM_PEAK_GBPS})
        else:
            rec["bound"] = "mfma"
        out["pdist2_bf16" if bf else "pdist2_fp32"] = rec
    log(f"[word-embed] {out}")
    return out


def multiphase_leg(eng, args):
    """Multi-phase co-clustering, batched: N objects of three phases
    (Euclidean d=3 with 8-32 points; dense 64-bin histograms, cost (i-j)^2;
    sparse 1024-bin histograms with 5-40 nonzeros, cost |i-j|), K=32, centroid
    supports 16 / 64 / 1024 (a sparse-histogram centroid has s = V bins, as the
    reference forces, clustering_io.c:182-185).  Per phase: the kernel forms
    tuned, an FP64 and an FP32 update from the same state (objective gate),
    then a timed FP32 update replayed as a graph and one launched eagerly for
    the plan kernel's time.  Centroids start from the label-wise averages."""
    from d2_kmeans_amd import synth
    from d2_kmeans_amd.data import BADMMOptions, Centroids
    K, S0 = 32, 16
    opts = BADMMOptions(10, 2.0, 10)
    t0 = time.perf_counter()
    data, meta = synth.multiphase_histograms(n=args.multiphase_n)
    log(f"[multiphase] generated N={args.multiphase_n} in {time.perf_counter() - t0:.1f} s")
    labels = meta["classes"].astype(np.int32) % K
    N = data.size
    cents = []
    for k, ph in enumerate(data.phases):
        if k == 0:
            cents.append(synth.seed_centroids(ph, K, S0, labels=labels))
            continue
        V = ph.vocab_size
        H = np.zeros((K, V))
        obj = np.repeat(np.arange(N), ph.p_str)
        bins = (np.tile(np.arange(V), N) if ph.p_supp_sym is None else ph.p_supp_sym)
        np.add.at(H, (labels[obj], bins), ph.p_w)
        H += 1e-6
        H /= H.sum(1, keepdims=True)
        cents.append(Centroids(K, V, 0, H.reshape(-1)))
    eng.set_data(data)
    names = ["euclid_d3_s16", "dense_hist64", "sparse_hist1024"]
    out = {"workload": f"3-phase co-clustering: N={N}, K={K}; phase 0 Euclidean d=3, m in [8,32], "
                       "s=16; phase 1 dense 64-bin histograms, cost (i-j)^2, s=64; phase 2 sparse "
                       "1024-bin histograms, 5-40 nonzeros, cost |i-j|, s=V=1024; labels = mixture "
                       "component, centroids = label-wise averages",
           "phases": {}}
    t_total, upd_total = 0.0, 0
    for k, ph in enumerate(data.phases):
        c = cents[k]
        elems = float(c.s) * float(ph.col)

        def reset(lsw_all):
            eng.set_centroids(k, c)
            eng.set_labels(labels, np.ones(N, np.int8) if lsw_all else np.zeros(N, np.int8))

        reset(True)
        eng.badmm(k, opts, precision=32)
        rep = eng.tune(k, 32, reps=2)
        reset(True)
        st64 = eng.badmm(k, opts, precision=64)
        reset(True)
        st32 = eng.badmm(k, opts, precision=32)
        err = abs(st32["obj"] - st64["obj"]) / abs(st64["obj"])
        reset(False)
        eng.synchronize()
        t1 = time.perf_counter()
        stg = eng.badmm(k, opts, precision=32)
        eng.synchronize()
        dt = time.perf_counter() - t1
        eng.set_graph(False)
        reset(False)
        ste = eng.badmm(k, opts, precision=32)
        eng.set_graph(True)
        plan_ms = ste["ms_plan"] / ste["n_plan_launches"]
        t_total += dt
        upd_total += N * (opts.maxIters + 1)
        out["phases"][names[k]] = {
            "plan_elements": int(elems), "ms_per_update": 1e3 * dt,
            "plan_updates_per_s": N * (opts.maxIters + 1) / dt, "graph": stg["graph"],
            "plan_forms": stg["plan_forms"], "plan_kernel_ms_per_pass": plan_ms,
            "algorithmic_GBps": 16.0 * elems / (plan_ms / 1e3) / 1e9,
            "obj_fp64": st64["obj"], "obj_fp32": st32["obj"], "obj_rel_err": err,
            "tuner": rep.strip()}
        log(f"[multiphase] phase {k}: {out['phases'][names[k]]}")
    out["plan_updates_per_s"] = upd_total / t_total
    out["ms_per_update_all_phases"] = 1e3 * t_total
    out["max_obj_rel_err"] = max(v["obj_rel_err"] for v in out["phases"].values())
    log(f"[multiphase] {out['plan_updates_per_s']:.4g} plan updates/s over the three phases")
    return out


def cluster_leg(eng, args, data=None, c0=None):
    """The whole clustering loop (d2_clustering, clustering.c:409-525) on the
    image-like workload at full size: Elkan labeling with exact transport
    LPs, B-ADMM, label switching and bound relaxation, a few outer
    iterations from one starting centroid per stroke class, in FP64 and in
    FP32 on the device.  Reported: time per stage, label changes and exact
    distances per outer iteration, whether both precisions give the same
    labels, and the final labels of a random sample checked against the
    FP64 oracle's brute force against the final centroids."""
    from d2_kmeans_amd import synth
    from d2_kmeans_amd.data import BADMMOptions
    from oracle import oracle as O
    K, s = 10, 40
    opts = BADMMOptions(10, 2.0, 10)
    if data is None:
        data, meta = synth.grid_strokes(n=args.n)
        c0 = synth.seed_centroids(data.phases[0], K, s, labels=meta["classes"])
    N = data.size
    eng.set_data(data)
    runs = {}
    for precision in (64, 32):
        eng.set_centroids(0, c0)
        eng.set_labels(np.full(N, -1, np.int32), np.zeros(N, np.int8))
        eng.synchronize()
        t0 = time.perf_counter()
        st = eng.clustering(K, args.cluster_iters, opts, use_triangle=True, precision=precision)
        eng.synchronize()
        dt = time.perf_counter() - t0
        lab = eng.get_labels()[0]
        runs[precision] = (st, dt, lab, eng.get_centroids(0))
        log(f"[cluster] fp{precision}: {dt:.2f} s, changes {st['changes']}, "
            f"distances {st['dist_count']}, labeling {st['ms_labeling']:.0f} ms, "
            f"B-ADMM {st['ms_badmm']:.0f} ms")
    st64, dt64, lab64, c64 = runs[64]
    st32, dt32, lab32, c32 = runs[32]
    rng = np.random.default_rng(synth.SEED_BASE + 50)
    idx = np.sort(rng.choice(N, min(N, args.cluster_check), replace=False))
    sub = data.subset(idx)
    log(f"[cluster] FP64 oracle brute force on {len(idx)} sampled objects x {K} centroids")
    lo64, _, _, _ = O.labeling(sub, [c64], label=np.full(len(idx), -1, np.int32))
    lo32, _, _, _ = O.labeling(sub, [c32], label=np.full(len(idx), -1, np.int32))

    def run_out(st, dt):
        its = st["iterations"]
        return {"seconds": dt, "iterations": its, "changes": st["changes"],
                "dist_count": st["dist_count"],
                "lps_per_object": [d / N for d in st["dist_count"]],
                "ms_labeling": st["ms_labeling"], "ms_badmm": st["ms_badmm"],
                "final_mean_cost": st["final_cost"]}

    out = {"workload": f"grid-strokes D2 (N={N}, K={K}, s={s}), {args.cluster_iters} outer "
                       "iterations from one object per stroke class, Elkan labeling, B-ADMM "
                       "maxIters=10",
           "fp64": run_out(st64, dt64), "fp32": run_out(st32, dt32),
           "labels_fp32_equal_fp64": bool((lab32 == lab64).all()),
           "label_agreement_fp32_fp64": float(np.mean(lab32 == lab64)),
           "final_cost_rel_err_fp32": abs(st32["final_cost"] - st64["final_cost"]) /
                                      abs(st64["final_cost"]),
           "oracle_check": {"objects": int(len(idx)),
                            "fp64_labels_equal": bool((lo64 == lab64[idx]).all()),
                            "fp32_labels_equal": bool((lo32 == lab32[idx]).all())}}
    log(f"[cluster] {out}")
    return out


def headline_leg(eng, args):
    """B-ADMM plan updates/s on the image-like workload, gated against FP64."""
    from d2_kmeans_amd import synth
    from d2_kmeans_amd.data import BADMMOptions
    K, s = 10, 40
    opts = BADMMOptions(10, 2.0, 10)
    log(f"[bench] generating grid-strokes workload N={args.n}")
    data, meta = synth.grid_strokes(n=args.n)
    labels = meta["classes"]
    ph = data.phases[0]
    c0 = synth.seed_centroids(ph, K, s, labels=labels)
    eng.set_data(data)

    def reset(precision, lsw_all=True):
        eng.set_centroids(0, c0)
        eng.set_labels(labels, np.ones(ph.size, np.int8) if lsw_all else np.zeros(ph.size, np.int8))

    # ---- plan-kernel tuner: one update to fill the plans, then every
    # feasible kernel form per object-size class timed (plans restored)
    tune_report = None
    if not args.no_tune:
        reset(32)
        eng.badmm(0, opts, precision=32)
        tune_report = eng.tune(0, 32, reps=3)
        log(f"[bench] tuner:\n{tune_report}")

    # ---- gate: FP32 vs FP64 from the same state
    gate = None
    if not args.no_gate:
        log("[bench] gate: FP64 and FP32 updates from the same state")
        reset(64)
        st64 = eng.badmm(0, opts, precision=64)
        c64 = eng.get_centroids(0)
        reset(32)
        st32 = eng.badmm(0, opts, precision=32)
        c32 = eng.get_centroids(0)
        gate = {
            "obj_fp64": st64["obj"], "obj_fp32": st32["obj"],
            "obj_rel_err": abs(st32["obj"] - st64["obj"]) / abs(st64["obj"]),
            "centroid_supp_max_rel": float(np.max(np.abs(c32.p_supp - c64.p_supp)) / np.max(np.abs(c64.p_supp))),
            "centroid_w_max_abs": float(np.max(np.abs(c32.p_w - c64.p_w))),
        }
        log(f"[bench] gate: {gate}")

    # ---- timed engine steps (FP32 plans, FP64 reductions)
    reset(32)
    for i in range(args.warmup):
        eng.badmm(0, opts, precision=32)
        log(f"[bench] warmup {i} done")
    eng.set_labels(labels, np.zeros(ph.size, np.int8))  # steady state: no label switches
    passes = opts.maxIters + 1

    def timed_region():
        eng.synchronize()
        plan_ms, plan_launches, totals = 0.0, 0, []
        t0 = time.perf_counter()
        for i in range(args.steps):
            st = eng.badmm(0, opts, precision=32)
            plan_ms += st["ms_plan"]
            plan_launches += st["n_plan_launches"]
            totals.append(st["ms_total"])
        eng.synchronize()
        return 1000.0 * (time.perf_counter() - t0) / args.steps, plan_ms, plan_launches, totals, st

    # the product's default: the inner passes of an update replayed as one
    # HIP graph.  HIP cannot time graph nodes with events, so the plan
    # kernel's duration comes from the same number of updates launched one by
    # one on the same stream right after (the same kernels and shapes)
    ms_per_step, _, _, totals, st = timed_region()
    graph_used = st["graph"]
    value = ph.size * passes / (ms_per_step / 1000.0)
    log(f"[bench] {args.steps} steps (graph={graph_used}): {ms_per_step:.3f} ms/step, "
        f"device {np.mean(totals):.3f} ms")
    eng.set_graph(False)
    ms_eager, plan_ms, plan_launches, totals_e, st_e = timed_region()
    eng.set_graph(True)
    log(f"[bench] {args.steps} steps launched one by one: {ms_eager:.3f} ms/step, device "
        f"{np.mean(totals_e):.3f} ms, plan kernel {plan_ms / plan_launches:.4f} ms per pass")
    elems = st["plan_elems"]
    avg_plan_ms = plan_ms / plan_launches
    bytes_per_launch = 16.0 * elems  # Z and Y read + written once, FP32
    achieved = bytes_per_launch / (avg_plan_ms / 1000.0) / 1e9
    # the practical ceiling for this access pattern on this GPU: two arrays of
    # the plan's size streamed in place (read and written), no arithmetic
    ceiling, ceiling_report = eng.stream_probe(elems, 5)
    log(f"[bench] two-array in-place stream over {int(elems)} elements: best "
        f"{ceiling:.1f} GB/s\n{ceiling_report}")

    # ---- plain one-thread-per-object form (the 20x denominator)
    plain = None
    if not args.no_plain:
        reset(32)
        eng.badmm(0, opts, precision=32, variant="plain")
        eng.set_labels(labels, np.zeros(ph.size, np.int8))
        t1 = time.perf_counter()
        stp = eng.badmm(0, opts, precision=32, variant="plain")
        eng.synchronize()
        pw = time.perf_counter() - t1
        plain = {"value": ph.size * passes / pw, "ms_per_step": 1000 * pw,
                 "plan_kernel_ms_avg": stp["ms_plan"] / stp["n_plan_launches"]}
        log(f"[bench] plain form: {plain}")

    if args.copy_calib:
        eng.get_plan(0, 32, "Z")
        log(f"[bench] copy calibration: k_f32_to_f64 over {int(elems)} elements "
            f"({4 * elems / 1e9:.3f} GB read, {8 * elems / 1e9:.3f} GB written)")

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline: {args.cpu_sample} objects on 1 core, "
            f"{args.cpu_sample_all} on every core")
        cpu = cpu_baseline(data, labels, c0, opts, args.cpu_sample, args.cpu_sample_all)
        log(f"[bench] cpu baseline: {cpu}")

    return dict(value=value, ms_per_step=ms_per_step, ph=ph, K=K, s=s, elems=elems,
                data=data, c0=c0,
                achieved=achieved, avg_plan_ms=avg_plan_ms, bytes_per_launch=bytes_per_launch,
                ceiling=ceiling,
                gate=gate, plain=plain, cpu=cpu, plan_forms=st["plan_forms"], graph=graph_used,
                tune_report=tune_report,
                eager={"value": ph.size * passes / (ms_eager / 1000.0), "ms_per_step": ms_eager})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--cpu-sample", type=int, default=12000,
                    help="objects of the one-core CPU baseline")
    ap.add_argument("--cpu-sample-all", type=int, default=200_000,
                    help="objects of the all-cores CPU baseline (split over the cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-gate", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="rule-based plan-kernel forms only")
    ap.add_argument("--skip-badmm", action="store_true",
                    help="skip the headline leg (for profiling the other legs alone)")
    ap.add_argument("--no-assign", action="store_true")
    ap.add_argument("--no-word-embed", action="store_true")
    ap.add_argument("--no-multiphase", action="store_true")
    ap.add_argument("--no-cluster", action="store_true")
    ap.add_argument("--cluster-iters", type=int, default=5)
    ap.add_argument("--cluster-check", type=int, default=200)
    ap.add_argument("--multiphase-n", type=int, default=100_000)
    ap.add_argument("--assign-n", type=int, default=3000)
    ap.add_argument("--assign-qps", type=float, default=1000.0)
    ap.add_argument("--assign-batch", type=int, default=0,
                    help="queries per launch group (0: 16 with >= 13 hardware queues, else 48)")
    ap.add_argument("--assign-streams", type=int, default=0,
                    help="server streams (0: 12 with >= 13 hardware queues, else 6)")
    ap.add_argument("--hw-queues", type=int, default=HW_QUEUES,
                    help="GPU_MAX_HW_QUEUES for this process (0: keep the environment's)")
    ap.add_argument("--assign-plain-n", type=int, default=300)
    ap.add_argument("--assign-check", type=int, default=100)
    ap.add_argument("--assign-train", type=int, default=4000)
    ap.add_argument("--assign-train-iters", type=int, default=6)
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one FP32->FP64 plan copy kernel (reads 4 B, writes 8 B per "
                         "plan element) to scale rocprof byte counters")
    args = ap.parse_args()
    hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    if args.assign_batch <= 0:
        args.assign_batch = 16 if hwq >= 13 else 48
    if args.assign_streams <= 0:
        args.assign_streams = 12 if hwq >= 13 else 6
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from d2_kmeans_amd.engine import Engine

    eng = Engine(0)
    h = None if args.skip_badmm else headline_leg(eng, args)
    clus = None if args.no_cluster else cluster_leg(eng, args, *((h["data"], h["c0"]) if h else ()))
    assign = None if args.no_assign else assign_leg(eng, args)
    word = None if args.no_word_embed else word_embed_leg(eng)
    multi = None if args.no_multiphase else multiphase_leg(eng, args)

    hd = h or {}
    ph = hd.get("ph")
    out = {
        "metric": METRIC, "value": hd.get("value"), "unit": "plan updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": hd.get("ms_per_step"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "grid-strokes D2, 1 Euclidean phase d=2, 16x16 grid, "
                               f"N={args.n}, m in [20,80], K=10, s=40, B-ADMM maxIters=10 "
                               "(11 passes), rhoCoeff=2.0, supports every 10 passes",
                   "objects": args.n, "plan_elements": int(hd["elems"]) if h else None,
                   "K": 10, "s": 40},
        "roofline": None if h is None else {
            "bound": "hbm", "achieved": hd["achieved"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": hd["achieved"] / HBM_PEAK_GBPS, "traffic": profiled_traffic(),
            "kernel": "k_badmm_plan_reg / k_badmm_plan64 / k_badmm_plan (FP32), forms by "
                      "object-size class m<=M:form/workgroups-per-CU(objects): " + hd["plan_forms"],
            "kernel_ms_avg": hd["avg_plan_ms"],
            "kernel_timing": "HIP events around every plan launch of the same number of "
                             "updates launched one by one right after the timed region "
                             "(graph nodes cannot be timed with HIP events)",
            "algorithmic_bytes_per_launch": hd["bytes_per_launch"],
            "stream_ceiling": {
                "GBps": hd["ceiling"], "frac": hd["achieved"] / hd["ceiling"],
                "what": "best of dword and float4 grid-stride kernels streaming two FP32 arrays "
                        "of the plan's size in place (16 B per element), measured in the same run "
                        "(d2mi_stream_probe); frac = achieved / this"}},
        "inner_loop_graph": hd.get("graph"),
        "eager_launches": hd.get("eager"),
        "tuner": hd.get("tune_report"),
        "gate_obj_rel_err": None if not hd.get("gate") else hd["gate"]["obj_rel_err"],
        "gate": hd.get("gate"),
        "plain_form": hd.get("plain"),
        "speedup_vs_plain": None if not hd.get("plain") else hd["value"] / hd["plain"]["value"],
        "cpu_baseline": hd.get("cpu"),
        "clustering_loop": clus,
        "assign": assign,
        "word_embed": word,
        "multiphase": multi,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
