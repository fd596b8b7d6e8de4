"""The reference-compatible ABI of libd2mi.so (include/d2mi_clustering.h):
the reference's own structs go in unchanged.  The B-ADMM plug point is
checked against the reference's build (oracle/_ref) on the same structs, the
driver and the CLI against the FP64 oracle's closed loop."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from d2_kmeans_amd import _lib, synth
from d2_kmeans_amd.data import BADMMOptions, Centroids, D2Data, read_d2
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_dp = C.POINTER(C.c_double)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def lib(engine):
    L = C.CDLL(_lib.LIB_PATH)
    L.d2_centroid_sphBregman.argtypes = [C.POINTER(O.RefMph), C.POINTER(O.RefVarMph), C.c_int,
                                         C.POINTER(O.RefSph), C.POINTER(O.RefSph)]
    L.d2_allocate.argtypes = [C.POINTER(O.RefMph), C.c_int, C.c_size_t, C.POINTER(C.c_int),
                              C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.d2_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(O.RefMph)]
    L.d2_clustering.argtypes = [C.c_int, C.c_int, C.POINTER(O.RefMph), C.POINTER(O.RefMph),
                                C.c_int, C.c_char, C.c_char_p]
    L.d2_free.argtypes = [C.POINTER(O.RefMph)]
    L.d2_match_by_distmat.restype = C.c_double
    L.d2_match_by_distmat.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_size_t]
    L._dpdist2.argtypes = [C.c_int, C.c_size_t, C.c_size_t, _dp, _dp, _dp]
    return L


def set_opts(L, opts):
    clu = O.RefOpts.in_dll(L, "badmm_clu_options")
    clu.maxIters, clu.rhoCoeff, clu.updatePerLoops = opts.maxIters, opts.rhoCoeff, opts.updatePerLoops
    C.c_void_p.in_dll(L, "p_badmm_options").value = C.addressof(clu)


def ours_badmm(L, phase, label, lsw, cent, opts, precision):
    """d2_centroid_sphBregman through libd2mi.so with the reference's structs."""
    os.environ["D2MI_PRECISION"] = str(precision)
    set_opts(L, opts)
    keep = O._Keep()
    sph = O._ref_sph(phase, keep, s_override=cent.s)
    arr = (O.RefSph * 1)(sph)
    lab = keep.add(np.ascontiguousarray(label, np.int32).copy())
    m = O.RefMph(1, phase.size, phase.size, O.ptr(lab, O._ip), cent.K, arr)
    lsw_b = keep.add(np.ascontiguousarray(lsw, np.int8).copy())
    vm = O.RefVarMph(1, None, None, lsw_b.ctypes.data_as(O._cp), O.RefTrieq())
    csph = O._ref_cent_sph(cent, phase.metric, keep)
    rc = L.d2_centroid_sphBregman(C.byref(m), C.byref(vm), 0, C.byref(csph), C.byref(csph))
    os.environ.pop("D2MI_PRECISION", None)
    assert rc == 0


@pytest.mark.parametrize("precision,tol", [(64, 1e-9), (32, 1e-4)])
def test_centroid_sphBregman_drop_in(lib, precision, tol):
    say(f"d2_centroid_sphBregman drop-in fp{precision}")
    data, _ = synth.gaussian_mixture(n=800, seed=41)
    ph = data.phases[0]
    c0 = O.centroid_rands(ph, 6, 6)
    lab, lsw, _, _ = O.labeling(data, [c0])
    opts = BADMMOptions(10, 2.0, 10)
    ours = c0.copy()
    ours_badmm(lib, ph, lab, lsw, ours, opts, precision)
    ref = c0.copy()
    if O.ref_available():
        O.ref_badmm(ph, lab, lsw, ref, np.zeros(6 * ph.col), opts)
    else:
        O.badmm(ph, lab, lsw, ref, np.zeros(6 * ph.col), opts)
    assert np.max(np.abs(ours.p_w - ref.p_w)) <= tol * np.max(np.abs(ref.p_w))
    assert np.max(np.abs(ours.p_supp - ref.p_supp)) <= tol * np.max(np.abs(ref.p_supp))


def test_solver_and_pdist2_entry_points(lib):
    say("d2_match_by_distmat / _dpdist2")
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 3))
    B = rng.standard_normal((9, 3))
    Cm = np.empty(54)
    lib._dpdist2(3, 6, 9, O.ptr(A), O.ptr(B), O.ptr(Cm))
    np.testing.assert_allclose(Cm, O.pdist2(A, B), rtol=1e-14)
    wX = rng.dirichlet(np.ones(6))
    wY = rng.dirichlet(np.ones(9))
    v = lib.d2_match_by_distmat(6, 9, O.ptr(Cm), O.ptr(wX), O.ptr(wY), None, None, 0)
    assert abs(v - O.transport(Cm, wX, wY)) < 1e-12


@pytest.mark.parametrize("kind", ["euclid", "word_embed", "dense_hist", "sparse_hist"])
def test_compute_distance_every_metric(lib, kind):
    """d2_compute_distance (clustering.c:33-128) through the C ABI with the
    reference's structs, object i of the data against centroid j, for each
    metric the reference handles (Euclidean :46-62, word embeddings :63-80,
    dense histograms :81-90, sparse histograms :91-101), against the
    oracle's exact transport distance."""
    from d2_kmeans_amd.data import Centroids
    from cases import hist_phase
    say(f"d2_compute_distance {kind}")
    if kind == "euclid":
        data, _ = synth.gaussian_mixture(n=40, seed=5)
        ph = data.phases[0]
        cent = O.centroid_rands(ph, 3, 6)
    elif kind == "word_embed":
        data, _ = synth.word_embed_docs(n=40, V=200, d=8, m_lo=3, m_hi=20, seed=3, s=6)
        ph = data.phases[0]
        cent = O.centroid_rands(ph, 3, 6)
    else:
        ph = hist_phase("dense" if kind == "dense_hist" else "sparse", 30,
                        16 if kind == "dense_hist" else 64, seed=4)
        rng = np.random.default_rng(2)
        V = ph.vocab_size
        cent = Centroids(3, V, 0, rng.dirichlet(np.ones(V), size=3).reshape(-1))
    lib.d2_compute_distance.restype = C.c_double
    lib.d2_compute_distance.argtypes = [C.POINTER(O.RefMph), C.c_size_t, C.POINTER(O.RefMph),
                                        C.c_size_t, C.c_int, C.c_void_p, C.c_size_t]
    keep = O._Keep()
    a = O.RefMph(1, ph.size, ph.size, None, 0, (O.RefSph * 1)(O._ref_sph(ph, keep)))
    csph = O._ref_cent_sph(cent, ph.metric, keep)
    csph.dist_mat = O.ptr(ph.dist_mat)
    csph.vocab_size = ph.vocab_size
    b = O.RefMph(1, cent.K, cent.K, None, 0, (O.RefSph * 1)(csph))
    one = D2Data([ph])
    for i in range(0, ph.size, 7):
        for j in range(cent.K):
            got = lib.d2_compute_distance(C.byref(a), i, C.byref(b), j, -1, None, 0)
            want = O.distance(one, i, [cent], j)
            assert np.isfinite(got)
            assert abs(got - want) <= 1e-9 * max(1.0, want), (i, j, got, want)


def test_clustering_driver_on_reference_data(lib, tmp_path):
    """d2_allocate + d2_read + d2_clustering (phase 0 of data/mountaindat.d2,
    first 300 objects, K=4, s=6) against the oracle's loop from the same
    initial centroids (glibc-rand initialisation, identical on both sides)."""
    say("d2_clustering driver on mountaindat")
    fn = os.path.join(GOLD, "mountaindat_head.d2").encode()
    m = O.RefMph()
    strides = (C.c_int * 2)(6, 8)
    dims = (C.c_int * 2)(3, 3)
    types = (C.c_int * 2)(0, 0)
    assert lib.d2_allocate(C.byref(m), 2, 300, strides, dims, types) == 0
    assert lib.d2_read(fn, None, C.byref(m)) == 0
    set_opts(lib, BADMMOptions(10, 2.0, 10))
    os.environ["D2MI_PRECISION"] = "64"
    c = O.RefMph()
    c.ph = None
    rc = lib.d2_clustering(4, 20, C.byref(m), C.byref(c), 0, b"\x01",
                           str(tmp_path / "run").encode())
    os.environ.pop("D2MI_PRECISION", None)
    assert rc == 0
    labels = np.ctypeslib.as_array(m.label, shape=(300,)).copy()
    supp = np.ctypeslib.as_array(c.ph[0].p_supp, shape=(4 * 6 * 3,)).copy()
    data = read_d2(os.path.join(GOLD, "mountaindat_head.d2"), 300, [3, 3], strides=[6, 8])
    c0 = O.centroid_rands(data.phases[0], 4, 6)
    cents = [c0, O.centroid_rands(data.phases[1], 4, 8)]
    lab_o, it, _, cost = O.clustering(data, cents, 4, 20, BADMMOptions(10, 2.0, 10),
                                      use_triangle=True, selected_phase=0)
    np.testing.assert_array_equal(labels, lab_o)
    np.testing.assert_allclose(supp, cents[0].p_supp, rtol=1e-7, atol=1e-8)
    assert os.path.exists(str(tmp_path / "run_c.d2")) or it < 10
    lib.d2_free(C.byref(m))
    lib.d2_free(C.byref(c))


def _cli(kind):
    if kind == "ours":
        exe = os.path.join(ROOT, "d2_kmeans_amd", "d2mi")
    else:
        exe = O.REF_MAIN
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    return exe


def _centroids_phase0(path, K, s, d):
    """Phase 0 of a two-phase Euclidean _c.d2 whose phase 1 was not
    clustered (written as "dim\\n0\\n", clustering_io.c:237-241); weights
    renormalised as d2_read does (clustering_io.c:149-151)."""
    toks = open(path).read().split()
    q = 0
    w, x = [], []
    for _ in range(K):
        for ph in range(2):
            dim, m = int(toks[q]), int(toks[q + 1])
            q += 2
            if m == 0:
                continue
            assert ph == 0 and m == s and dim == d
            wk = np.array([float(t) for t in toks[q:q + m]])
            w.append(wk / wk.sum())
            q += m
            x.append(np.array([float(t) for t in toks[q:q + m * d]]))
            q += m * d
    return Centroids(K, s, d, np.concatenate(w), np.concatenate(x))


@pytest.mark.parametrize("kind", ["ours", "reference_main"])
def test_cli_end_to_end(tmp_path, kind):
    """The command line with the reference's flags (main.cc:53-282), run as
    this package's d2mi and as the reference's own main.cc linked against
    libd2mi.so (built by oracle/Makefile.ref), in FP64: the clustering run's
    labels equal the oracle's closed loop from the same glibc-rand initial
    centroids (B-ADMM at the program's defaults, maxIters = 100,
    centroid_Bregman.c:17), its written centroids agree to the 6 printed
    decimals, and --eval of the written centroids gives the oracle's
    brute-force labels against the centroids read back from that file."""
    exe = _cli(kind)
    say(f"cli {kind}")
    src = os.path.join(GOLD, "mountaindat_head.d2")
    out = str(tmp_path / "m")
    env = dict(os.environ, D2MI_PRECISION="64")
    args = ["-i", src, "-p", "2", "-n", "300", "-d", "3,3", "-s", "6,8", "--phase_only", "0",
            "-k", "4"]
    r = subprocess.run([exe, *args, "-m", "15", "-o", out], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
    lab1 = np.loadtxt(out + ".label", dtype=int)
    data = read_d2(src, 300, [3, 3], strides=[6, 8])
    cents = [O.centroid_rands(data.phases[0], 4, 6, seed=0),
             O.centroid_rands(data.phases[1], 4, 8, seed=0)]
    lab_o, _, _, _ = O.clustering(data, cents, 4, 15, BADMMOptions(100, 2.0, 10),
                                  use_triangle=True, selected_phase=0)
    np.testing.assert_array_equal(lab1, lab_o)
    cw = _centroids_phase0(out + "_c.d2", 4, 6, 3)
    np.testing.assert_allclose(cw.p_supp, cents[0].p_supp, atol=1.01e-6)
    r2 = subprocess.run([exe, *args, "--eval", out + "_c.d2", "-o", out + "_eval"],
                        capture_output=True, text=True, timeout=300, env=env)
    assert r2.returncode == 0, (r2.stdout + r2.stderr)[-2000:]
    lab2 = np.loadtxt(out + "_eval.label", dtype=int)
    lab_e, _, _, _ = O.labeling(data, [cw, None], label=np.full(300, -1, np.int32),
                                selected_phase=0)
    np.testing.assert_array_equal(lab2, lab_e)
