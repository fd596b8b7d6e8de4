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
from d2_kmeans_amd.data import BADMMOptions, read_d2
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


def test_cli_end_to_end(tmp_path):
    """The d2 command line with the reference's flags (main.cc:53-282):
    cluster, then --eval the written centroids; labels agree."""
    say("cli")
    cli = os.path.join(ROOT, "d2_kmeans_amd", "d2mi")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    src = os.path.join(GOLD, "mountaindat_head.d2")
    out = str(tmp_path / "m")
    r = subprocess.run([cli, "-i", src, "-p", "2", "-n", "300", "-d", "3,3", "-s", "6,8",
                        "--phase_only", "0", "-k", "4", "-m", "15", "-o", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lab1 = np.loadtxt(out + ".label", dtype=int)
    assert lab1.shape == (300,) and set(lab1) <= set(range(4))
    r2 = subprocess.run([cli, "-i", src, "-p", "2", "-n", "300", "-d", "3,3", "-s", "6,8",
                         "--phase_only", "0", "-k", "4", "--eval", out + "_c.d2", "-o",
                         out + "_eval"], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr[-2000:]
    lab2 = np.loadtxt(out + "_eval.label", dtype=int)
    # centroids are written with 6 decimals (clustering_io.c:223), so a
    # handful of near-boundary objects may move; most labels must agree
    assert np.mean(lab1 == lab2) > 0.97
