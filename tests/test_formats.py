"""The reference's file formats and its own caller, on the CPU.

* Text output: d2_read, then d2_write (``_c.d2``), d2_write_labels
  (``.label``), d2_write_split as read and pre-processed (``.ind`` and the
  shards) and d2_write_labels_serial (``.label_o``) run through libd2mi.so
  and through the reference build (oracle/_ref), each in its own process; every
  file must be byte-identical (clustering_io.c:14-480).  Inputs: the first
  300 objects of the reference's data/mountaindat.d2 (two Euclidean phases),
  a three-phase file (Euclidean, dense and sparse histograms), a two-phase
  file (Euclidean, dense histogram) and a word-embedding file, generated
  here with fixed seeds.
* The reference's command-line program (src/app/main.cc + util.cc),
  unchanged, compiles against the reference's own header and links against
  libd2mi.so alone.

Both need /root/reference (the reference build, the program's sources) and
are skipped without it.  No GPU call is made: none of these entry points
reaches the device."""
import filecmp
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from d2_kmeans_amd import _lib
from d2_kmeans_amd.data import (D2_EUCLIDEAN_L2, D2_HISTOGRAM, D2_SPARSE_HISTOGRAM,
                                D2_WORD_EMBED, read_d2)
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = "/root/reference"

needs_ref = pytest.mark.skipif(not (os.path.isdir(REF) and O.ref_available()),
                               reason="needs the reference sources and oracle/_ref")


def _fmt(xs):
    return "".join(f"{x:.6f} " for x in xs)


def write_multiphase(path, n=60, seed=11, sparse=True):
    """Phase 0 Euclidean d=2 (3-9 points), phase 1 a dense 8-bin histogram
    (some bins empty), phase 2 (``sparse``) a sparse 16-bin histogram (2-6
    bins)."""
    rng = np.random.default_rng(seed)
    with open(path, "w") as f:
        for _ in range(n):
            m = int(rng.integers(3, 10))
            f.write(f"2\n{m}\n{_fmt(rng.uniform(0.1, 1.0, m))}\n")
            for _ in range(m):
                f.write(_fmt(rng.normal(0, 3, 2)) + "\n")
            h = rng.uniform(0.05, 1.0, 8) * (rng.uniform(size=8) < 0.7)
            h[int(rng.integers(8))] += 0.5
            f.write(f"0 8 {_fmt(h)}\n")
            if not sparse:
                continue
            k = int(rng.integers(2, 7))
            bins = np.sort(rng.choice(16, k, replace=False)) + 1
            f.write(f"0\n{k}\n{_fmt(rng.uniform(0.1, 1.0, k))}\n")
            f.write("".join(f"{b} " for b in bins) + "\n")
    for ph, V in ((1, 8), (2, 16))[:2 if sparse else 1]:
        i = np.arange(V)
        D = (i[:, None] - i[None, :]) ** 2 if ph == 1 else np.abs(i[:, None] - i[None, :])
        with open(f"{path}.hist{ph}", "w") as f:
            f.write(f"{V}\n" + "\n".join(_fmt(r) for r in D.astype(float)) + "\n")


def write_words(path, n=50, V=30, d=4, seed=12):
    """Word-embedding documents of 2-9 words (id 0 on disk is the last word,
    clustering_io.c:164-166) with a d=4 vocabulary side file."""
    rng = np.random.default_rng(seed)
    with open(path, "w") as f:
        for i in range(n):
            m = int(rng.integers(2, 10))
            ids = rng.choice(V, m, replace=False) + 1
            if i % 7 == 0:
                ids[0] = 0
            f.write(f"{d}\n{m}\n{_fmt(rng.uniform(0.1, 1.0, m))}\n")
            f.write("".join(f"{x} " for x in ids) + "\n")
    with open(f"{path}.vocab0", "w") as f:
        f.write(f"{d} {V}\n")
        for _ in range(V):
            f.write(_fmt(rng.normal(0, 1, d)) + "\n")


def _run(lib, spec, out):
    spec_path = out + ".json"
    json.dump(spec, open(spec_path, "w"))
    r = subprocess.run([sys.executable, os.path.join(HERE, "io_roundtrip.py"), lib, spec_path,
                        out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]


def _files(root):
    out = []
    for dp, _, fs in os.walk(root):
        for f in fs:
            out.append(os.path.relpath(os.path.join(dp, f), root))
    return sorted(out)


def _cases(tmp):
    mp = os.path.join(tmp, "mp.d2")
    write_multiphase(mp)
    hd = os.path.join(tmp, "hd.d2")
    write_multiphase(hd, seed=13, sparse=False)
    wd = os.path.join(tmp, "wd.d2")
    write_words(wd)
    return {
        "mountaindat": dict(file=os.path.join(GOLD, "mountaindat_head.d2"), n=300,
                            strides=[6, 8], dims=[3, 3], types=[D2_EUCLIDEAN_L2] * 2, splits=3),
        # the reference cannot pre-process a sparse histogram phase (its
        # weight pointer stays NULL, clustering_io.c:386-455): shards as read only
        "multiphase": dict(file=mp, side=[".hist1", ".hist2"], n=60, strides=[5, 8, 4],
                           dims=[2, 0, 0],
                           types=[D2_EUCLIDEAN_L2, D2_HISTOGRAM, D2_SPARSE_HISTOGRAM], splits=4,
                           pre=False),
        "dense_hist": dict(file=hd, side=[".hist1"], n=60, strides=[5, 8], dims=[2, 0],
                           types=[D2_EUCLIDEAN_L2, D2_HISTOGRAM], splits=4),
        "words": dict(file=wd, side=[".vocab0"], n=50, strides=[5], dims=[4],
                      types=[D2_WORD_EMBED], splits=2),
    }


@needs_ref
@pytest.mark.parametrize("case", ["mountaindat", "multiphase", "dense_hist", "words"])
def test_written_files_identical_to_reference(tmp_path, case):
    spec = _cases(str(tmp_path))[case]
    ours, ref = str(tmp_path / "ours"), str(tmp_path / "ref")
    _run(_lib.LIB_PATH, spec, ours)
    _run(O.LIB_REF, spec, ref)
    fo, fr = _files(ours), _files(ref)
    assert fo == fr
    produced = [f for f in fo if not f.endswith("data.d2") and "hist" not in f
                and "vocab" not in f]
    # every output the reference writes is there: _c.d2, .label, .label_o,
    # .ind and the shards of both split modes
    want = ["run_c.d2", "run.label", "run.label_o", "raw/data.d2.ind", "raw/data.d2.0"]
    if spec.get("pre", True):
        want += ["pre/data.d2.ind", "pre/data.d2.0"]
    for w in want:
        assert w in produced, w
    for f in produced:
        assert filecmp.cmp(os.path.join(ours, f), os.path.join(ref, f), shallow=False), f


@needs_ref
def test_split_shards_read_back(tmp_path):
    """The raw shards, read back in .ind order, are the objects that went in
    (weights and supports to the 6 printed decimals), and .label_o is the
    shard-order label file put back in input order."""
    spec = _cases(str(tmp_path))["mountaindat"]
    out = str(tmp_path / "ours")
    _run(_lib.LIB_PATH, spec, out)
    ind = np.loadtxt(os.path.join(out, "raw", "data.d2.ind"), dtype=np.int64)
    assert sorted(ind) == list(range(300))
    orig = read_d2(spec["file"], 300, [3, 3], strides=[6, 8])
    got_w, got_s = [], []
    for k in range(spec["splits"]):
        part = read_d2(os.path.join(out, "raw", f"data.d2.{k}"), 300, [3, 3], strides=[6, 8])
        for q in range(part.size):
            w, s = part.phases[0].object(q)
            got_w.append(w)
            got_s.append(s)
    assert len(got_w) == 300
    for q, i in enumerate(ind):
        w, s = orig.phases[0].object(int(i))
        np.testing.assert_allclose(got_w[q], w, atol=2e-6)
        np.testing.assert_allclose(got_s[q], s, atol=1e-6)
    lab = np.loadtxt(os.path.join(out, "run.label"), dtype=int)
    lab_o = np.loadtxt(os.path.join(out, "run.label_o"), dtype=int)
    np.testing.assert_array_equal(lab_o[ind], lab)


@needs_ref
def test_reference_main_links_against_libd2mi(tmp_path):
    """src/app/main.cc and util.cc, unchanged, compiled with the reference's
    own include/d2/clustering.h, link against libd2mi.so and nothing else of
    the reference (no libad2c, no Mosek wrapper, no MPI).  Its
    --prepare_batches path (main.cc:183-194: d2_read, d2_write_split,
    d2_free; no device work) then writes the same shards, as read and
    pre-processed (-Q), as the reference build's d2_write_split."""
    exe = str(tmp_path / "d2_main")
    pkg = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-O2", f"-I{REF}/include", f"-I{REF}/src/app", f"{REF}/src/app/main.cc",
           f"{REF}/src/app/util.cc", f"-L{pkg}", "-ld2mi", f"-Wl,-rpath,{pkg}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    nm = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True)
    undefined = {l.split()[-1].split("@")[0] for l in nm.stdout.splitlines() if l.strip()}
    assert {"d2_clustering", "d2_read", "d2_write_split", "d2_assignment"} <= undefined
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libd2mi.so" in ldd and "libad2c" not in ldd and "mosek" not in ldd

    spec = _cases(str(tmp_path))["mountaindat"]
    ref = str(tmp_path / "ref")
    _run(O.LIB_REF, spec, ref)
    for sub, extra in (("raw", []), ("pre", ["-Q"])):
        d = tmp_path / ("main_" + sub)
        d.mkdir()
        fn = str(d / "data.d2")
        with open(spec["file"]) as src, open(fn, "w") as dst:
            dst.write(src.read())
        r = subprocess.run([exe, "-i", fn, "-p", "2", "-n", "300", "-d", "3,3", "-s", "6,8",
                            "--prepare_batches", "3", *extra], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
        for f in ["data.d2.ind"] + [f"data.d2.{k}" for k in range(3)]:
            assert filecmp.cmp(str(d / f), os.path.join(ref, sub, f), shallow=False), (sub, f)
