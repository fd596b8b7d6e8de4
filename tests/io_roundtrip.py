"""Drives the reference's text-I/O entry points through ONE library, in a
process of its own (the product and the reference build export the same
names, so they never share a process here).

    python tests/io_roundtrip.py LIB SPEC.json OUTDIR

SPEC names the input file (with its side files beside it), the phases
(n, strides, dims, types) and the number of shards.  In OUTDIR the script
writes, through LIB only:

* ``run_c.d2``       d2_write of everything read (clustering_io.c:204-250)
* ``run.label``      d2_write_labels of a fixed label pattern (:252-275)
* ``raw/data.d2.*``  d2_write_split, as read (:329-385)
* ``pre/data.d2.*``  d2_write_split, pre-processed (:386-470)
* ``run.label_o``    d2_write_labels_serial over raw/data.d2.ind (:277-324)

The test compares the folders of two libraries byte for byte.
"""
import ctypes as C
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle.oracle import RefMph  # noqa: E402


def main(libpath, spec_path, outdir):
    spec = json.load(open(spec_path))
    L = C.CDLL(libpath, mode=os.RTLD_LAZY | os.RTLD_GLOBAL)
    L.d2_allocate.argtypes = [C.POINTER(RefMph), C.c_int, C.c_size_t, C.POINTER(C.c_int),
                              C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.d2_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(RefMph)]
    L.d2_write.argtypes = [C.c_char_p, C.POINTER(RefMph)]
    L.d2_write_labels.argtypes = [C.c_char_p, C.POINTER(RefMph)]
    L.d2_write_labels_serial.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(RefMph)]
    L.d2_write_split.argtypes = [C.c_char_p, C.POINTER(RefMph), C.c_int, C.c_char]
    s_ph = len(spec["dims"])
    src = spec["file"]
    os.makedirs(outdir, exist_ok=True)
    fns = {}
    for sub in ("raw", "pre"):
        os.makedirs(os.path.join(outdir, sub), exist_ok=True)
        fns[sub] = os.path.join(outdir, sub, "data.d2")
        shutil.copy(src, fns[sub])
        for side in spec.get("side", []):
            shutil.copy(src + side, fns[sub] + side)
    ints = lambda xs: (C.c_int * s_ph)(*xs)  # noqa: E731
    m = RefMph()
    rc = L.d2_allocate(C.byref(m), s_ph, spec["n"], ints(spec["strides"]), ints(spec["dims"]),
                       ints(spec["types"]))
    assert rc == 0
    assert L.d2_read(fns["raw"].encode(), None, C.byref(m)) == 0
    out = os.path.join(outdir, "run")
    L.d2_write((out + "_c.d2").encode(), C.byref(m))
    for i in range(m.size):
        m.label[i] = (7 * i + 3) % 5
    L.d2_write_labels((out + ".label").encode(), C.byref(m))
    splits = spec["splits"]
    assert L.d2_write_split(fns["raw"].encode(), C.byref(m), splits, b"\0") == 0
    if spec.get("pre", True):
        assert L.d2_write_split(fns["pre"].encode(), C.byref(m), splits, b"\1") == 0
    L.d2_write_labels_serial((fns["raw"] + ".ind").encode(), out.encode(), C.byref(m))
    sys.stdout.flush()


if __name__ == "__main__":
    main(*sys.argv[1:4])
