"""The C-ABI library loads and exports every function include/d2mi.h
declares (no compute call: CPU only)."""
import ctypes
import os
import re

from d2_kmeans_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(d2mi_\w+|d2_\w+|_dpdist2)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared()
    assert {"d2mi_create", "d2mi_badmm", "d2mi_pdist2", "d2mi_set_phase"} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    engine_api = {n for n in declared() if n.startswith("d2mi_")}
    assert engine_api <= set(_lib.SIGNATURES), engine_api - set(_lib.SIGNATURES)


def test_reference_globals_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for g in ("d2_alg_type", "p_badmm_options", "badmm_clu_options", "badmm_cen_options",
              "world_rank", "nprocs"):
        ctypes.c_int.in_dll(lib, g)
