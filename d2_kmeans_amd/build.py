"""Builds libd2mi.so (HIP kernels for gfx950 + the C ABI of include/d2mi.h)
in-tree with hipcc.  ``python -m d2_kmeans_amd.build``."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libd2mi.so")
SOURCES = ["context.hip", "badmm.hip", "pdist2.hip", "transport.hip", "labeling.hip",
           "serve.hip", "compat.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("D2MI_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 ops (v_pk_*_f32) stay off: plain float code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _stale(src: str, obj: str) -> bool:
    """True when ``obj`` is older than its source or any header it may use."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps += [os.path.join(ROOT, "include", h) for h in os.listdir(os.path.join(ROOT, "include"))]
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = False, jobs: int = 4, out: str = LIB, extra=(),
          changed_only: bool = False) -> str:
    """Compiles every source (only out-of-date ones with ``changed_only``)
    and links ``out``; ``extra`` hipcc flags (for example -D defines) apply
    to every translation unit."""
    objdir = os.path.join(PKG, "_obj" if out == LIB else "_obj_" + os.path.basename(out))
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        if changed_only and not _stale(os.path.join(CSRC, src), obj):
            continue
        cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs)
    _wait(procs)
    tmp = out + ".tmp"
    # -Bsymbolic-functions: the library's own calls (d2_write from
    # d2_clustering, ...) bind to its own definitions even in a process that
    # also loaded another libad2c-like library globally
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic-functions",
           "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    if out == LIB:
        build_cli(verbose)
    return out


CLI = os.path.join(PKG, "d2mi")


def build_cli(verbose: bool = False) -> str:
    """The d2 command line (reference flags) linked against libd2mi.so."""
    cmd = ["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(CSRC, "d2_cli.cpp"), f"-L{PKG}", "-ld2mi", "-Wl,-rpath,$ORIGIN",
           "-o", CLI + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(CLI + ".tmp", CLI)
    return CLI


def _wait(procs):
    while procs:
        src, p = procs.pop(0)
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            sys.stderr.write(text)
            raise RuntimeError(f"hipcc failed on {src}")
        # keep real warnings visible, drop the host pass's note on the GPU-only feature
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    print(build(verbose=True, changed_only="--changed" in sys.argv))
