"""Builds libknnj.so (the gfx950 kernels + C ABI) in-tree with hipcc.

    python -m knn_sweet_amd.build          # or knn_sweet_amd.build.build()

Every HIP source under csrc/ is compiled for gfx950 only, with
 - -ffp-contract=off: no implicit fused multiply-add; the kernels write
   __builtin_fmaf where they want one, and the binary64 rerank must follow
   the oracle's unfused operation sequence;
 - -target-feature -packed-fp32-ops: no packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libknnj.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(os.path.dirname(HERE), 'include')}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _headers_of(src: str, seen=None) -> set:
    """Local headers `src` includes, recursively (quoted includes only)."""
    seen = set() if seen is None else seen
    with open(src) as f:
        for line in f:
            line = line.strip()
            if line.startswith('#include "'):
                h = os.path.normpath(os.path.join(os.path.dirname(src), line.split('"')[1]))
                if os.path.exists(h) and h not in seen:
                    seen.add(h)
                    _headers_of(h, seen)
    return seen


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(f) > t for f in [src, __file__, *_headers_of(src)])


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every source (incremental: only those whose object is older
    than the source or a header it includes) and links libknnj.so."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    # the slowest translation units first
    srcs.sort(key=lambda p: -os.path.getsize(p) - (1 << 20 if "k_join_mfma_d" in p else 0))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda p: _compile(p, incremental), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


CLI = os.path.join(HERE, "knnjoin")


def build_cli() -> str:
    """The reference-compatible command-line driver, linked against libknnj.so."""
    src = os.path.join(HERE, "cli", "knnjoin.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(os.path.dirname(HERE), 'include')}", src,
           f"-L{HERE}", "-lknnj", "-Wl,-rpath,$ORIGIN", "-o", CLI + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"knnjoin build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
