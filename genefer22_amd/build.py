"""Build recipe of libgfe.so (the HIP engine, its C ABI and the PRP driver) for gfx950.

    python -m genefer22_amd.build        # or genefer22_amd.build.build()

Compiles in-tree with hipcc; the .so sits next to this file and is what genefer22_amd loads.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgfe.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GMP_INC = os.environ.get("GFE_GMP_INC", "/opt/conda/include")
GMP_LIB = os.environ.get("GFE_GMP_LIB", "/usr/lib/x86_64-linux-gnu/libgmp.so.10")
GMP_LINK = ["-L" + os.path.dirname(GMP_LIB), "-l:" + os.path.basename(GMP_LIB)]

SOURCES = ["gfe_engine.hip", "gfe_batch.hip", "genefer_driver.cpp", "gfe_serve.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed fp32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmd = [HIPCC, *FLAGS, f"-I{GMP_INC}", *srcs, *GMP_LINK, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err[-6000:]}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
