"""Build recipe of libgfe.so (the HIP engine, its C ABI and the PRP driver) for gfx950.

    python -m genefer22_amd.build        # or genefer22_amd.build.build()

Compiles in-tree with hipcc (one object per source under _build/, in parallel, then one link); the
.so sits next to this file and is what genefer22_amd loads.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgfe.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GMP_INC = os.environ.get("GFE_GMP_INC", "/opt/conda/include")
GMP_LIB = os.environ.get("GFE_GMP_LIB", "/usr/lib/x86_64-linux-gnu/libgmp.so.10")
GMP_LINK = ["-L" + os.path.dirname(GMP_LIB), "-l:" + os.path.basename(GMP_LIB)]

SOURCES = ["gfe_engine.hip", "gfe_batch.hip", "gfe_spec_engine.hip", "gfe_spec_batch.hip", "gfe_fused.hip", "genefer_driver.cpp",
           "gfe_serve.cpp", "gfe_noalloc.hip"]
OBJDIR = os.path.join(HERE, "_build")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed fp32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    return r.returncode, err


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source to an object (in parallel), then link libgfe.so."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmds = []
    for s in srcs:
        obj = os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o")
        cmds.append([HIPCC, *FLAGS, f"-I{GMP_INC}", "-c", os.path.join(CSRC, s), "-o", obj])
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_run, cmds))
    for cmd, (rc, err) in zip(cmds, results):
        if verbose:
            print(" ".join(cmd), flush=True)
        if rc != 0:
            raise RuntimeError(f"hipcc failed ({rc}) on {cmd[-3]}:\n{err[-6000:]}")
        if verbose and err.strip():
            print(err, file=sys.stderr)
    objs = [c[-1] for c in cmds]
    rc, err = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, *GMP_LINK, "-o", LIB + ".tmp"])
    if rc != 0:
        raise RuntimeError(f"link failed ({rc}):\n{err[-6000:]}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
