#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own CPU transforms.

The reference sources are compiled in place by `make -C oracle ref` into oracle/_ref/libgfref.so
(oracle/ref_shim.cpp drives the reference's transform interface and gint exactly as its PRP loop
does).  This script runs here, where the reference checkout exists; the GPU tests read only the
JSON it writes.  Workloads come from genefer22_amd.workloads (seeded).

    python tests/golden/make_golden.py [--jobs 8] [--only c1,c3t,...]

Outputs (all values are the reference's, hex RES64 / old64):
  ref_c1.json  : C1 (n=12, b in [1024, 2^20)), 8 full quick tests (RES64, old64, GL passed)
  ref_c2.json  : C2 (n=16, b in [1e8, 1.5e9]), 2 candidates, RES64 every 1000 squarings to 200,000
                 and the SHA-256 of the canonical digits at 200,000 (int32 little-endian, digit 0 first)
  ref_c3t.json : C3 (n=15, b in [1e8, 2e9)), all 512 candidates, RES64 after 2048 squarings
  ref_c3f.json : C3, full quick tests of the first --c3f-count candidates (default all 512); written
                 after every finished candidate and resumed from what the file already holds
  ref_c1i32.json : C1 full quick tests through the reference's exact integer transform (impl "i32",
                 src/transformCPUi32.h), the first 4 bases of ref_c1.json
  ref_c3i32.json : C3, 16 candidates, RES64 after 2048 squarings through the "i32" transform
  ref_c4.json  : C4 (n=22, b in [1e7, 3.5e7]), 1 candidate, RES64 every 2000 squarings to 100,000
                 and the SHA-256 of the canonical digits at 100,000
  ref_c4max.json : C4 at the top of its range (b = 35,000,000, n = 22), RES64 every 2000 squarings to
                 20,000 and the SHA-256 of the canonical digits there
  ref_c5.json  : C5 serving mix, the first 48 requests, RES64 after their 4096 squarings
  ref_c5all.json : C5, all 2000 requests (the same (n, b) for any arrival rate), RES64 and old64
                 after their 4096 squarings; written after every finished request and resumed from
                 what the file already holds
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from genefer22_amd import workloads as W  # noqa: E402


def _ref(b, n, threads, check_error=False, impl=""):
    import oracle as O
    return O.RefEngine(b, n, threads=threads, check_error=check_error, impl=impl)


def job_full(b, n, threads, impl=""):
    t0 = time.time()
    r = _ref(b, n, threads, impl=impl)
    es, res, gl = r.prp(iters=0, b_gl=0, do_gl=True)
    return {"b": b, "n": n, "esize": es, "res64": f"{res.res64:016X}", "old64": f"{res.old64:016X}",
            "is_prp": res.is_one, "gl_ok": gl, "impl": r.impl, "seconds": round(time.time() - t0, 1)}


def job_trunc(b, n, iters, threads, check_error=False, impl=""):
    r = _ref(b, n, threads, check_error, impl)
    es, res, _ = r.prp(iters=iters, b_gl=0, do_gl=False)
    return {"b": b, "n": n, "iters": iters, "res64": f"{res.res64:016X}", "old64": f"{res.old64:016X}",
            "impl": r.impl, "max_error": r.error()}


def job_trace(b, n, iters, every, threads, check_error=False):
    t0 = time.time()
    r = _ref(b, n, threads, check_error)
    tr = r.prp_trace(iters, every)
    canon = r.get_digits(canonical=True).astype("<i4")
    return {"b": b, "n": n, "iters": iters, "every": every,
            "sha256": hashlib.sha256(canon.tobytes()).hexdigest(),
            "res64": [f"{a:016X}" for a, _ in tr], "old64": [f"{o:016X}" for _, o in tr],
            "impl": r.impl, "max_error": r.error(), "seconds": round(time.time() - t0, 1)}


def _full_resumable(ex, name, cfg_name, bases, n, src):
    """Full quick tests of `bases`, the file rewritten (in base order) after each finished test and
    the tests the file already holds skipped, so a long run can be stopped and resumed."""
    from concurrent.futures import as_completed
    path = os.path.join(HERE, name)
    done = {}
    if os.path.exists(path):
        with open(path) as f:
            for r in json.load(f)["results"]:
                done[r["b"]] = r
    futs = {ex.submit(job_full, b, n, 1): b for b in bases if b not in done}
    for fut in as_completed(futs):
        done[futs[fut]] = fut.result()
        write(name, {"source": src, "config": cfg_name,
                     "results": [done[b] for b in bases if b in done]}, quiet=True)
        print(f"{name}: {len(done)}/{len(bases)}", flush=True)


def _trunc_resumable(ex, name, cfg_name, reqs, src):
    """Truncated runs of the serving requests `reqs`, keyed by request id, the file rewritten (in
    request order) after each finished run and the runs it already holds skipped."""
    from concurrent.futures import as_completed
    path = os.path.join(HERE, name)
    done = {}
    if os.path.exists(path):
        with open(path) as f:
            for r in json.load(f)["results"]:
                done[r["rid"]] = r
    # the large n first, so the pool ends on short runs
    todo = sorted((q for q in reqs if q.rid not in done), key=lambda q: -q.n)
    futs = {ex.submit(job_trunc, q.b, q.n, q.iters, 1): q for q in todo}
    for k, fut in enumerate(as_completed(futs)):
        q = futs[fut]
        done[q.rid] = dict(fut.result(), rid=q.rid)
        if k % 16 == 15 or k == len(futs) - 1:
            write(name, {"source": src, "config": cfg_name,
                         "results": [done[r.rid] for r in reqs if r.rid in done]}, quiet=True)
            print(f"{name}: {len(done)}/{len(reqs)}", flush=True)


def write(name, obj, quiet=False):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1)
    if not quiet:
        print(f"wrote {name}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--only", default="c1,c1i32,c3i32,c2,c3t,c3f,c4,c4max,c5")
    ap.add_argument("--c3f-count", type=int, default=512)
    args = ap.parse_args()
    only = set(args.only.split(","))
    ex = ProcessPoolExecutor(max_workers=args.jobs)
    src = "reference CPU transforms (oracle/_ref/libgfref.so), PRP loop of src/genefer.h:479-552"

    if "c1" in only:
        bases = W.config_bases(W.C1, 8)
        futs = [ex.submit(job_full, b, 12, 1) for b in bases]
        write("ref_c1.json", {"source": src, "config": "C1", "results": [f.result() for f in futs]})
    if "c3t" in only:
        bases = W.config_bases(W.C3, 512)
        futs = [ex.submit(job_trunc, b, 15, 2048, 1) for b in bases]
        write("ref_c3t.json", {"source": src, "config": "C3 truncated", "results": [f.result() for f in futs]})
    if "c5" in only:
        reqs = W.serving_requests(48, 100.0)
        futs = [ex.submit(job_trunc, q.b, q.n, q.iters, 1, True) for q in reqs]
        write("ref_c5.json", {"source": src, "config": "C5 serving sample (first 48 requests, seed 5, rate 100/s)",
                              "results": [dict(f.result(), rid=q.rid) for f, q in zip(futs, reqs)]})
    if "c5all" in only:
        _trunc_resumable(ex, "ref_c5all.json", "C5 serving workload (all 2000 requests, seed 5)",
                         W.serving_requests(2000, 24.0), src)
    if "c2" in only:
        bases = W.config_bases(W.C2, 2)
        futs = [ex.submit(job_trace, b, 16, 200_000, 1000, 4) for b in bases]
        write("ref_c2.json", {"source": src, "config": "C2", "results": [f.result() for f in futs]})
    if "c1i32" in only:
        bases = W.config_bases(W.C1, 4)
        futs = [ex.submit(job_full, b, 12, 1, "i32") for b in bases]
        write("ref_c1i32.json", {"source": src.replace("transforms", "transform_i32"), "config": "C1 (i32)",
                                 "results": [f.result() for f in futs]})
    if "c3i32" in only:
        bases = W.config_bases(W.C3, 16)
        futs = [ex.submit(job_trunc, b, 15, 2048, 1, False, "i32") for b in bases]
        write("ref_c3i32.json", {"source": src.replace("transforms", "transform_i32"),
                                 "config": "C3 truncated (i32)", "results": [f.result() for f in futs]})
    if "c3f" in only:
        _full_resumable(ex, "ref_c3f.json", "C3 full tests", W.config_bases(W.C3, args.c3f_count), 15, src)
    if "c4" in only:
        b = W.config_bases(W.C4, 1)[0]
        r = job_trace(b, 22, 100_000, 2000, args.jobs, True)
        write("ref_c4.json", {"source": src, "config": "C4", "results": [r]})
    if "c4max" in only:
        r = job_trace(W.C4["hi"] - 1, 22, 20_000, 2000, args.jobs, True)
        write("ref_c4max.json", {"source": src, "config": "C4, largest b", "results": [r]})
    ex.shutdown()


if __name__ == "__main__":
    main()
