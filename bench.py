#!/usr/bin/env python3
"""Headline benchmark of the gfx950 squaring engine for b^(2^n) + 1.

One JSON line on stdout (progress on stderr):
  * value      : ns per squaring of one candidate at n = 22 (N = 4,194,304 digits), b seeded in
                 [1e7, 3.5e7] (prime set NTT3 31-bit), residue resident in HBM, W warm-up then K timed
                 squarings (one step = one squaring = column pass, row pass, column pass, CRT + carry),
                 timed with HIP events on the engine's stream;
  * roofline   : the dominant kernel's algorithmic bytes (one read + one write of the rns x N uint32
                 residues, 8 rns N bytes per launch) over its average launch duration, against 8 TB/s;
  * passes     : every pass's algorithmic HBM rate (and the whole squaring's) against 8 TB/s;
  * valu       : the same kernel's butterfly rate and estimated VALU-slot use (integer-bound path);
  * job_n22    : the n = 22 job end to end (config.job_n22): 100,000 PRP squarings with Gerbicz-Li
                 updates every 2000, engine set-up and read-out included, host clock; RES64 against
                 the reference's value at 100,000;
  * batched_n15: candidates/hour at n = 15, 512 seeded candidates in one batch vs one candidate per
                 launch, from timed iteration blocks (full-test length = esize squarings); the timed
                 (autotuned) plan is then run on all 512 candidates for 2048 squarings and every
                 RES64 / old64 compared with the reference's (the run fails on a mismatch); with
                 --c3-full the whole workload (512 full tests) runs end to end on the host clock;
  * serving    : p50/p99 time-to-result of seeded Poisson requests of mixed n (12..20) through the
                 serving loop, and candidates/hour (each truncated to its first 4096 squarings)
                 against the serial plain path: the 2000-request C5 workload at about 80 % of capacity
                 (serving) and 40 requests at 2/s (serving_light: requests seldom overlap, so
                 time-to-result is a lone request's service time); every served RES64 / old64 is
                 compared with the reference's for the same request (serving.parity, from
                 tests/golden/ref_c5all.json when present; the run fails on a mismatch);
  * parity     : RES64 after P PRP iterations on the GPU vs the reference's own CPU transform;
  * cpu_baseline: the reference's CPU transform (compiled from its sources into oracle/_ref) timed on
                 this host's cores on a bounded sample of the same workload;
  * small_configs: the plain path at n = 12, 15, 16 (C1, C3, C2) beside the reference's CPU
                 transform on one core, and C3 candidates/hour of one CPU process per core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from genefer22_amd import workloads as W  # noqa: E402

METRIC = "ns per squaring at n=22; candidates/hour at n=15 batched vs plain, bit-exact"
HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz, one wave-instruction per CU per clock
# (integer multiplies, min and add3 issue at this rate on gfx950; 32-bit adds somewhat faster).
VALU_PEAK_LANE_INSTR = 256 * 64 * 2.4e9

# Per-launch counters of the dominant kernel for plans profiled with rocprofv3 on MI355X: HBM bytes
# = FETCH_SIZE x 1.998 + WRITE_SIZE x 1.000 (the copy kernel's calibration) and VALU instructions per
# radix-2 butterfly = SQ_INSTS_VALU x 64 / butterflies.  These are profile constants, not live
# measurements: each entry names the kernel (as rocprofv3 prints it) and the summary under profiles/
# it comes from, and the line reports them with that source.  Keyed by the plan fields the kernel
# depends on: the row pass by (a, row_loge, row_ng, flags).
MEASURED = {
    ("row", 8, 5, 1, 8): {"kernel": "gfe::ks_row_sq<gfe::ModS, 5, 14>", "fetch_mib": 144.9, "write_mib": 48.0,
                          "valu_instr_per_butterfly": 11.9, "profile": "profiles/n22_spec_plan_kernels.md"},
    ("row", 7, 5, 1, 40): {"kernel": "gfe::ks_row_sq<gfe::ModS, 5, 14>", "fetch_mib": 96.8, "write_mib": 32.0,
                           "valu_instr_per_butterfly": 11.92, "profile": "profiles/n22_limb_plan_kernels.md"},
}
MEASURED[("row", 7, 5, 1, 42)] = {"kernel": "gfe::ks_row_sq<gfe::ModS, 5, 14>", "fetch_mib": 96.8, "write_mib": 32.0,
                                  "valu_instr_per_butterfly": 11.92, "profile": "profiles/n22_limb_fused_kernels.md"}


def measured(kernel: str, plan) -> dict:
    a, cl, cng, rl, rng, fl = plan
    key = (kernel, a, rl, rng, fl) if kernel == "row" else (kernel, a, cl, cng, fl)
    m = dict(MEASURED.get(key, {}))
    if m:
        m["traffic"] = int((m["fetch_mib"] + m["write_mib"]) * 2**20)
    return m


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_n22(G, steps: int, warmup: int, plan_arg: str | None, tune: bool):
    n = 22
    b = W.config_bases(W.C4, 1)[0]
    t = G.Transform(b, n)
    if plan_arg:
        t.set_plan([int(v) for v in plan_arg.split(",")])
    elif tune:
        log(f"[bench] n=22 b={b}: tuning")
        t.tune(6)
    plan = t.info()
    log(f"[bench] plan: {plan} {t.get_plan()}")
    # HBM calibration: register copies of N int32 (k_copy, 2 x 16 MiB per launch) for the
    # FETCH_SIZE / WRITE_SIZE scale of the profiles
    for _ in range(8):
        t.copy(1, 2)
        t.copy(2, 1)
    t.sync()
    t.set(1)
    dup_pattern = 0x9B4D_63A1_5C27_E8F0
    t.time_squarings(max(warmup, 1), dup_pattern)
    ms = t.time_squarings(steps, dup_pattern)
    kreps = 200
    _, km = t.time_squarings(kreps, dup_pattern, per_kernel=True)
    fused = bool(t.get_plan()[5] & 2)
    limb = bool(t.get_plan()[5] & 32)
    if limb and fused:  # the normalisation runs inside the inverse column pass: no pass of its own
        names = ["col_fwd", "row", "col_inv_norm", "norm"]
    elif fused:
        names = ["col_fwd_all", "row", "col_inv_norm", "fold"]
    else:
        names = ["col_fwd", "row", "col_inv", "norm"]
    rns = 3 if "NTT3" in plan else 2
    N = 1 << n
    # the transform: rns primes x N residues, or for the two-digit-limb transform 4 primes x N/2 limbs
    tn, tp = (n - 1, 4) if limb else (n, rns)
    words = tp << tn
    dom = max(range(4), key=lambda i: km[i])
    ng = t.get_plan()[2]
    if limb and fused:  # digits (+ range carries) in, residues out | row | residues in, digits + carries out
        slots = 16 * (N // 2) // ng
        alg = {0: 4 * N + 4 * words + slots, 1: 8 * words, 2: 4 * words + 4 * N + slots, 3: 0}
    elif fused:  # digits in, residues out | row | residues in, digits out | segment carries
        alg = {0: 4 * N + 4 * rns * N, 1: 8 * rns * N, 2: 4 * rns * N + 4 * N + 8 * N // ng, 3: 16 * N // ng}
    else:
        alg = {0: 4 * N + 4 * words, 1: 8 * words, 2: 8 * words, 3: 4 * words + 4 * N}
    alg_bytes = alg[dom]
    achieved = alg_bytes / (km[dom] * 1e-3) / 1e9
    # achieved algorithmic HBM rate of every pass (one read + one write of what the pass must move)
    passes = {names[i]: {"us": round(km[i] * 1e3, 2), "alg_bytes": alg[i],
                         "GBps": round(alg[i] / (km[i] * 1e-3) / 1e9, 1),
                         "frac": round(alg[i] / (km[i] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)} for i in range(4) if km[i] > 0}
    sq_bytes = sum(alg.values())
    passes["squaring"] = {"us": round(ms / steps * 1e3, 2), "alg_bytes": sq_bytes,
                          "GBps": round(sq_bytes / (ms / steps * 1e-3) / 1e9, 1),
                          "frac": round(sq_bytes / (ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # butterflies of the dominant pass (row: (n - a) levels fwd + inv, N/2 per level per prime)
    a = t.get_plan()[0]
    levels = {0: a, 1: 2 * (tn - a), 2: a}.get(dom, 0)
    bfly = tp * levels * ((1 << tn) // 2)
    bfly_rate = bfly / (km[dom] * 1e-3) if levels else 0.0
    meas = measured(names[dom], t.get_plan())
    ipb = meas.get("valu_instr_per_butterfly")
    res = {
        "b": b, "plan": plan, "plan_tuple": list(t.get_plan()), "limb": limb, "ms_per_step": ms / steps, "ns_per_squaring": ms / steps * 1e6,
        "kernel_us": {names[i]: round(km[i] * 1e3, 2) for i in range(4)},
        "passes": passes,
        "roofline": {
            "bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": meas.get("traffic"),
            "alg_bytes_per_launch": alg_bytes, "kernel_us": round(km[dom] * 1e3, 2),
            "duration_source": f"HIP events around {kreps} back-to-back launches of the kernel alone on the engine's stream",
            "traffic_source": (f"profile constant: {meas['profile']} ({meas['kernel']}, FETCH_SIZE x 1.998 + WRITE_SIZE)"
                               if meas else None),
            "note": "the kernel is bound by integer VALU issue (see valu), not by HBM",
        },
        "valu": {
            "kernel": names[dom], "butterflies": bfly, "butterflies_per_s": round(bfly_rate),
            "valu_instr_per_butterfly": ipb,
            "valu_instr_source": f"profile constant: {meas['profile']} (SQ_INSTS_VALU)" if meas else None,
            "achieved_lane_instr_per_s": round(bfly_rate * ipb) if ipb else None,
            "peak_lane_instr_per_s": VALU_PEAK_LANE_INSTR,
            "frac": round(bfly_rate * ipb / VALU_PEAK_LANE_INSTR, 4) if ipb else None,
        },
    }
    t.close()
    return res


def bench_batched_n15(G, count: int, block: int, tune: bool = True, plan=None, full: bool = False):
    """Candidates/hour at n = 15: `count` seeded candidates in one batch vs one candidate per launch.

    Both legs time squaring iterations with HIP events on resident state; candidates/hour =
    3600 s / (seconds per squaring x mean esize) x candidates per iteration."""
    n = 15
    bases = W.config_bases(W.C3, count)
    mean_es = sum(W.esize(b, n) for b in bases) / count
    log(f"[bench] n=15: {count} candidates, mean esize {mean_es:.0f}")
    bt = G.Batch(n, count)
    for i, b in enumerate(bases):
        if bt.submit(b, 0, i) is None:
            raise RuntimeError("batch refused a candidate")
    bt.step()  # the first 64-iteration block of every candidate's program
    if plan:
        bt.set_plan(plan)
        plan_b = bt.info()
    else:
        plan_b = bt.tune(2) if tune else bt.info()
    log(f"[bench] batch plan: {plan_b}")
    bt.time_rounds(block)
    per_iter_b = bt.time_rounds(4 * block) / (4 * block)
    plan_tuple_b = list(bt.get_plan())
    bt.close()
    parity = batch_parity_c3(G, bases, plan_tuple_b)
    log(f"[bench] n=15 parity under the timed plan: {parity}")
    if not parity["match"]:
        raise RuntimeError(f"n=15 batch residues under plan {plan_tuple_b} differ from the reference: {parity}")
    e2e = batch_end_to_end_c3(G, bases, plan_tuple_b) if full else None
    plain_ms = []
    for b in bases[:4]:
        t = G.Transform(b, n)
        if tune:
            t.tune(4)
        t.set(1)
        t.time_squarings(block)
        plain_ms.append(t.time_squarings(4 * block) / (4 * block))
        t.close()
    per_sq_plain = sum(plain_ms) / len(plain_ms)
    cand_h_batched = 3600.0 / (per_iter_b * 1e-3 * mean_es) * count
    cand_h_plain = 3600.0 / (per_sq_plain * 1e-3 * mean_es)
    return {
        "candidates": count, "mean_esize": round(mean_es), "block": block, "batch_plan": plan_b, "batch_plan_tuple": plan_tuple_b,
        "batched_ms_per_iteration": round(per_iter_b, 4), "plain_ms_per_squaring": round(per_sq_plain, 5),
        "candidates_per_hour_batched": round(cand_h_batched, 1),
        "candidates_per_hour_plain": round(cand_h_plain, 1),
        "speedup": round(cand_h_batched / cand_h_plain, 2),
        "candidates_per_hour_note": "from timed 64-round blocks x mean esize; the Gerbicz-Li products and "
                                    "register operations are not in it (end_to_end measures them)",
        "parity": parity,
        "end_to_end": e2e,
    }


def _run_batch(G, bases, iters, plan):
    bt = G.Batch(15, len(bases))
    bt.set_plan(plan)
    for i, b in enumerate(bases):
        if bt.submit(b, iters, i) is None:
            raise RuntimeError("batch refused a candidate")
    got = {}
    last = time.perf_counter()
    while len(got) < len(bases):
        bt.step()
        for r in bt.poll():
            got[r["tag"]] = r
        if time.perf_counter() - last > 30.0:  # a long run reports its progress
            last = time.perf_counter()
            log(f"[bench] n=15 batch: {len(got)}/{len(bases)} candidates done")
    bt.close()
    return [got[i] for i in range(len(bases))]


def batch_parity_c3(G, bases, plan):
    """Every candidate's RES64 / old64 after 2048 PRP squarings under the timed batch plan against
    the reference's CPU transform (tests/golden/ref_c3t.json)."""
    ref = {r["b"]: r for r in json.load(open(os.path.join(ROOT, "tests", "golden", "ref_c3t.json")))["results"]}
    t0 = time.perf_counter()
    res = _run_batch(G, bases, 2048, plan)
    wall = time.perf_counter() - t0
    checked = [r for r in res if r["b"] in ref]
    bad = [r["b"] for r in checked if (f"{r['res64']:016X}", f"{r['old64']:016X}") != (ref[r["b"]]["res64"], ref[r["b"]]["old64"])]
    return {"plan_tuple": list(plan), "squarings": 2048, "candidates": len(res), "checked": len(checked),
            "mismatches": bad[:8], "match": not bad and len(checked) == len(res), "wall_s": round(wall, 2),
            "reference": "tests/golden/ref_c3t.json (the reference's CPU transform)"}


def batch_end_to_end_c3(G, bases, plan):
    """The whole C3 workload: every candidate's full quick test (PRP loop, Gerbicz-Li updates and
    check, read-out) in one batch under the timed plan, on the host clock; RES64 against the
    reference's full tests where tests/golden/ref_c3f.json holds them."""
    path = os.path.join(ROOT, "tests", "golden", "ref_c3f.json")
    ref = {r["b"]: r for r in json.load(open(path))["results"]} if os.path.exists(path) else {}
    t0 = time.perf_counter()
    res = _run_batch(G, bases, 0, plan)
    wall = time.perf_counter() - t0
    checked = [r for r in res if r["b"] in ref]
    bad = [r["b"] for r in checked if f"{r['res64']:016X}" != ref[r["b"]]["res64"]]
    return {"candidates": len(res), "wall_s": round(wall, 1),
            "candidates_per_hour_measured": round(len(res) / wall * 3600.0, 1),
            "gl_ok": sum(1 for r in res if r["gl_ok"] == 1), "checked_against_reference": len(checked),
            "mismatches": bad[:8], "match": not bad}


def engine_kind(G, b: int, n: int) -> tuple:
    """What a plain engine of (b, n) runs: n, its plan (digit or limb transform, fused or not) and
    the prime set the reference's rule selects."""
    return (n, tuple(G.engine_plan(b, n)), tuple(G.select(b, n)))


def plain_ms_by_kind(G, reqs, cache: dict) -> dict:
    """ms per squaring of one candidate on the plain path, per engine kind of the requests (one
    representative b each, the library's default plan), added to `cache`."""
    for q in reqs:
        k = engine_kind(G, q.b, q.n)
        if k in cache:
            continue
        t = G.Transform(q.b, q.n)
        t.set(1)
        t.time_squarings(32)
        cache[k] = t.time_squarings(256) / 256
        t.close()
    return cache


def bench_serving(G, count: int, rate: float, iters: int, ms_cache: dict):
    """Serving: `count` seeded requests (n uniform in 12..20, b up to the reference's limit per n,
    Poisson arrivals at `rate`/s, the first `iters` PRP squarings each) replayed on the wall clock
    through the native serving loop; p50/p99 time-to-result and requests/hour, against the same
    requests run one after another on the plain one-candidate path (per-n ms/squaring measured on
    the library's default plans) and against each request's own plain run time (p50 of
    iters x ms/squaring of its engine kind -- n, plan and prime set, which differ between bases of one
    n: the latency floor of one request alone on the plain path, set-up and read-out excluded)."""
    from genefer22_amd import serve
    reqs = W.serving_requests(count, rate, iters=iters)
    t0 = time.perf_counter()
    results, summary, stats = serve.run(reqs)
    wall = time.perf_counter() - t0
    parity = serving_parity(reqs, results)
    if parity and not parity["match"]:
        raise RuntimeError(f"served residues differ from the reference: {parity}")
    plain_ms_by_kind(G, reqs, ms_cache)
    ms_of = [ms_cache[engine_kind(G, q.b, q.n)] for q in reqs]
    serial = sum(q.iters * m for q, m in zip(reqs, ms_of)) / 1e3
    alone = sorted(q.iters * m / 1e3 for q, m in zip(reqs, ms_of))
    ms_by_n = {}
    for q, m in zip(reqs, ms_of):
        ms_by_n.setdefault(q.n, []).append(m)
    return {
        "requests": count, "arrival_rate_per_s": rate, "squarings_per_request": iters,
        "p50_s": summary["p50_s"], "p99_s": summary["p99_s"], "makespan_s": summary["makespan_s"],
        "p50_admit_wait_s": summary["p50_admit_wait_s"],
        "plain_alone_p50_s": round(serve.percentile(alone, 50), 5),
        "plain_alone_p99_s": round(serve.percentile(alone, 99), 5),
        "requests_per_hour": summary["requests_per_hour"],
        "candidates_per_hour": summary["requests_per_hour"],
        "truncation": f"each candidate runs the first {iters} squarings of its PRP loop (full tests at n = 19-20 "
                      f"take 13-27 M squarings each)",
        "plain_serial_requests_per_hour": round(count / serial * 3600.0, 1),
        "speedup_vs_plain_serial": round(serial / summary["makespan_s"], 2),
        "plain_ms_per_squaring_by_n": {int(k): round(sum(v) / len(v), 5) for k, v in sorted(ms_by_n.items())},
        "plain_kinds_measured": len(ms_cache),
        "pools_created": stats["pools_created"], "bytes_peak": stats["bytes_peak"], "wall_s": round(wall, 2),
        "lone_runs": stats.get("solo", 0),
        "parity": parity,
    }


def serving_parity(reqs, results, path=None):
    """Every served RES64 / old64 against the reference CPU transform's for the same request
    (tests/golden/ref_c5all.json, by request id: the (n, b) of a request id do not depend on the
    arrival rate); None when the fixture is absent."""
    path = path or os.path.join(ROOT, "tests", "golden", "ref_c5all.json")
    if not os.path.exists(path):
        return None
    ref = {r["rid"]: r for r in json.load(open(path))["results"]}
    checked, bad = 0, []
    for q, g in zip(reqs, results):
        r = ref.get(q.rid)
        if r is None or (r["n"], r["b"], r["iters"]) != (q.n, q.b, q.iters):
            continue
        checked += 1
        if (f"{g['res64']:016X}", f"{g['old64']:016X}") != (r["res64"], r["old64"]):
            bad.append(q.rid)
    return {"checked": checked, "of": len(reqs), "mismatches": bad[:16], "match": not bad,
            "reference": "tests/golden/ref_c5all.json (the reference's CPU transform)"}


def parity_n22(G, O, b: int, iters: int, threads: int):
    """RES64 after `iters` PRP iterations: GPU driver vs the reference CPU transform."""
    g = G.run_prp(b, 22, iters=iters, b_gl=-1)
    ref = O.RefEngine(b, 22, threads=threads, check_error=True)
    es, r, _ = ref.prp(iters=iters, b_gl=-1)
    err = ref.error()
    return {"iters": iters, "gpu_res64": f"{g['res64']:016X}", "ref_res64": f"{r.res64:016X}",
            "match": g["res64"] == r.res64, "ref_impl": ref.impl, "ref_max_error": err}


def bench_c4_job(G, b: int, iters: int, b_gl: int):
    """The n = 22 job as the PRP loop runs it (src/genefer.h:479-552): engine set-up, the first `iters`
    squarings with the Gerbicz-Li product updated every b_gl squarings, and the residue read-out,
    timed on the host clock; RES64 compared with the reference's value at 100,000 squarings
    (tests/golden/ref_c4.json, written from the reference's CPU transform)."""
    t0 = time.perf_counter()
    g = G.run_prp(b, 22, iters=iters, b_gl=b_gl)
    wall = time.perf_counter() - t0
    out = {"b": b, "iters": g["iters"], "b_gl": b_gl, "wall_s": round(wall, 3),
           "ns_per_squaring_job": round(wall / g["iters"] * 1e9, 1), "res64": f"{g['res64']:016X}"}
    path = os.path.join(ROOT, "tests", "golden", "ref_c4.json")
    if iters == 100_000 and os.path.exists(path):
        ref = json.load(open(path))["results"][0]
        if ref["b"] == b:
            out["ref_res64"] = ref["res64"][-1]
            out["match"] = out["res64"] == ref["res64"][-1]
    return out


def cpu_baseline(O, b: int, threads: int, budget_s: float):
    """The reference's CPU transform at n = 22 (same b), squarings timed for about budget_s."""
    ref = O.RefEngine(b, 22, threads=threads)
    ref.set(1)
    for _ in range(3):
        ref.square_dup(False)
    count, t0 = 0, time.perf_counter()
    while True:
        ref.square_dup(count & 1)
        count += 1
        el = time.perf_counter() - t0
        if el >= budget_s or count >= 20000:
            break
    return {"value": round(el / count * 1e9, 1), "unit": "ns/squaring", "cores": threads, "kind": "reference",
            "sample": f"{count} squarings of one n=22 candidate (b={b}) with the reference's '{ref.impl}' CPU "
                      f"transform, {threads} thread(s), {el:.1f} s"}


def bench_small_configs(G, O, cores: int, budget_s: float = 2.0):
    """The plain one-candidate path at C1 (n = 12), C3 (n = 15) and C2 (n = 16), first seeded b of
    each: GPU ns per squaring (library default plan, HIP events) beside the reference's CPU
    transform on one core (about budget_s of squarings each, as the reference's bench times
    squareDup, src/genefer.h:1285-1307); for C3 also full-test candidates/hour of `cores`
    single-threaded CPU processes against the batched GPU leg."""
    out = {}
    for name, cfg in (("C1", W.C1), ("C3", W.C3), ("C2", W.C2)):
        n = cfg["n"]
        b = W.config_bases(cfg, 1)[0]
        t = G.Transform(b, n)
        t.set(1)
        t.time_squarings(64)
        gpu_ns = t.time_squarings(1024) / 1024 * 1e6
        t.close()
        ref = O.RefEngine(b, n, threads=1)
        ref.set(1)
        for _ in range(3):
            ref.square_dup(False)
        count, t0 = 0, time.perf_counter()
        while True:
            ref.square_dup(count & 1)
            count += 1
            el = time.perf_counter() - t0
            if el >= budget_s or count >= 200_000:
                break
        cpu_ns = el / count * 1e9
        out[name] = {"n": n, "b": b, "gpu_ns_per_squaring": round(gpu_ns, 1),
                     "cpu_ns_per_squaring_1core": round(cpu_ns, 1), "cpu_impl": ref.impl, "cpu_squarings": count}
    bases = W.config_bases(W.C3, 512)
    mean_es = sum(W.esize(b, 15) for b in bases) / len(bases)
    out["C3"]["cpu_candidates_per_hour"] = round(cores * 3600.0 / (out["C3"]["cpu_ns_per_squaring_1core"] * 1e-9 * mean_es), 1)
    out["C3"]["cpu_processes"] = cores
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the n=15 batched leg")
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--c3-full", action="store_true",
                    help="also run the whole C3 workload end to end (every candidate's full test, about 6 minutes)")
    ap.add_argument("--batch-plan", default=None, help="fixed n=15 batch plan a,cl,cng,rl,rng,flags (default: autotuned)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests of the C5 serving replay")
    ap.add_argument("--serve-rate", type=float, default=24.0,
                    help="arrivals/s of the C5 replay (about 80 %% of the serving loop's capacity for this mix)")
    ap.add_argument("--light-requests", type=int, default=40, help="requests of the light-load serving replay")
    ap.add_argument("--light-rate", type=float, default=2.0,
                    help="arrivals/s of the light-load serving replay (about a fifth of one plain engine's capacity "
                         "for this mix, so that requests seldom overlap and time-to-result shows a lone request's)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--job-iters", type=int, default=100_000,
                    help="squarings of the timed n=22 PRP job with Gerbicz-Li updates every 2000 (0: skip)")
    ap.add_argument("--plan", default=None, help="fixed n=22 pass plan a,cl,cng,rl,rng,flags (default: the library's)")
    ap.add_argument("--no-tune", action="store_true", help="time the library's default batch plan instead of the autotuned one")
    ap.add_argument("--tune-n22", action="store_true",
                    help="autotune the n=22 plan (default: the library's default plan, which the tuner picks within "
                         "run-to-run noise; a fixed plan keeps the line comparable with the committed profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): candidates are independent and shard "
              f"across GPUs with no exchange; run one process per GPU instead.", file=sys.stderr)
        sys.exit(2)

    import genefer22_amd as G
    G.lib()
    res = bench_n22(G, args.steps, args.warmup, args.plan, args.tune_n22)
    log(f"[bench] n=22: {res['ns_per_squaring']:.0f} ns/squaring; kernels {res['kernel_us']}")

    job = bench_c4_job(G, res["b"], args.job_iters, 2000) if args.job_iters > 0 else None
    if job:
        log(f"[bench] n=22 job: {job}")

    bplan = [int(v) for v in args.batch_plan.split(",")] if args.batch_plan else None
    batched = None if args.no_batch else bench_batched_n15(G, args.batch, 64, not args.no_tune, bplan, args.c3_full)
    if batched:
        log(f"[bench] n=15: {batched}")

    serving = serving_light = None
    if not args.no_serve:
        ms_cache = {}
        serving = bench_serving(G, args.serve_requests, args.serve_rate, 4096, ms_cache)
        log(f"[bench] serving (saturated): {serving}")
        serving_light = bench_serving(G, args.light_requests, args.light_rate, 4096, ms_cache)
        log(f"[bench] serving (light load): {serving_light}")

    threads = min(16, os.cpu_count() or 1)
    parity = cpu = small = None
    if not args.no_cpu:
        import oracle as O
        if O.RefEngine.available():
            parity = parity_n22(G, O, res["b"], 200, threads)
            log(f"[bench] parity: {parity}")
            cpu = cpu_baseline(O, res["b"], threads, args.cpu_seconds)
            log(f"[bench] cpu baseline: {cpu}")
            small = bench_small_configs(G, O, threads)
            log(f"[bench] plain path at C1/C3/C2 beside the CPU: {small}")
        else:
            log("[bench] oracle/_ref/libgfref.so missing: no CPU baseline")

    line = {
        "metric": METRIC,
        "value": round(res["ns_per_squaring"], 1),
        "unit": "ns/squaring",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded b; residue 2^k mod b^N+1 from set(1) and a fixed dup pattern)",
        "config": {
            "workload": "one candidate b^(2^22)+1, N=4194304, b seeded in [1e7,3.5e7] ("
                        + ("2^21 limbs of two digits, four 31-bit primes" if res["limb"] else "NTT3, 31-bit primes") + ")",
            "n": 22, "b": res["b"], "plan": res["plan"], "plan_tuple": res["plan_tuple"], "kernel_us": res["kernel_us"],
            "job_n22": job, "batched_n15": batched, "serving": serving, "serving_light": serving_light, "parity": parity,
            "small_configs": small,
        },
        "roofline": res["roofline"],
        "passes": res["passes"],
        "valu": res["valu"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
