"""Serving: candidate requests of mixed n through the native serving loop (include/gfe_serve.h).

`run` replays a list of requests (seeded Poisson arrivals from genefer22_amd.workloads) on the
wall clock through gfe_serve_replay and summarises time-to-result (p50/p99) and throughput;
`plain_serial_seconds` prices the same requests on the one-candidate-per-launch path from measured
per-n squaring times.
"""
from __future__ import annotations

import numpy as np

from . import Transform, serve_replay
from .workloads import Request


def percentile(xs, q: float) -> float:
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else float("nan")


def summarize(results: list[dict]) -> dict:
    """Latency and throughput of a replay (times are seconds on the replay clock)."""
    ttr = [r["t_done"] - r["t_arrival"] for r in results]
    wait = [r["t_admit"] - r["t_arrival"] for r in results]
    t_first = min(r["t_arrival"] for r in results)
    t_last = max(r["t_done"] for r in results)
    span = t_last - t_first
    squarings = sum(r["iters"] for r in results)
    return {
        "requests": len(results),
        "makespan_s": round(span, 4),
        "p50_s": round(percentile(ttr, 50), 5),
        "p99_s": round(percentile(ttr, 99), 5),
        "max_s": round(max(ttr), 5),
        "p50_admit_wait_s": round(percentile(wait, 50), 5),
        "p99_admit_wait_s": round(percentile(wait, 99), 5),
        "requests_per_hour": round(len(results) / span * 3600.0, 1),
        "squarings": int(squarings),
    }


def run(requests: list[Request], device: int = 0, hbm_budget: int = 0, prepare: bool = True):
    """Replays `requests` (arrival times in seconds from the start) through the serving loop.
    With `prepare`, a pool of every n in the mix is created before the clock starts, as a server
    does at start-up.  Returns (results in request order, summary, server statistics)."""
    rq = [(q.n, q.b, q.iters, q.arrival) for q in requests]
    ns = sorted({q.n for q in requests}) if prepare else []
    results, stats = serve_replay(rq, device=device, hbm_budget=hbm_budget, prepare=ns)
    return results, summarize(results), stats


def plain_ms_per_squaring(ns, b_of_n: dict, reps: int = 256, device: int = 0) -> dict:
    """ms per squaring of one candidate on the plain path (one engine, tuned), per n."""
    out = {}
    for n in ns:
        t = Transform(b_of_n[n], n, device=device)
        t.tune(2)
        t.set(1)
        t.time_squarings(32)
        out[n] = t.time_squarings(reps) / reps
        t.close()
    return out


def plain_serial_seconds(requests: list[Request], ms_by_n: dict) -> float:
    """Seconds to run the requests one after another on the plain path (squarings only)."""
    return sum(q.iters * ms_by_n[q.n] for q in requests) / 1e3
