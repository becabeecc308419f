"""Synthetic, seeded workloads of the squaring engine (candidates b^(2^n) + 1).

Bases follow the reference's command-line rules (src/main.cpp:244-253): b even, 1024 <= b <= 2e9,
b not a power of two.  Every generator is deterministic in its seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

MASK64 = (1 << 64) - 1


def splitmix64(state: int):
    """SplitMix64 stream (Steele, Lea, Flood 2014)."""
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def valid_base(b: int) -> bool:
    return b % 2 == 0 and 1024 <= b <= 2_000_000_000 and (b & (b - 1)) != 0


def seeded_bases(seed: int, lo: int, hi: int, count: int) -> list[int]:
    """`count` even bases uniform in [lo, hi) (hi exclusive), valid for the reference."""
    out, g = [], splitmix64(seed)
    while len(out) < count:
        b = lo + 2 * (next(g) % ((hi - lo) // 2))
        if valid_base(b):
            out.append(b)
    return out


def esize(b: int, n: int) -> int:
    """Bits of E = b^(2^n): squarings of a full test (exact for n <= 18; for larger n the float
    estimate floor(2^n log2 b) + 1, off only when 2^n log2 b is within 1e-6 of an integer)."""
    if n <= 18:
        return (b ** (1 << n)).bit_length()
    return int(math.floor((1 << n) * math.log2(b))) + 1


# Configurations of the measurement plan (one GPU):
#   C1 plain n=12, b in [1024, 2^20), full test;  C2 plain n=16, b in [1e8, 1.5e9], 200,000 squarings;
#   C3 batched 512 x n=15, b in [1e8, 2e9);  C4 plain n=22, b in [1e7, 3.5e7], 1e5 squarings;
#   C5 serving, n uniform in 12..20, b up to the reference's CPU limit per n.
C1 = dict(n=12, lo=1024, hi=1 << 20, seed=12)
C2 = dict(n=16, lo=100_000_000, hi=1_500_000_001, seed=16)
C3 = dict(n=15, lo=100_000_000, hi=2_000_000_000, seed=15)
C4 = dict(n=22, lo=10_000_000, hi=35_000_001, seed=22)

# b limit per n of the serving mix (the reference's CPU limits, src/genefer.h:1401-1415).
SERVE_B_MAX = {12: 2_000_000_000, 13: 2_000_000_000, 14: 2_000_000_000, 15: 2_000_000_000,
               16: 1_500_000_000, 17: 1_000_000_000, 18: 1_000_000_000, 19: 70_000_000, 20: 55_000_000}


def config_bases(cfg: dict, count: int) -> list[int]:
    return seeded_bases(cfg["seed"], cfg["lo"], cfg["hi"], count)


@dataclass
class Request:
    rid: int
    arrival: float   # seconds after the start of the run
    n: int
    b: int
    iters: int       # PRP squarings to run (0 = full test)


def serving_requests(count: int, rate_per_s: float, seed: int = 5, iters: int = 4096,
                     n_lo: int = 12, n_hi: int = 20) -> list[Request]:
    """Poisson arrivals at `rate_per_s`; n uniform in [n_lo, n_hi]; b uniform even up to the limit
    of its n (from 1024).  Each request runs the first `iters` squarings of its PRP loop (full tests
    at n = 19-20 take 13-27 M squarings each, beyond a bounded benchmark)."""
    g = splitmix64(seed)
    t, out = 0.0, []
    for rid in range(count):
        u = (next(g) >> 11) / float(1 << 53)
        t += -math.log(1.0 - u) / rate_per_s
        n = n_lo + next(g) % (n_hi - n_lo + 1)
        while True:
            b = 1024 + 2 * (next(g) % ((SERVE_B_MAX[n] - 1024) // 2))
            if valid_base(b):
                break
        out.append(Request(rid, t, n, b, iters))
    return out
