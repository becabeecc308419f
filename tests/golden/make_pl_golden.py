#!/usr/bin/env python3
"""Writes tests/golden/pl_n12.json: the Pietrzak-Li proof of one n = 12 candidate, from integer
arithmetic alone (no transform).

The PRP residues follow from the exponent bits (src/genefer.h:479-552: r <- r^2, doubled when bit i
of E = b^N is set, checkpoint k at i = k B_PL), and the proof's mu_k, the weights w and the proof
key follow src/genefer.h:620-734: mu_0 = ckpt[0]; for k = 1..depth, i = 2^(depth-k),
mu_k = ckpt[i]^w[0] * prod_{j = i, 2i, .. < L/2} ckpt[i + 2j]^w[j]; q_k = hash32(mu_k);
v1 = prod_k mu_k^q_k; w[i/2 + j] = w[j] q_k for j = 0, i, 2i, .. (i > 1); key = hash64(v1).
hash32 / hash64 are the gint hashes of canonical digits (oracle.hash32 / hash64, checked against
the reference's gint).  Everything is Python integers mod F = b^N + 1, so the values pin the proof
against mathematics rather than against any engine.  Takes a few minutes (41k squarings of a
41k-bit number).

    python tests/golden/make_pl_golden.py
"""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

B, N_LOG, DEPTHS = 1026, 12, (3, 5, 7)


def proof(ck: dict, depth: int, n: int, b: int, F: int):
    h32 = lambda v: O.hash32(O.int_to_digits(v, n, b), n, b)
    h64 = lambda v: O.hash64(O.int_to_digits(v, n, b), n, b)
    L = 1 << depth
    mus = [ck[0]]
    q = h32(ck[0])
    v1 = pow(ck[0], q, F)
    w = [0] * (L // 2)
    w[0] = q
    for k in range(1, depth + 1):
        i = 1 << (depth - k)
        mu = pow(ck[i], w[0], F)
        for j in range(i, L // 2, i):
            mu = mu * pow(ck[i + 2 * j], w[j], F) % F
        mus.append(mu)
        q = h32(mu)
        v1 = v1 * pow(mu, q, F) % F
        if i > 1:
            for j in range(0, L // 2, i):
                w[i // 2 + j] = w[j] * q
    return [h64(m) for m in mus], h64(v1)


def main():
    O.build()
    b, n = B, N_LOG
    N = 1 << n
    F = b ** N + 1
    E = b ** N
    es = E.bit_length()
    bpl = {d: O.b_pietrzak_li(es, d) for d in DEPTHS}
    ck = {d: {} for d in DEPTHS}
    t0 = time.time()
    r = 1
    for i in range(es - 1, -1, -1):
        r = r * r
        if (E >> i) & 1:
            r <<= 1
        r %= F
        for d in DEPTHS:
            if i % bpl[d] == 0:
                ck[d][i // bpl[d]] = r
    res = O.residue(O.int_to_digits(r, n, b), n, b)
    out = {"b": b, "n": n, "esize": es, "res64": f"{res.res64:016X}", "old64": f"{res.old64:016X}", "proofs": {}}
    for d in DEPTHS:
        mus, key = proof(ck[d], d, n, b, F)
        out["proofs"][str(d)] = {"B_PL": bpl[d], "mu_hash64": [f"{h:016X}" for h in mus], "pkey": f"{key:016X}"}
    with open(os.path.join(HERE, "pl_n12.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote pl_n12.json in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
