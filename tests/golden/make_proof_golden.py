#!/usr/bin/env python3
"""Writes tests/golden/ref_proof.json and the certificate files ref_cert_<b>_<n>.cert from the
reference's own command-line program.

`make -C oracle cli` compiles the reference's main.cpp and CPU transforms in place into
oracle/_ref/genefer.  For each case this script runs, in a scratch directory, the three roles of
the reference's proof protocol (src/main.cpp:297-315, src/genefer.h:750-1062):

  genefer -p -b B -n N -f g   proof mode: PRP with Pietrzak-Li checkpoints, the Gerbicz-Li check,
                              the proof file g.proof (depth 7, hard-wired at src/main.cpp:237) and
                              the proof key;
  genefer -s -b B -n N -f g   server mode: reads g.proof, recomputes the proof key, writes the
                              certificate g.cert and prints the check key (its three random values
                              come from std::rand seeded with the time, so the certificate differs
                              from run to run; the check key is what a check of that certificate
                              must print);
  genefer -c -b B -n N -f g   check mode: reads g.cert and prints the check key.

Recorded per case: RES64 / old64 / proof key (results.txt), the proof file's SHA-256 and size, the
hash64 of each of its residue records (mu_0 .. mu_7), the certificate's B, and the check key printed
by the server and by the check.  The certificate itself is copied next to the JSON so that a check
on the GPU engine can read the very file the reference wrote.  Runs here, where the reference
checkout exists; the tests read only what it writes.

    python tests/golden/make_proof_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

CLI = os.path.join(ROOT, "oracle", "_ref", "genefer")
# (b, n): the integer-arithmetic proof fixture's candidate (tests/golden/pl_n12.json), the first
# C1 base, and a reference KAT base at n = 13 (src/main.cpp:431)
CASES = [(1026, 12), (146950, 12), (16580478, 13)]
DEPTH = 7


def run(args, cwd):
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=3600)
    if r.returncode != 0:
        raise RuntimeError(f"genefer {' '.join(args)} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return open(os.path.join(cwd, "results.txt")).read().splitlines()[-1]


def field(line, name):
    m = re.search(name + r" = ([0-9A-F]{16})", line)
    if not m:
        raise RuntimeError(f"no {name} in {line!r}")
    return m.group(1)


def proof_records(data: bytes):
    version, depth = struct.unpack_from("<ii", data, 0)
    off, recs = 8, []
    for _ in range(depth + 1):
        size, base = struct.unpack_from("<II", data, off)
        d = np.frombuffer(data, dtype="<i4", count=size, offset=off + 8).copy()
        recs.append((base, d))
        off += 8 + 4 * size
    assert off + 4 == len(data), "proof file: records and CRC trailer"
    return version, depth, recs


def main():
    if not os.path.exists(CLI):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "cli"], check=True)
    O.build()
    out = {"source": "the reference's command-line program (oracle/_ref/genefer, CPU build of src/main.cpp), "
                     "proof / server / check modes", "depth": DEPTH, "cases": []}
    for b, n in CASES:
        with tempfile.TemporaryDirectory() as d:
            p = run(["-p", "-b", str(b), "-n", str(n), "-f", "g"], d)
            data = open(os.path.join(d, "g.proof"), "rb").read()
            version, depth, recs = proof_records(data)
            assert version == 1 and depth == DEPTH
            s = run(["-s", "-b", str(b), "-n", str(n), "-f", "g"], d)
            cert = open(os.path.join(d, "g.cert"), "rb").read()
            c = run(["-c", "-b", str(b), "-n", str(n), "-f", "g"], d)
            name = f"ref_cert_{b}_{n}.cert"
            shutil.copyfile(os.path.join(d, "g.cert"), os.path.join(HERE, name))
        case = {"b": b, "n": n, "res64": field(p, "res64"), "old64": field(p, "old64"), "pkey": field(p, "pkey"),
                "proof_sha256": hashlib.sha256(data).hexdigest(), "proof_bytes": len(data),
                "mu_hash64": [f"{O.hash64(r, n, b):016X}" for base, r in recs],
                "server_pkey": field(s, "pkey"), "ckey": field(s, "ckey"), "check_ckey": field(c, "ckey"),
                "cert": name, "cert_B": struct.unpack_from("<i", cert, 4)[0],
                "cert_sha256": hashlib.sha256(cert).hexdigest()}
        assert case["server_pkey"] == case["pkey"] and case["check_ckey"] == case["ckey"]
        out["cases"].append(case)
        print(f"b={b} n={n}: pkey {case['pkey']} ckey {case['ckey']}", flush=True)
    with open(os.path.join(HERE, "ref_proof.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote ref_proof.json", flush=True)


if __name__ == "__main__":
    main()
