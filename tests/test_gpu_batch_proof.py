"""Proof batches (gfe_batch_create_proof) and proof requests of the serving loop: the reference's
proof mode (src/genefer.h:750-766) for many candidates at once, the Pietrzak-Li checkpoints kept
in each slot's HBM registers and the proof's powers run as squaring and multiplying rounds of the
batch's blocks.  Checked against the reference's own program (tests/golden/ref_proof.json: its
proof files at depth 7, written by the CPU build of src/main.cpp), against the integer fixture
(tests/golden/pl_n12.json, depths 3 and 5) and against the single engine's proof of the same
candidate (gfe_run_proof), alone and in a mix."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture(name):
    path = os.path.join(HERE, "golden", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return json.load(open(path))


def h(x):
    return f"{x:016X}"


def run(bt, bases):
    for i, b in enumerate(bases):
        assert bt.submit(b, 0, i) is not None
    got = {}
    while len(got) < len(bases):
        bt.step()
        for r in bt.poll():
            got[r["tag"]] = r
    return [got[i] for i in range(len(bases))]


def test_proof_batch_matches_the_reference_program(gfe):
    """Both n = 12 candidates of the reference's proof fixture in one batch, beside two candidates
    without a fixture (checked against the single engine's proof)."""
    fx = fixture("ref_proof.json")
    depth = fx["depth"]
    cases = [c for c in fx["cases"] if c["n"] == 12]
    extra = [2000000, 999999618]
    bt = gfe.Batch(12, 4, depth=depth)
    out = run(bt, [c["b"] for c in cases] + extra)
    bt.close()
    for c, r in zip(cases, out):
        assert (h(r["res64"]), h(r["old64"]), r["gl_ok"]) == (c["res64"], c["old64"], 1), c["b"]
        assert r["depth"] == depth
        assert [h(x) for x in r["mu_hash64"]] == c["mu_hash64"], c["b"]
        assert h(r["pkey"]) == c["pkey"], c["b"]
    for b, r in zip(extra, out[len(cases):]):
        alone = gfe.run_proof(b, 12, depth)
        assert (r["res64"], r["pkey"], r["mu_hash64"]) == (alone["res64"], alone["pkey"], alone["mu_hash64"]), b


@pytest.mark.parametrize("depth", [3, 5])
def test_proof_batch_matches_integer_arithmetic(gfe, depth):
    fx = fixture("pl_n12.json")
    want = fx["proofs"][str(depth)]
    bt = gfe.Batch(fx["n"], 2, depth=depth)
    out = run(bt, [fx["b"], fx["b"]])
    bt.close()
    for r in out:
        assert (h(r["res64"]), h(r["old64"]), r["gl_ok"]) == (fx["res64"], fx["old64"], 1)
        assert [h(x) for x in r["mu_hash64"]] == want["mu_hash64"]
        assert h(r["pkey"]) == want["pkey"]


def test_serving_proof_requests_in_a_mix(gfe):
    """Proof requests of n = 12 and 13 (depth 7) arrive among quick tests; each proof equals the
    reference program's, and the quick tests their alone results."""
    fx = fixture("ref_proof.json")
    depth = fx["depth"]
    reqs = []
    for k, c in enumerate(fx["cases"]):
        reqs.append((c["n"], c["b"], 0, 0.01 * k, depth))
        reqs.append((12, 1534 + 2 * k, 4096, 0.01 * k + 0.005, 0))
    results, stats = gfe.serve_replay(reqs)
    assert stats["completed"] == len(reqs)
    for (n, b, iters, _, d), r in zip(reqs, results):
        assert (r["n"], r["b"]) == (n, b)
        if d:
            c = next(c for c in fx["cases"] if c["b"] == b)
            assert (h(r["res64"]), r["gl_ok"], r["depth"]) == (c["res64"], 1, depth), b
            assert [h(x) for x in r["mu_hash64"]] == c["mu_hash64"], b
            assert h(r["pkey"]) == c["pkey"], b
        else:
            alone = gfe.run_batch(n, [b], iters=iters)[0]
            assert (r["res64"], r["old64"], r["depth"]) == (alone["res64"], alone["old64"], 0), b
