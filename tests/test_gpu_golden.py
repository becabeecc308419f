"""GPU parity against fixtures written by the reference's own CPU transforms (tests/golden/
make_golden.py), on the measurement configurations C1-C5.  Bit-exact RES64 and old64, and on the
long traces the SHA-256 of the whole canonical residue."""
import hashlib
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def golden(name):
    path = os.path.join(HERE, "golden", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return json.load(open(path))["results"]


def h(x):
    return f"{x:016X}"


def canonical_sha256(gfe, digits, n, b):
    """SHA-256 of the canonical digits in [0, b) (int32 little-endian, digit 0 first), as the
    fixture writer hashes the reference's canonical register."""
    canon = gfe.residue(digits, n, b, canonical=True)["digits"]
    return hashlib.sha256(canon.astype("<i4").tobytes()).hexdigest()


def test_c1_full_quick_tests(gfe):
    """C1: n = 12, b in [1024, 2^20): full quick tests, one engine per candidate."""
    for r in golden("ref_c1.json"):
        g = gfe.run_quick(r["b"], 12)
        assert (h(g["res64"]), h(g["old64"])) == (r["res64"], r["old64"]), r["b"]
        assert g["gl_ok"] == 1 and g["esize"] == r["esize"]


def test_c1_full_quick_tests_batched(gfe):
    rs = golden("ref_c1.json")
    out = gfe.run_batch(12, [r["b"] for r in rs])
    for r, g in zip(rs, out):
        assert (h(g["res64"]), h(g["old64"]), g["gl_ok"]) == (r["res64"], r["old64"], 1), r["b"]


def test_c2_checkpoints_to_200000(gfe):
    """C2: n = 16, b in [1e8, 1.5e9]: RES64 at every 1000th of the first 200,000 squarings, and the
    whole canonical residue at 200,000 (SHA-256)."""
    for r in golden("ref_c2.json"):
        tr = gfe.run_prp_trace(r["b"], 16, r["iters"], r["every"])
        assert [h(a) for a, _ in tr] == r["res64"], r["b"]
        assert [h(o) for _, o in tr] == r["old64"]
        g = gfe.run_prp(r["b"], 16, iters=r["iters"], b_gl=-1, digits=True)
        assert h(g["res64"]) == r["res64"][-1], r["b"]
        assert canonical_sha256(gfe, g["digits"], 16, r["b"]) == r["sha256"], r["b"]


def test_c3_all_512_truncated_in_one_batch(gfe):
    """C3: the 512 seeded n = 15 candidates in one batch, RES64 after 2048 squarings."""
    rs = golden("ref_c3t.json")
    assert len(rs) == 512
    out = gfe.run_batch(15, [r["b"] for r in rs], iters=2048)
    bad = [(r["b"], h(g["res64"]), r["res64"]) for r, g in zip(rs, out) if (h(g["res64"]), h(g["old64"])) != (r["res64"], r["old64"])]
    assert not bad, bad[:5]


def test_c3_full_tests_batched(gfe):
    """C3: full quick tests (PRP + Gerbicz-Li) of the first 64 seeded candidates in one batch (the
    whole fixture, up to all 512, runs in bench.py --c3-full)."""
    rs = golden("ref_c3f.json")[:64]
    out = gfe.run_batch(15, [r["b"] for r in rs])
    for r, g in zip(rs, out):
        assert (h(g["res64"]), h(g["old64"])) == (r["res64"], r["old64"]), r["b"]
        assert g["gl_ok"] == 1 and g["iters"] == r["esize"]


def test_c4_checkpoints_to_100000(gfe):
    """C4: n = 22, b in [1e7, 3.5e7]: RES64 at every 2000th of the first 100,000 squarings; and the
    same residue at 100,000 with the Gerbicz-Li product updated every 2000 iterations (the loop's
    residue does not depend on the cadence, src/genefer.h:529-535), whole canonical residue included
    (SHA-256)."""
    for r in golden("ref_c4.json"):
        tr = gfe.run_prp_trace(r["b"], 22, r["iters"], r["every"])
        assert [h(a) for a, _ in tr] == r["res64"], r["b"]
        g = gfe.run_prp(r["b"], 22, iters=r["iters"], b_gl=2000, digits=True)
        assert h(g["res64"]) == r["res64"][-1], r["b"]
        assert canonical_sha256(gfe, g["digits"], 22, r["b"]) == r["sha256"], r["b"]


def test_c4_top_of_range(gfe):
    """C4 at its largest b (35,000,000, n = 22): the default plan is the fused two-digit-limb plan,
    whose CRT has the least slack at the top of the range; RES64 at every 2000th of the first 20,000
    squarings and the whole canonical residue there (SHA-256) equal the reference's."""
    for r in golden("ref_c4max.json"):
        t = gfe.Transform(r["b"], 22)
        assert t.get_plan()[5] == 42, t.get_plan()
        t.close()
        tr = gfe.run_prp_trace(r["b"], 22, r["iters"], r["every"])
        assert [h(a) for a, _ in tr] == r["res64"], r["b"]
        g = gfe.run_prp(r["b"], 22, iters=r["iters"], b_gl=2000, digits=True)
        assert h(g["res64"]) == r["res64"][-1], r["b"]
        assert canonical_sha256(gfe, g["digits"], 22, r["b"]) == r["sha256"], r["b"]


def test_c5_serving_sample(gfe):
    """C5: the first 48 serving requests (n in 12..20), RES64 after each one's 4096 squarings,
    batched per n."""
    rs = golden("ref_c5.json")
    by_n = {}
    for r in rs:
        by_n.setdefault(r["n"], []).append(r)
    for n, group in sorted(by_n.items()):
        out = gfe.run_batch(n, [r["b"] for r in group], iters=group[0]["iters"])
        for r, g in zip(group, out):
            assert (h(g["res64"]), h(g["old64"])) == (r["res64"], r["old64"]), (n, r["b"])
