"""Mixed batches (gfe_batch_create_mixed): candidates of n = 12 and n = 13 in one batch, every
squaring block one launch of the single-workgroup squaring with each workgroup at its slot's n
(below 128 active slots of n = 13 those run the four-step passes after it).  Each candidate's
residue equals its result in a batch of its own n, and the n = 12 full tests equal the reference's
(tests/golden/ref_c1.json)."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def run_mixed(gfe, cands, iters, capacity):
    bt = gfe.Batch(13, capacity, mixed=True)
    for i, (n, b) in enumerate(cands):
        assert bt.submit(b, iters, i, n=n) is not None
    got = {}
    while len(got) < len(cands):
        bt.step()
        for r in bt.poll():
            got[r["tag"]] = r
    bt.close()
    return [got[i] for i in range(len(cands))]


def alone(gfe, cands, iters):
    out = {}
    for n in (12, 13):
        bases = [b for m, b in cands if m == n]
        if bases:
            for b, r in zip(bases, gfe.run_batch(n, bases, iters=iters)):
                out[(n, b)] = r
    return [out[c] for c in cands]


def test_mixed_n_in_one_batch_equals_alone(gfe):
    """A few candidates of each n: n = 12 in the mixed launch, n = 13 on the four-step passes."""
    cands = [(12, 1534), (13, 999999618), (12, 30000000), (13, 20000000), (12, 999999618), (13, 1999999266),
             (12, 1999999266), (13, 146950)]
    got = run_mixed(gfe, cands, 2048, 8)
    want = alone(gfe, cands, 2048)
    for (n, b), g, w in zip(cands, got, want):
        assert g["n"] == n and g["b"] == b
        assert (g["res64"], g["old64"]) == (w["res64"], w["old64"]), (n, b)


def test_mixed_launch_with_many_n13_slots(gfe):
    """128+ active slots of n = 13 beside slots of n = 12: the whole block in one launch."""
    cands = [(13, 1000000 + 2 * k) for k in range(130)] + [(12, 5000000 + 2 * k) for k in range(10)]
    got = run_mixed(gfe, cands, 256, 140)
    want = alone(gfe, cands, 256)
    bad = [(n, b) for (n, b), g, w in zip(cands, got, want) if (g["res64"], g["old64"]) != (w["res64"], w["old64"])]
    assert not bad, bad[:5]


def test_mixed_full_tests_match_the_reference(gfe):
    path = os.path.join(HERE, "golden", "ref_c1.json")
    if not os.path.exists(path):
        pytest.skip("ref_c1.json not generated")
    rs = json.load(open(path))["results"][:4]
    cands = [(12, r["b"]) for r in rs] + [(13, 1000000), (13, 123456790)]
    got = run_mixed(gfe, cands, 0, 6)
    for r, g in zip(rs, got):
        assert (f"{g['res64']:016X}", f"{g['old64']:016X}", g["gl_ok"]) == (r["res64"], r["old64"], 1), r["b"]
    for (n, b), g in zip(cands[4:], got[4:]):
        w = gfe.run_batch(n, [b])[0]
        assert (g["res64"], g["gl_ok"]) == (w["res64"], 1), b
