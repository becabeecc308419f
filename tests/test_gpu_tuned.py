"""The autotuners pick any admissible plan at run time, so every plan they may pick is pinned:

* the batch tuner's choice on C3 candidates (n = 15, b in [1e8, 2e9)) is run again on 64
  candidates for 2048 squarings and compared with the reference's RES64 / old64
  (tests/golden/ref_c3t.json, the reference's CPU transform);
* every plan the engine tuner (gfe_tune) times at n = 15, and the compile-time plans plus a seeded
  sample of the others at n = 22, give the default plan's canonical residue after squarings of a
  random balanced residue (the default plan is pinned against the reference elsewhere);
* every compile-time batch plan and a seeded sample of the others at n = 15 give the default
  batch plan's RES64 / old64 after 128 PRP squarings (two 64-round blocks, so the carries deferred
  across a block boundary are covered).
"""
import json
import os
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _ref_c3t():
    with open(os.path.join(HERE, "golden", "ref_c3t.json")) as f:
        return json.load(f)["results"]


def _run_batch(gfe, bt, bases, iters):
    for i, b in enumerate(bases):
        assert bt.submit(b, iters, i) is not None
    got = {}
    while len(got) < len(bases):
        bt.step()
        for r in bt.poll():
            got[r["tag"]] = r
    return [got[i] for i in range(len(bases))]


def test_tuned_batch_plan_matches_reference():
    import genefer22_amd as gfe
    ref = _ref_c3t()[:64]
    bases = [r["b"] for r in ref]
    bt = gfe.Batch(15, len(bases))
    for i, b in enumerate(bases):
        bt.submit(b, 0, i)
    bt.step()
    bt.tune(1)
    plan = bt.get_plan()
    bt.close()
    print(f"tuned batch plan {plan}", flush=True)
    bt = gfe.Batch(15, len(bases))
    bt.set_plan(plan)
    res = _run_batch(gfe, bt, bases, 2048)
    bt.close()
    for r, want in zip(res, ref):
        assert (r["b"], f"{r['res64']:016X}", f"{r['old64']:016X}") == (want["b"], want["res64"], want["old64"]), plan


def _plan_sample(plans, keep_spec, others, seed):
    spec = [p for p in plans if p[5] & 8] if keep_spec else []
    rest = [p for p in plans if not (p[5] & 8)]
    random.Random(seed).shuffle(rest)
    return spec + rest[:others]


@pytest.mark.parametrize("n,b,others", [(15, 1619951882, None), (22, 20204692, 100)])
def test_every_engine_plan_gives_the_default_residue(gfe, n, b, others):
    plans = gfe.candidate_plans(n)
    if others is not None:
        plans = _plan_sample(plans, True, others, seed=n)
    rs = np.random.RandomState(n)
    d0 = rs.randint(-(b // 2), b // 2 + 1, size=1 << n).astype(np.int32)
    pattern = 0b10110
    t = gfe.Transform(b, n)
    t.set_digits(d0, 0)
    for i in range(8):
        t.square_dup((pattern >> i) & 1)
    want = gfe.residue(t.get_digits(0), n, b)
    bad = []
    for p in plans:
        t.set_plan(p)
        t.set_digits(d0, 0)
        for i in range(8):
            t.square_dup((pattern >> i) & 1)
        got = gfe.residue(t.get_digits(0), n, b)
        if (got["res64"], got["hash64"]) != (want["res64"], want["hash64"]):
            bad.append(p)
    t.close()
    print(f"n={n}: {len(plans)} plans checked", flush=True)
    assert not bad, bad[:10]


def test_every_batch_plan_gives_the_default_residue(gfe):
    n = 15
    ref = _ref_c3t()
    bases = [ref[0]["b"], ref[1]["b"]]          # one b above 1e9, one below
    assert bases[0] > 10**9 > bases[1]
    bt = gfe.Batch(n, 2)
    bt.set_graphs(False)
    want = [(r["res64"], r["old64"]) for r in _run_batch(gfe, bt, bases, 128)]
    plans = _plan_sample(gfe.candidate_plans(n), True, 300, seed=7)
    bad = []
    for p in plans:
        bt.set_plan(p)
        got = [(r["res64"], r["old64"]) for r in _run_batch(gfe, bt, bases, 128)]
        if got != want:
            bad.append(p)
    bt.close()
    print(f"n={n}: {len(plans)} batch plans checked", flush=True)
    assert not bad, bad[:10]
