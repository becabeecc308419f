"""GPU parity of the batched engine: every candidate's result in a batch equals its own
single-candidate run, the reference's KATs and the oracle, with mixed prime sets in one batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_quick_n12_kats(gfe):
    """Full quick tests of the reference's n=12 KATs (src/main.cpp:416-425) in one batch, all four
    prime sets mixed; RES64 equal to the KATs and the Gerbicz-Li check passes."""
    kats = {23445612: None, 45686464: None, 999999618: None, 1999999266: None,
            23448336: 0x81CFF004ACE85CAA, 45687570: 0x6C3EA0FBFFEE2C34,
            1000000000: 0x4E43A5B93273C649, 2000000000: 0xE5F3EA34F6C68EBD}
    bases = list(kats)
    out = gfe.run_batch(12, bases)
    for b, r in zip(bases, out):
        assert r["b"] == b and r["gl_ok"] == 1, r
        if kats[b] is not None:
            assert r["res64"] == kats[b], (b, hex(r["res64"]))
    # same residues as one candidate per engine
    for b, r in zip(bases[:4], out[:4]):
        single = gfe.run_quick(b, 12)
        assert (single["res64"], single["old64"]) == (r["res64"], r["old64"])


@pytest.mark.parametrize("n", [13, 15])
def test_batch_truncated_matches_oracle(gfe, oracle, n):
    """First k PRP squarings of several candidates in one batch (slot refill: capacity < count)
    against the oracle's restatement."""
    rs = np.random.RandomState(n)
    bases = [int(x) * 2 for x in rs.randint(50_000_000, 1_000_000_000, 6)] + [1999999266, 20000]
    iters = 48 if n == 13 else 24
    out = gfe.run_batch(n, bases, iters=iters, capacity=5)
    for b, r in zip(bases, out):
        e = oracle.Engine(b, n)
        oracle.prp(e, oracle.exponent(b, n), iters=iters, B_GL=-1)
        want = e.residue(0)
        assert r["iters"] == iters
        assert r["res64"] == want.res64 and r["old64"] == want.old64, (b, hex(r["res64"]), hex(want.res64))


def test_batch_alone_equals_mixed(gfe):
    """A candidate's result does not depend on what else is in the batch."""
    mix = [1000000000, 1534, 30000000, 1999999266, 777777778, 4096 + 2]
    alone = gfe.run_batch(14, [mix[0]], iters=300)[0]
    mixed = gfe.run_batch(14, mix, iters=300)[0]
    assert (alone["res64"], alone["old64"], alone["hash64"]) == (mixed["res64"], mixed["old64"], mixed["hash64"])


def test_batch_api_submit_step_poll(gfe):
    bt = gfe.Batch(12, 3)
    slots = [bt.submit(b, iters=100, tag=i) for i, b in enumerate([1000000000, 1534, 2000000000])]
    assert sorted(slots) == [0, 1, 2]
    assert bt.submit(1026, iters=10) is None  # full
    done = []
    while len(done) < 3:
        bt.step()
        done += bt.poll()
    assert sorted(r["tag"] for r in done) == [0, 1, 2]
    for r in done:
        single = gfe.run_prp(r["b"], 12, iters=100, b_gl=-1)
        assert single["res64"] == r["res64"]
    bt.close()


@pytest.mark.parametrize("plan", [(7, 4, 8, 4, 8, 1), (6, 3, 16, 4, 4, 0), (4, 4, 32, 3, 4, 1), (7, 4, 8, 4, 8, 3),
                                  (6, 3, 16, 4, 4, 2), (6, 3, 16, 4, 4, 4), (7, 4, 8, 4, 8, 5),
                                  (6, 3, 16, 4, 8, 8), (7, 4, 8, 4, 16, 8), (6, 3, 16, 4, 8, 24)])
def test_batch_plan_variants(gfe, plan):
    """Batched squarings under fixed plans (incl. factored row twiddles) equal single runs."""
    bases = [1000000000, 1534, 30000000, 1999999266]
    bt = gfe.Batch(13, len(bases))
    bt.set_plan(plan)
    assert bt.get_plan() == plan
    for i, b in enumerate(bases):
        bt.submit(b, iters=200, tag=i)
    done = []
    while len(done) < len(bases):
        bt.step()
        done += bt.poll()
    bt.close()
    for r in done:
        assert r["res64"] == gfe.run_prp(r["b"], 13, iters=200, b_gl=-1)["res64"], (plan, r["b"])


def test_batch_graphs_equal_eager(gfe):
    """Squaring blocks replayed as captured HIP graphs (hundreds of replays, with the quick test's
    register operations launched eagerly between them) give the results of kernel-by-kernel
    launches, and the full quick test equals the plain driver's."""
    cases = [(1534, 0), (1000000000, 700), (1999999266, 700), (30000000, 700)]
    out = {}
    for graphs in (True, False):
        bt = gfe.Batch(12, len(cases))
        bt.set_graphs(graphs)
        for i, (b, it) in enumerate(cases):
            bt.submit(b, iters=it, tag=i)
        done = []
        while len(done) < len(cases):
            bt.step()
            done += bt.poll()
        bt.close()
        out[graphs] = sorted((r["tag"], r["res64"], r["old64"], r["hash64"]) for r in done)
    assert out[True] == out[False]
    full = gfe.run_quick(1534, 12)
    assert out[True][0][1] == full["res64"]
    for tag, (b, it) in enumerate(cases[1:], start=1):
        assert out[True][tag][1] == gfe.run_prp(b, 12, iters=it, b_gl=-1)["res64"], b


@pytest.mark.parametrize("n,count", [(12, 40), (13, 160)])
def test_batch_fused_squaring_by_occupancy(gfe, n, count):
    """With the single-workgroup squaring in the plan: at n = 13 blocks of 128 or more active slots
    run it and blocks of fewer run the four-step passes (the tail of the run, as candidates finish
    one by one); at n = 12 every block runs it.  Every residue equals its single run."""
    from genefer22_amd import workloads
    bases = workloads.seeded_bases(1300 + n, 1024, 2_000_000_000, count)
    bt = gfe.Batch(n, count)
    assert bt.get_plan()[5] & 16
    iters = [200 + 3 * i for i in range(count)]
    for i, b in enumerate(bases):
        bt.submit(b, iters=iters[i], tag=i)
    done = []
    while len(done) < count:
        bt.step()
        done += bt.poll()
    bt.close()
    by_tag = {r["tag"]: r for r in done}
    for i in (0, 1, count // 2, count - 2, count - 1):
        alone = gfe.run_prp(bases[i], n, iters=iters[i], b_gl=-1)
        assert (by_tag[i]["res64"], by_tag[i]["old64"]) == (alone["res64"], alone["old64"]), (i, bases[i])


def test_batch_graph_cache_turnover(gfe):
    """Candidates that finish highest slot first shrink the grid depth block after block, so each
    block needs a graph of its own depth and the cache of executable graphs turns over many times
    while earlier launches are still queued; every residue equals its single run."""
    from genefer22_amd import workloads
    n, cap = 14, 24
    bases = workloads.seeded_bases(1414, 1024, 2_000_000_000, cap)
    iters = [64 * (cap - i) + 5 for i in range(cap)]
    bt = gfe.Batch(n, cap)
    assert not bt.get_plan()[5] & 16
    bt.set_graphs(True)
    for i, b in enumerate(bases):
        assert bt.submit(b, iters=iters[i], tag=i) == i
    done = []
    while len(done) < cap:
        bt.step()
        done += bt.poll()
    bt.close()
    by_tag = {r["tag"]: r for r in done}
    for i, b in enumerate(bases):
        alone = gfe.run_prp(b, n, iters=iters[i], b_gl=-1)
        assert (by_tag[i]["res64"], by_tag[i]["old64"]) == (alone["res64"], alone["old64"]), (i, b)


def test_batch_default_small_n_squares_in_one_workgroup(gfe):
    """A batch of 40 candidates at n = 12 defaults to the single-workgroup squaring (plan bit 16:
    one launch per block of rounds, one workgroup per slot); its residues equal single runs on the
    four-step engine."""
    from genefer22_amd import workloads
    bases = workloads.seeded_bases(1212, 1024, 2_000_000_000, 40)
    bt = gfe.Batch(12, 40)
    assert bt.get_plan()[5] & 16
    for i, b in enumerate(bases):
        bt.submit(b, iters=300, tag=i)
    done = []
    while len(done) < len(bases):
        bt.step()
        done += bt.poll()
    bt.close()
    by_tag = {r["tag"]: r for r in done}
    for i in (0, 7, 19, 33, 39):
        alone = gfe.run_prp(bases[i], 12, iters=300, b_gl=-1)
        assert (by_tag[i]["res64"], by_tag[i]["old64"]) == (alone["res64"], alone["old64"]), bases[i]


def test_batch_plan_changes_with_pending_carries(gfe):
    """n = 19 (rows of 2^11 points and more: squaring rounds leave their carries for the next round's
    column pass): plan changes between blocks — to another compile-time plan, to the all-prime
    fused plan whose passes use the carry array for segment carries, and back — apply or clear
    the pending carries, and every residue equals its single run."""
    n = 19
    bases = [1000000000, 1534, 30000000, 1999999266]
    iters = 5 * 64 + 17
    bt = gfe.Batch(n, len(bases))
    base_plan = tuple(bt.get_plan())
    assert base_plan[5] & 8 and n - base_plan[0] >= 11

    def first_ok(cands):
        for p in cands:
            try:
                bt.set_plan(p)
                return p
            except gfe.GfeError:
                continue
        raise AssertionError("no admissible plan")

    for i, b in enumerate(bases):
        bt.submit(b, iters=iters, tag=i)
    schedule = [
        [(a, cl, cng, rl, rng, 8) for a in (7, 6) for cl in (4, 3) for cng in (16, 8) for rl in (4, 5) for rng in (1, 2)],
        [(a, cl, cng, rl, rng, 2) for a in (9, 8, 10) for cl in (4, 3) for cng in (16, 8) for rl in (4,) for rng in (1, 2, 4)],
        [base_plan],
    ]
    done = []
    bt.step()
    for cands in schedule:
        first_ok(cands)
        bt.step()
        done += bt.poll()
    while len(done) < len(bases):
        bt.step()
        done += bt.poll()
    bt.close()
    for r in done:
        alone = gfe.run_prp(r["b"], n, iters=iters, b_gl=-1)
        assert (r["res64"], r["old64"]) == (alone["res64"], alone["old64"]), r["b"]
