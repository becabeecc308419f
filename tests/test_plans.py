"""Host logic of libgfe.so that needs no device: pass-plan defaults, admissibility and the
autotuner's candidate list, and the prime-set selection of create_ocl (src/transform_ocl.cpp:35-50)
against the oracle's restatement, including the bounds between prime sets."""
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def G():
    import genefer22_amd
    if not os.path.exists(genefer22_amd.LIB_PATH):
        from genefer22_amd import build
        build.build()
    genefer22_amd.lib()
    return genefer22_amd


def threads(n, plan):
    a, cl, cng, rl, rng, _ = plan
    return cng * (1 << a) // (1 << cl), rng * (1 << (n - a)) // (1 << rl)


@pytest.mark.parametrize("n", range(12, 24))
def test_default_plans_are_admissible_candidates(G, n):
    plans = G.candidate_plans(n)
    assert len(plans) == len(set(plans)) > 100
    for cap in (0, 1, 31, 32, 512):
        p = G.default_plan(n, cap)
        assert G.plan_admissible(n, p), (n, cap, p)
        assert p in plans, (n, cap, p)
        assert p[5] & 8, "the defaults use the compile-time passes"
        # the single-workgroup squaring is a batch default for n <= 13 and 32+ slots only
        assert bool(p[5] & 16) == (cap >= 32 and n <= 13), (n, cap, p)


@pytest.mark.parametrize("n", [12, 15, 19, 22, 23])
def test_candidate_plans_are_well_formed(G, n):
    for p in G.candidate_plans(n):
        a, cl, cng, rl, rng, fl = p
        assert G.plan_admissible(n, p)
        assert 0 <= fl < 32 and (fl & 6) != 6, p            # fused and persistent exclude each other
        if fl & 8:
            assert fl in (8, 24), p                          # compile-time passes: no other variant
            assert 5 <= a <= 11 and 6 <= n - a <= 14, p      # instantiated shapes
        if fl & 16:
            assert n <= 13, p
        tc, tr = threads(n, p)
        assert 64 <= tc <= 1024 and 64 <= tr <= 1024, p
        if fl & 8:
            assert (cl < 5 or tc <= 512) and (rl < 5 or tr <= 512), p
        assert cng <= 1 << (n - a) and rng <= 1 << a, p


def test_malformed_plans_are_refused(G):
    n = 22
    good = G.default_plan(n)
    assert G.plan_admissible(n, good)
    a, cl, cng, rl, rng, fl = good
    bad = [
        (a, cl, cng, rl, rng, 6),       # fused and persistent together
        (a, cl, cng, rl, rng, 9),       # compile-time passes with factored twiddles
        (a, cl, cng, rl, rng, 32),      # unknown flag bit
        (a, cl, cng, rl, rng, -1),
        (a, cl, cng, rl, rng, 24),      # single-workgroup squaring above n = 13
        (7, cl, cng, rl, rng, 8),       # 15 row levels: no instantiated row pass
        (a, cl, 4, rl, rng, 8),         # fewer than 8 columns per workgroup
        (a, cl, cng, rl, 1 << 9, 8),    # more rows per workgroup than rows
        (n, 4, 1, 4, 1, 0),             # no row levels left
        (a, 6, cng, rl, rng, 8),        # 64 points per thread
    ]
    for p in bad:
        assert not G.plan_admissible(n, p), p
    with pytest.raises(ValueError):
        G.plan_admissible(n, (1, 2, 3))
    for n_bad in (11, 24):
        with pytest.raises(G.GfeError):
            G.default_plan(n_bad)


def test_prime_set_selection_matches_the_oracle(G, oracle):
    rs = np.random.RandomState(11)
    for _ in range(400):
        n = int(rs.randint(12, 24))
        b = 2 * int(rs.randint(512, 1_000_000_001))
        assert G.select(b, n) == oracle.select(b, n), (b, n)


@pytest.mark.parametrize("n", [12, 15, 16, 22])
def test_prime_set_bounds(G, oracle, n):
    """Around each bound of the rule (b^2 < P1 P2 / 2 / N for NTT2 with the 31-bit, then the 32-bit
    primes; b <= 10^9 for NTT3 with the 31-bit primes) both restatements agree and switch sets."""
    N = 1 << n
    S = [127 * 2**24 + 1, 63 * 2**25 + 1]
    U = [125 * 2**25 + 1, 243 * 2**24 + 1]
    for P1, P2 in (S, U):
        lim = (P1 * P2 // 2) // N          # b^2 must stay below this
        r = math.isqrt(lim - 1)            # largest b with b^2 < lim
        for b in (r - 1, r, r + 1, r + 2):
            if b < 1024:
                continue
            assert G.select(b, n) == oracle.select(b, n), (b, n)
        below, above = G.select(r, n), G.select(r + 1, n)
        assert below[0] == 2 and (above != below), (n, r, below, above)
    for b in (999_999_998, 1_000_000_000, 1_000_000_002):
        sel = G.select(b, n)
        assert sel == oracle.select(b, n)
        assert sel == (3, b > 1_000_000_000), (b, n, sel)


def test_ntt3_in_31bit_primes_is_exact_for_every_b():
    """NTT3 computes in the 31-bit primes for every b <= 2e9: balanced digits (|d| <= b/2 + 4) are
    residues of each 31-bit prime, and the largest squared-and-doubled coefficient of the negacyclic
    product, N (b/2 + 4)^2 2, stays below half the three primes' product for every n the engine
    admits.  The same bound fails for the two-prime set, which keeps the reference's rule."""
    S = [127 * 2**24 + 1, 63 * 2**25 + 1, 15 * 2**27 + 1]
    b_max = 2_000_000_000
    d_max = b_max // 2 + 4
    assert d_max < min(S)
    for n in range(12, 24):
        assert (1 << n) * d_max * d_max * 2 < S[0] * S[1] * S[2] // 2, n
    assert (1 << 12) * d_max * d_max * 2 > S[0] * S[1] // 2


@pytest.mark.parametrize("is32", [0, 1])
def test_host_twiddle_tables(G, is32):
    """The engine's twiddle tables from Montgomery-product recurrences equal direct powers
    psi^bitrev(j) 2^32 mod p, psi a primitive 2N-th root (psi^N = -1); built on the calling thread
    alone (the fallback when no thread can be started) and on several threads they agree."""
    n = 10
    N = 1 << n
    primes = ([127 * 2**24 + 1, 63 * 2**25 + 1, 15 * 2**27 + 1], [125 * 2**25 + 1, 243 * 2**24 + 1, 235 * 2**24 + 1])[is32]
    gens = ([3, 5, 31], [3, 7, 3])[is32]
    serial = G.host_twiddles(n, is32, max_threads=0)
    assert np.array_equal(serial, G.host_twiddles(n, is32, max_threads=3))
    assert np.array_equal(serial, G.host_twiddles(n, is32, max_threads=1))
    for k, (p, g) in enumerate(zip(primes, gens)):
        psi = pow(g, (p - 1) // (2 * N), p)
        assert pow(psi, N, p) == p - 1
        for j in range(N):
            e = int(format(j, f"0{n}b")[::-1], 2)
            assert int(serial[k, j]) == (pow(psi, e, p) << 32) % p, (k, j)
    with pytest.raises(G.GfeError):
        G.host_twiddles(24, is32)


def test_engine_plan_takes_the_limb_transform_where_it_is_exact(G):
    """gfe_engine_plan: the two-digit-limb transform (flag 32) exactly where the limb rule admits
    (b, n) -- the rule restated in tests/test_limb.py -- with its inverse column pass fused to the
    normalisation (flag 2) exactly where the fused rule admits it too, and the digit transform's
    default elsewhere; C4's bases always get the fused limb plan."""
    from test_limb import limb_fused_ok, limb_ok
    from genefer22_amd import workloads as W
    for b in W.config_bases(W.C4, 8):
        p = G.engine_plan(b, 22)
        assert p[5] == 42 and G.plan_admissible(22, p), (b, p)
        assert p[:5] + (8,) == G.default_plan(21), p  # the 2^21-point transform's compile-time plan
    rs = np.random.RandomState(3)
    for n in range(12, 24):
        for b in [1026, 1 << 23 | 2, 8388610, 30000000, 60000000, 200000000, 249999998, 1999999266] + \
                 [int(x) * 2 for x in rs.randint(512, 1_000_000_000, 12)]:
            p = G.engine_plan(b, n)
            want = limb_ok(b, n)
            assert bool(p[5] & 32) == want, (b, n, p)
            if want:
                assert bool(p[5] & 2) == limb_fused_ok(b, n, p[2]), (b, n, p)
            else:
                assert p == G.default_plan(n), (b, n, p)


def test_limb_plans_need_the_compile_time_passes(G):
    base = G.default_plan(21)
    assert G.plan_admissible(22, base[:5] + (40,))
    for fl in (32, 33, 34, 36, 56, 41, 43, 44, 46):  # runtime passes, fused / persistent columns,
        assert not G.plan_admissible(22, base[:5] + (fl,)), fl  # one-workgroup squaring, factored rows
    assert G.plan_admissible(22, base[:5] + (42,))  # compile-time passes, inverse columns fused with the norm
    # the fused kernel runs the four primes' column groups in one workgroup: 4 T <= 1024 threads
    assert not G.plan_admissible(22, (7, 3, 32, 5, 1, 42)) and G.plan_admissible(22, (7, 3, 32, 5, 1, 40))
