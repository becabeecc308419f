"""GPU parity of the HIP squaring engine against the oracle (C restatement of the reference's
exact path) and the reference's known-answer RES64 values.  Bit-exact: canonical residues equal."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (b, n, expected prime set) -- one case per prime set of create_ocl (src/transform_ocl.cpp:35-50)
CASES = [
    (1534, 12, (2, False)),        # NTT2, 31-bit primes
    (30000000, 12, (2, True)),     # NTT2, 32-bit primes (2.34e7 <= b < 4.57e7)
    (999999618, 12, (3, False)),   # NTT3, 31-bit primes
    (1999999266, 12, (3, True)),   # NTT3, 32-bit primes
]


def prp_bits(b, n, iters):
    E = b ** (1 << n)
    es = E.bit_length()
    return [((E >> i) & 1) for i in range(es - 1, es - 1 - iters, -1)]


def canon(O, d, n, b):
    return O.unbalance(d, n, b)


@pytest.mark.parametrize("b,n,sel", CASES)
def test_prime_set_selection(gfe, oracle, b, n, sel):
    assert gfe.select(b, n) == sel
    assert oracle.select(b, n) == sel


@pytest.mark.parametrize("b,n,sel", CASES)
def test_transform_set(gfe, b, n, sel):
    """NTT3 computes in the 31-bit primes for every b (balanced digits keep it exact); the info
    string names the reference's set beside it when they differ.  NTT2 keeps the reference's set."""
    rns, is32 = sel
    info = gfe.Transform(b, n).info()
    want = f"NTT{rns}{'u' if is32 and rns == 2 else 's'}"
    assert info.startswith(want), info
    assert ("reference set NTT3u" in info) == (rns == 3 and is32), info


@pytest.mark.parametrize("b,n,sel", CASES)
def test_square_dup_matches_oracle(gfe, oracle, b, n, sel):
    """r0 after k PRP iterations (squareDup with the exponent bits) equals the oracle's."""
    iters = 300
    bits = prp_bits(b, n, iters)
    t = gfe.Transform(b, n)
    o = oracle.Engine(b, n)
    t.set(1); o.set(1)
    for k, bit in enumerate(bits):
        t.square_dup(bit)
        o.square_dup(bit)
        if k in (0, 1, 2, 9, 99, iters - 1):
            got = t.get_digits(0)
            assert np.abs(got).max() <= b // 2 + 4, "digits left the balanced bound"
            assert np.array_equal(canon(oracle, got, n, b), canon(oracle, o.get_digits(0), n, b)), f"iteration {k}"


@pytest.mark.parametrize("b,n", [(1000000000, 13), (40000, 14), (123456790, 16), (2000000000, 16), (8000000, 18), (30000000, 20)])
def test_larger_n_matches_oracle(gfe, oracle, b, n):
    iters = {13: 40, 14: 30, 16: 12, 18: 5, 20: 3}[n]
    bits = prp_bits(b, n, iters)
    t = gfe.Transform(b, n)
    o = oracle.Engine(b, n)
    t.set(1); o.set(1)
    for bit in bits:
        t.square_dup(bit); o.square_dup(bit)
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))


@pytest.mark.parametrize("b,n", [(1534, 12), (999999618, 12), (1999999266, 12)])
def test_mul_copy_multiplicand(gfe, oracle, b, n):
    """initMultiplicand / mul / copy / squareMul on random balanced residues."""
    rs = np.random.RandomState(b % 1000)
    N = 1 << n
    a = oracle.balance(rs.randint(0, b, N).astype(np.int32), n, b)
    c = oracle.balance(rs.randint(0, b, N).astype(np.int32), n, b)
    t = gfe.Transform(b, n, num_regs=3)
    o = oracle.Engine(b, n, num_regs=3)
    t.set_digits(a, 0); o.set_digits(a, 0)
    t.set_digits(c, 1); o.set_digits(c, 1)
    t.mul(1); o.mul(1)
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))
    t.copy(2, 0); o.copy(2, 0)
    t.square_mul(7); o.square_mul(7)
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))
    assert np.array_equal(canon(oracle, t.get_digits(2), n, b), canon(oracle, o.get_digits(2), n, b))


def test_known_answer_quick_n12(gfe):
    """Full quick test (PRP + Gerbicz-Li) against the reference's KATs (src/main.cpp:424-425)."""
    for b, want in [(1000000000, 0x4E43A5B93273C649), (2000000000, 0xE5F3EA34F6C68EBD)]:
        r = gfe.run_quick(b, 12)
        assert r["res64"] == want, hex(r["res64"])
        assert r["gl_ok"] == 1


def _plans(n):
    for a in sorted({4, n // 2, n - 4}):
        for cl in (3, 4, 5):
            for rl in (3, 4, 5):
                for cng in (1, 8, 32):
                    for rng in (1, 4):
                        # bit 0 factored row twiddles, bit 1 fused columns, bit 2 pipelined columns,
                        # bit 3 compile-time column and row passes, bit 4 squaring in one workgroup
                        for flags in (0, 1, 2, 3, 4, 5, 8, 24):
                            yield (a, cl, cng, rl, rng, flags)


@pytest.mark.parametrize("b,n", [(1534, 12), (1999999266, 12), (999999618, 14), (30000000, 17)])
def test_every_plan_variant(gfe, oracle, b, n):
    """Every admissible pass plan (column/row split, points per thread, sub-transforms per
    workgroup, table or factored row twiddles, separate or fused all-prime column passes with the
    CRT and carry, pipelined persistent column passes) gives the oracle's residue, digits within
    the balanced bound."""
    iters = 12
    bits = prp_bits(b, n, iters)
    o = oracle.Engine(b, n)
    o.set(1)
    for bit in bits:
        o.square_dup(bit)
    want = canon(oracle, o.get_digits(0), n, b)
    t = gfe.Transform(b, n)
    tried = 0
    for plan in _plans(n):
        try:
            t.set_plan(plan)
        except gfe.GfeError:
            continue
        assert t.get_plan() == plan
        t.set(1)
        for bit in bits:
            t.square_dup(bit)
        got = t.get_digits(0)
        assert np.abs(got).max() <= b // 2 + 4, plan
        assert np.array_equal(canon(oracle, got, n, b), want), plan
        tried += 1
    assert tried >= 20
    with pytest.raises(gfe.GfeError):
        t.set_plan((n, 4, 1, 4, 1, 0))  # no row levels left


@pytest.mark.parametrize("flags", [1, 2, 3, 4, 5, 8, 24])
def test_multiplicand_plan_variants(gfe, oracle, flags):
    """initMultiplicand + mul (row pass forward-only and product modes) under factored twiddles,
    fused column passes and compile-time passes."""
    # the single-workgroup squaring (bit 4) is for n <= 13
    b, n, shape = (1000000000, 13, (6, 3, 16, 4, 8)) if flags & 16 else (1000000000, 14, (7, 4, 8, 4, 8))
    t = gfe.Transform(b, n)
    t.set_plan(shape + (flags,))
    o = oracle.Engine(b, n)
    t.set(3); o.set(3)
    for _ in range(5):
        t.square_dup(False); o.square_dup(False)
    t.copy(1, 0); o.copy(1, 0)
    t.square_dup(True); o.square_dup(True)
    t.mul(1); o.mul(1)
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))


@pytest.mark.parametrize("b,n", [(20204692, 16), (1999999266, 15)])
def test_pending_carries_reach_every_use(gfe, oracle, b, n):
    """With compile-time passes a squaring leaves its workgroup carries for the next squaring's
    column pass; every other use of the registers applies them first: copy, read-out, context
    save, initMultiplicand of register 0, a plan change, and the timed squarings of
    time_squarings (per-kernel leg included).  Each state is checked against the oracle."""
    t = gfe.Transform(b, n)
    plan = t.get_plan()
    assert plan[5] & 8
    o = oracle.Engine(b, n)
    want = lambda reg=0: canon(oracle, o.get_digits(reg), n, b)
    got = lambda reg=0: canon(oracle, t.get_digits(reg), n, b)
    t.set(3); o.set(3)
    for k in range(4):
        t.square_dup(k & 1); o.square_dup(k & 1)
    t.copy(1, 0); o.copy(1, 0)                       # copy while pending
    t.square_dup(False); o.square_dup(False)
    data = t.save_context(2)                          # context while pending
    u = gfe.Transform(b, n)
    assert u.read_context(data, 2)
    assert np.array_equal(canon(oracle, u.get_digits(0), n, b), want(0))
    u.close()
    t.square_dup(True); o.square_dup(True)
    t.init_multiplicand(0); o.init_multiplicand(0)    # forward transform of r0 while pending
    t.square_dup(False); o.square_dup(False)
    t.mul(); o.mul()
    assert np.array_equal(got(0), want(0)) and np.array_equal(got(1), want(1))
    t.square_dup(False); o.square_dup(False)
    alt = next(p for p in _plans(n) if p[5] == 0 and p != plan and _admissible(gfe, b, n, p))
    t.set_plan(alt)                                   # plan change while pending
    t.square_dup(True); o.square_dup(True)
    t.set_plan(plan)
    t.square_dup(False); o.square_dup(False)
    pattern = 0b1011
    t.time_squarings(6, pattern)                      # timed squarings, then the per-kernel leg
    t.time_squarings(3, pattern, per_kernel=True)    # 3 timed squarings, then each pass timed alone
    for i in list(range(6)) + list(range(3)):
        o.square_dup((pattern >> (i % 64)) & 1)
    o.square_dup(False)                               # the per-pass leg squares once (dup = 0)
    assert np.array_equal(got(0), want(0))
    assert np.abs(t.get_digits(0)).max() <= b // 2 + 4
    t.close()


def _admissible(gfe, b, n, plan):
    t = gfe.Transform(b, n)
    try:
        t.set_plan(plan)
        return True
    except gfe.GfeError:
        return False
    finally:
        t.close()


def test_ntt3_bound_at_its_extreme(gfe, oracle):
    """NTT3 computes in the 31-bit primes for every b: at n = 23 and b = 1,999,999,998 (the
    reference's rule names the 32-bit set) every digit at the balanced maximum b/2 makes the
    largest coefficient, N (b/2)^2 = 2^82.8, doubled by dup: still below half the 31-bit
    primes' product (2^91.9).  Two squarings equal the oracle's, which computes in the
    reference's set."""
    n, b = 23, 1999999998
    assert gfe.select(b, n) == (3, True)
    d = np.full(1 << n, b // 2, dtype=np.int32)
    t = gfe.Transform(b, n)
    o = oracle.Engine(b, n)
    assert t.info().startswith("NTT3s")
    t.set_digits(d, 0); o.set_digits(d, 0)
    for dup in (True, False):
        t.square_dup(dup); o.square_dup(dup)
        assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b)), dup
    t.close()


@pytest.mark.parametrize("b,n", [(2000000, 21), (150000, 23), (30000000, 23)])
def test_largest_n_from_full_size_digits(gfe, oracle, b, n):
    """The largest transforms the reference's command line admits (n = 21, 23; NTT2 with 32-bit and
    31-bit primes, NTT3) from a seeded full-size balanced residue (every digit nonzero, carries
    across every workgroup and the negacyclic wrap), two squarings, one doubled: equal to the
    oracle."""
    rs = np.random.RandomState(n * 1000 + b % 997)
    d = oracle.balance(rs.randint(0, b, 1 << n).astype(np.int32), n, b)
    t = gfe.Transform(b, n)
    o = oracle.Engine(b, n)
    assert (t.get_plan()[5] & 8) and o.rns == gfe.select(b, n)[0]
    t.set_digits(d, 0); o.set_digits(d, 0)
    for dup in (True, False):
        t.square_dup(dup); o.square_dup(dup)
    got = t.get_digits(0)
    assert np.abs(got.astype(np.int64)).max() <= b // 2 + 4
    assert np.array_equal(canon(oracle, got, n, b), canon(oracle, o.get_digits(0), n, b))
    t.close()


def test_set_rejects_values_beyond_the_balanced_bound(gfe, oracle):
    """Registers hold balanced digits: set(a) accepts a <= b/2 + 4 and refuses larger a (a value
    that would wrap int32 or exceed a 31-bit prime's residues would otherwise come back wrong)."""
    b, n = 1026, 12
    t = gfe.Transform(b, n)
    t.set(b // 2 + 4)
    d = t.get_digits(0)
    assert d[0] == b // 2 + 4 and not d[1:].any()
    for a in (b // 2 + 5, 2**31, 2**32 - 1):
        with pytest.raises(gfe.GfeError):
            t.set(a)
    t.close()


def test_rebase_equals_a_fresh_engine(gfe, oracle):
    """gfe_rebase moves an engine to another b of its n and prime set: the squarings then equal a
    fresh engine's for that b (the limb transform's and the digit transform's bases alike), the
    registers keep the last value's digits (carries still pending from its last squaring included),
    and a b of another prime set is refused."""
    n = 16
    t = gfe.Transform(20204692, n)
    assert t.get_plan()[5] & 32
    bits = [1, 0, 1, 1, 0, 0, 1, 0, 1, 1] * 3
    last = None  # (b, the oracle's digits) of the value the engine holds
    for b in (33554434, 1500000000, 20204692):  # limb transform; NTT3 beyond the limb bound; back
        t.rebase(b)
        if last is not None:
            pb, want = last
            assert np.array_equal(canon(oracle, t.get_digits(0), n, pb), want), (pb, b)
        fresh = gfe.Transform(b, n)
        assert t.get_plan() == fresh.get_plan() and t.info() == fresh.info(), b
        o = oracle.Engine(b, n)
        t.set(3); o.set(3)
        for bit in bits:
            t.square_dup(bit); o.square_dup(bit)
        assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b)), b
        for bit in bits[:3]:  # leaves the last squaring's carries pending for the rebase
            t.square_dup(bit); o.square_dup(bit)
        last = (b, canon(oracle, o.get_digits(0), n, b))
        fresh.close()
    with pytest.raises(gfe.GfeError):
        t.rebase(4098)  # NTT2 at n = 16: another prime set
    t.close()
