"""GPU parity of the two-digit-limb transform (limbs d_2j + b d_2j+1 over four 31-bit primes,
genefer22_amd/csrc/gfe_limb.h) against the oracle's squaring, which knows nothing of limbs:
canonical residues equal after PRP squarings, products, squareMul, every admissible limb plan (the
inverse column pass fused with the normalisation or not), and switching between the limb and the
digit transform mid-run.  Digits stay within the balanced
bound.  The C4 fixtures (n = 22, reference CPU transform) run through the limb plan by default
(tests/test_gpu_golden.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIMB = 32  # plan flag of the limb transform


def prp_bits(b, n, iters):
    E = b ** (1 << n)
    es = E.bit_length()
    return [((E >> i) & 1) for i in range(es - 1, es - 1 - iters, -1)]


def canon(O, d, n, b):
    return O.unbalance(d, n, b)


# (b, n): NTT3 candidates inside the limb bound (b >= 2^23, N l_max^2 < P/4), the largest b of n = 12
# included; the lone-request sizes n = 18 (rows of 16 points per thread, two per workgroup) and
# n = 20 at the top of the serving mix's range (2^8-point columns); C4's range at n = 22
CASES = [(50000000, 12), (249999998, 12), (100000000, 13), (60000000, 15), (30000000, 18), (54999998, 20),
         (20204692, 22)]


@pytest.mark.parametrize("b,n", CASES)
def test_engine_starts_with_the_limb_plan(gfe, b, n):
    t = gfe.Transform(b, n)
    assert t.get_plan()[5] & LIMB, t.info()
    assert "limbs of two digits" in t.info()
    assert t.info().startswith("NTT3s")  # the reference's set, as reported and in context payloads


@pytest.mark.parametrize("b,n", CASES)
def test_limb_squarings_match_oracle(gfe, oracle, b, n):
    iters = {12: 200, 13: 120, 15: 40, 18: 16, 20: 6, 22: 3}[n]
    bits = prp_bits(b, n, iters)
    t = gfe.Transform(b, n)
    o = oracle.Engine(b, n)
    t.set(1); o.set(1)
    for k, bit in enumerate(bits):
        t.square_dup(bit)
        o.square_dup(bit)
        if k in (0, 1, 5, iters // 2, iters - 1):
            got = t.get_digits(0)
            assert np.abs(got).max() <= b // 2 + 4, f"digits left the balanced bound at {k}"
            assert np.array_equal(canon(oracle, got, n, b), canon(oracle, o.get_digits(0), n, b)), f"iteration {k}"


@pytest.mark.parametrize("b,n", [(50000000, 12), (249999998, 12), (100000000, 13)])
def test_limb_extreme_digits(gfe, oracle, b, n):
    """Every digit at +-(b/2 + 4), the registers' bound: the largest limbs and coefficients."""
    N = 1 << n
    h = b // 2 + 4
    for pattern in (np.full(N, h), np.full(N, -h), np.where(np.arange(N) % 2 == 0, h, -h),
                    np.where(np.arange(N) % 3 == 0, -h, h)):
        d = pattern.astype(np.int32)
        t = gfe.Transform(b, n)
        o = oracle.Engine(b, n)
        t.set_digits(d, 0); o.set_digits(d, 0)
        for dup in (1, 0, 1):
            t.square_dup(dup); o.square_dup(dup)
        got = t.get_digits(0)
        assert np.abs(got).max() <= b // 2 + 4
        assert np.array_equal(canon(oracle, got, n, b), canon(oracle, o.get_digits(0), n, b))
        t.close()


@pytest.mark.parametrize("b,n", [(50000000, 12), (100000000, 13)])
def test_limb_mul_copy_multiplicand(gfe, oracle, b, n):
    """initMultiplicand / mul / copy / squareMul through the limb transform."""
    rs = np.random.RandomState(n)
    N = 1 << n
    a = oracle.balance(rs.randint(0, b, N).astype(np.int32), n, b)
    c = oracle.balance(rs.randint(0, b, N).astype(np.int32), n, b)
    t = gfe.Transform(b, n, num_regs=3)
    o = oracle.Engine(b, n, num_regs=3)
    assert t.get_plan()[5] & LIMB
    t.set_digits(a, 0); o.set_digits(a, 0)
    t.set_digits(c, 1); o.set_digits(c, 1)
    t.mul(1); o.mul(1)
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))
    t.init_multiplicand(1); o.init_multiplicand(1)
    t.mul(); o.mul()
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))
    t.copy(2, 0); o.copy(2, 0)
    t.square_mul(7); o.square_mul(7)
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))
    assert np.array_equal(canon(oracle, t.get_digits(2), n, b), canon(oracle, o.get_digits(2), n, b))


@pytest.mark.parametrize("b,n", [(100000000, 14), (60000000, 15)])
def test_every_limb_plan(gfe, oracle, b, n):
    """Every compile-time plan of the N/2-point transform with the limb flag gives the oracle's
    residue (and so does the digit transform's default plan on the same engine)."""
    iters = 10
    bits = prp_bits(b, n, iters)
    o = oracle.Engine(b, n)
    o.set(1)
    for bit in bits:
        o.square_dup(bit)
    want = canon(oracle, o.get_digits(0), n, b)
    t = gfe.Transform(b, n)
    plans = [p[:5] + (p[5] | LIMB,) for p in gfe.candidate_plans(n - 1) if p[5] == 8]
    plans.append(gfe.default_plan(n))
    tried = 0
    for plan in plans:
        t.set_plan(plan)
        assert t.get_plan() == tuple(plan)
        t.set(1)
        for bit in bits:
            t.square_dup(bit)
        got = t.get_digits(0)
        assert np.abs(got).max() <= b // 2 + 4, plan
        assert np.array_equal(canon(oracle, got, n, b), want), plan
        tried += 1
    assert tried >= 10


@pytest.mark.parametrize("b,n", [(100000000, 14), (60000000, 15)])
def test_every_fused_limb_plan(gfe, oracle, b, n):
    """Every admissible plan with the inverse column pass fused to the limb normalisation (flags
    2 | 8 | 32: ranges of col_ng limbs, their carries added by the next forward column pass) gives
    the oracle's residue, and a squaring's digits stay balanced once the carries are in."""
    iters = 10
    bits = prp_bits(b, n, iters)
    o = oracle.Engine(b, n)
    o.set(1)
    for bit in bits:
        o.square_dup(bit)
    want = canon(oracle, o.get_digits(0), n, b)
    t = gfe.Transform(b, n)
    assert t.get_plan()[5] == 42, t.info()
    plans = [p[:5] + (42,) for p in gfe.candidate_plans(n - 1) if p[5] == 8]
    tried = 0
    for plan in plans:
        if not gfe.plan_admissible(n, plan):
            continue
        t.set_plan(plan)
        t.set(1)
        for bit in bits:
            t.square_dup(bit)
        got = t.get_digits(0)
        assert np.abs(got).max() <= b // 2 + 4, plan
        assert np.array_equal(canon(oracle, got, n, b), want), plan
        tried += 1
    assert tried >= 8


def test_switching_between_limb_and_digit_transforms(gfe, oracle):
    """The fused limb plan, the plain limb plan and the digit transform in turn: each hands the
    next its pending carries in their own form (range slots, boundary pairs, workgroup carries)."""
    b, n = 100000000, 13
    bits = prp_bits(b, n, 60)
    t = gfe.Transform(b, n)
    fused_plan, digit_plan = t.get_plan(), gfe.default_plan(n)
    assert fused_plan[5] == 42 and not digit_plan[5] & LIMB
    limb_plan = fused_plan[:5] + (40,)
    o = oracle.Engine(b, n)
    t.set(1); o.set(1)
    for k, bit in enumerate(bits):
        if k % 7 == 0:
            t.set_plan((fused_plan, digit_plan, limb_plan)[(k // 7) % 3])
        t.square_dup(bit); o.square_dup(bit)
        if k % 13 == 5:
            t.copy(1, 0)
            assert np.array_equal(canon(oracle, t.get_digits(1), n, b), canon(oracle, o.get_digits(0), n, b))
    assert np.array_equal(canon(oracle, t.get_digits(0), n, b), canon(oracle, o.get_digits(0), n, b))


def test_limb_plan_refused_outside_its_bound(gfe):
    """b = 1999999266 at n = 12 exceeds the limb bound; the batch never runs limb plans."""
    limb12 = gfe.Transform(50000000, 12).get_plan()
    limb13 = gfe.Transform(100000000, 13).get_plan()
    assert limb12[5] == 42 and limb13[5] == 42
    t = gfe.Transform(1999999266, 12)
    assert not t.get_plan()[5] & LIMB
    with pytest.raises(gfe.GfeError):
        t.set_plan(limb12)
    # b = 249999998 at n = 12: limbs, but the fused normalisation's larger range-start limbs would
    # leave its bound
    u = gfe.Transform(249999998, 12)
    assert u.get_plan()[5] == 40
    with pytest.raises(gfe.GfeError):
        u.set_plan(u.get_plan()[:5] + (42,))
    u.close()
    bt = gfe.Batch(13, 4)
    with pytest.raises(gfe.GfeError):
        bt.set_plan(limb13)
    bt.close()
