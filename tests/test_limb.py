"""The two-digit-limb transform's CRT and carry (genefer22_amd/csrc/gfe_limb.h), emulated on the CPU
step by step -- 32-bit Montgomery products, 64-bit wrap-around sums, IEEE doubles with correctly
rounded fma -- and checked against Python integers: the pieces h0 + h1 B + h2 B^2 equal the
coefficient at the bound's edges, and a whole normalisation (threads of 4 limbs, workgroups of 1024
limbs, the boundary carries with the negacyclic wrap) turns the exact negacyclic square of the limbs
into balanced digits whose value is the square mod b^N + 1.  The constants follow make_limb_crt
(gfe_engine.hip); the GPU tests check the device code against the oracle on the same inputs."""
import random
from fractions import Fraction

import pytest

PRIMES = [2130706433, 2113929217, 2013265921, 27 * 2 ** 26 + 1]
P = PRIMES[0] * PRIMES[1] * PRIMES[2] * PRIMES[3]
M64 = (1 << 64) - 1


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def rint(x):
    return float(round(x))  # round half to even, as rint in the default rounding mode


def i64(x):
    x &= M64
    return x - (1 << 64) if x >> 63 else x


def bal_piece(x, B):
    r = x % B
    if r > B // 2:
        r -= B
    return r, (x - r) // B


def limb_ok(b, n):
    """Host rule of gfe_engine.hip (limb_ok), the NTT3 condition of the reference's selection included."""
    N = 1 << n
    if b * b < (PRIMES[0] * PRIMES[1] // 2) // N or b < (1 << 23):
        return False
    if b * b < (4194304001 * 4076863489 // 2) // N:  # NTT2 on the 32-bit primes
        return False
    lmax = (b // 2 + 4) * (b + 1)
    return lmax * lmax < (P // 4) >> n


def make_consts(b, n):
    B = b * b
    c = {"B": B, "b": b, "w": [], "zfac": []}
    for p in PRIMES:
        Q = P // p
        w0, x = bal_piece(Q, B)
        w1, w2 = bal_piece(x, B)
        c["w"].append((w0, w1, w2))
    P0, x = bal_piece(P, B)
    P1, P2 = bal_piece(x, B)
    c["P"] = (P0, P1, P2)
    return c


def split_B(s64, sd, c):
    """s = qi B + h with qi = rint(sd / B), no further balancing (|h| <= B/2 + 2^38)."""
    B = c["B"]
    qi = int(rint(sd * (1.0 / B)))
    h = i64(s64 - qi * B)
    return h, qi


def pieces(z, c, dup):
    """limb_pieces after the Montgomery products: z_k = x (P/p_k)^-1 mod p_k."""
    ts = 0.0
    for zk, p in zip(z, PRIMES):
        ts = fma(float(zk), 1.0 / p, ts)
    t = int(rint(ts))
    P0, P1, P2 = c["P"]
    s0 = ((-t) * P0) & M64
    s0d = -float(t) * float(P0)
    for zk, w in zip(z, c["w"]):
        s0 = (s0 + zk * (w[0] & M64)) & M64
        s0d = fma(float(zk), float(w[0]), s0d)
    h0, q0 = split_B(s0, s0d, c)
    s1 = (q0 - t * P1) & M64
    s1d = fma(-float(t), float(P1), float(q0))
    for zk, w in zip(z, c["w"]):
        s1 = (s1 + zk * (w[1] & M64)) & M64
        s1d = fma(float(zk), float(w[1]), s1d)
    h1, q1 = split_B(s1, s1d, c)
    h2 = q1 - t * P2 + sum(zk * w[2] for zk, w in zip(z, c["w"]))
    if dup:
        h0, h1, h2 = 2 * h0, 2 * h1, 2 * h2
    return h0, h1, h2


def div_bal(f, b):
    """Balanced division (div_bal): f = q b + r, r in (-b/2, b/2]."""
    r = f % b
    if r > b // 2:
        r -= b
    return r, (f - r) // b


def limb_digits(v, b):
    """limb_digits as the device computes it -- rounded double quotients, one correction step each
    in 32-bit arithmetic -- checked against the exact balanced divisions: v = d0 + b d1 + B c."""
    assert abs(v) < 2 ** 62
    binv = 1.0 / b
    hi, lo = b // 2, -((b - 1) // 2)
    q = rint(float(v) * binv)
    assert abs(q) < 2 ** 51
    ql = int(q) & 0xFFFFFFFF
    r = ((v & 0xFFFFFFFF) - ql * b) & 0xFFFFFFFF
    r = r - (1 << 32) if r >> 31 else r
    a1 = (r > hi) - (r < lo)
    d0 = r - a1 * b
    ql = (ql + a1) & 0xFFFFFFFF
    q = q + float(a1)
    c2 = int(rint(q * binv))
    r2 = (ql - c2 * b) & 0xFFFFFFFF
    r2 = r2 - (1 << 32) if r2 >> 31 else r2
    a2 = (r2 > hi) - (r2 < lo)
    d1 = r2 - a2 * b
    c = c2 + a2
    e0, e1 = div_bal(v, b)
    e1, ec = div_bal(e1, b)
    assert (d0, d1, c) == (e0, e1, ec), (v, b)
    return d0, d1, c


def absorb(d, cin, b, e1=None):
    c = cin
    for j in range(len(d) - 1):
        if e1 is not None and j == 2:
            c += e1
        if c:
            d[j], c = div_bal(d[j] + c, b)
    d[-1] += c


def normalise(coeffs, b, n, dup):
    """norm_limb_body over every workgroup, then carry_limb_body at every range boundary."""
    c = make_consts(b, n)
    M = 1 << (n - 1)
    z = [[x % p * pow(P // p, -1, p) % p for p in PRIMES] for x in coeffs]
    hs = [pieces(zj, c, False) for zj in z]
    if dup:
        hs = [(2 * a, 2 * b_, 2 * e) for a, b_, e in hs]
    digits = [0] * (2 * M)
    nwg = M // 1024
    wl = [(0, 0)] * nwg
    for g in range(nwg):
        carries = []
        outs = []
        for t in range(256):
            j0 = (g * 256 + t) * 4
            h = hs[j0:j0 + 4]
            if t > 0:
                prev = hs[j0 - 4:j0]
                pb1, pb2, pl2 = prev[3][1], prev[2][2], prev[3][2]
            else:
                pb1 = pb2 = pl2 = 0
            d = []
            cr = 0
            for j in range(4):
                lj = h[j][0] + (pb1 + pb2 if j == 0 else h[j - 1][1] + (pl2 if j == 1 else h[j - 2][2]))
                d0, d1, cr = limb_digits(lj + cr, b)
                d += [d0, d1]
            carries.append(cr)
            outs.append(d)
        for t in range(256):
            if t > 0:
                absorb(outs[t], carries[t - 1], b)
            j0 = (g * 256 + t) * 4
            digits[2 * j0:2 * j0 + 8] = outs[t]
        last = hs[(g + 1) * 1024 - 4:(g + 1) * 1024]
        wl[g] = (last[3][1] + last[2][2] + carries[255], last[3][2])
    for w in range(nwg):
        e0, e1 = wl[w - 1] if w > 0 else (-wl[nwg - 1][0], -wl[nwg - 1][1])
        seg = digits[2048 * w:2048 * w + 8]
        absorb(seg, e0, b, e1)
        digits[2048 * w:2048 * w + 8] = seg
    return digits


def value(d, b):
    return sum(di * b ** i for i, di in enumerate(d))


def negacyclic_square(l):
    M = len(l)
    out = [0] * M
    for i in range(M):
        li = l[i]
        if li == 0:
            continue
        for j in range(M):
            k = i + j
            if k < M:
                out[k] += li * l[j]
            else:
                out[k - M] -= li * l[j]
    return out


@pytest.mark.parametrize("b,n", [(35000000, 22), (10000000, 22), (1 << 23, 22), (43000000, 22),
                                 (60000000, 20), (20000000, 23), (250000000, 12)])
def test_pieces_at_the_bound(b, n):
    """h0 + h1 B + h2 B^2 = x with |h0|, |h1| <= B/2 at the largest coefficients the bound allows,
    their negatives, zero and random values."""
    if not limb_ok(b, n):
        pytest.skip("outside the limb transform's range")
    c = make_consts(b, n)
    B = c["B"]
    lmax = (b // 2 + 4) * (b + 1)
    xmax = (1 << (n - 1)) * lmax * lmax  # one (undoubled) coefficient of M products
    rng = random.Random(b ^ n)
    xs = [xmax, -xmax, 0, 1, -1, B // 2, -(B // 2), B * B // 2 + 1] + [rng.randint(-xmax, xmax) for _ in range(400)]
    for x in xs:
        z = [x % p * pow(P // p, -1, p) % p for p in PRIMES]
        h0, h1, h2 = pieces(z, c, False)
        assert h0 + h1 * B + h2 * B * B == x
        assert abs(h0) <= B // 2 + 2 ** 38 and abs(h1) <= B // 2 + 2 ** 38  # split_B does not rebalance
        assert abs(h2) <= (1 << n) // 4 + 2


def test_limb_rule_admits_c4_and_refuses_outside():
    assert limb_ok(35000000, 22) and limb_ok(10000000, 22)
    assert not limb_ok(50000000, 22)       # N l_max^2 >= P/4
    assert not limb_ok(4000000, 22)        # b < 2^23
    assert not limb_ok(1999999990, 15)     # far outside the bound


@pytest.mark.parametrize("b,n,dup,seed", [(50000000, 12, False, 1), (249999998, 12, True, 2),
                                          (100000000, 13, True, 3)])
def test_normalisation_gives_the_square(b, n, dup, seed):
    """Balanced digits in, the exact negacyclic square of their limbs through the emulated
    normalisation, digits out: the value is (2) r^2 mod b^N + 1 and every digit stays balanced."""
    assert limb_ok(b, n)
    N = 1 << n
    rng = random.Random(seed)
    half = b // 2
    d = [rng.randint(-half - 4, half + 4) for _ in range(N)]
    d[0], d[1], d[-1] = half + 4, -(half + 4), half + 4  # the extremes the registers allow
    limbs = [d[2 * j] + b * d[2 * j + 1] for j in range(N // 2)]
    sq = negacyclic_square(limbs)
    out = normalise(sq, b, n, dup)
    mod = b ** N + 1
    r = value(d, b) % mod
    want = (r * r * (2 if dup else 1)) % mod
    assert value(out, b) % mod == want
    assert max(abs(x) for x in out) <= half + 4


# --- the fused inverse column pass and normalisation (norm_limb_rows, carry_limb_slot_body, and the
# range carries col_fwd_pass adds to the next transform's limbs)

def limb_fused_ok(b, n, ng):
    """Host rule of gfe_engine.hip (limb_fused_ok): range-start limbs counted at their larger size."""
    if not limb_ok(b, n) or ng < 8:
        return False
    B = b * b
    h2max = P // (4 * B * B) + 1
    L = (b // 2 + 4) * (b + 1)
    L0 = (b + 4) * (b + 1)
    L1 = L + h2max + 2
    ranges = (1 << (n - 1)) // ng
    return 2 * ranges * (L0 * L0 + L1 * L1 + (ng - 2) * L * L) < P // 4


def normalise_fused(coeffs, b, n, dup, ng, S):
    """norm_limb_rows over every range of ng limbs (threads of S limbs): the digits without the
    range carries, and slots[w] = the carry into range w as (u0, u1, v)."""
    c = make_consts(b, n)
    M = 1 << (n - 1)
    z = [[x % p * pow(P // p, -1, p) % p for p in PRIMES] for x in coeffs]
    hs = [pieces(zj, c, dup) for zj in z]
    digits = [0] * (2 * M)
    nrange = M // ng
    slots = [None] * nrange
    tpr = ng // S
    for w in range(nrange):
        outs, carries = [], []
        for jt in range(tpr):
            j0 = w * ng + jt * S
            h = hs[j0:j0 + S]
            if jt > 0:
                prev = hs[j0 - S:j0]
                pb1, pb2, pl2 = prev[S - 1][1], prev[S - 2][2], prev[S - 1][2]
            else:
                pb1 = pb2 = pl2 = 0
            d, cr = [], 0
            for j in range(S):
                lj = h[j][0] + (pb1 + pb2 if j == 0 else h[j - 1][1] + (pl2 if j == 1 else h[j - 2][2]))
                d0, d1, cr = limb_digits(lj + cr, b)
                d += [d0, d1]
            # the carry a thread hands to the next within its range: the device adds it with one
            # correction per digit (absorb_small_carry), exact while |cr| <= 8
            assert abs(cr) <= 2, (w, jt, cr)
            outs.append(d)
            carries.append(cr)
        for jt in range(tpr):
            if jt > 0:
                absorb(outs[jt], carries[jt - 1], b)
            j0 = w * ng + jt * S
            digits[2 * j0:2 * j0 + 2 * S] = outs[jt]
        last = hs[(w + 1) * ng - S:(w + 1) * ng]
        e0 = last[S - 1][1] + last[S - 2][2] + carries[tpr - 1]
        e1 = last[S - 1][2]
        u0, u1, u2 = limb_digits(e0, b)
        nxt, sg = (w + 1, 1) if w + 1 < nrange else (0, -1)
        slots[nxt] = (sg * u0, sg * u1, sg * (u2 + e1))
        assert abs(u2) <= 2 and abs(u2 + e1) < 2 ** 31
    return digits, slots


def flush_slots(digits, slots, b, ng):
    """carry_limb_slot_body over every range."""
    out = list(digits)
    for w, (u0, u1, v) in enumerate(slots):
        seg = out[2 * ng * w:2 * ng * w + 8]
        absorb(seg, u0 + u1 * b, b, v)
        out[2 * ng * w:2 * ng * w + 8] = seg
    return out


def fused_limbs(digits, slots, b, ng):
    """The limbs the next forward column pass forms: digit pairs, range carries added unnormalised."""
    limbs = [digits[2 * j] + b * digits[2 * j + 1] for j in range(len(digits) // 2)]
    for w, (u0, u1, v) in enumerate(slots):
        limbs[w * ng] += u0 + b * u1
        limbs[w * ng + 1] += v
    return limbs


def test_fused_rule():
    assert limb_fused_ok(20204692, 22, 16)  # C4's first seeded b
    assert limb_fused_ok(35000000, 22, 16)
    assert not limb_fused_ok(50000000, 22, 16)   # outside the limb transform altogether
    assert limb_ok(249999998, 12) and not limb_fused_ok(249999998, 12, 8)  # limbs, but not fused
    # the fused bound is stricter than the plain one: a b near the plain limit fails it
    edge = [b for b in range(43000000, 45000000, 50000) if limb_ok(b, 22)]
    assert edge and not limb_fused_ok(edge[-1], 22, 8)


@pytest.mark.parametrize("b,n,dup,ng,S,seed", [(50000000, 12, True, 16, 4, 4), (200000000, 12, False, 8, 2, 5),
                                               (100000000, 13, True, 32, 8, 6), (60000000, 12, False, 16, 4, 7)])
def test_fused_normalisation_and_carries(b, n, dup, ng, S, seed):
    """Digits of the square through the fused normalisation: flushed, they are balanced and equal
    the square; as limbs with the range carries added (the next column pass), they keep the value
    and stay within the sizes limb_fused_ok counts; a second squaring from those limbs is exact."""
    if not limb_fused_ok(b, n, ng):
        pytest.skip("outside the fused bound")
    N, M = 1 << n, 1 << (n - 1)
    rng = random.Random(seed)
    half = b // 2
    d = [rng.choice([-half - 4, half + 4]) if rng.random() < 0.3 else rng.randint(-half - 4, half + 4) for _ in range(N)]
    mod = b ** N + 1
    r = value(d, b) % mod
    want = (r * r * (2 if dup else 1)) % mod
    digits, slots = normalise_fused(negacyclic_square([d[2 * j] + b * d[2 * j + 1] for j in range(M)]), b, n, dup, ng, S)
    assert max(abs(x) for x in digits) <= half + 4
    flushed = flush_slots(digits, slots, b, ng)
    assert value(flushed, b) % mod == want
    assert max(abs(x) for x in flushed) <= half + 4
    limbs = fused_limbs(digits, slots, b, ng)
    B = b * b
    assert sum(l * B ** j for j, l in enumerate(limbs)) % mod == want
    L = (half + 4) * (b + 1)
    h2max = P // (4 * B * B) + 1
    for j, l in enumerate(limbs):
        bound = (b + 4) * (b + 1) if j % ng == 0 else L + h2max + 2 if j % ng == 1 else L
        assert abs(l) <= bound, (j, l, bound)
    digits2, slots2 = normalise_fused(negacyclic_square(limbs), b, n, False, ng, S)
    assert value(flush_slots(digits2, slots2, b, ng), b) % mod == want * want % mod


@pytest.mark.parametrize("b", [1 << 23, (1 << 23) + 1, 20204692, 43999999, 100000001, 249999998])
def test_limb_digits_at_the_edges(b):
    """The device's two-digit split (rounded double quotients, one 32-bit correction each) equals the
    exact balanced division for values at quotient ties, at the balanced digits' ends and up to
    2^61 in size (limb_digits asserts the equality)."""
    B = b * b
    rng = random.Random(b)
    vs = [0, 1, -1, b // 2, -(b // 2), (b - 1) // 2 + 1, B // 2, -(B // 2), 2 ** 61 - 1, -(2 ** 61) + 1]
    for k in (1, 3, 2 ** 20, 2 ** 30, 2 ** 37):
        for off in (-1, 0, 1):
            vs += [k * B + off, -k * B + off, k * b + b // 2 + off, -k * b - b // 2 + off,
                   k * B + (B // 2) + off, k * B + b * (b // 2) + b // 2 + off]
    vs += [rng.randint(-(2 ** 61), 2 ** 61) for _ in range(2000)]
    for v in vs:
        if abs(v) < 2 ** 61:
            d0, d1, c = limb_digits(v, b)
            assert d0 + b * d1 + B * c == v
