"""The normalisation's fast path (crt_digit_fast in genefer22_amd/csrc/gfe_device.h): for N b < 2^49
one rounded double quotient plus an exact 32-bit remainder gives the balanced digit and the carry.
Restated here with IEEE doubles (Python floats) and correctly rounded fused multiply-adds
(Fraction), and checked against exact integer division at the edges of every admissible
configuration: coefficients at the bound 2 N (b/2 + 4)^2, the smallest NTT3 bases, N b just under
2^49, carries at +-N b / 2."""
import math
import random
from fractions import Fraction

S = [127*2**24+1, 63*2**25+1, 15*2**27+1]; U = [125*2**25+1, 243*2**24+1, 235*2**24+1]
def fma(a, b, c): return float(Fraction(a) * Fraction(b) + Fraction(c))
def rint(x):
    r = round(x)  # Python round: half to even == rint default mode
    return float(r)
def lo32(q):
    hi = math.floor(q * 2.0**-32); return int(fma(hi, -2.0**32, q)) & 0xFFFFFFFF
def u32(x): return x & 0xFFFFFFFF
def i32(x): x &= 0xFFFFFFFF; return x - (1 << 32) if x >= 1 << 31 else x
def digit(ps, rns, x, dup, b, cd, cl):
    p1, p2, p3 = ps
    M = p1 * p2 * (p3 if rns == 3 else 1)
    xm = x % M
    r1 = xm % p1; y2 = ((xm % p2 - r1) * pow(p1, -1, p2)) % p2
    binv = 1.0 / b
    if rns == 2:
        lo = r1 + p1 * y2
        neg = (y2 > (p2 - 1) // 2) or (y2 == (p2 - 1) // 2 and r1 > (p1 - 1) // 2)
        xx = lo - p1 * p2 if neg else lo
        assert xx == x
        xl = u32(xx); xd = float(xx)
    else:
        y3 = ((xm % p3 - r1 - p1 * y2) * pow(p1 * p2, -1, p3)) % p3
        neg = (y3 > (p3 - 1) // 2) or (y3 == (p3 - 1) // 2 and ((y2 > (p2 - 1) // 2) or (y2 == (p2 - 1) // 2 and r1 > (p1 - 1) // 2)))
        Y3 = y3 - p3 if neg else y3
        assert r1 + p1 * y2 + p1 * p2 * Y3 == x
        xl = u32(r1 + p1 * y2 + u32(p1 * p2) * u32(Y3))
        xd = fma(float(p1 * p2), float(Y3), fma(float(p1), float(y2), float(r1)))
    if dup: xl = u32(xl << 1); xd *= 2.0
    fd = xd + cd; fl = u32(xl + cl)
    qd = rint(fd * binv); ql = lo32(qd)
    r = i32(fl - ql * b)
    if r > b // 2: r -= b; qd += 1.0; ql = u32(ql + 1)
    elif r < -((b - 1) // 2): r += b; qd -= 1.0; ql = u32(ql - 1)
    return r, qd, ql


def _check(ps, rns, n, b, trials, seed):
    rnd = random.Random(seed)
    N = 1 << n
    bound = 2 * N * (b // 2 + 4) ** 2
    cmax = N * b // 2
    for t in range(trials):
        if t % 4 == 0:
            x = rnd.choice([bound, -bound, bound - rnd.randrange(b), -bound + rnd.randrange(b)])
        else:
            x = rnd.randrange(-bound, bound + 1)
        dup = rnd.random() < 0.5
        if dup:
            x //= 2
        c = rnd.randrange(-cmax, cmax + 1) if t % 3 else rnd.choice([cmax, -cmax, 0])
        r, qd, ql = digit(ps, rns, x, dup, b, float(c), u32(c))
        q, d = divmod((2 * x if dup else x) + c, b)
        if d > b // 2:
            d -= b
            q += 1
        assert (r, int(qd), ql) == (d, q, u32(q)), (rns, n, b, x, dup, c)


def test_fast_normalisation_edges():
    for n in (12, 16, 22, 24):
        N = 1 << n
        for ps, lim in ((S, 1_000_000_000), (U, 2_000_000_000)):
            b2max = int(math.isqrt(ps[0] * ps[1] // 2 // N)) - 2
            b2max -= b2max % 2
            bmax = min((1 << 49) // N - 2, lim)
            bmax -= bmax % 2
            for rns, b in ((2, 1024), (2, b2max), (3, b2max + 2), (3, bmax)):
                _check(ps, rns, n, b, 200, n * 7 + rns + b % 97)
