// gfe_limb.h -- the two-digit-limb transform: CRT and carry of limbs over four 31-bit primes.
//
// b^N + 1 = B^M + 1 with B = b^2 and M = N / 2, so a residue is also the M limbs
// l_j = d_2j + b d_2j+1 in base B, and a squaring mod b^N + 1 is a negacyclic convolution of length
// M of the limbs.  With balanced digits |l_j| <= (b/2 + 4)(b + 1) and a (doubled) coefficient is at
// most N l_max^2; four primes p1 p2 p3 p4 (P ~ 2^123.6) keep that below P / 4 for b up to about
// 4.4e7 at n = 22 (limb_ok on the host).  Against the three-prime transform of length N this
// halves the transform length and multiplies the primes by 4/3: 36 % fewer butterflies and a third
// less transform data per squaring.
//
// CRT (explicit form).  With z_k = x (P/p_k)^-1 mod p_k (one Montgomery product per prime: the
// inverse transform's raw output times a constant that also undoes the transform scale and the
// pointwise product's 2^-32), x = sum_k z_k (P/p_k) - t P where t = round(sum_k z_k / p_k), since
// |x| < P/4 puts the fractional part within 1/4 of an integer.  P/p_k and P are split in balanced
// base-B pieces on the host (w0 + w1 B + w2 B^2), so x = h0 + h1 B + h2 B^2 with |h0|, |h1| <= B/2
// follows from two balanced divisions by B of sums known exactly mod 2^64 and as double estimates
// (error < 2^35 << B, B >= 2^46), and h2 (|h2| <= N/4 + 1) exactly.  The limb of the product is then
// L_j = h0_j + h1_(j-1) + h2_(j-2) (|L_j| <~ B + N/4), split into two balanced digits with a carry
// of a few units: every quantity stays within 64 bits.
#pragma once

#include "gfe_kernels.h"

namespace gfe {

constexpr int LIMB_PRIMES = 4;
constexpr uint32_t LIMB_P4 = 27u * (1u << 26) + 1u;  // 1811939329, generator 13; 2^26 | p4 - 1
constexpr uint32_t LIMB_G4 = 13u;
constexpr int LIMB_S = 4;                             // limbs per normalisation thread
constexpr uint32_t LIMB_RANGE = 256 * LIMB_S;         // limbs per normalisation workgroup (2048 digits)

struct LimbCrt {
    uint32_t p[4], q[4];
    uint32_t zc[4], zc_q[4];   // mulp operands of 2^32 M^-1 (P/p_k)^-1 mod p_k and their w q factors
    double inv_p[4];           // 1 / p_k
    int64_t w0[4], w1[4], w2[4];  // P/p_k = w0 + w1 B + w2 B^2 (balanced pieces)
    uint32_t w0l[4], w0h[4], w1l[4], w1h[4];  // low and high words of w0, w1 (two's complement)
    int32_t w2s[4];            // w2 (|w2| <= 2: P/p_k < 2^93 <= 2 B^2)
    double w0d[4], w1d[4];
    int64_t P0, P1, P2;        // P = P0 + P1 B + P2 B^2
    double P0d, P1d;
    uint64_t B;                // b^2
    double invB;
    uint32_t b;
    int32_t pad;
    double binv;
};

// The integer-valued double v (|v| < 2^51) as an int64: v + 1.5 2^52 has v's two's complement in
// the low bits of its significand, so one add and an integer subtract recover it.
__device__ __forceinline__ int64_t i64_of(double v) {
    return (int64_t)(__double_as_longlong(v + 0x1.8p52) - 0x4338000000000000ll);
}

// Division by B of the exact integer s, given as s mod 2^64 and an estimate sd with |sd - s| < 2^38:
// s = q B + h with q = rint(sd / B), so |h| <= B/2 + 2^38 (B >= 2^46: within B/2 (1 + 2^-8)), not
// balanced further -- every later step is exact for any h of that size.
// *qd receives q as a double (exact: the pieces of P/p_k are at most B/2, so |q| <= 2^32 + 2).
__device__ __forceinline__ int64_t split_B(uint64_t s64, double sd, const LimbCrt &c, int64_t *q, double *qd) {
    const double qr = rint(sd * c.invB);
    const int64_t qi = i64_of(qr);
    *q = qi;
    *qd = qr;
    return (int64_t)(s64 - (uint64_t)qi * c.B);
}

// z times the 64-bit constant (lo, hi) added to the 64-bit accumulator (acc, acch) mod 2^64: the
// low product into acc (one 32x32+64 multiply-add), the high product's low word into acch.
__device__ __forceinline__ void mac64(uint64_t &acc, uint32_t &acch, uint32_t z, uint32_t lo, uint32_t hi) {
    acc += (uint64_t)z * lo;
    acch += z * hi;
}

// Base-B pieces of the limb coefficient whose raw inverse-transform residues are a[0..3] (times 2
// when dup).
__device__ __forceinline__ void limb_pieces(const LimbCrt &c, const uint32_t (&a)[4], bool dup, int64_t &h0,
                                            int64_t &h1, int64_t &h2) {
    uint32_t z[4];
    double zd[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const ModS m{c.p[k], c.q[k]};
        z[k] = m.mulp(a[k], c.zc[k], c.zc_q[k]);
        zd[k] = (double)z[k];
    }
    double ts = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) ts = fma(zd[k], c.inv_p[k], ts);
    const int32_t t = (int32_t)rint(ts);
    const double td = (double)t;
    uint64_t s0 = (uint64_t)(-(int64_t)t) * (uint64_t)c.P0;
    uint32_t s0h = 0;
    double s0d = -td * c.P0d;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        mac64(s0, s0h, z[k], c.w0l[k], c.w0h[k]);
        s0d = fma(zd[k], c.w0d[k], s0d);
    }
    int64_t q0;
    double q0d;
    h0 = split_B(s0 + ((uint64_t)s0h << 32), s0d, c, &q0, &q0d);
    uint64_t s1 = (uint64_t)q0 - (uint64_t)(int64_t)t * (uint64_t)c.P1;
    uint32_t s1h = 0;
    double s1d = fma(-td, c.P1d, q0d);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        mac64(s1, s1h, z[k], c.w1l[k], c.w1h[k]);
        s1d = fma(zd[k], c.w1d[k], s1d);
    }
    int64_t q1;
    double q1d;
    h1 = split_B(s1 + ((uint64_t)s1h << 32), s1d, c, &q1, &q1d);
    int64_t s2 = q1 - (int64_t)t * c.P2;
#pragma unroll
    for (int k = 0; k < 4; ++k) s2 += (int64_t)(int32_t)z[k] * c.w2s[k];  // z < 2^31
    h2 = s2;
    if (dup) {  // 2x = 2 h0 + 2 h1 B + 2 h2 B^2: pieces up to B stay within the limb bounds below
        h0 *= 2; h1 *= 2; h2 *= 2;
    }
}

// Two balanced digits of the limb value v and the carry out: v = d0 + b d1 + B c, digits in
// [-(b-1)/2, b/2].  v is exact with |v| < 2^62, so (double)v is within 2^9 of it and each rounded
// quotient below is within 1 of the balanced one (b >= 2^23): one correction step per digit, in
// 32-bit arithmetic, since the remainders are below b in size; the first quotient's low word comes
// from its double as in i64_of.
__device__ __forceinline__ void limb_digits(int64_t v, uint32_t b, double binv, int32_t &d0, int32_t &d1, int64_t &c) {
    const int32_t hi = (int32_t)(b / 2), lo = -(int32_t)((b - 1) / 2), bs = (int32_t)b;
    double q = rint((double)v * binv);                                     // |q| < 2^39
    uint32_t ql = (uint32_t)__double_as_longlong(q + 0x1.8p52);
    int32_t r = (int32_t)((uint32_t)v - ql * b);
    const int32_t a1 = (r > hi) - (r < lo);                                // -1, 0 or 1
    d0 = r - a1 * bs;
    ql += (uint32_t)a1;
    q += (double)a1;
    int32_t c2 = (int32_t)rint(q * binv);                                  // |c2| < 2^17
    int32_t r2 = (int32_t)(ql - (uint32_t)c2 * b);
    const int32_t a2 = (r2 > hi) - (r2 < lo);
    d1 = r2 - a2 * bs;
    c = (int64_t)(c2 + a2);
}

// Adds the carry c of the thread below into balanced digits d (c = 0 leaves them): while |c| is at
// most 8, far below b / 2, every digit needs at most one correction by b and the carry on stays in
// {-1, 0, 1}, so 32-bit compares do what div_bal does; a larger c takes absorb_carry.
template <int S>
__device__ __forceinline__ void absorb_small_carry(int32_t (&d)[S], int64_t c, uint32_t b, double binv) {
    if (c > 8 || c < -8) {
        absorb_carry<S>(d, c, b, binv);
        return;
    }
    const int32_t hi = (int32_t)(b / 2), lo = -(int32_t)((b - 1) / 2), bs = (int32_t)b;
    int32_t k = (int32_t)c;
#pragma unroll
    for (int j = 0; j < S - 1; ++j) {
        if (k != 0) {
            const int32_t f = d[j] + k;
            k = (f > hi) - (f < lo);
            d[j] = f - k * bs;
        }
    }
    d[S - 1] += k;
}

// Adds the carries of a limb normalisation range boundary into the next range's first S digits:
// e0 at digit 0, e1 at digit 2 (the limb after), digit by digit while a carry remains.
template <int S>
__device__ __forceinline__ void absorb_carry2(int32_t (&d)[S], int64_t e0, int64_t e1, uint32_t b, double binv) {
    int64_t c = e0;
#pragma unroll
    for (int j = 0; j < S - 1; ++j) {
        if (j == 2) c += e1;
        if (c != 0) {
            const int64_t f = (int64_t)d[j] + c;
            d[j] = div_bal((uint64_t)f, (double)f, b, binv, &c);
        }
    }
    d[S - 1] += (int32_t)c;
}

// Normalisation of limbs [blk 1024, (blk + 1) 1024): scale + CRT to base-B pieces, the pieces of
// the limbs below added in (the previous thread's through LDS; the previous workgroup's by
// carry_limb_body after the kernel), two digits per limb with the small carry, the thread carries
// handed up through LDS.  wl[2 blk], wl[2 blk + 1] receive what the next range's first two limbs
// still need: e0 (into its limb 0) and e1 (into its limb 1).  blockDim.x must be 256.
__device__ __forceinline__ void norm_limb_body(const LimbCrt &c, const uint32_t *__restrict__ ws, size_t M,
                                               int32_t *__restrict__ dst, int64_t *__restrict__ wl, bool dup,
                                               unsigned blk) {
    constexpr int S = LIMB_S;
    __shared__ int64_t xs[3][256];
    const int tid = threadIdx.x;
    const size_t j0 = ((size_t)blk * 256 + tid) * S;
    uint32_t a[4][S];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 v = *reinterpret_cast<const uint4 *>(ws + (size_t)k * M + j0);
        a[k][0] = v.x; a[k][1] = v.y; a[k][2] = v.z; a[k][3] = v.w;
    }
    int64_t h0[S], h1[S], h2[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const uint32_t r[4] = {a[0][j], a[1][j], a[2][j], a[3][j]};
        limb_pieces(c, r, dup, h0[j], h1[j], h2[j]);
    }
    xs[0][tid] = h1[S - 1];
    xs[1][tid] = h2[S - 2];
    xs[2][tid] = h2[S - 1];
    __syncthreads();
    int64_t pb1 = 0, pb2 = 0, pl2 = 0;  // previous thread: h1 of its last limb, h2 of its last two
    if (tid > 0) { pb1 = xs[0][tid - 1]; pb2 = xs[1][tid - 1]; pl2 = xs[2][tid - 1]; }
    const uint32_t b = c.b;
    const double binv = c.binv;
    int32_t d[2 * S];
    int64_t cr = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int64_t lj = h0[j] + (j == 0 ? pb1 + pb2 : h1[j - 1] + (j == 1 ? pl2 : h2[j - 2]));
        limb_digits(lj + cr, b, binv, d[2 * j], d[2 * j + 1], cr);
    }
    // the thread's carry-out (a few units) goes into the next thread's digits
    __syncthreads();
    xs[0][tid] = cr;
    __syncthreads();
    if (tid > 0) absorb_carry<2 * S>(d, xs[0][tid - 1], b, binv);
    int32_t *o = dst + 2 * j0;
#pragma unroll
    for (int j = 0; j < 2 * S; j += 4) *reinterpret_cast<int4 *>(o + j) = make_int4(d[j], d[j + 1], d[j + 2], d[j + 3]);
    if (tid == 255) {
        wl[2 * (size_t)blk] = h1[S - 1] + h2[S - 2] + cr;
        wl[2 * (size_t)blk + 1] = h2[S - 1];
    }
}

// Range w's first 8 digits receive the boundary carries of range w - 1 (of the last range, negated,
// for w = 0: B^M = -1); the pair is cleared once applied.
__device__ __forceinline__ void carry_limb_body(int32_t *dc, int64_t *wl, unsigned nwg, uint32_t b, unsigned w) {
    const unsigned src = (w == 0) ? nwg - 1 : w - 1;
    int64_t e0 = wl[2 * (size_t)src], e1 = wl[2 * (size_t)src + 1];
    if (e0 == 0 && e1 == 0) return;
    wl[2 * (size_t)src] = 0;
    wl[2 * (size_t)src + 1] = 0;
    if (w == 0) { e0 = -e0; e1 = -e1; }
    int32_t *p = dc + (size_t)w * 2 * LIMB_RANGE;
    int32_t e[8];
    const int4 v0 = *reinterpret_cast<const int4 *>(p), v1 = *reinterpret_cast<const int4 *>(p + 4);
    e[0] = v0.x; e[1] = v0.y; e[2] = v0.z; e[3] = v0.w; e[4] = v1.x; e[5] = v1.y; e[6] = v1.z; e[7] = v1.w;
    absorb_carry2<8>(e, e0, e1, b, 1.0 / (double)b);
    *reinterpret_cast<int4 *>(p) = make_int4(e[0], e[1], e[2], e[3]);
    *reinterpret_cast<int4 *>(p + 4) = make_int4(e[4], e[5], e[6], e[7]);
}

// ---------------------------------------------------------------------------------------------
// Fused inverse column pass and normalisation (ks_col_inv_norm_limb): a workgroup owns NG
// consecutive columns of the 2^a x C matrix of limbs for all four primes, so after its inverse
// column transforms (col_inv_to_lds) it holds 2^a rows of NG consecutive limbs each -- limbs
// k C + blk NG + j, j < NG, for row k -- and normalises every row as a range of its own.  What a
// range hands on to the next (e0 into its limb 0, e1 into its limb 1: the pieces of the range's
// last limbs and the digit carry) is split into digits, e0 = u0 + u1 b + u2 B, and left in
// slots[next] = (u0, u1, u2 + e1), negated for the range after the last (B^M = -1).  The next
// forward column pass adds these to the first two limbs' digit pairs without normalising them
// (col_fwd_pass, ls): such a limb is at most (b + 4)(b + 1) or a balanced limb plus a few units
// more than the largest h2 -- about twice a balanced limb -- which the host's bound for these plans
// counts (limb_fused_ok).  Every other use of the register runs carry_limb_slot_body first.

struct LimbColNormArgs {
    const uint32_t *ws;   // [4][M] residues after the row pass (read only)
    const uint32_t *tw;   // [4][M] twiddle tables
    PrimeArgs pr;
    int32_t *dst;         // register 0: 2M digits
    int4 *slots;          // M / ng: the carry into each range of ng limbs
    const LimbCrt *crt;
    int n, a, ng;         // n = log2 M
    int dup;
};

// The 64-bit value of the lane one below within each row of 16 lanes (DPP row_shr:1, a VALU move
// per word); lane 0 of a row reads 0.
__device__ __forceinline__ int64_t lane_below(int64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(uint64_t)v, 0x111, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)((uint64_t)v >> 32), 0x111, 0xF, 0xF, false);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The 32-bit value of the lane one below, as lane_below.
__device__ __forceinline__ int32_t lane_below32(int32_t v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);
}

// The normalisation half: rows of NG limbs whose four primes' residues are at
// NL[p rows NG + k NG + j]; thread t takes S consecutive limbs of row t / (NG / S).  The NG / S
// threads of a row are consecutive lanes of one 16-lane DPP row (NG / S <= 16), so the pieces of
// the previous thread's last limbs and its digit carry come across by DPP moves, with no LDS
// round trip and no workgroup barrier.  blockDim.x = rows NG / S.
template <int S>
__device__ __forceinline__ void norm_limb_rows(const LimbCrt &c, const uint32_t *NL, uint32_t NG, uint32_t rows,
                                               uint32_t C, unsigned blk, int32_t *__restrict__ dst,
                                               int4 *__restrict__ slots, uint32_t nrange, bool dup) {
    static_assert(S >= 2 && S % 2 == 0, "two or more limbs per thread");
    const uint32_t tid = threadIdx.x, tpr = NG / S;
    const uint32_t k = tid / tpr, jt = tid % tpr, j0 = jt * S;
    uint32_t a[4][S];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t *src = NL + (size_t)p * rows * NG + k * NG + j0;
        if constexpr (S % 4 == 0) {
#pragma unroll
            for (int j = 0; j < S; j += 4) {
                const uint4 v = *reinterpret_cast<const uint4 *>(src + j);
                a[p][j] = v.x; a[p][j + 1] = v.y; a[p][j + 2] = v.z; a[p][j + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < S; j += 2) {
                const uint2 v = *reinterpret_cast<const uint2 *>(src + j);
                a[p][j] = v.x; a[p][j + 1] = v.y;
            }
        }
    }
    int64_t h0[S], h1[S], h2[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const uint32_t r[4] = {a[0][j], a[1][j], a[2][j], a[3][j]};
        limb_pieces(c, r, dup, h0[j], h1[j], h2[j]);
    }
    // the row's previous thread: h1 of its last limb, h2 of its last two (a row's first thread: none)
    const int64_t pb1 = lane_below(h1[S - 1]), pb2 = lane_below(h2[S - 2]), pl2 = lane_below(h2[S - 1]);
    const bool first = jt == 0;
    const uint32_t b = c.b;
    const double binv = c.binv;
    int32_t d[2 * S];
    int64_t cr = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int64_t lj = h0[j] + (j == 0 ? (first ? 0 : pb1 + pb2) : h1[j - 1] + (j == 1 ? (first ? 0 : pl2) : h2[j - 2]));
        limb_digits(lj + cr, b, binv, d[2 * j], d[2 * j + 1], cr);
    }
    // |cr| <= 2: a limb is below 2B + 2^40 + N/2 + 2 in size (doubled pieces), so its carry over
    // B >= 2^46 after two balanced digits is below 2.52 + |carry in| / B
    const int32_t cin = lane_below32((int32_t)cr);  // |cr| < 2^17 + 1 (limb_digits) in any case
    if (!first) absorb_small_carry<2 * S>(d, cin, b, binv);
    int32_t *o = dst + 2 * ((size_t)k * C + (size_t)blk * NG + j0);
#pragma unroll
    for (int j = 0; j < 2 * S; j += 4) *reinterpret_cast<int4 *>(o + j) = make_int4(d[j], d[j + 1], d[j + 2], d[j + 3]);
    if (jt == tpr - 1) {
        const int64_t e0 = h1[S - 1] + h2[S - 2] + cr, e1 = h2[S - 1];
        int32_t u0, u1;
        int64_t u2;
        limb_digits(e0, b, binv, u0, u1, u2);
        uint32_t w = k * (C / NG) + blk + 1;
        int32_t sg = 1;
        if (w == nrange) { w = 0; sg = -1; }
        slots[w] = make_int4(sg * u0, sg * u1, sg * (int32_t)(u2 + e1), 0);
    }
}

// Range w of 2^rlog limbs (digits from 2w 2^rlog) receives slots[w] = (u0, u1, v): u0 + u1 b at its
// digit 0 and v at its digit 2, normalised over its first 8 digits as carry_limb_body does.
__device__ __forceinline__ void carry_limb_slot_body(int32_t *dc, const int4 *slots, int rlog, uint32_t b, unsigned w) {
    const int4 v = slots[w];
    if (v.x == 0 && v.y == 0 && v.z == 0) return;
    int32_t *p = dc + ((size_t)w << (rlog + 1));
    int32_t e[8];
    const int4 v0 = *reinterpret_cast<const int4 *>(p), v1 = *reinterpret_cast<const int4 *>(p + 4);
    e[0] = v0.x; e[1] = v0.y; e[2] = v0.z; e[3] = v0.w; e[4] = v1.x; e[5] = v1.y; e[6] = v1.z; e[7] = v1.w;
    absorb_carry2<8>(e, (int64_t)v.x + (int64_t)v.y * (int64_t)b, (int64_t)v.z, b, 1.0 / (double)b);
    *reinterpret_cast<int4 *>(p) = make_int4(e[0], e[1], e[2], e[3]);
    *reinterpret_cast<int4 *>(p + 4) = make_int4(e[4], e[5], e[6], e[7]);
}

}  // namespace gfe
