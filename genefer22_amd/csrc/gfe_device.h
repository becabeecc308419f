// gfe_device.h -- device-side modular arithmetic, CRT and carry for the gfx950 squaring engine.
//
// Residues are plain integers in [0, p) (no Montgomery form between operations).
//  * products by a constant (twiddles, CRT constants) take the constant in Montgomery form
//    w 2^32 mod p, so that Montgomery reduction returns the plain product a w; with the constant's
//    quotient factor w q precomputed (mulp) that is three 32-bit multiplies (mul_hi, mul_lo, mul_hi,
//    one VALU issue slot each on gfx950) and a correction -- the issue cost of Shoup's method with
//    4-byte instead of 8-byte twiddles;
//  * data x data products (the pointwise square / product) use the same reduction; the resulting
//    factor 2^-32 is folded into the per-prime scale applied before the CRT.
//  * S primes (p < 2^31) correct with unsigned min() tricks; U primes (p > 2^31, 2p >= 2^32) with
//    compares.
// The reference's arithmetic for the same primes is ocl/kernel.cl:94-126 (Montgomery on data kept
// in Montgomery form); the residues computed here are the same integers mod p.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfe {

// ---------------------------------------------------------------------------------------------
// Modular arithmetic

struct ModS {  // p < 2^31
    uint32_t p, q;  // q = p^-1 mod 2^32
    __device__ __forceinline__ uint32_t add(uint32_t a, uint32_t b) const {
        const uint32_t s = a + b;
        return min(s, s - p);
    }
    __device__ __forceinline__ uint32_t sub(uint32_t a, uint32_t b) const {
        const uint32_t d = a - b;
        return min(d, d + p);
    }
    // a * w mod p; a is any 32-bit value, wm = w 2^32 mod p.
    __device__ __forceinline__ uint32_t mulw(uint32_t a, uint32_t wm) const { return mont(a, wm); }
    // a * b * 2^-32 mod p, a < 2^32, b < p.
    __device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b) const {
        const uint64_t t = (uint64_t)a * b;
        const uint32_t m = (uint32_t)t * q;
        const uint32_t r = (uint32_t)(t >> 32) - __umulhi(m, p);
        return min(r, r + p);
    }
    // a * w mod p with wm = w 2^32 mod p and wq = wm q mod 2^32: the Montgomery quotient is
    // a wq mod 2^32, so the product needs no 64-bit multiply (three full-rate 32-bit multiplies
    // instead of a half-rate v_mad_u64_u32 and two).
    __device__ __forceinline__ uint32_t mulp(uint32_t a, uint32_t wm, uint32_t wq) const {
        const uint32_t r = __umulhi(a, wm) - __umulhi(a * wq, p);
        return min(r, r + p);
    }
};

struct ModU {  // 2^31 < p < 2^32
    uint32_t p, q;
    // x - y mod p for x, y < p: one subtraction whose borrow selects the corrected value
    // (v_sub_co + v_add + v_cndmask; a + b may not fit 32 bits, so the sum is formed as a - (p - b)).
    __device__ __forceinline__ uint32_t subr(uint32_t x, uint32_t y) const {
        uint32_t d;
        const bool borrow = __builtin_sub_overflow(x, y, &d);
        return borrow ? d + p : d;
    }
    __device__ __forceinline__ uint32_t add(uint32_t a, uint32_t b) const { return subr(a, p - b); }
    __device__ __forceinline__ uint32_t sub(uint32_t a, uint32_t b) const { return subr(a, b); }
    __device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b) const {
        const uint64_t t = (uint64_t)a * b;
        const uint32_t m = (uint32_t)t * q;
        return subr((uint32_t)(t >> 32), __umulhi(m, p));
    }
    // a * w mod p with wm = w 2^32 mod p (Montgomery form of w).
    __device__ __forceinline__ uint32_t mulw(uint32_t a, uint32_t wm) const { return mont(a, wm); }
    // As ModS::mulp (wq = wm q mod 2^32).
    __device__ __forceinline__ uint32_t mulp(uint32_t a, uint32_t wm, uint32_t wq) const {
        return subr(__umulhi(a, wm), __umulhi(a * wq, p));
    }
};

// Workgroups are dispatched to the 8 XCDs round-robin by linear workgroup id, and each XCD has its
// own L2.  When a grid's x extent is a multiple of 8 the XCD of workgroup (x, y, z) is x mod 8;
// this maps x to block (x mod 8) (nx / 8) + x / 8, so that every XCD runs one contiguous range of
// blocks and neighbouring blocks that share cache lines meet in the same L2.
__device__ __forceinline__ unsigned xcd_block(unsigned x, unsigned nx) {
    return (nx & 7u) ? x : (x & 7u) * (nx >> 3) + (x >> 3);
}

// Residue of a signed digit.
__device__ __forceinline__ uint32_t to_res(int32_t d, uint32_t p) {
    return (d < 0) ? (uint32_t)d + p : (uint32_t)d;
}

// ---------------------------------------------------------------------------------------------
// CRT and carry

// Per-candidate normalisation constants (host-computed).
struct CrtConst {
    uint32_t p[3];
    uint32_t q[3];
    uint32_t scale[3];     // multiplies the raw inverse-transform output by 2^32 N^-1 (Montgomery form)
    uint32_t inv_p1_m2;    // 1/p1 mod p2 (Montgomery form)
    uint32_t p1_m3;        // p1 mod p3 (Montgomery form)
    uint32_t inv_p12_m3;   // 1/(p1 p2) mod p3 (Montgomery form)
    uint64_t p12;       // p1 p2
    double p12_d;       // (double)(p1 p2)
    // the fast normalisation (crt_digit_fast): Montgomery quotient factors w q of the constants above,
    uint32_t scale_q[3], inv_p1_m2_q, p1_m3_q, inv_p12_m3_q;
    uint32_t p12_lo;    // p1 p2 mod 2^32
    int32_t fast;       // 1 when N b < 2^49 (one rounded quotient is within 1 of the true one)
    double p1_d;        // (double)p1
};

// Balanced division of the exact integer f (given as f mod 2^64 and an approximation f_d with
// |f_d - f| small relative to b) by b: returns d in (-b/2, b/2] with f = q b + d, q in *quot.
__device__ __forceinline__ int32_t div_bal(uint64_t f64, double f_d, uint32_t b, double binv, int64_t *quot) {
    int64_t qi = (int64_t)rint(f_d * binv);
    int64_t r = (int64_t)(f64 - (uint64_t)qi * (uint64_t)b);
    const int64_t k = (int64_t)rint((double)r * binv);
    r -= k * (int64_t)b;
    qi += k;
    const int64_t hi = (int64_t)(b / 2), lo = -(int64_t)((b - 1) / 2);
    if (r > hi) { r -= b; qi += 1; }
    if (r < lo) { r += b; qi -= 1; }
    *quot = qi;
    return (int32_t)r;
}

// Exact coefficient from residues (already scaled), as (x mod 2^64, approximate x).
template <int RNS, class M1, class M2, class M3>
__device__ __forceinline__ void garner(const CrtConst &c, const M1 &m1, const M2 &m2, const M3 &m3,
                                       uint32_t a1, uint32_t a2, uint32_t a3, uint64_t *x64, double *xd) {
    // y2 = (a2 - a1) / p1 mod p2
    const uint32_t a1m2 = (a1 >= c.p[1]) ? a1 - c.p[1] : a1;
    const uint32_t y2 = m2.mulw(m2.sub(a2, a1m2), c.inv_p1_m2);
    const uint64_t lo = (uint64_t)a1 + (uint64_t)c.p[0] * y2;  // in [0, p1 p2)
    if (RNS == 2) {
        const bool neg = (y2 > (c.p[1] - 1) / 2) || ((y2 == (c.p[1] - 1) / 2) && (a1 > (c.p[0] - 1) / 2));
        const uint64_t x = neg ? lo - c.p12 : lo;
        *x64 = x;
        *xd = (double)(int64_t)x;
    } else {
        // y3 = (a3 - y1 - p1 y2) / (p1 p2) mod p3
        const uint32_t a1m3 = (a1 >= c.p[2]) ? a1 - c.p[2] : a1;  // p1 < 2 p3
        const uint32_t t = m3.sub(m3.sub(a3, a1m3), m3.mulw(y2, c.p1_m3));
        const uint32_t y3 = m3.mulw(t, c.inv_p12_m3);
        const bool neg = (y3 > (c.p[2] - 1) / 2) ||
                         ((y3 == (c.p[2] - 1) / 2) && ((y2 > (c.p[1] - 1) / 2) ||
                                                       ((y2 == (c.p[1] - 1) / 2) && (a1 > (c.p[0] - 1) / 2))));
        const int64_t Y3 = neg ? (int64_t)y3 - (int64_t)c.p[2] : (int64_t)y3;
        *x64 = lo + c.p12 * (uint64_t)Y3;
        *xd = (double)lo + c.p12_d * (double)Y3;
    }
}

// Low 32 bits of an integer-valued double q, |q| < 2^52 (all steps exact).
__device__ __forceinline__ uint32_t lo32_of(double q) {
    const double hi = floor(q * 0x1p-32);
    return (uint32_t)fma(hi, -0x1p32, q);
}

// One digit of the normalisation when N b < 2^49 (c.fast): the exact coefficient x of the digit
// (Garner mixed radix r1 + p1 y2 + p1 p2 Y3, centred), times 2 when dup, plus the carry, divided by b
// with a balanced remainder.  x enters as its low 32 bits and a double estimate; |x| <= 2 N (b/2+4)^2
// makes the estimate's error below b/8 (NTT3: the fused products' rounding, at most 2^11 absolute,
// against b >= 2^19; NTT2: x is converted exactly from int64), so the rounded quotient is within
// one of the true one and the exact 32-bit remainder (|r| < 0.9 b < 2^31) decides the last step.
// The carry (|c| < N b / 2 < 2^48) travels as an exact double cd and its low 32 bits cl.
template <int RNS, class M1, class M2, class M3>
__device__ __forceinline__ int32_t crt_digit_fast(const CrtConst &c, const M1 &m1, const M2 &m2, const M3 &m3,
                                                  uint32_t a1, uint32_t a2, uint32_t a3, bool dup, uint32_t b,
                                                  double binv, double &cd, uint32_t &cl) {
    const uint32_t r1 = m1.mulp(a1, c.scale[0], c.scale_q[0]);
    const uint32_t r2 = m2.mulp(a2, c.scale[1], c.scale_q[1]);
    const uint32_t r1m2 = (r1 >= c.p[1]) ? r1 - c.p[1] : r1;
    const uint32_t y2 = m2.mulp(m2.sub(r2, r1m2), c.inv_p1_m2, c.inv_p1_m2_q);
    uint32_t xl;
    double xd;
    if (RNS == 2) {
        const uint64_t lo = (uint64_t)r1 + (uint64_t)c.p[0] * y2;
        const bool neg = (y2 > (c.p[1] - 1) / 2) || ((y2 == (c.p[1] - 1) / 2) && (r1 > (c.p[0] - 1) / 2));
        const int64_t x = neg ? (int64_t)(lo - c.p12) : (int64_t)lo;
        xl = (uint32_t)x;
        xd = (double)x;
    } else {
        const uint32_t r3 = m3.mulp(a3, c.scale[2], c.scale_q[2]);
        const uint32_t r1m3 = (r1 >= c.p[2]) ? r1 - c.p[2] : r1;
        const uint32_t t = m3.sub(m3.sub(r3, r1m3), m3.mulp(y2, c.p1_m3, c.p1_m3_q));
        const uint32_t y3 = m3.mulp(t, c.inv_p12_m3, c.inv_p12_m3_q);
        const bool neg = (y3 > (c.p[2] - 1) / 2) ||
                         ((y3 == (c.p[2] - 1) / 2) && ((y2 > (c.p[1] - 1) / 2) ||
                                                       ((y2 == (c.p[1] - 1) / 2) && (r1 > (c.p[0] - 1) / 2))));
        const int32_t Y3 = neg ? (int32_t)(y3 - c.p[2]) : (int32_t)y3;
        xl = r1 + c.p[0] * y2 + c.p12_lo * (uint32_t)Y3;
        xd = fma(c.p12_d, (double)Y3, fma(c.p1_d, (double)y2, (double)r1));
    }
    if (dup) { xl <<= 1; xd *= 2.0; }
    const double fd = xd + cd;
    const uint32_t fl = xl + cl;
    double qd = rint(fd * binv);
    uint32_t ql = lo32_of(qd);
    int32_t r = (int32_t)(fl - ql * b);
    if (r > (int32_t)(b / 2)) { r -= (int32_t)b; qd += 1.0; ql += 1u; }
    else if (r < -(int32_t)((b - 1) / 2)) { r += (int32_t)b; qd -= 1.0; ql -= 1u; }
    cd = qd;
    cl = ql;
    return r;
}

// The exact int64 value of an integer-valued double carry (|c| < 2^52).
__device__ __forceinline__ int64_t carry_of(double cd) {
    const double hi = floor(cd * 0x1p-32);
    const uint32_t lo = (uint32_t)fma(hi, -0x1p32, cd);
    return (int64_t)(((uint64_t)(uint32_t)(int32_t)hi << 32) | lo);
}

}  // namespace gfe
