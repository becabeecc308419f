// exponent.h -- the top bits of the PRP exponent E = b^(2^n) without computing E (host code, GMP).
//
// A PRP loop that runs only its first k squarings reads bits esize-1 .. esize-k of E
// (src/genefer.h:479-552).  E itself takes n squarings of numbers up to esize bits: 92 ms at n = 20
// and 0.48 s at n = 22 on one core, more than the squarings themselves for a short request.  The
// top bits come from n squarings of P-bit approximations instead, with a lower and an upper bound
// carried side by side (lo <= E_i / 2^s <= hi after each squaring and truncation, floor for lo and
// ceiling for hi): after n squarings hi - lo is below 2^(n + 2 - P) of their size, so with
// P = k + n + 128 the two bounds agree in their top k + 8 bits unless a multiple of the last kept
// bit's weight falls between them (about one case in 2^100); then, and for small E, E is computed.
#pragma once

#include <gmp.h>

#include <cstdint>

namespace gfe {

// T = floor(E / 2^S) with at least k bits, exact; *S the shift, *esize = bits of E.
inline void exponent_top(uint32_t b, uint32_t n, uint64_t k, mpz_t T, uint64_t *S, uint64_t *esize) {
    const uint64_t P = k + n + 128;
    // E has about 2^n log2(b) bits; small ones are computed outright
    const double approx_bits = (double)(uint64_t(1) << n) * (double)(64 - __builtin_clzll((uint64_t)b));
    if (approx_bits <= (double)(4 * P)) {
        mpz_ui_pow_ui(T, b, uint64_t(1) << n);
        *S = 0;
        *esize = mpz_sizeinbase(T, 2);
        return;
    }
    mpz_t lo, hi;
    mpz_inits(lo, hi, nullptr);
    mpz_set_ui(lo, b);
    mpz_set_ui(hi, b);
    uint64_t s = 0;
    for (uint32_t i = 0; i < n; ++i) {
        mpz_mul(lo, lo, lo);
        mpz_mul(hi, hi, hi);
        s *= 2;
        const uint64_t bits = mpz_sizeinbase(lo, 2);
        if (bits > P) {
            const uint64_t d = bits - P;
            mpz_fdiv_q_2exp(lo, lo, d);
            mpz_cdiv_q_2exp(hi, hi, d);
            s += d;
        }
    }
    const uint64_t bl = mpz_sizeinbase(lo, 2), bh = mpz_sizeinbase(hi, 2);
    bool ok = bl == bh && bl > k + 8;
    if (ok) {
        const uint64_t t = bl - (k + 8);
        mpz_fdiv_q_2exp(lo, lo, t);
        mpz_fdiv_q_2exp(hi, hi, t);
        ok = mpz_cmp(lo, hi) == 0;
        if (ok) {
            mpz_set(T, lo);
            *S = s + t;
            *esize = bl + s;
        }
    }
    if (!ok) {
        mpz_ui_pow_ui(T, b, uint64_t(1) << n);
        *S = 0;
        *esize = mpz_sizeinbase(T, 2);
    }
    mpz_clears(lo, hi, nullptr);
}

}  // namespace gfe
