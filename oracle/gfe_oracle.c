/* gfe_oracle.c -- CPU restatement of genefer's exact-integer squaring engine (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg (CPU timings and the RES64 check beside them) may load it; the product (genefer22_amd/) never links, loads or calls it.
 *
 * What it restates (reference files, path:line):
 *   - prime sets, generators and Montgomery constants: src/transformGPU.h:37-115, 129-164;
 *   - prime-set selection rule: src/transform_ocl.cpp:35-50;
 *   - twiddles w[s+j] = r_{4s}^{bitrev(j,2s)+1}: src/transformGPU.h:1046-1078;
 *   - the "z-transform" (recursive factorisation of x^N+1 down to degree-1 residues, no DWT
 *     weights) with forward butterfly FWD2, inverse BCK2, pair square SQR2/SQR2N and pair
 *     product MUL2/MUL2N: ocl/kernel.cl:235-273, src/transformCPUi32.h:580-741.  The
 *     reference groups two levels per radix-4 step; here the same levels run one at a time;
 *   - NORM = (N/2)^-1 and the Garner CRT: ocl/kernel.cl:2318-2335, src/transformCPUi32.h:482-503;
 *   - the full carry with the negacyclic wrap x^N = -1 (baseMod / baseModMul):
 *     src/transformCPUi32.h:743-798.  The reference divides by b with a two-step Barrett
 *     reduction (src/transformCPUi32.h:423-480, valid for 32 < b < 2^31); here the division is
 *     exact 128-bit integer division with the same truncation toward zero;
 *   - gint canonical form, RES64/old64, hash64/hash32: src/gint.h:40-225.
 * Modular products use a plain 64-bit remainder (no Montgomery form): the residues are the same
 * integers mod p.
 */
#include "gfe_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

static const uint32_t k_primes[6] = {
    127u * (1u << 24) + 1u, 63u * (1u << 25) + 1u, 15u * (1u << 27) + 1u,
    125u * (1u << 25) + 1u, 243u * (1u << 24) + 1u, 235u * (1u << 24) + 1u};
static const uint32_t k_gen[6] = {3, 5, 31, 3, 7, 3};

uint32_t gfo_prime(int idx) { return k_primes[idx]; }
uint32_t gfo_generator(int idx) { return k_gen[idx]; }

static uint32_t mulm(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)(((uint64_t)a * b) % p); }
static uint32_t addm(uint32_t a, uint32_t b, uint32_t p) { uint64_t s = (uint64_t)a + b; return (uint32_t)(s >= p ? s - p : s); }
static uint32_t subm(uint32_t a, uint32_t b, uint32_t p) { return a >= b ? a - b : a + (p - b); }
static uint32_t powm(uint32_t a, uint64_t e, uint32_t p) {
    uint32_t r = 1 % p;
    while (e) { if (e & 1) r = mulm(r, a, p); a = mulm(a, a, p); e >>= 1; }
    return r;
}
static uint32_t invm(uint32_t a, uint32_t p) { return powm(a, p - 2, p); }
static uint32_t mf(uint32_t a, uint32_t p) { return (uint32_t)((((uint64_t)a) << 32) % p); } /* a*2^32 mod p */

/* src/transformGPU.h:37-105 (constants), derived here from their closed forms. */
void gfo_prime_constants(int idx, uint32_t out[9]) {
    const uint32_t p = k_primes[idx], g = k_gen[idx];
    uint32_t q = 1; /* Newton iteration for p^-1 mod 2^32 */
    for (int i = 0; i < 5; ++i) q *= 2u - p * q;
    const uint32_t R = (uint32_t)((((uint64_t)1) << 32) % p);
    const uint32_t RSQ = mulm(R, R, p);
    const uint32_t I = powm(g, (p - 1) / 4, p), S8 = powm(g, (p - 1) / 8, p);
    out[0] = p;
    out[1] = q;
    out[2] = R;
    out[3] = RSQ;
    out[4] = mf(g, p);
    out[5] = mf(I, p);
    out[6] = mulm(mf(I, p), R, p);
    out[7] = mf(S8, p);
    out[8] = mf(mulm(I, S8, p), p);
}

/* src/transformGPU.h:67-75, 107-115. */
void gfo_garner_constants(int is32, uint32_t out32[5], uint64_t out64[2]) {
    const int o = is32 ? 3 : 0;
    const uint32_t p1 = k_primes[o], p2 = k_primes[o + 1], p3 = k_primes[o + 2];
    out32[0] = mf(invm(p2 % p1, p1), p1);
    out32[1] = mf(invm(p3 % p1, p1), p1);
    out32[2] = mf(invm(p3 % p2, p2), p2);
    const u128 P = (u128)p1 * p2 * p3, P_2 = P / 2;
    out32[3] = (uint32_t)P;
    out32[4] = (uint32_t)P_2;
    out64[0] = (uint64_t)(P >> 32);
    out64[1] = (uint64_t)(P_2 >> 32);
}

/* ZPT::norm, src/transformGPU.h:163 with m = N/2 (1009-1011). */
uint32_t gfo_norm(int idx, uint32_t n) {
    const uint32_t p = k_primes[idx], m = 1u << (n - 1);
    return p - (p - 1) / m;
}

/* src/transform_ocl.cpp:35-50. */
void gfo_select(uint32_t b, uint32_t n, int *rns, int *is32) {
    const uint64_t bb = (uint64_t)b * b, N = (uint64_t)1 << n;
    if (bb < ((uint64_t)k_primes[0] * k_primes[1] / 2) / N) { *rns = 2; *is32 = 0; }
    else if (bb < ((uint64_t)k_primes[3] * k_primes[4] / 2) / N) { *rns = 2; *is32 = 1; }
    else if (b <= 1000000000u) { *rns = 3; *is32 = 0; }
    else { *rns = 3; *is32 = 1; }
}

/* transform::bitRev, src/transform.h:101-106: reverses log2(m) bits of i. */
static size_t bitrev(size_t i, size_t m) {
    size_t r = 0;
    for (size_t k = m, j = i; k > 1; k /= 2, j /= 2) r = (2 * r) | (j % 2);
    return r;
}

/* src/transformGPU.h:1046-1078 (plain values). */
void gfo_twiddles(int idx, uint32_t n, uint32_t *w) {
    const uint32_t p = k_primes[idx], g = k_gen[idx];
    const size_t N = (size_t)1 << n;
    w[0] = 0;
    for (size_t s = 1; s < N / 2; s *= 2) {
        const uint32_t r = powm(g, (p - 1) / (uint32_t)(4 * s), p);
        for (size_t j = 0; j < s; ++j) w[s + j] = powm(r, bitrev(j, 2 * s) + 1, p);
    }
}

struct gfo_engine {
    uint32_t b, n;
    size_t N;
    int rns, is32, num_regs;
    uint32_t p[3], norm[3];
    uint32_t inv21, inv31, inv32;   /* 1/p2 mod p1, 1/p3 mod p1, 1/p3 mod p2 */
    uint32_t *w[3], *wi[3];   /* forward roots and their inverses, N/2 entries each */
    int32_t *regs;            /* num_regs x N digits */
    uint32_t *zp;             /* rns x N multiplicand (transform domain) */
    uint32_t *z;              /* rns x N work */
};

gfo_engine *gfo_create(uint32_t b, uint32_t n, int num_regs, int rns, int is32) {
    if (n < 2 || n > 24 || num_regs < 1) return NULL;
    gfo_engine *e = (gfo_engine *)calloc(1, sizeof(gfo_engine));
    e->b = b; e->n = n; e->N = (size_t)1 << n; e->num_regs = num_regs;
    if (rns == 0) gfo_select(b, n, &e->rns, &e->is32); else { e->rns = rns; e->is32 = is32; }
    for (int k = 0; k < e->rns; ++k) {
        const int idx = (e->is32 ? 3 : 0) + k;
        e->p[k] = k_primes[idx];
        e->norm[k] = gfo_norm(idx, n);
        e->w[k] = (uint32_t *)malloc(e->N / 2 * sizeof(uint32_t));
        e->wi[k] = (uint32_t *)malloc(e->N / 2 * sizeof(uint32_t));
        gfo_twiddles(idx, n, e->w[k]);
        e->wi[k][0] = 0;
        /* the inverse of the root of block j at level s is minus the root of block s-1-j
         * ("inverse of roots is wi[s + j] or w[s + s - j - 1]", ocl/kernel.cl:516) */
        for (size_t s = 1; s < e->N / 2; s *= 2)
            for (size_t j = 0; j < s; ++j) e->wi[k][s + j] = (e->p[k] - e->w[k][s + (s - 1 - j)]) % e->p[k];
    }
    e->inv21 = invm(e->p[1] % e->p[0], e->p[0]);
    if (e->rns == 3) {
        e->inv31 = invm(e->p[2] % e->p[0], e->p[0]);
        e->inv32 = invm(e->p[2] % e->p[1], e->p[1]);
    }
    e->regs = (int32_t *)calloc((size_t)num_regs * e->N, sizeof(int32_t));
    e->zp = (uint32_t *)calloc((size_t)e->rns * e->N, sizeof(uint32_t));
    e->z = (uint32_t *)calloc((size_t)e->rns * e->N, sizeof(uint32_t));
    return e;
}

void gfo_destroy(gfo_engine *e) {
    if (!e) return;
    for (int k = 0; k < e->rns; ++k) { free(e->w[k]); free(e->wi[k]); }
    free(e->regs); free(e->zp); free(e->z); free(e);
}

int gfo_rns(const gfo_engine *e) { return e->rns; }
int gfo_is32(const gfo_engine *e) { return e->is32; }

/* set_int, ocl/kernel.cl:126 / Zp(int32_t), src/transformCPUi32.h:37. */
static uint32_t to_res(int32_t d, uint32_t p) { return d < 0 ? (uint32_t)((int64_t)d + p) : (uint32_t)d; }

/* Forward levels 0 .. n-2 (FWD2, ocl/kernel.cl:235-239). */
static void forward(const gfo_engine *e, int k, uint32_t *z) {
    const uint32_t p = e->p[k];
    const size_t N = e->N;
    for (size_t s = 1; s < N / 2; s *= 2) {
        const size_t t = N / (2 * s);
        for (size_t j = 0; j < s; ++j) {
            const uint32_t w = e->w[k][s + j];
            uint32_t *blk = z + 2 * j * t;
            for (size_t i = 0; i < t; ++i) {
                const uint32_t u = blk[i], v = mulm(blk[i + t], w, p);
                blk[i] = addm(u, v, p);
                blk[i + t] = subm(u, v, p);
            }
        }
    }
}

/* Inverse levels n-2 .. 0 (BCK2, ocl/kernel.cl:241-245), without the 1/2 per level. */
static void backward(const gfo_engine *e, int k, uint32_t *z) {
    const uint32_t p = e->p[k];
    const size_t N = e->N;
    for (size_t s = N / 4; s >= 1; s /= 2) {
        const size_t t = N / (2 * s);
        for (size_t j = 0; j < s; ++j) {
            const uint32_t wi = e->wi[k][s + j];
            uint32_t *blk = z + 2 * j * t;
            for (size_t i = 0; i < t; ++i) {
                const uint32_t u = blk[i], v = blk[i + t];
                blk[i] = addm(u, v, p);
                blk[i + t] = mulm(subm(u, v, p), wi, p);
            }
        }
        if (s == 1) break;
    }
}

/* Pair square / product after level n-2: block j (elements 4j..4j+3) was split with root
 * w = w[N/4 + j] into residues mod (x^2 - w) and mod (x^2 + w).  SQR2 / SQR2N / MUL2 / MUL2N,
 * ocl/kernel.cl:247-273; square_22 / mul_22, src/transformCPUi32.h:580-611. */
static void pairs(const gfo_engine *e, int k, uint32_t *z, const uint32_t *zp) {
    const uint32_t p = e->p[k];
    const size_t N = e->N;
    for (size_t j = 0; j < N / 4; ++j) {
        const uint32_t w = e->w[k][N / 4 + j], mw = (p - w) % p;
        for (int h = 0; h < 2; ++h) {
            const uint32_t c = h ? mw : w;
            uint32_t *q = z + 4 * j + 2 * h;
            const uint32_t z0 = q[0], z1 = q[1];
            if (zp) {
                const uint32_t y0 = zp[4 * j + 2 * h], y1 = zp[4 * j + 2 * h + 1];
                q[0] = addm(mulm(z0, y0, p), mulm(mulm(z1, y1, p), c, p), p);
                q[1] = addm(mulm(z0, y1, p), mulm(z1, y0, p), p);
            } else {
                q[0] = addm(mulm(z0, z0, p), mulm(mulm(z1, z1, p), c, p), p);
                q[1] = mulm(addm(z0, z0, p), z1, p);
            }
        }
    }
}

/* Garner CRT to the centred integer in (-P/2, P/2): ocl/kernel.cl:2318-2335. */
static i128 garner(const gfo_engine *e, const uint32_t *r) {
    if (e->rns == 2) {
        const uint32_t p1 = e->p[0], p2 = e->p[1];
        const uint32_t u12 = mulm(subm(r[0], r[1] % p1, p1), e->inv21, p1);
        const u128 P = (u128)p1 * p2;
        const u128 x = (u128)r[1] + (u128)u12 * p2;
        return (x > P / 2) ? (i128)x - (i128)P : (i128)x;
    } else {
        const uint32_t p1 = e->p[0], p2 = e->p[1], p3 = e->p[2];
        const uint32_t u13 = mulm(subm(r[0], r[2] % p1, p1), e->inv31, p1);
        const uint32_t u23 = mulm(subm(r[1], r[2] % p2, p2), e->inv32, p2);
        const uint32_t u123 = mulm(subm(u13, u23 % p1, p1), e->inv21, p1);
        const u128 P = (u128)p1 * p2 * p3;
        const u128 x = (u128)u123 * ((u128)p2 * p3) + (u128)u23 * p3 + r[2];
        return (x > P / 2) ? (i128)x - (i128)P : (i128)x;
    }
}

/* Full carry with the negacyclic wrap (baseMod, src/transformCPUi32.h:743-772):
 * truncated division as reduce96/reduce64 do (remainder has the sign of f). */
static void carry_full(int32_t *d, const i128 *x, size_t N, uint32_t b) {
    i128 f = 0;
    for (size_t i = 0; i < N; ++i) {
        f += x[i];
        d[i] = (int32_t)(f % (i128)b);
        f /= (i128)b;
    }
    while (f != 0) {
        f = -f; /* f x^N = -f */
        for (size_t i = 0; i < N; ++i) {
            f += d[i];
            d[i] = (int32_t)(f % (i128)b);
            f /= (i128)b;
            if (f == 0) return;
        }
    }
}

static void transform_product(gfo_engine *e, int32_t *dst, const int32_t *src, int use_zp, int dup) {
    const size_t N = e->N;
    i128 *x = (i128 *)malloc(N * sizeof(i128));
    for (int k = 0; k < e->rns; ++k) {
        uint32_t *z = e->z + (size_t)k * N;
        for (size_t i = 0; i < N; ++i) z[i] = to_res(src[i], e->p[k]);
        forward(e, k, z);
        pairs(e, k, z, use_zp ? e->zp + (size_t)k * N : NULL);
        backward(e, k, z);
        for (size_t i = 0; i < N; ++i) z[i] = mulm(z[i], e->norm[k], e->p[k]);
    }
    for (size_t i = 0; i < N; ++i) {
        uint32_t r[3] = {0, 0, 0};
        for (int k = 0; k < e->rns; ++k) r[k] = e->z[(size_t)k * N + i];
        x[i] = garner(e, r);
        if (dup) x[i] += x[i];
    }
    carry_full(dst, x, N, e->b);
    free(x);
}

void gfo_set(gfo_engine *e, uint32_t a) {
    memset(e->regs, 0, e->N * sizeof(int32_t));
    e->regs[0] = (int32_t)a;
}

void gfo_square_dup(gfo_engine *e, int dup) { transform_product(e, e->regs, e->regs, 0, dup); }

/* squareMul = squareDup(false) then baseModMul (src/transformCPUi32.h:774-798, 897-901). */
void gfo_square_mul(gfo_engine *e, int32_t a) {
    gfo_square_dup(e, 0);
    const size_t N = e->N;
    i128 *x = (i128 *)malloc(N * sizeof(i128));
    for (size_t i = 0; i < N; ++i) x[i] = (i128)e->regs[i] * a;
    carry_full(e->regs, x, N, e->b);
    free(x);
}

void gfo_init_multiplicand(gfo_engine *e, int src) {
    const size_t N = e->N;
    const int32_t *d = e->regs + (size_t)src * N;
    for (int k = 0; k < e->rns; ++k) {
        uint32_t *z = e->zp + (size_t)k * N;
        for (size_t i = 0; i < N; ++i) z[i] = to_res(d[i], e->p[k]);
        forward(e, k, z);
    }
}

void gfo_mul(gfo_engine *e) { transform_product(e, e->regs, e->regs, 1, 0); }

void gfo_copy(gfo_engine *e, int dst, int src) {
    memmove(e->regs + (size_t)dst * e->N, e->regs + (size_t)src * e->N, e->N * sizeof(int32_t));
}

void gfo_get_digits(const gfo_engine *e, int reg, int32_t *out) {
    memcpy(out, e->regs + (size_t)reg * e->N, e->N * sizeof(int32_t));
}

void gfo_set_digits(gfo_engine *e, int reg, const int32_t *in) {
    memcpy(e->regs + (size_t)reg * e->N, in, e->N * sizeof(int32_t));
}

void gfo_schoolbook_mul(const int32_t *a, const int32_t *c, int32_t *out, uint32_t n, uint32_t b) {
    const size_t N = (size_t)1 << n;
    i128 *x = (i128 *)calloc(N, sizeof(i128));
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            const i128 t = (i128)a[i] * c[j];
            if (i + j < N) x[i + j] += t; else x[i + j - N] -= t;
        }
    carry_full(out, x, N, b);
    free(x);
}

/* gint::unbalance, src/gint.h:40-102. */
void gfo_unbalance(int32_t *d, uint32_t n, uint32_t b) {
    const size_t N = (size_t)1 << n;
    const int32_t base = (int32_t)b;
    int64_t f = 0;
    for (size_t i = 0; i < N; ++i) {
        f += d[i];
        int32_t r = (int32_t)(f % base);
        if (r < 0) r += base;
        d[i] = r;
        f -= r;
        f /= base;
    }
    while (f != 0) {
        f = -f;
        for (size_t i = 0; i < N; ++i) {
            f += d[i];
            int32_t r = (int32_t)(f % base);
            if (r < 0) r += base;
            d[i] = r;
            f -= r;
            f /= base;
            if (f == 0) break;
        }
        if (f == 1) {
            int minus_one = 1;
            for (size_t i = 0; i < N; ++i) if (d[i] != 0) { minus_one = 0; break; }
            if (minus_one) { d[0] = -1; break; }
        }
    }
}

/* gint::balance, src/gint.h:104-149. */
void gfo_balance(int32_t *d, uint32_t n, uint32_t b) {
    const size_t N = (size_t)1 << n;
    const int32_t base = (int32_t)b;
    int64_t f = 0;
    for (size_t i = 0; i < N; ++i) {
        f += d[i];
        int32_t r = (int32_t)(f % base);
        if (r > base / 2) r -= base;
        if (r <= -base / 2) r += base;
        d[i] = r;
        f -= r;
        f /= base;
    }
    while (f != 0) {
        f = -f;
        for (size_t i = 0; i < N; ++i) {
            f += d[i];
            int32_t r = (int32_t)(f % base);
            if (r > base / 2) r -= base;
            if (r <= -base / 2) r += base;
            d[i] = r;
            f -= r;
            f /= base;
            if (f == 0) break;
        }
    }
}

/* gint::isOne, src/gint.h:169-202 (non-CYCLO): RES64 = canonical value mod 2^64,
 * old64 = low bytes of the 8 top digits. */
int gfo_is_one(const int32_t *d, uint32_t n, uint32_t b, uint64_t *res64, uint64_t *old64) {
    const size_t N = (size_t)1 << n;
    int one = (d[0] == 1);
    if (one) for (size_t i = 1; i < N; ++i) one &= (d[i] == 0);
    uint64_t r64 = 0, bb = 1;
    for (size_t i = 0; i < N; ++i) { r64 += (uint64_t)(uint32_t)d[i] * bb; bb *= b; }
    *res64 = r64;
    uint64_t old = 0;
    for (size_t i = 8; i != 0; --i) old = (old << 8) | (uint8_t)d[N - i];
    *old64 = old;
    return one;
}

static uint64_t rotl64(uint64_t x, unsigned r) { return (x << r) | (x >> ((-r) & 63)); }

/* gint::gethash64 / gethash32, src/gint.h:204-225. */
uint64_t gfo_hash64(const int32_t *d, uint32_t n) {
    const size_t N = (size_t)1 << n;
    uint64_t h = 0;
    for (size_t i = 0; i < N; ++i) {
        const uint32_t a = (uint32_t)d[i];
        h += a;
        h ^= rotl64(a + 0xc39d8a0552b073e8ull, (unsigned)((17 * (uint64_t)a + 5) % 64));
    }
    return h;
}

uint32_t gfo_hash32(const int32_t *d, uint32_t n) {
    const uint64_t h = gfo_hash64(d, n);
    const uint32_t v = (uint32_t)h ^ (uint32_t)(h >> 32);
    return v < 2 ? 2 : v;
}
