// gfe_engine.hip -- host side of the gfx950 squaring engine and its C ABI (include/gfe.h).
//
// Replaces the reference's engines<RNS, is32> / transformGPUs<RNS, is32> (src/transformGPU.h:
// 210-1195) and create_ocl (src/transform_ocl.cpp:20-52).  Device state of one engine:
//   regs  : num_regs x N int32 digits (register 0 = r0), balanced, |d| <= b/2 + 4
//   ws    : rns x N uint32 residues (transform-domain work space of one squaring)
//   zp    : rns x N uint32 residues (multiplicand, transform domain)
//   carry : N / 8 int64: workgroup carries of the normalisation, or segment carries of the fused
//           inverse column pass
//   tw    : rns x N uint32 twiddles (Montgomery form)
// One squaring = k_col_fwd, k_row, k_col_inv, k_norm, k_carry_wg on the engine's stream, or with
// a fused plan k_col_fwd_all, k_row, k_col_inv_norm, k_fold.

#include "gfe_engine.h"
#include "gfe_limb.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gfe.h"

namespace gfe {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

// ------------------------------------------------------------------------------------------------
// Single-engine kernels of the small passes

struct CarryArgs {
    int32_t *dst;
    int64_t *carry;
    uint32_t b;
    unsigned nwg;   // workgroups of the carry pass (each 256 x 8 digits)
};

__global__ void __launch_bounds__(256) k_carry_wg(CarryArgs A) {
    const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < A.nwg) carry_wg_body(A.dst, A.carry, A.nwg, A.b, w);
}

struct LimbNormArgs {
    const uint32_t *ws;     // [4][N/2] residues of the limb transform
    int32_t *dst;           // register 0
    int64_t *wl;            // 2 x workgroups: boundary carries
    const LimbCrt *crt;
    size_t M;
    int dup;
};

__global__ void __launch_bounds__(256) k_norm_limb(LimbNormArgs A) {
    norm_limb_body(*A.crt, A.ws, A.M, A.dst, A.wl, A.dup != 0, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_carry_limb(int32_t *dst, int64_t *wl, unsigned nwg, uint32_t b) {
    const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nwg) carry_limb_body(dst, wl, nwg, b, w);
}

__global__ void __launch_bounds__(256) k_carry_limb_slots(int32_t *dst, const int4 *slots, unsigned nrange, int rlog,
                                                          uint32_t b) {
    const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nrange) carry_limb_slot_body(dst, slots, rlog, b, w);
}

struct FoldArgs {
    int32_t *dst;
    const int64_t *seg;
    size_t nseg;
    int ng;
    uint32_t b;
};

__global__ void __launch_bounds__(256) k_fold(FoldArgs A) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < A.nseg) fold_body(A.dst, A.seg, A.nseg, A.ng, A.b, j);
}

struct MulScalarArgs {
    int32_t *dst;
    int64_t *carry;
    uint32_t b;
    int32_t a;
};

__global__ void __launch_bounds__(256) k_mulscalar(MulScalarArgs A) { mulscalar_body(A.dst, A.carry, A.b, A.a, blockIdx.x); }

// Register copy, 16 bytes per thread (the reference's copy kernel, ocl/kernel.cl:2501-2506).
__global__ void __launch_bounds__(256) k_copy(int4 *__restrict__ dst, const int4 *__restrict__ src, size_t n4) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}

// r = a (digit 0 = a, others 0).
__global__ void __launch_bounds__(256) k_set(int32_t *dst, size_t N, int32_t a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = (i == 0) ? a : 0;
}


// ------------------------------------------------------------------------------------------------
// Host modular arithmetic

static const uint32_t k_primes[6] = {127u * (1u << 24) + 1u, 63u * (1u << 25) + 1u, 15u * (1u << 27) + 1u,
                                     125u * (1u << 25) + 1u, 243u * (1u << 24) + 1u, 235u * (1u << 24) + 1u};
static const uint32_t k_gen[6] = {3, 5, 31, 3, 7, 3};

// Prime k of a set: 0 = the 31-bit primes, 1 = the 32-bit primes, LIMB_SET = the four primes of the
// two-digit-limb transform (the 31-bit primes and LIMB_P4).
constexpr int LIMB_SET = 2;
uint32_t prime_of(int is32, int k) {
    if (is32 == LIMB_SET) return k < 3 ? k_primes[k] : LIMB_P4;
    return k_primes[(is32 ? 3 : 0) + k];
}
static uint32_t gen_of(int set, int k) {
    if (set == LIMB_SET) return k < 3 ? k_gen[k] : LIMB_G4;
    return k_gen[(set ? 3 : 0) + k];
}

static uint32_t hmul(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)(((uint64_t)a * b) % p); }
static uint32_t hpow(uint32_t a, uint64_t e, uint32_t p) {
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = hmul(r, a, p);
        a = hmul(a, a, p);
        e >>= 1;
    }
    return r;
}
static uint32_t hinv(uint32_t a, uint32_t p) { return hpow(a % p, p - 2, p); }
static uint32_t pinv32(uint32_t p) {
    uint32_t q = 1;
    for (int i = 0; i < 5; ++i) q *= 2u - p * q;
    return q;
}

// Operand of ModS::mulw / ModU::mulw for the constant w: its Montgomery form w 2^32 mod p.
uint32_t mulw_const(uint32_t w, uint32_t p) { return (uint32_t)((((uint64_t)w) << 32) % p); }

uint32_t psi_mont(int is32, int k, uint32_t n) {
    const uint32_t p = prime_of(is32, k), g = gen_of(is32, k);
    return mulw_const(hpow(g, (p - 1) >> (n + 1), p), p);
}

uint32_t prime_qinv(uint32_t p) { return pinv32(p); }

// create_ocl's rule, src/transform_ocl.cpp:35-50.
void select_primes(uint32_t b, uint32_t n, int *rns, int *is32) {
    const uint64_t bb = (uint64_t)b * b, N = (uint64_t)1 << n;
    if (bb < ((uint64_t)k_primes[0] * k_primes[1] / 2) / N) { *rns = 2; *is32 = 0; }
    else if (bb < ((uint64_t)k_primes[3] * k_primes[4] / 2) / N) { *rns = 2; *is32 = 1; }
    else if (b <= 1000000000u) { *rns = 3; *is32 = 0; }
    else { *rns = 3; *is32 = 1; }
}

// The reference takes the 32-bit primes for NTT3 when b > 10^9 because its digits may reach b
// (src/transform_ocl.cpp:42-50).  Registers here hold balanced digits, |d| <= b/2 + 4 <= 10^9 + 4
// (gfe_set_digits enforces it on input), so every digit is a residue of the 31-bit primes
// (p >= 2,013,265,921) and the squared (or doubled) coefficient, at most N (b/2 + 4)^2 2 <= 2^84
// for n <= 23, stays below half their product (2^91.9): NTT3 computes in the 31-bit primes for
// every b <= 2 10^9, whose lazy reductions make the passes cheaper.  The reference's set still
// decides what gfe_select reports and the residues of the context payload.
int transform_is32(int rns, int is32) { return rns == 3 ? 0 : is32; }

// Host Montgomery product a b 2^-32 mod p (a, b < p; q = p^-1 mod 2^32): t - m p has a zero low
// word, so the difference of the high words is exact.
static inline uint32_t hmont(uint32_t a, uint32_t b, uint32_t p, uint32_t q) {
    const uint64_t t = (uint64_t)a * b;
    const uint32_t m = (uint32_t)t * q;
    const int64_t r = (int64_t)(t >> 32) - (int64_t)(((uint64_t)m * p) >> 32);
    return (uint32_t)(r < 0 ? r + p : r);
}

// Twiddles Psi[j] = psi^bitrev_n(j), j in [0, N), psi = g^((p-1)/2N), in mulw form (Psi[0] = 1 is
// the j = 0 factor of factored row twiddles).  The powers psi^e 2^32 mod p follow by Montgomery
// products with psi 2^32 mod p, and e's bit reversal by a reversed increment.
static void make_twiddles(int is32, int k, uint32_t n, uint32_t *out) {
    const uint32_t p = prime_of(is32, k), g = gen_of(is32, k), q = pinv32(p);
    const size_t N = (size_t)1 << n;
    const uint32_t psim = mulw_const(hpow(g, (p - 1) >> (n + 1), p), p);
    uint32_t w = mulw_const(1, p);
    size_t r = 0;
    for (size_t e = 0; e < N; ++e) {
        out[r] = w;
        w = hmont(w, psim, p, q);
        size_t bit = N >> 1;
        while (r & bit) { r ^= bit; bit >>= 1; }
        r |= bit;
    }
}

// Tables of `count` (set, prime) pairs, table i at out + i N, one host thread per table.  When a
// thread cannot be started (std::system_error on thread or process limits) the tables not handed
// to a thread are built on the calling thread instead; `max_threads` = 0 builds all of them here.
void make_twiddle_tables(const int *sets, const int *primes, int count, uint32_t n, uint32_t *out, int max_threads) {
    const size_t N = (size_t)1 << n;
    std::vector<std::thread> th;
    int started = 0;
    try {
        th.reserve((size_t)count);
        for (; started < count && started < max_threads; ++started)
            th.emplace_back(make_twiddles, sets[started], primes[started], n, out + (size_t)started * N);
    } catch (const std::exception &) {
        // `started` counts the threads that did start; the rest are built below
    }
    for (int i = started; i < count; ++i) make_twiddles(sets[i], primes[i], n, out + (size_t)i * N);
    for (auto &t : th) t.join();
}

void set_lds_attr(const void *fn) {
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

// Carry rounds of the fused inverse column pass: the first balanced division leaves quotients
// |q| <= Q1 = (vmax + b/2) / b, with vmax = 2 N (b/2 + 4)^2 the largest coefficient of a doubled
// square; a round turns |q| <= Q into |q| <= 1 + floor(Q / b).  Returns the rounds until |q| <= 1.
int carry_rounds(uint32_t b, uint32_t n) {
    const long double bb = (long double)b;
    const long double half = bb / 2 + 4;
    const long double vmax = 2.0L * (long double)(1ull << n) * half * half;
    long double Q = floorl((vmax + bb / 2) / bb) + 1;
    int r = 1;
    while (Q > 1) {
        Q = 1 + floorl(Q / bb);
        ++r;
    }
    return r;
}

CrtConst make_crt(int rns, int is32, uint32_t n, uint32_t b) {
    CrtConst c;
    memset(&c, 0, sizeof(c));
    for (int k = 0; k < 3; ++k) {
        const uint32_t p = prime_of(is32, k);
        c.p[k] = p;
        c.q[k] = pinv32(p);
        // output of the inverse transform = N 2^-32 (product); multiply by 2^32 / N
        const uint32_t s = hmul((uint32_t)((((uint64_t)1) << 32) % p), hinv((uint32_t)(((uint64_t)1 << n) % p), p), p);
        c.scale[k] = mulw_const(s, p);
    }
    const uint32_t p1 = c.p[0], p2 = c.p[1], p3 = c.p[2];
    c.inv_p1_m2 = mulw_const(hinv(p1 % p2, p2), p2);
    c.p1_m3 = mulw_const(p1 % p3, p3);
    c.inv_p12_m3 = mulw_const(hinv(hmul(p1 % p3, p2 % p3, p3), p3), p3);
    c.p12 = (uint64_t)p1 * p2;
    c.p12_d = (double)c.p12;
    for (int k = 0; k < 3; ++k) c.scale_q[k] = c.scale[k] * c.q[k];
    c.inv_p1_m2_q = c.inv_p1_m2 * c.q[1];
    c.p1_m3_q = c.p1_m3 * c.q[2];
    c.inv_p12_m3_q = c.inv_p12_m3 * c.q[2];
    c.p12_lo = (uint32_t)c.p12;
    c.p1_d = (double)p1;
    // one rounded quotient is within 1 of the true one when N b < 2^49 (crt_digit_fast)
    c.fast = (b != 0 && ((uint64_t)b << n) < (1ull << 49)) ? 1 : 0;
    (void)rns;
    return c;
}

// Balanced base-B piece of x: returns r in (-B/2, B/2] with x = r (mod B) and leaves (x - r) / B in x.
static int64_t bal_piece(__int128 &x, __int128 B) {
    __int128 r = x % B;
    if (r > B / 2) r -= B;
    else if (r <= -(B / 2)) r += B;
    x = (x - r) / B;
    return (int64_t)r;
}

// Constants of the limb transform's CRT and carry (gfe_limb.h) for (b, n): transform length
// M = 2^(n-1).
LimbCrt make_limb_crt(uint32_t n, uint32_t b) {
    LimbCrt c;
    memset(&c, 0, sizeof(c));
    __int128 P = 1;
    for (int k = 0; k < 4; ++k) P *= prime_of(LIMB_SET, k);
    const __int128 B = (__int128)b * b;
    for (int k = 0; k < 4; ++k) {
        const uint32_t p = prime_of(LIMB_SET, k);
        c.p[k] = p;
        c.q[k] = pinv32(p);
        const __int128 Q = P / p;
        const uint32_t u = hinv((uint32_t)(Q % p), p);  // (P/p)^-1 mod p
        // raw inverse-transform output = M 2^-32 (coefficient); times 2^32 M^-1 (P/p)^-1
        const uint32_t s = hmul((uint32_t)((((uint64_t)1) << 32) % p), hinv((uint32_t)(((uint64_t)1 << (n - 1)) % p), p), p);
        c.zc[k] = mulw_const(hmul(s, u, p), p);
        c.zc_q[k] = c.zc[k] * c.q[k];
        c.inv_p[k] = 1.0 / (double)p;
        __int128 x = Q;
        c.w0[k] = bal_piece(x, B);
        c.w1[k] = bal_piece(x, B);
        c.w2[k] = (int64_t)x;
        c.w0d[k] = (double)c.w0[k];
        c.w1d[k] = (double)c.w1[k];
        c.w0l[k] = (uint32_t)(uint64_t)c.w0[k];
        c.w0h[k] = (uint32_t)((uint64_t)c.w0[k] >> 32);
        c.w1l[k] = (uint32_t)(uint64_t)c.w1[k];
        c.w1h[k] = (uint32_t)((uint64_t)c.w1[k] >> 32);
        c.w2s[k] = (int32_t)c.w2[k];  // |w2| <= 2 for b >= 2^23 (limb_ok)
    }
    __int128 x = P;
    c.P0 = bal_piece(x, B);
    c.P1 = bal_piece(x, B);
    c.P2 = (int64_t)x;
    c.P0d = (double)c.P0;
    c.P1d = (double)c.P1;
    c.B = (uint64_t)B;
    c.invB = 1.0 / (double)c.B;
    c.b = b;
    c.binv = 1.0 / (double)b;
    return c;
}

// ------------------------------------------------------------------------------------------------
// Plans

static Plan runtime_default_plan(uint32_t n);

// Default plan: the first admissible compile-time plan of a preference list drawn from timings on
// MI355X (n = 22: 2^8-point columns of 16 points per thread, 2^14-point rows of 32 points per
// thread; n = 15: 2^6 x 2^9 with 8 and 16 points per thread), else the runtime-geometry default.
Plan default_plan(uint32_t n) {
    const int nn = (int)n;
    std::vector<int> as;
    auto add_a = [&](int a) {
        if (a >= SPEC_COL_MIN && a <= SPEC_COL_MAX && std::find(as.begin(), as.end(), a) == as.end()) as.push_back(a);
    };
    if (nn >= 20) add_a(nn - 14);
    add_a(nn / 2 - 1);
    add_a(nn / 2);
    add_a((nn + 1) / 2);
    for (int a = SPEC_COL_MIN; a <= SPEC_COL_MAX; ++a) add_a(a);
    const int cols[][2] = {{4, 16}, {3, 8}, {3, 16}, {4, 32}, {5, 16}, {3, 32}, {4, 8}};
    const int rows[][2] = {{5, 1}, {4, 2}, {4, 1}, {4, 4}, {4, 8}, {4, 16}, {5, 2}, {4, 32}, {4, 64}};
    for (int a : as) {
        const bool big_rows = (nn - a) >= 13;
        for (const auto &r : rows) {
            if (!big_rows && r[0] == 5) continue;
            for (const auto &c : cols) {
                Plan P;
                P.a = a; P.col_loge = c[0]; P.col_ng = c[1]; P.row_loge = r[0]; P.row_ng = r[1]; P.norm_s = 8;
                P.spec = 1;
                if (a <= 6 && c[0] == 4) continue;  // short columns: 8 points per thread
                if (plan_valid(P, n)) return P;
            }
        }
    }
    return runtime_default_plan(n);
}

// Batches of small candidates square each candidate in one workgroup (one launch per block of
// rounds): on MI355X that is 1.4x-1.9x the four-step passes' throughput at n = 12-13 once the
// batch holds 64 or more candidates, while a lone candidate is faster spread over the four-step
// passes' many workgroups.
Plan default_batch_plan(uint32_t n, uint32_t capacity) {
    Plan P = default_plan(n);
    if (capacity >= 32) {
        Plan F = P;
        F.fuse_sq = 1;
        if (plan_valid(F, n)) return F;
    }
    return P;
}

bool limb_ok(uint32_t b, uint32_t n) {
    int rns = 0, is32 = 0;
    select_primes(b, n, &rns, &is32);
    if (rns != 3 || n < 12 || n > 23 || b < (1u << 23)) return false;
    unsigned __int128 P = 1;
    for (int k = 0; k < 4; ++k) P *= prime_of(LIMB_SET, k);
    const uint64_t lmax = ((uint64_t)(b / 2) + 4) * ((uint64_t)b + 1);
    const unsigned __int128 l2 = (unsigned __int128)lmax * lmax;
    return l2 < (P / 4) >> n;
}

// The limb transform's plans are the compile-time plans of its length N/2 = 2^(n-1).
Plan default_limb_plan(uint32_t n) {
    Plan P = default_plan(n - 1);
    P.limb = 1;
    if (!plan_valid(P, n)) {
        P = Plan();
        return P;  // limb = 0: no limb plan for this n
    }
    return P;
}

// Fused plans leave the range carries of the last normalisation in the first two limbs of each
// range of ng limbs (col_fwd_pass): limb 0's digits take (u0, u1) with |u0|, |u1| <= b/2, limb 1's
// low digit u2 + e1 with |u2| <= 2 and |e1| <= h2max = P / (4 B^2) + 1.  A coefficient of the
// doubled product is at most 2 sum_i |l_i|^2 (Cauchy-Schwarz), which over M / ng ranges is
// 2 (M / ng) (L0^2 + L1^2 + (ng - 2) L^2) with L the balanced limb, L0 = (b + 4)(b + 1) and
// L1 = L + h2max + 2; the explicit CRT needs it below P / 4.
bool limb_fused_ok(uint32_t b, uint32_t n, int ng) {
    if (!limb_ok(b, n) || ng < 8) return false;
    unsigned __int128 P = 1;
    for (int k = 0; k < 4; ++k) P *= prime_of(LIMB_SET, k);
    const unsigned __int128 B = (unsigned __int128)b * b;
    const unsigned __int128 h2max = P / (4 * B * B) + 1;
    const unsigned __int128 L = ((unsigned __int128)(b / 2) + 4) * ((unsigned __int128)b + 1);
    const unsigned __int128 L0 = ((unsigned __int128)b + 4) * ((unsigned __int128)b + 1);
    const unsigned __int128 L1 = L + h2max + 2;
    const unsigned __int128 per_range = L0 * L0 + L1 * L1 + (unsigned __int128)(ng - 2) * L * L;
    const unsigned __int128 ranges = ((unsigned __int128)1 << (n - 1)) / (unsigned)ng;
    return 2 * ranges * per_range < P / 4;
}

// Device bytes of an engine: registers, work space, multiplicand and twiddles of rns primes, the
// carry buffer (N bytes) and, with limb_tables, the limb transform's twiddles (4 x N/2 words).
size_t engine_bytes(int rns, size_t N, uint32_t num_regs, bool limb_tables) {
    return (size_t)num_regs * N * 4 + 3 * (size_t)rns * N * 4 + N + (limb_tables ? 8 * N : 0);
}

// An engine's first plan: the limb transform where it is exact (with the inverse column pass and
// normalisation fused where that bound holds), else the digit transform's default.
Plan engine_plan_for(uint32_t b, uint32_t n) {
    Plan P = default_plan(n);
    if (!limb_ok(b, n)) return P;
    const Plan L = default_limb_plan(n);
    if (!L.limb) return P;
    Plan F = L;
    F.fused = 1;
    return (plan_valid(F, n) && limb_fused_ok(b, n, F.col_ng)) ? F : L;
}

std::vector<Plan> candidate_limb_plans(uint32_t n) {
    std::vector<Plan> out;
    if (n < 12) return out;
    for (Plan P : candidate_plans(n - 1)) {
        if (!P.spec || P.fuse_sq) continue;
        P.limb = 1;
        if (plan_valid(P, n)) out.push_back(P);
        P.fused = 1;
        if (plan_valid(P, n)) out.push_back(P);
    }
    return out;
}

static Plan runtime_default_plan(uint32_t n) {
    Plan P;
    P.a = (int)n / 2;
    if (n >= 22) P.a = (int)n - 12;
    P.col_loge = 4;
    P.row_loge = 4;
    const int E = 1 << P.col_loge;
    const int G = 1 << P.a;
    P.col_ng = std::min(16, std::max(1, 1024 * E / G));
    while (P.col_ng * G / E < 64 && P.col_ng < (1 << (n - P.a))) P.col_ng *= 2;
    const int Gr = 1 << (n - P.a), Er = 1 << P.row_loge;
    const int R = 1 << P.a;
    P.row_ng = std::min(R, std::max(1, 256 / (Gr / Er)));
    P.norm_s = 8;
    return P;
}

bool plan_valid(const Plan &P, uint32_t n) {
    if (P.limb) {  // compile-time passes of the N/2-point transform
        if (!P.spec || P.fuse_sq || P.pipe || P.row_tw || n < 12) return false;
        Plan Q = P;
        Q.limb = 0;
        Q.fused = 0;
        if (!plan_valid(Q, n - 1)) return false;
        return !P.fused || limb_col_norm_ok(P.a, P.col_loge, P.col_ng);
    }
    if (P.fuse_sq) {
        if ((int)n < FUSED_MIN_N || (int)n > FUSED_MAX_N || P.fused || P.pipe) return false;
        Plan Q = P;
        Q.fuse_sq = 0;
        return plan_valid(Q, n);
    }
    const int gc = P.a, gr = (int)n - P.a;
    if (P.spec) {
        if (P.fused || P.pipe || P.row_tw) return false;
        if (!spec_col_ok(gc, P.col_loge, P.col_ng) || !spec_row_ok(gr, P.row_loge)) return false;
        if (P.col_ng > (1 << gr) || P.row_ng > (1 << gc)) return false;
        const int Tc = P.col_ng * (1 << gc) / (1 << P.col_loge);
        const int Tr = P.row_ng * (1 << gr) / (1 << P.row_loge);
        const int maxc = P.col_loge >= 5 ? 512 : 1024, maxr = P.row_loge >= 5 ? 512 : 1024;
        if (Tc < 64 || Tc > maxc || Tr < 64 || Tr > maxr) return false;
        // the ROW_FWD / ROW_MUL passes of the multiplicand keep the runtime-geometry row kernel
        const size_t lr_rt = (size_t)P.row_ng * ((1 << gr) + ((1 << gr) >> P.row_loge)) * 4;
        if (col_pass_lds_words(gc, P.col_loge, P.col_ng) * 4 > 160 * 1024 ||
            row_pass_lds_words(gr, P.row_loge, P.row_ng) * 4 > 160 * 1024 || lr_rt > 160 * 1024)
            return false;
        return true;
    }
    if (P.pipe && P.col_ng * ((1 << gc) + ((1 << gc) >> P.col_loge)) * 4 + (4 << gc) > 160 * 1024) return false;
    if (P.fused && (P.col_ng < 8 || P.col_ng > 64 || P.col_loge > 4 ||
                    P.col_ng * (1 << gc) / (1 << P.col_loge) > 512)) return false;
    if (gc < P.col_loge || gr < P.row_loge) return false;
    const int Tc = P.col_ng * (1 << gc) / (1 << P.col_loge);
    const int Tr = P.row_ng * (1 << gr) / (1 << P.row_loge);
    if (Tc < 32 || Tc > 1024 || Tr < 32 || Tr > 1024) return false;
    if (P.col_loge >= 5 && Tc > 512) return false;
    if (P.col_ng > (1 << gr) || P.row_ng > (1 << gc)) return false;
    const size_t lc = (size_t)P.col_ng * ((1 << gc) + ((1 << gc) >> P.col_loge)) * 4;
    const size_t lr = (size_t)P.row_ng * ((1 << gr) + ((1 << gr) >> P.row_loge)) * 4;
    if (lc > 160 * 1024 || lr > 160 * 1024) return false;
    return true;
}

// Plans the autotuners time: column levels a in [4, n-4], 8..32 points per thread, 1..64
// columns or rows per workgroup, 64..1024 threads, table or factored row twiddles, separate or
// fused (all-prime) column passes.
// Workgroups per group of a persistent launch: each takes ceil(nblk / grid) blocks, and groups x
// grid stays within the resident workgroups whenever nblk allows.
unsigned persistent_grid(unsigned nblk, unsigned groups, unsigned resident) {
    const unsigned per = std::max(1u, resident / std::max(1u, groups));
    const unsigned u = (nblk + per - 1) / per;
    return (nblk + u - 1) / u;
}

std::vector<Plan> candidate_plans(uint32_t n) {
    std::vector<Plan> plans;
    for (int a = 4; a <= (int)n - 4; ++a)
        for (int cl = 3; cl <= 5; ++cl)
            for (int cng = 1; cng <= 64; cng *= 2)
                for (int rl = 3; rl <= 5; ++rl)
                    for (int rng = 1; rng <= 64; rng *= 2)
                        for (int f = 0; f < 4; ++f)
                            for (int tw = 0; tw < 2; ++tw) {
                                if (f == 3 && tw) continue;
                                Plan P;
                                P.a = a; P.col_loge = cl; P.col_ng = cng; P.row_loge = rl; P.row_ng = rng;
                                P.norm_s = 8; P.fused = (f == 1); P.pipe = (f == 2); P.spec = (f == 3); P.row_tw = tw;
                                const int Tc = cng * (1 << a) / (1 << cl), Tr = rng * (1 << (n - a)) / (1 << rl);
                                if (Tc < 64 || Tr < 64 || Tc > 1024 || Tr > 1024) continue;
                                if (plan_valid(P, n)) plans.push_back(P);
                                if (P.spec && (int)n >= FUSED_MIN_N && (int)n <= FUSED_MAX_N) {
                                    Plan F = P;
                                    F.fuse_sq = 1;
                                    if (plan_valid(F, n)) plans.push_back(F);
                                }
                            }
    return plans;
}

void plan_to_array(const Plan &P, int32_t out[6]) {
    out[0] = P.a; out[1] = P.col_loge; out[2] = P.col_ng; out[3] = P.row_loge; out[4] = P.row_ng;
    out[5] = P.row_tw | (P.fused << 1) | (P.pipe << 2) | (P.spec << 3) | (P.fuse_sq << 4) | (P.limb << 5);
}

bool plan_from_array(const int32_t in[6], uint32_t n, Plan *out) {
    Plan P;
    const int fl = in[5] & 15;
    const bool limb_fused = (in[5] & 32) && fl == 10;  // the limb transform's fused inverse columns
    if (in[5] < 0 || in[5] > 63 || (fl & 6) == 6 || ((fl & 8) && fl != 8 && !limb_fused)) return false;
    P.a = in[0]; P.col_loge = in[1]; P.col_ng = in[2]; P.row_loge = in[3]; P.row_ng = in[4];
    P.row_tw = fl & 1; P.fused = (fl >> 1) & 1; P.pipe = (fl >> 2) & 1; P.spec = (fl >> 3) & 1;
    P.fuse_sq = (in[5] >> 4) & 1;
    P.limb = (in[5] >> 5) & 1;
    P.norm_s = 8;
    auto pow2 = [](int v) { return v >= 1 && (v & (v - 1)) == 0; };
    if (P.a < 1 || P.a >= (int)n - P.limb || P.col_loge < 3 || P.col_loge > 5 || P.row_loge < 3 || P.row_loge > 5) return false;
    if (!pow2(P.col_ng) || !pow2(P.row_ng)) return false;
    if (!plan_valid(P, n)) return false;
    *out = P;
    return true;
}

std::string plan_str(const Plan &P, uint32_t n) {
    std::ostringstream ss;
    const uint32_t tn = n - (uint32_t)P.limb;
    if (P.limb) ss << "limbs of two digits (2^" << tn << " limbs, four 31-bit primes); ";
    ss << "cols 2^" << P.a << " (E=" << (1 << P.col_loge) << ", " << P.col_ng << "/wg) x rows 2^" << (tn - P.a)
       << " (E=" << (1 << P.row_loge) << ", " << P.row_ng << "/wg" << (P.row_tw ? ", factored twiddles" : "")
       << "), "
       << (P.limb && P.fused ? "inverse columns fused with the normalisation (ranges of " + std::to_string(P.col_ng) +
                                   " limbs, " + std::to_string((1 << P.col_loge) / 4) + " limbs/thread)"
           : P.fused       ? std::string("all-prime columns with CRT + carry")
           : P.limb        ? std::string("norm 4 limbs/thread")
                           : std::string("norm 8 digits/thread"))
       << (P.pipe ? ", pipelined columns and norm" : "") << (P.spec ? ", compile-time passes" : "")
       << (P.fuse_sq ? ", squaring in one workgroup" : "");
    return ss.str();
}

// ------------------------------------------------------------------------------------------------
// Kernel dispatch

typedef void (*ColFn)(ColArgs);
typedef void (*RowFn)(RowArgs);
typedef void (*NormFn)(NormArgs);

template <class M> static ColFn col_fwd_fn(int loge) {
    switch (loge) {
        case 3: return k_col_fwd<M, 3>;
        case 4: return k_col_fwd<M, 4>;
        default: return k_col_fwd<M, 5>;
    }
}
template <class M> static ColFn col_inv_fn(int loge) {
    switch (loge) {
        case 3: return k_col_inv<M, 3>;
        case 4: return k_col_inv<M, 4>;
        default: return k_col_inv<M, 5>;
    }
}
template <class M, int MODE, int TWF> static RowFn row_fn_t(int loge) {
    switch (loge) {
        case 3: return k_row<M, 3, MODE, TWF>;
        case 4: return k_row<M, 4, MODE, TWF>;
        default: return k_row<M, 5, MODE, TWF>;
    }
}
template <class M, int MODE> static RowFn row_fn(int loge, int twf) {
    return twf ? row_fn_t<M, MODE, 1>(loge) : row_fn_t<M, MODE, 0>(loge);
}

typedef void (*ColAllFn)(ColAllArgs);

struct KernelSet {
    ColFn col_fwd, col_inv;
    RowFn row_sq, row_mul, row_fwd;
    NormFn norm;
    ColAllFn col_fwd_all, col_inv_norm;
    ColFn col_fwd_p, col_inv_p;
    NormFn norm_p;
};

template <class M> static void pipe_fns(int loge, ColFn *f, ColFn *i) {
    switch (loge) {
        case 3: *f = k_col_fwd_p<M, 3>; *i = k_col_inv_p<M, 3>; break;
        case 4: *f = k_col_fwd_p<M, 4>; *i = k_col_inv_p<M, 4>; break;
        default: *f = k_col_fwd_p<M, 5>; *i = k_col_inv_p<M, 5>; break;
    }
}

template <int RNS, class M> static void fused_fns(int loge, ColAllFn *f, ColAllFn *i) {
    if (loge == 3) { *f = k_col_fwd_all<RNS, M, 3>; *i = k_col_inv_norm<RNS, M, 3>; }
    else { *f = k_col_fwd_all<RNS, M, 4>; *i = k_col_inv_norm<RNS, M, 4>; }
}

static KernelSet kernels_for(int rns, int is32, const Plan &P) {
    KernelSet K;
    if (is32) {
        K.col_fwd = col_fwd_fn<ModU>(P.col_loge);
        K.col_inv = col_inv_fn<ModU>(P.col_loge);
        K.row_sq = row_fn<ModU, ROW_SQUARE>(P.row_loge, P.row_tw);
        K.row_mul = row_fn<ModU, ROW_MUL>(P.row_loge, P.row_tw);
        K.row_fwd = row_fn<ModU, ROW_FWD>(P.row_loge, P.row_tw);
        K.norm = (rns == 2) ? k_norm<2, ModU, 8> : k_norm<3, ModU, 8>;
        if (rns == 2) fused_fns<2, ModU>(P.col_loge, &K.col_fwd_all, &K.col_inv_norm);
        else fused_fns<3, ModU>(P.col_loge, &K.col_fwd_all, &K.col_inv_norm);
        pipe_fns<ModU>(P.col_loge, &K.col_fwd_p, &K.col_inv_p);
        K.norm_p = (rns == 2) ? k_norm_p<2, ModU, 8> : k_norm_p<3, ModU, 8>;
    } else {
        K.col_fwd = col_fwd_fn<ModS>(P.col_loge);
        K.col_inv = col_inv_fn<ModS>(P.col_loge);
        K.row_sq = row_fn<ModS, ROW_SQUARE>(P.row_loge, P.row_tw);
        K.row_mul = row_fn<ModS, ROW_MUL>(P.row_loge, P.row_tw);
        K.row_fwd = row_fn<ModS, ROW_FWD>(P.row_loge, P.row_tw);
        K.norm = (rns == 2) ? k_norm<2, ModS, 8> : k_norm<3, ModS, 8>;
        if (rns == 2) fused_fns<2, ModS>(P.col_loge, &K.col_fwd_all, &K.col_inv_norm);
        else fused_fns<3, ModS>(P.col_loge, &K.col_fwd_all, &K.col_inv_norm);
        pipe_fns<ModS>(P.col_loge, &K.col_fwd_p, &K.col_inv_p);
        K.norm_p = (rns == 2) ? k_norm_p<2, ModS, 8> : k_norm_p<3, ModS, 8>;
    }
    return K;
}

static std::once_flag g_attr_once;

// Kernels may use up to 160 KiB of LDS (gfx950).
static void set_lds_attributes() {
    std::call_once(g_attr_once, []() {
        const int lds = 160 * 1024;
        const int loges[3] = {3, 4, 5};
        for (int L : loges) for (int twf = 0; twf < 2; ++twf) {
            Plan P;
            P.col_loge = L;
            P.row_loge = L;
            P.row_tw = twf;
            for (int is32 = 0; is32 < 2; ++is32) {
                KernelSet K = kernels_for(3, is32, P);
                (void)hipFuncSetAttribute((const void *)K.col_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                (void)hipFuncSetAttribute((const void *)K.col_inv, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                (void)hipFuncSetAttribute((const void *)K.row_sq, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                (void)hipFuncSetAttribute((const void *)K.row_mul, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                (void)hipFuncSetAttribute((const void *)K.row_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                for (int r = 2; r <= 3; ++r) {
                    const KernelSet F = kernels_for(r, is32, P);
                    (void)hipFuncSetAttribute((const void *)F.col_fwd_all, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                    (void)hipFuncSetAttribute((const void *)F.col_inv_norm, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                }
                (void)hipFuncSetAttribute((const void *)K.col_fwd_p, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                (void)hipFuncSetAttribute((const void *)K.col_inv_p, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            }
        }
        (void)hipGetLastError();
    });
}

void engine_set_lds_attributes() {
    set_lds_attributes();
    spec_set_lds_attributes();
    fused_set_lds_attributes();
}

// ------------------------------------------------------------------------------------------------
// Engine

Engine::~Engine() {
    if (dev_ok) {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(regs); (void)hipFree(ws); (void)hipFree(zp); (void)hipFree(carry);
        (void)hipFree(tw); (void)hipFree(crt);
        if (tw_limb) (void)hipFree(tw_limb);
        if (limb_crt) (void)hipFree(limb_crt);
        if (stream) (void)hipStreamDestroy(stream);
        for (auto &ev : events) (void)hipEventDestroy(ev);
    }
}

int Engine::init(uint32_t b_, uint32_t n_, uint32_t num_regs_, int device_) {
    b = b_; n = n_; num_regs = num_regs_; device = device_;
    N = (size_t)1 << n;
    select_primes(b, n, &rns, &ref_is32);
    is32 = transform_is32(rns, ref_is32);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(GFE_ENODEV, "no HIP device");
    if (device < 0 || device >= ndev) return set_error(GFE_ENODEV, "device index out of range");
    if (hipSetDevice(device) != hipSuccess) return set_error(GFE_EHIP, "hipSetDevice failed");
    set_lds_attributes();
    spec_set_lds_attributes();
    fused_set_lds_attributes();
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return set_error(GFE_EHIP, "stream");
    plan = default_plan(n);
    rounds = carry_rounds(b, n);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 256;
    if (hipMalloc(&regs, (size_t)num_regs * N * 4) != hipSuccess || hipMalloc(&ws, (size_t)rns * N * 4) != hipSuccess ||
        hipMalloc(&zp, (size_t)rns * N * 4) != hipSuccess || hipMalloc(&carry, (N / 8) * 8) != hipSuccess ||

        hipMalloc(&tw, (size_t)rns * N * 4) != hipSuccess || hipMalloc(&crt, sizeof(CrtConst)) != hipSuccess) {
        dev_ok = true;
        return set_error(GFE_ENOMEM, "device allocation failed");
    }
    dev_ok = true;
    {
        std::vector<uint32_t> htw((size_t)rns * N);
        const int sets[3] = {is32, is32, is32}, primes[3] = {0, 1, 2};
        make_twiddle_tables(sets, primes, rns, n, htw.data());
        if (hipMemcpy(tw, htw.data(), (size_t)rns * N * 4, hipMemcpyHostToDevice) != hipSuccess)
            return set_error(GFE_EHIP, "twiddle upload");
    }
    const CrtConst c = make_crt(rns, is32, n, b);
    if (hipMemcpy(crt, &c, sizeof(c), hipMemcpyHostToDevice) != hipSuccess) return set_error(GFE_EHIP, "crt upload");
    for (int k = 0; k < 3; ++k) { pr.p[k] = prime_of(is32, k); pr.q[k] = pinv32(pr.p[k]); }
    pr.p[3] = pr.q[3] = 0;
    can_limb = limb_ok(b, n) && default_limb_plan(n).limb;
    if (can_limb) {
        // the two-digit-limb transform: four tables of N/2 twiddles and its CRT constants
        const size_t M = N / 2;
        if (hipMalloc(&tw_limb, 4 * M * 4) != hipSuccess || hipMalloc(&limb_crt, sizeof(LimbCrt)) != hipSuccess)
            return set_error(GFE_ENOMEM, "device allocation failed");
        std::vector<uint32_t> htw(4 * M);
        const int sets[4] = {LIMB_SET, LIMB_SET, LIMB_SET, LIMB_SET}, primes[4] = {0, 1, 2, 3};
        make_twiddle_tables(sets, primes, 4, n - 1, htw.data());
        if (hipMemcpy(tw_limb, htw.data(), 4 * M * 4, hipMemcpyHostToDevice) != hipSuccess)
            return set_error(GFE_EHIP, "twiddle upload");
        const LimbCrt lc = make_limb_crt(n, b);
        if (hipMemcpy(limb_crt, &lc, sizeof(lc), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(GFE_EHIP, "crt upload");
        for (int k = 0; k < 4; ++k) {
            pr4.p[k] = prime_of(LIMB_SET, k);
            pr4.q[k] = pinv32(pr4.p[k]);
            bm4[k] = mulw_const(b % pr4.p[k], pr4.p[k]);
        }
        plan = engine_plan_for(b, n);
    }
    if (hipMemsetAsync(regs, 0, (size_t)num_regs * N * 4, stream) != hipSuccess) return set_error(GFE_EHIP, "memset");
    for (int i = 0; i < 12; ++i) { hipEvent_t ev; if (hipEventCreate(&ev) != hipSuccess) return set_error(GFE_EHIP, "event"); events.push_back(ev); }
    if (hipStreamSynchronize(stream) != hipSuccess) return set_error(GFE_EHIP, "init sync");
    return GFE_OK;
}

void Engine::bind(uint32_t b_, uint32_t n_, int device_, hipStream_t stream_, int32_t *digits, uint32_t *ws_,
                  uint32_t *tw_, int64_t *carry_, CrtConst *crt_, const Plan &plan_) {
    b = b_; n = n_; num_regs = 1; device = device_;
    N = (size_t)1 << n;
    select_primes(b, n, &rns, &ref_is32);
    is32 = transform_is32(rns, ref_is32);
    stream = stream_;
    regs = digits; ws = ws_; zp = nullptr; tw = tw_; carry = carry_; crt = crt_;
    plan = plan_;
    rounds = carry_rounds(b, n);
    cus = 256;
    for (int k = 0; k < 3; ++k) { pr.p[k] = prime_of(is32, k); pr.q[k] = pinv32(pr.p[k]); }
    carry_pending = false;
    dev_ok = false;  // nothing here is the engine's to free
}

size_t Engine::lds_col() const {
    return (size_t)plan.col_ng * ((1u << plan.a) + ((1u << plan.a) >> plan.col_loge)) * 4;
}
size_t Engine::lds_row() const {
    const uint32_t G = 1u << (tn() - plan.a);
    return (size_t)plan.row_ng * (G + (G >> plan.row_loge)) * 4;
}

// Device-resident workgroups of a kernel configuration (cached).
unsigned Engine::resident_blocks(const void *fn, int threads, size_t lds) {
    const auto key = std::make_tuple(fn, threads, lds);
    const auto it = resident.find(key);
    if (it != resident.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, lds) != hipSuccess || nb < 1) nb = 1;
    (void)hipGetLastError();
    const unsigned r = (unsigned)nb * (unsigned)std::max(cus, 1);
    resident[key] = r;
    return r;
}

// Forward column pass from register `src` into `dst` (ws or zp).
void Engine::launch_col_fwd(uint32_t src, uint32_t *dst) {
    // the fused limb plans' forward pass stages the four primes' digits once (ks_col_fwd_limb_all)
    // and takes only the range carries
    const bool staged = plan.limb && plan.fused && plan.spec && spec_col_fwd_limb_all(plan.col_loge, plan.a) != nullptr;
    if (staged && carry_pending && pend_kind != 2) flush_carry();
    // the pending carries enter this pass when it reads register 0 for a squaring or product and
    // knows their form (the digit transform: k_norm's; the limb transform: either limb form)
    const bool absorb = carry_pending && src == 0 && defer_carry() && (pend_kind != 0) == (plan.limb != 0) &&
                        (pend_kind != 2 || plan.col_ng == (1 << pend_rlog));
    if (carry_pending && !absorb) flush_carry();
    if (plan.limb) {
        ColArgs A;
        A.src = regs + (size_t)src * N; A.ws = dst; A.tw = tw_limb; A.pr = pr4;
        A.n = tn(); A.a = plan.a; A.ng = plan.col_ng;
        for (int k = 0; k < 4; ++k) A.bm[k] = bm4[k];
        if (absorb) {  // the caller overwrites register 0 (square, mul)
            if (pend_kind == 2) {
                A.ls = reinterpret_cast<const int4 *>(carry);
                A.rlog = pend_rlog;
            } else {
                A.wc = carry;
            }
            A.b = b;
            carry_pending = false;
        }
        const uint32_t C = 1u << (tn() - plan.a);
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        if (staged) {
            hipLaunchKernelGGL(spec_col_fwd_limb_all(plan.col_loge, plan.a), dim3(C / plan.col_ng, 1, 1), dim3(4 * T),
                               limb_col_fwd_all_lds_bytes(plan.a, plan.col_loge, plan.col_ng), stream, A);
            return;
        }
        hipLaunchKernelGGL(spec_col_fwd_limb(plan.col_loge, plan.a), dim3(C / plan.col_ng, 4, 1), dim3(T),
                           col_pass_lds_words(plan.a, plan.col_loge, plan.col_ng) * 4, stream, A);
        return;
    }
    if (plan.spec) {
        ColArgs A;
        A.src = regs + (size_t)src * N; A.ws = dst; A.tw = tw; A.pr = pr;
        A.n = (int)n; A.a = plan.a; A.ng = plan.col_ng;
        if (absorb) {  // the caller overwrites register 0 (square, mul)
            A.wc = carry; A.b = b;
            carry_pending = false;
        }
        const uint32_t C = 1u << (n - plan.a);
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        hipLaunchKernelGGL(spec_col_fwd(is32, plan.col_loge, plan.a), dim3(C / plan.col_ng, rns, 1), dim3(T),
                           col_pass_lds_words(plan.a, plan.col_loge, plan.col_ng) * 4, stream, A);
        return;
    }
    const KernelSet K = kernels_for(rns, is32, plan);
    if (plan.pipe && !plan.fused) {
        ColArgs A;
        A.src = regs + (size_t)src * N; A.ws = dst; A.tw = tw; A.pr = pr;
        A.n = (int)n; A.a = plan.a; A.ng = plan.col_ng;
        const unsigned nblk = (1u << (n - plan.a)) / plan.col_ng;
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        const size_t lds = lds_col() + ((size_t)4 << plan.a);
        const unsigned g = persistent_grid(nblk, rns, resident_blocks((const void *)K.col_fwd_p, T, lds));
        hipLaunchKernelGGL(K.col_fwd_p, dim3(g, rns, 1), dim3(T), lds, stream, A);
        return;
    }
    if (plan.fused) {
        ColAllArgs A;
        memset(&A, 0, sizeof(A));
        A.src = regs + (size_t)src * N; A.ws = dst; A.tw = tw; A.pr = pr;
        A.n = (int)n; A.a = plan.a; A.ng = plan.col_ng;
        const uint32_t C = 1u << (n - plan.a);
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        hipLaunchKernelGGL(K.col_fwd_all, dim3(C / plan.col_ng, 1, 1), dim3(T), lds_col(), stream, A);
        return;
    }
    ColArgs A;
    A.src = regs + (size_t)src * N; A.ws = dst; A.tw = tw; A.pr = pr;
    A.n = (int)n; A.a = plan.a; A.ng = plan.col_ng;
    const uint32_t C = 1u << (n - plan.a);
    const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
    hipLaunchKernelGGL(K.col_fwd, dim3(C / plan.col_ng, rns, 1), dim3(T), lds_col(), stream, A);
}

// Inverse column pass; with a fused plan it also normalises register 0 (dup: doubled).
void Engine::launch_col_inv(int dup) {
    if (plan.spec && plan.limb && plan.fused) {  // with the normalisation: register 0 and the range carries
        LimbColNormArgs A;
        A.ws = ws; A.tw = tw_limb; A.pr = pr4; A.dst = regs; A.slots = reinterpret_cast<int4 *>(carry);
        A.crt = static_cast<const LimbCrt *>(limb_crt);
        A.n = tn(); A.a = plan.a; A.ng = plan.col_ng; A.dup = dup;
        const uint32_t C = 1u << (tn() - plan.a);
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        hipLaunchKernelGGL(spec_col_inv_norm_limb(plan.col_loge, plan.a), dim3(C / plan.col_ng, 1, 1), dim3(4 * T),
                           limb_col_norm_lds_bytes(plan.a, plan.col_loge, plan.col_ng), stream, A);
        carry_pending = false;  // register 0 was rewritten; launch_norm records its carries
        return;
    }
    if (plan.spec) {
        ColArgs A;
        A.src = nullptr; A.ws = ws; A.tw = cur_tw(); A.pr = cur_pr();
        A.n = tn(); A.a = plan.a; A.ng = plan.col_ng;
        const uint32_t C = 1u << (tn() - plan.a);
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        hipLaunchKernelGGL(spec_col_inv(plan.limb ? 0 : is32, plan.col_loge, plan.a), dim3(C / plan.col_ng, tp(), 1),
                           dim3(T), col_pass_lds_words(plan.a, plan.col_loge, plan.col_ng) * 4, stream, A);
        return;
    }
    const KernelSet K = kernels_for(rns, is32, plan);
    if (plan.pipe && !plan.fused) {
        ColArgs A;
        A.src = nullptr; A.ws = ws; A.tw = tw; A.pr = pr;
        A.n = (int)n; A.a = plan.a; A.ng = plan.col_ng;
        const unsigned nblk = (1u << (n - plan.a)) / plan.col_ng;
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        const size_t lds = lds_col() + ((size_t)4 << plan.a);
        const unsigned g = persistent_grid(nblk, rns, resident_blocks((const void *)K.col_inv_p, T, lds));
        hipLaunchKernelGGL(K.col_inv_p, dim3(g, rns, 1), dim3(T), lds, stream, A);
        return;
    }
    if (plan.fused) {
        ColAllArgs A;
        memset(&A, 0, sizeof(A));
        A.ws = ws; A.tw = tw; A.dst = regs; A.seg = carry; A.crt = crt; A.pr = pr; A.b = b;
        A.dup = dup; A.rounds = rounds; A.n = (int)n; A.a = plan.a; A.ng = plan.col_ng;
        const uint32_t C = 1u << (n - plan.a);
        const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
        hipLaunchKernelGGL(K.col_inv_norm, dim3(C / plan.col_ng, 1, 1), dim3(T), lds_col(), stream, A);
        return;
    }
    ColArgs A;
    A.src = nullptr; A.ws = ws; A.tw = tw; A.pr = pr;
    A.n = (int)n; A.a = plan.a; A.ng = plan.col_ng;
    const uint32_t C = 1u << (n - plan.a);
    const int T = plan.col_ng * (1 << plan.a) / (1 << plan.col_loge);
    hipLaunchKernelGGL(K.col_inv, dim3(C / plan.col_ng, rns, 1), dim3(T), lds_col(), stream, A);
}

void Engine::launch_row(int mode, uint32_t *data) {
    const KernelSet K = kernels_for(rns, plan.limb ? 0 : is32, plan);
    RowArgs A;
    A.ws = data; A.zp = zp; A.tw = cur_tw(); A.pr = cur_pr();
    A.n = tn(); A.a = plan.a; A.ng = plan.row_ng;
    const uint32_t R = 1u << plan.a, G = 1u << (tn() - plan.a);
    const int T = plan.row_ng * G / (1 << plan.row_loge);
    if (plan.spec && mode == ROW_SQUARE) {
        hipLaunchKernelGGL(spec_row_square(plan.limb ? 0 : is32, plan.row_loge, tn() - plan.a), dim3(R / plan.row_ng, tp(), 1),
                           dim3(T), row_pass_lds_words(tn() - plan.a, plan.row_loge, plan.row_ng) * 4, stream, A);
        return;
    }
    RowFn f = (mode == ROW_SQUARE) ? K.row_sq : (mode == ROW_MUL) ? K.row_mul : K.row_fwd;
    hipLaunchKernelGGL(f, dim3(R / plan.row_ng, tp(), 1), dim3(T), lds_row(), stream, A);
}

// Normalisation of register 0 (fused plan: the cross-segment carries of k_col_inv_norm).
void Engine::launch_norm(int dup) {
    if (plan.limb && plan.fused) {  // done by the inverse column pass: its range carries are pending
        pend_kind = 2;
        pend_rlog = __builtin_ctz((unsigned)plan.col_ng);
        carry_pending = true;
        return;
    }
    if (plan.limb) {
        LimbNormArgs A;
        A.ws = ws; A.dst = regs; A.wl = carry; A.crt = static_cast<const LimbCrt *>(limb_crt);
        A.M = N / 2; A.dup = dup;
        const unsigned nwg = (unsigned)(A.M / LIMB_RANGE);
        hipLaunchKernelGGL(k_norm_limb, dim3(nwg, 1, 1), dim3(256), 0, stream, A);
        pend_kind = 1;
        if (defer_carry()) carry_pending = true;
        else launch_carry_limbs();
        return;
    }
    if (plan.fused) {
        FoldArgs A;
        A.dst = regs; A.seg = carry; A.nseg = N / (size_t)plan.col_ng; A.ng = plan.col_ng; A.b = b;
        hipLaunchKernelGGL(k_fold, dim3((unsigned)((A.nseg + 255) / 256), 1, 1), dim3(256), 0, stream, A);
        return;
    }
    const KernelSet K = kernels_for(rns, is32, plan);
    NormArgs A;
    A.ws = ws; A.dst = regs; A.carry = carry; A.crt = crt; A.b = b; A.n = (int)n; A.dup = dup;
    const unsigned nwg = (unsigned)(N / (plan.norm_s * 256));
    if (plan.pipe) {
        const unsigned g = persistent_grid(nwg, 1, resident_blocks((const void *)K.norm_p, 256, 0));
        hipLaunchKernelGGL(K.norm_p, dim3(g, 1, 1), dim3(256), 0, stream, A);
    } else {
        hipLaunchKernelGGL(K.norm, dim3(nwg, 1, 1), dim3(256), 0, stream, A);
    }
    pend_kind = 0;
    if (defer_carry()) carry_pending = true;
    else launch_carry();
}

void Engine::flush_carry() {
    if (!carry_pending) return;
    carry_pending = false;
    launch_carry();
}

void Engine::launch_carry() {
    if (pend_kind == 2) launch_carry_slots();
    else if (pend_kind == 1) launch_carry_limbs();
    else launch_carry_digits();
}

void Engine::launch_carry_slots() {
    const unsigned nrange = (unsigned)((N / 2) >> pend_rlog);
    hipLaunchKernelGGL(k_carry_limb_slots, dim3((nrange + 255) / 256, 1, 1), dim3(256), 0, stream, regs,
                       reinterpret_cast<const int4 *>(carry), nrange, pend_rlog, b);
}

void Engine::launch_carry_limbs() {
    const unsigned nwg = (unsigned)(N / 2 / LIMB_RANGE);
    hipLaunchKernelGGL(k_carry_limb, dim3((nwg + 255) / 256, 1, 1), dim3(256), 0, stream, regs, carry, nwg, b);
}

void Engine::launch_carry_digits() {
    CarryArgs A;
    A.dst = regs; A.carry = carry; A.b = b;
    A.nwg = (unsigned)(N / (8 * 256));
    hipLaunchKernelGGL(k_carry_wg, dim3((A.nwg + 255) / 256, 1, 1), dim3(256), 0, stream, A);
}

void Engine::launch_fused(int dup) {
    FusedArgs A;
    memset(&A, 0, sizeof(A));
    A.regs = regs; A.tw = tw; A.crt = crt; A.slots = nullptr; A.act = nullptr; A.dupf = nullptr;
    A.rounds = 1; A.dup0 = dup; A.b = b; A.rns = rns; A.is32 = is32; A.prS = pr; A.prU = pr;
    A.tws = (uint32_t)N;  // the primes' tables lie N entries apart
    hipLaunchKernelGGL(fused_square_fn((int)n), dim3(1), dim3(fused_threads((int)n)), fused_lds_bytes((int)n), stream, A);
}

void Engine::square(int dup) {
    if (plan.fuse_sq) {
        flush_carry();
        launch_fused(dup);
        return;
    }
    launch_col_fwd(0, ws);
    launch_row(ROW_SQUARE, ws);
    launch_col_inv(dup);
    launch_norm(dup);
}

void Engine::mul() {
    launch_col_fwd(0, ws);
    launch_row(ROW_MUL, ws);
    launch_col_inv(0);
    launch_norm(0);
}

void Engine::init_multiplicand(uint32_t src) {
    flush_carry();  // register 0 is read but not rewritten
    launch_col_fwd(src, zp);
    launch_row(ROW_FWD, zp);
}

void Engine::mul_scalar(int32_t a) {
    flush_carry();
    MulScalarArgs A;
    A.dst = regs; A.carry = carry; A.b = b; A.a = a;
    const unsigned nwg = (unsigned)(N / (8 * 256));
    hipLaunchKernelGGL(k_mulscalar, dim3(nwg, 1, 1), dim3(256), 0, stream, A);
    launch_carry_digits();
}

void Engine::copy(uint32_t dst, uint32_t src) {
    const size_t n4 = N / 4;
    const int T = 256;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n4 + T - 1) / T), 1, 1), dim3(T), 0, stream,
                       reinterpret_cast<int4 *>(regs + (size_t)dst * N), reinterpret_cast<const int4 *>(regs + (size_t)src * N),
                       n4);
}

void Engine::set(uint32_t a) {
    const int T = 256;
    hipLaunchKernelGGL(k_set, dim3((unsigned)((N + T - 1) / T), 1, 1), dim3(T), 0, stream, regs, N, (int32_t)a);
}

int Engine::check(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(GFE_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return GFE_OK;
}

}  // namespace gfe

// ================================================================================================
// C ABI

using gfe::Engine;
using gfe::set_error;

struct gfe_engine {
    Engine eng;
};

#define GFE_GUARD(e)                                                         \
    do {                                                                     \
        if (!(e)) return set_error(GFE_EINVAL, "null engine");               \
        if (hipSetDevice((e)->eng.device) != hipSuccess)                     \
            return set_error(GFE_EHIP, "hipSetDevice failed");               \
    } while (0)

extern "C" {

const char *gfe_last_error(void) { return gfe::g_last_error.c_str(); }

int gfe_select(uint32_t b, uint32_t n, int *rns, int *is32) {
    if (!rns || !is32) return set_error(GFE_EINVAL, "null output");
    gfe::select_primes(b, n, rns, is32);
    return GFE_OK;
}

int gfe_create(uint32_t b, uint32_t n, uint32_t num_regs, int device, gfe_engine **out) {
    if (!out) return set_error(GFE_EINVAL, "null output");
    *out = nullptr;
    if (n < 12 || n > 23) return set_error(GFE_EINVAL, "n must be in [12, 23]");
    if (b < 1024 || b > 2000000000u) return set_error(GFE_EINVAL, "b must be in [1024, 2000000000]");
    if (num_regs < 1 || num_regs > 4096) return set_error(GFE_EINVAL, "num_regs must be in [1, 4096]");
    gfe_engine *e = new gfe_engine;
    const int rc = e->eng.init(b, n, num_regs, device);
    if (rc != GFE_OK) {
        const std::string msg = gfe::g_last_error;
        delete e;
        return set_error(rc, msg);
    }
    *out = e;
    return GFE_OK;
}

int gfe_destroy(gfe_engine *e) {
    if (!e) return GFE_OK;
    delete e;
    return GFE_OK;
}

int gfe_set(gfe_engine *e, uint32_t a) {
    GFE_GUARD(e);
    if ((uint64_t)a > (uint64_t)(e->eng.b / 2) + 4) return set_error(GFE_EINVAL, "set: a exceeds b/2 + 4");
    e->eng.flush_carry();
    e->eng.set(a);
    return e->eng.check("set");
}

int gfe_square_dup(gfe_engine *e, int dup) {
    GFE_GUARD(e);
    e->eng.square(dup != 0);
    return e->eng.check("squareDup");
}

int gfe_square_mul(gfe_engine *e, int32_t a) {
    GFE_GUARD(e);
    // transformGPUs::squareMul, src/transformGPU.h:1163-1171 (a <= 2 -> squareDup(a == 2))
    if (a <= 2) {
        e->eng.square(a == 2);
    } else {
        e->eng.square(0);
        e->eng.mul_scalar(a);
    }
    return e->eng.check("squareMul");
}

int gfe_init_multiplicand(gfe_engine *e, uint32_t src) {
    GFE_GUARD(e);
    if (src >= e->eng.num_regs) return set_error(GFE_EINVAL, "register index out of range");
    e->eng.init_multiplicand(src);
    return e->eng.check("initMultiplicand");
}

int gfe_mul(gfe_engine *e) {
    GFE_GUARD(e);
    e->eng.mul();
    return e->eng.check("mul");
}

int gfe_copy(gfe_engine *e, uint32_t dst, uint32_t src) {
    GFE_GUARD(e);
    e->eng.flush_carry();
    Engine &g = e->eng;
    if (dst >= g.num_regs || src >= g.num_regs) return set_error(GFE_EINVAL, "register index out of range");
    if (dst == src) return GFE_OK;
    g.copy(dst, src);
    return g.check("copy");
}

int gfe_get_digits(gfe_engine *e, uint32_t reg, int32_t *out) {
    GFE_GUARD(e);
    e->eng.flush_carry();
    Engine &g = e->eng;
    if (!out) return set_error(GFE_EINVAL, "null output");
    if (reg >= g.num_regs) return set_error(GFE_EINVAL, "register index out of range");
    if (hipMemcpyAsync(out, g.regs + (size_t)reg * g.N, g.N * 4, hipMemcpyDeviceToHost, g.stream) != hipSuccess)
        return set_error(GFE_EHIP, "read");
    if (hipStreamSynchronize(g.stream) != hipSuccess) return set_error(GFE_EHIP, "read sync");
    return GFE_OK;
}

int gfe_set_digits(gfe_engine *e, uint32_t reg, const int32_t *in) {
    GFE_GUARD(e);
    e->eng.flush_carry();
    Engine &g = e->eng;
    if (!in) return set_error(GFE_EINVAL, "null input");
    if (reg >= g.num_regs) return set_error(GFE_EINVAL, "register index out of range");
    // digits must be bounded as the engine keeps them (|d| <= b/2 + 4); larger values are the
    // caller's to balance first (transform::setInt balances, src/transform.h:241-246)
    const int64_t bound = (int64_t)(g.b / 2) + 4;
    for (size_t i = 0; i < g.N; ++i)
        if ((int64_t)in[i] > bound || (int64_t)in[i] < -bound) return set_error(GFE_EINVAL, "digit out of balanced range");
    if (hipMemcpyAsync(g.regs + (size_t)reg * g.N, in, g.N * 4, hipMemcpyHostToDevice, g.stream) != hipSuccess)
        return set_error(GFE_EHIP, "write");
    if (hipStreamSynchronize(g.stream) != hipSuccess) return set_error(GFE_EHIP, "write sync");
    return GFE_OK;
}

size_t gfe_mem_size(const gfe_engine *e) {
    if (!e) return 0;
    const Engine &g = e->eng;
    return gfe::engine_bytes(g.rns, g.N, g.num_regs, g.tw_limb != nullptr);
}

size_t gfe_cache_size(const gfe_engine *e) {
    if (!e) return 0;
    const Engine &g = e->eng;
    return (size_t)g.rns * g.N * 4 + g.N * 4;
}

// transform::EKind of the reference's GPU transform (src/transform.h:45, src/transformGPU.h:947):
// the first int of a context payload.
static int engine_kind(const Engine &g) { return g.rns == 2 ? 6 : 7; }

size_t gfe_context_size(const gfe_engine *e, uint32_t num_regs) {
    if (!e) return 0;
    return sizeof(int) + (size_t)num_regs * e->eng.rns * e->eng.N * 4;
}

// The registers hold signed digits; the payload holds their plain residues prime by prime, as the
// reference's device registers do between squarings (set_int: d < 0 -> d + p).
int gfe_save(gfe_engine *e, uint32_t num_regs, void *buf, size_t size) {
    GFE_GUARD(e);
    e->eng.flush_carry();
    Engine &g = e->eng;
    if (num_regs == 0) num_regs = g.num_regs;
    if (num_regs > g.num_regs) return set_error(GFE_EINVAL, "num_regs");
    if (!buf || size < gfe_context_size(e, num_regs)) return set_error(GFE_EINVAL, "buffer too small");
    std::vector<int32_t> d((size_t)num_regs * g.N);
    if (hipMemcpyAsync(d.data(), g.regs, d.size() * 4, hipMemcpyDeviceToHost, g.stream) != hipSuccess)
        return set_error(GFE_EHIP, "save");
    if (hipStreamSynchronize(g.stream) != hipSuccess) return set_error(GFE_EHIP, "save sync");
    const int kind = engine_kind(g);
    memcpy(buf, &kind, sizeof(kind));
    uint32_t *z = reinterpret_cast<uint32_t *>((char *)buf + sizeof(int));
    for (uint32_t r = 0; r < num_regs; ++r)
        for (int k = 0; k < g.rns; ++k) {
            const uint32_t p = gfe::prime_of(g.ref_is32, k);
            const int32_t *dr = d.data() + (size_t)r * g.N;
            uint32_t *zr = z + ((size_t)r * g.rns + k) * g.N;
            for (size_t i = 0; i < g.N; ++i) zr[i] = dr[i] < 0 ? (uint32_t)dr[i] + p : (uint32_t)dr[i];
        }
    return GFE_OK;
}

// Balanced digits |d| <= b/2 (+1 for digit 0) of the same value mod b^N + 1: a carry pass with the
// wrap x^N = -1, then at most two more passes for the carry that wraps; what is left of it stays in
// digit 0.
static void balance_digits(int32_t *d, size_t N, uint32_t b) {
    const int64_t bb = b, half = b / 2;
    int64_t c = 0;
    for (int pass = 0; pass < 3; ++pass) {
        for (size_t i = 0; i < N; ++i) {
            const int64_t v = (int64_t)d[i] + c, t = v + half;
            const int64_t q = t >= 0 ? t / bb : -((-t + bb - 1) / bb);
            d[i] = (int32_t)(v - q * bb);
            c = q;
        }
        c = -c;
        if (c == 0) return;
    }
    d[0] = (int32_t)((int64_t)d[0] + c);
}

// Digits come back from the first prime's residue (ZPT::get_int, src/transformGPU.h:146); the
// other primes must hold the residues of the same digit.  The reference's registers hold digits in
// (-b, b) (its normalisation keeps the sign of the carry, ocl/kernel.cl:2286-2300); they are
// balanced here, as the engine keeps them.
int gfe_load(gfe_engine *e, uint32_t num_regs, const void *buf, size_t size) {
    GFE_GUARD(e);
    e->eng.flush_carry();
    Engine &g = e->eng;
    if (num_regs == 0) num_regs = g.num_regs;
    if (num_regs > g.num_regs) return set_error(GFE_EINVAL, "num_regs");
    if (!buf || size < sizeof(int)) return set_error(GFE_EINVAL, "short context");
    int kind;
    memcpy(&kind, buf, sizeof(kind));
    if (kind != engine_kind(g)) return set_error(GFE_EKIND, "context of another engine kind");
    if (size < gfe_context_size(e, num_regs)) return set_error(GFE_EINVAL, "short context");
    const uint32_t *z = reinterpret_cast<const uint32_t *>((const char *)buf + sizeof(int));
    const uint32_t p0 = gfe::prime_of(g.ref_is32, 0);
    std::vector<int32_t> d((size_t)num_regs * g.N);
    for (uint32_t r = 0; r < num_regs; ++r) {
        const uint32_t *z0 = z + (size_t)r * g.rns * g.N;
        int32_t *dr = d.data() + (size_t)r * g.N;
        for (size_t i = 0; i < g.N; ++i) {
            const uint32_t v = z0[i];
            if (v >= p0) return set_error(GFE_EINVAL, "context residue not reduced");
            dr[i] = (v >= p0 / 2) ? (int32_t)(int64_t(v) - p0) : (int32_t)v;
        }
        for (int k = 1; k < g.rns; ++k) {
            const uint32_t p = gfe::prime_of(g.ref_is32, k);
            const uint32_t *zk = z0 + (size_t)k * g.N;
            for (size_t i = 0; i < g.N; ++i) {
                const uint32_t want = dr[i] < 0 ? (uint32_t)dr[i] + p : (uint32_t)dr[i];
                if (zk[i] != want) return set_error(GFE_EINVAL, "context residues disagree between primes");
            }
        }
        balance_digits(dr, g.N, g.b);
    }
    if (hipMemcpyAsync(g.regs, d.data(), d.size() * 4, hipMemcpyHostToDevice, g.stream) != hipSuccess)
        return set_error(GFE_EHIP, "load");
    if (hipStreamSynchronize(g.stream) != hipSuccess) return set_error(GFE_EHIP, "load sync");
    return GFE_OK;
}

double gfe_get_error(const gfe_engine *) { return 0.0; }

int gfe_info(const gfe_engine *e, char *buf, size_t len) {
    if (!e || !buf || len == 0) return set_error(GFE_EINVAL, "info");
    const Engine &g = e->eng;
    std::ostringstream ss;
    ss << "NTT" << g.rns << (g.is32 ? "u" : "s");
    if (g.ref_is32 != g.is32) ss << " (reference set NTT" << g.rns << (g.ref_is32 ? "u" : "s") << ")";
    ss << " n=" << g.n << " b=" << g.b << "; " << gfe::plan_str(g.plan, g.n);
    snprintf(buf, len, "%s", ss.str().c_str());
    return GFE_OK;
}

int gfe_sync(gfe_engine *e) {
    GFE_GUARD(e);
    e->eng.flush_carry();
    if (hipStreamSynchronize(e->eng.stream) != hipSuccess) return set_error(GFE_EHIP, "sync");
    return e->eng.check("sync");
}

int gfe_time_squarings(gfe_engine *e, uint32_t count, uint64_t dup_bits, double *ms, double *kernel_ms) {
    GFE_GUARD(e);
    Engine &g = e->eng;
    if (!ms || count == 0) return set_error(GFE_EINVAL, "time_squarings");
    hipEvent_t e0 = g.events[0], e1 = g.events[1];
    (void)hipEventRecord(e0, g.stream);
    for (uint32_t i = 0; i < count; ++i) g.square((dup_bits >> (i % 64)) & 1);
    (void)hipEventRecord(e1, g.stream);
    if (hipEventSynchronize(e1) != hipSuccess) return set_error(GFE_EHIP, "event sync");
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t;
    if (kernel_ms) {
        // each pass alone: `count` back-to-back launches between one pair of events on the engine's
        // stream, so the average is the pass's own duration (launch gaps amortised).  Register 0
        // stays a valid residue: one squaring (dup = 0) leaves the product's residues in ws, the
        // normalisation rewrites register 0 from them identically each time, and the forward
        // column pass only reads it; the row and inverse column passes work on ws alone.
        auto time_k = [&](auto &&launch, double *out) -> int {
            (void)hipEventRecord(g.events[2], g.stream);
            for (uint32_t i = 0; i < count; ++i) launch();
            (void)hipEventRecord(g.events[3], g.stream);
            if (hipEventSynchronize(g.events[3]) != hipSuccess) return set_error(GFE_EHIP, "event sync");
            float d = 0;
            (void)hipEventElapsedTime(&d, g.events[2], g.events[3]);
            *out = d / count;
            return GFE_OK;
        };
        for (int k = 0; k < 5; ++k) kernel_ms[k] = 0.0;
        int rc = GFE_OK;
        if (g.plan.fuse_sq) {
            rc = time_k([&]() { g.launch_fused(0); }, &kernel_ms[0]);
        } else if (g.plan.limb && g.plan.fused) {
            // the fused inverse column pass reads the row pass's output and writes register 0, so it is
            // timed first, on the product of one squaring; the forward passes then only read register 0
            g.flush_carry();
            g.launch_col_fwd(0, g.ws);
            g.launch_row(gfe::ROW_SQUARE, g.ws);
            rc = time_k([&]() { g.launch_col_inv(0); }, &kernel_ms[2]);
            g.launch_norm(0);
            g.flush_carry();
            if (rc == GFE_OK) rc = time_k([&]() { g.launch_col_fwd(0, g.ws); }, &kernel_ms[0]);
            if (rc == GFE_OK) rc = time_k([&]() { g.launch_row(gfe::ROW_SQUARE, g.ws); }, &kernel_ms[1]);
        } else {
            g.flush_carry();
            g.launch_col_fwd(0, g.ws);
            g.launch_row(gfe::ROW_SQUARE, g.ws);
            g.launch_col_inv(0);
            if (rc == GFE_OK) rc = time_k([&]() { g.launch_norm(0); }, &kernel_ms[3]);
            g.flush_carry();
            if (rc == GFE_OK) rc = time_k([&]() { g.launch_col_fwd(0, g.ws); }, &kernel_ms[0]);
            if (rc == GFE_OK) rc = time_k([&]() { g.launch_row(gfe::ROW_SQUARE, g.ws); }, &kernel_ms[1]);
            if (rc == GFE_OK) rc = time_k([&]() { g.launch_col_inv(0); }, &kernel_ms[2]);
        }
        if (rc != GFE_OK) return rc;
    }
    return g.check("time_squarings");
}

// The autotuner: every admissible plan is timed on `reps` squarings of a pseudo-random balanced
// residue (register 0 is saved and restored); the fastest is kept.
int gfe_tune(gfe_engine *e, uint32_t reps) {
    GFE_GUARD(e);
    e->eng.flush_carry();
    Engine &g = e->eng;
    if (reps == 0) reps = 8;
    std::vector<int32_t> saved(g.N), trial(g.N);
    if (hipMemcpyAsync(saved.data(), g.regs, g.N * 4, hipMemcpyDeviceToHost, g.stream) != hipSuccess ||
        hipStreamSynchronize(g.stream) != hipSuccess)
        return set_error(GFE_EHIP, "tune save");
    uint64_t x = 0x9E3779B97F4A7C15ull ^ g.b;
    const int64_t half = g.b / 2;
    for (size_t i = 0; i < g.N; ++i) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        trial[i] = (int32_t)((int64_t)(x % g.b) - half + ((g.b & 1) ? 0 : 1));
    }
    std::vector<gfe::Plan> plans = gfe::candidate_plans(g.n);
    if (g.can_limb)
        for (const gfe::Plan &P : gfe::candidate_limb_plans(g.n))
            if (!P.fused || gfe::limb_fused_ok(g.b, g.n, P.col_ng)) plans.push_back(P);
    const gfe::Plan orig = g.plan;
    // ms for `count` squarings under plan P (after one untimed), < 0 on a launch error
    auto time_plan = [&](const gfe::Plan &P, uint32_t count, double *ms) -> int {
        g.plan = P;
        if (hipMemcpyAsync(g.regs, trial.data(), g.N * 4, hipMemcpyHostToDevice, g.stream) != hipSuccess)
            return set_error(GFE_EHIP, "tune upload");
        g.carry_pending = false;  // register 0 was just rewritten
        g.square(0);
        (void)hipEventRecord(g.events[0], g.stream);
        for (uint32_t i = 0; i < count; ++i) g.square(0);
        (void)hipEventRecord(g.events[1], g.stream);
        if (hipEventSynchronize(g.events[1]) != hipSuccess) return set_error(GFE_EHIP, "tune sync");
        float t = 0;
        (void)hipEventElapsedTime(&t, g.events[0], g.events[1]);
        *ms = (hipGetLastError() != hipSuccess) ? -1.0 : (double)t;
        return GFE_OK;
    };
    // every plan on `reps` squarings, then the fastest eight again on 8 x reps: a few squarings
    // rank the plans, the short list is decided on timings long enough to be stable
    std::vector<std::pair<double, size_t>> screen;
    for (size_t k = 0; k < plans.size(); ++k) {
        double ms = 0;
        if (time_plan(plans[k], reps, &ms) != GFE_OK) { g.plan = orig; return GFE_EHIP; }
        if (ms >= 0) screen.emplace_back(ms, k);
    }
    std::sort(screen.begin(), screen.end());
    gfe::Plan best = orig;
    double best_ms = 1e30;
    for (size_t j = 0; j < std::min<size_t>(8, screen.size()); ++j) {
        double ms = 0;
        if (time_plan(plans[screen[j].second], 8 * reps, &ms) != GFE_OK) { g.plan = orig; return GFE_EHIP; }
        if (ms >= 0 && ms < best_ms) { best_ms = ms; best = plans[screen[j].second]; }
    }
    g.plan = best;
    g.carry_pending = false;
    if (hipMemcpyAsync(g.regs, saved.data(), g.N * 4, hipMemcpyHostToDevice, g.stream) != hipSuccess ||
        hipStreamSynchronize(g.stream) != hipSuccess)
        return set_error(GFE_EHIP, "tune restore");
    return g.check("tune");
}

int gfe_get_plan(const gfe_engine *e, int32_t plan[6]) {
    if (!e || !plan) return set_error(GFE_EINVAL, "get_plan");
    gfe::plan_to_array(e->eng.plan, plan);
    return GFE_OK;
}

int gfe_default_plan(uint32_t n, uint32_t capacity, int32_t plan[6]) {
    if (!plan || n < 12 || n > 23) return set_error(GFE_EINVAL, "default_plan");
    gfe::plan_to_array(capacity ? gfe::default_batch_plan(n, capacity) : gfe::default_plan(n), plan);
    return GFE_OK;
}

int gfe_engine_plan(uint32_t b, uint32_t n, int32_t plan[6]) {
    if (!plan || n < 12 || n > 23 || b < 1024 || b > 2000000000u) return set_error(GFE_EINVAL, "engine_plan");
    gfe::plan_to_array(gfe::engine_plan_for(b, n), plan);
    return GFE_OK;
}

int gfe_plan_admissible(uint32_t n, const int32_t plan[6]) {
    gfe::Plan P;
    if (!plan || n < 12 || n > 23 || !gfe::plan_from_array(plan, n, &P))
        return set_error(GFE_EINVAL, "plan not admissible for this n");
    return GFE_OK;
}

int gfe_candidate_plans(uint32_t n, int32_t *plans, uint32_t max, uint32_t *count) {
    if (!count || n < 12 || n > 23) return set_error(GFE_EINVAL, "candidate_plans");
    const std::vector<gfe::Plan> c = gfe::candidate_plans(n);
    *count = (uint32_t)c.size();
    if (plans)
        for (uint32_t i = 0; i < std::min<uint32_t>(max, *count); ++i) gfe::plan_to_array(c[i], plans + 6 * (size_t)i);
    return GFE_OK;
}

int gfe_host_twiddles(uint32_t n, int is32, int max_threads, uint32_t *out) {
    if (!out || n < 1 || n > 23 || (is32 != 0 && is32 != 1) || max_threads < 0)
        return set_error(GFE_EINVAL, "host_twiddles");
    const int sets[3] = {is32, is32, is32}, primes[3] = {0, 1, 2};
    gfe::make_twiddle_tables(sets, primes, 3, n, out, max_threads);
    return GFE_OK;
}

// Moves the engine to base b of the same n: only the b-dependent state changes (the CRT constants,
// the carry rounds, the limb transform's constants and whether it applies, the plan); the twiddle
// tables and buffers stay.
int gfe_rebase(gfe_engine *e, uint32_t b) {
    GFE_GUARD(e);
    Engine &g = e->eng;
    if (b < 1024 || b > 2000000000u) return set_error(GFE_EINVAL, "b must be in [1024, 2000000000]");
    int rns = 0, ref_is32 = 0;
    gfe::select_primes(b, g.n, &rns, &ref_is32);
    if (rns != g.rns || gfe::transform_is32(rns, ref_is32) != g.is32)
        return set_error(GFE_EKIND, "rebase: b selects another prime set at this n");
    g.flush_carry();  // register 0's pending carries, in the old base
    if (hipStreamSynchronize(g.stream) != hipSuccess) return set_error(GFE_EHIP, "rebase sync");
    g.b = b;
    g.ref_is32 = ref_is32;
    g.rounds = gfe::carry_rounds(b, g.n);
    const gfe::CrtConst c = gfe::make_crt(g.rns, g.is32, g.n, b);
    if (hipMemcpy(g.crt, &c, sizeof(c), hipMemcpyHostToDevice) != hipSuccess) return set_error(GFE_EHIP, "crt upload");
    g.can_limb = g.tw_limb != nullptr && gfe::limb_ok(b, g.n) && gfe::default_limb_plan(g.n).limb;
    if (g.can_limb) {
        const gfe::LimbCrt lc = gfe::make_limb_crt(g.n, b);
        if (hipMemcpy(g.limb_crt, &lc, sizeof(lc), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(GFE_EHIP, "crt upload");
        for (int k = 0; k < 4; ++k) g.bm4[k] = gfe::mulw_const(b % g.pr4.p[k], g.pr4.p[k]);
    }
    g.plan = g.can_limb ? gfe::engine_plan_for(b, g.n) : gfe::default_plan(g.n);
    return g.check("rebase");
}

int gfe_set_plan(gfe_engine *e, const int32_t plan[6]) {
    GFE_GUARD(e);
    if (!plan) return set_error(GFE_EINVAL, "set_plan");
    gfe::Plan P;
    if (!gfe::plan_from_array(plan, e->eng.n, &P)) return set_error(GFE_EINVAL, "plan not admissible for this n");
    if (P.limb && !e->eng.can_limb) return set_error(GFE_EINVAL, "the limb transform is not exact for this (b, n)");
    if (P.limb && P.fused && !gfe::limb_fused_ok(e->eng.b, e->eng.n, P.col_ng))
        return set_error(GFE_EINVAL, "the fused limb normalisation is not exact for this (b, n) and column group");
    e->eng.flush_carry();
    if (hipStreamSynchronize(e->eng.stream) != hipSuccess) return set_error(GFE_EHIP, "set_plan sync");
    e->eng.plan = P;
    return GFE_OK;
}

}  // extern "C"
