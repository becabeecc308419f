// gfe_pass.h -- row and column passes of the four-step NTT with compile-time geometry.
//
// Same transform as the runtime-geometry bodies of gfe_kernels.h (row_body, col_fwd_chunks,
// col_inv_chunks): negacyclic Cooley-Tukey forward with Psi[m + j] = psi^bitrev(m + j), pointwise
// square / product, Gentleman-Sande inverse; the residues these passes write are identical.  What
// differs is how the work is laid out for gfx950:
//
//  * LOGG (levels of the pass) and LOGE (points per thread) are template parameters, so the chunk
//    loop unrolls completely and every LDS and global element address is a per-thread base plus a
//    compile-time offset (ds_read/ds_write immediate offsets, one VGPR address per chunk);
//  * a chunk's twiddles are loaded into registers before its LDS exchange, so the loads overlap
//    the barrier;
//  * twiddle products use ModS::mulp (Montgomery quotient from a precomputed w q: no 64-bit
//    multiply);
//  * for the 31-bit primes (2p < 2^32) butterfly outputs that only feed the next level's multiply
//    are left in [0, 2p) (forward), and the inverse's difference is formed as v - u + p in (0, 2p)
//    before its multiply: mulp accepts any 32-bit operand and returns a reduced residue.
//
// Chunks: a LOGG-level pass runs a partial first chunk of R = LOGG mod LOGE levels (when R > 0),
// then full chunks of LOGE levels.  In chunk ci the thread holds the E = 2^LOGE elements
// k = base + rr 2^YL (rr < E) of one sub-transform, with base = (hi << (LOGG - lam0)) + lo,
// hi = s >> YL, lo = s mod 2^YL (s = the thread's sub-transform index) and YL = LOGG - lam0 - LOGE.
// base and rr 2^YL have disjoint bits, so the padded LDS index k + (k >> LOGE) of element rr is
// (base + (base >> LOGE)) + (c + (c >> LOGE)) with c = rr 2^YL: a compile-time offset.
#pragma once

#include <type_traits>

#include "gfe_kernels.h"
#include "gfe_limb.h"

namespace gfe {

template <int I, int END, class F>
__device__ __forceinline__ void static_for(const F &f) {
    if constexpr (I < END) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, END>(f);
    }
}

// END - 1, END - 2, ..., I
template <int I, int END, class F>
__device__ __forceinline__ void static_for_down(const F &f) {
    if constexpr (I < END) {
        f(std::integral_constant<int, END - 1>{});
        static_for_down<I, END - 1>(f);
    }
}

template <class M> struct LazyOk : std::false_type {};
template <> struct LazyOk<ModS> : std::true_type {};

template <int LOGG, int LOGE>
struct Geo {
    static_assert(LOGG >= LOGE, "a pass spans at least one chunk");
    static constexpr int R = LOGG % LOGE;
    static constexpr int NCH = (LOGG + LOGE - 1) / LOGE;
    static constexpr int lam0(int ci) { return R == 0 ? ci * LOGE : (ci == 0 ? 0 : R + (ci - 1) * LOGE); }
    static constexpr int ne(int ci) { return (R != 0 && ci == 0) ? R : LOGE; }
    static constexpr int ylog(int ci) { return LOGG - lam0(ci) - LOGE; }
};

// Padded LDS offset of element rr of a chunk with stride 2^YL.
template <int YL, int LOGE>
__device__ __forceinline__ constexpr uint32_t chunk_off(int rr) {
    return ((uint32_t)rr << YL) + (((uint32_t)rr << YL) >> LOGE);
}

// Forward CT levels 0..NE-1 of the chunk in x[], twiddles w[2^mu + bi] (load_fwd_tw).  Outputs of
// every level but the last that are the next level's multiplied operands (index bit h/2 set) stay
// lazy for the 31-bit primes: u + v and u - v + p, both in [0, 2p).
template <class M, int LOGE, int NE>
__device__ __forceinline__ void fwd_chunk(const M &md, uint32_t (&x)[1 << LOGE], const uint32_t (&w)[1 << LOGE]) {
    constexpr int E = 1 << LOGE;
#pragma unroll
    for (int mu = 0; mu < NE; ++mu) {
        const int h = E >> (mu + 1);
#pragma unroll
        for (int bi = 0; bi < (1 << mu); ++bi) {
            const uint32_t wv = w[(1 << mu) + bi], wq = wv * md.q;
#pragma unroll
            for (int i = 0; i < h; ++i) {
                const int i0 = bi * 2 * h + i, i1 = i0 + h;
                const uint32_t u = x[i0], v = md.mulp(x[i1], wv, wq);
                if (LazyOk<M>::value && mu + 1 < NE && (i0 & (h >> 1)) != 0) {
                    x[i0] = u + v;
                    x[i1] = u - v + md.p;
                } else {
                    x[i0] = md.add(u, v);
                    x[i1] = md.sub(u, v);
                }
            }
        }
    }
}

// Inverse GS levels NE-1..0 of the chunk, twiddles w[2^mu + j] (load_inv_tw; block bi uses entry
// 2^mu - 1 - bi).  The difference feeding the multiply is lazy for the 31-bit primes.
template <class M, int LOGE, int NE>
__device__ __forceinline__ void inv_chunk(const M &md, uint32_t (&x)[1 << LOGE], const uint32_t (&w)[1 << LOGE]) {
    constexpr int E = 1 << LOGE;
#pragma unroll
    for (int mu = NE - 1; mu >= 0; --mu) {
        const int h = E >> (mu + 1);
#pragma unroll
        for (int bi = 0; bi < (1 << mu); ++bi) {
            const uint32_t wv = w[(2 << mu) - 1 - bi], wq = wv * md.q;
#pragma unroll
            for (int i = 0; i < h; ++i) {
                const int i0 = bi * 2 * h + i, i1 = i0 + h;
                const uint32_t u = x[i0], v = x[i1];
                x[i0] = md.add(u, v);
                const uint32_t d = LazyOk<M>::value ? v - u + md.p : md.sub(v, u);
                x[i1] = md.mulp(d, wv, wq);
            }
        }
    }
}

// Byte offset -> element pointer (keeps the 32-bit per-thread offset separate from the uniform
// base so that global accesses can use a scalar base address).
template <class T>
__device__ __forceinline__ T *at_bytes(T *base, uint32_t bytes) {
    return reinterpret_cast<T *>(reinterpret_cast<typename std::conditional<std::is_const<T>::value, const char, char>::type *>(base) + bytes);
}

// ---------------------------------------------------------------------------------------------
// Row pass: levels a..n-1 on NG rows of G = 2^LOGG contiguous residues, the pointwise square or
// product, inverse levels n-1..a (ROW_FWD: forward only).  Thread t -> row q = t / (G/E) of the
// block's rows (blk NG + q), sub-transform s = t mod (G/E).  LDS: row q at q (G + G / 2^PAD),
// element k at k + (k >> PAD), PAD = row_pad(LOGG, LOGE).  `data`, `zp` and `tw` point at one
// prime's N values.
//
// Bank conflicts (32 banks, half a wave at a time): with 16 points per thread and rows of 2^10+
// points, no padding k + (k >> P) serves both sides of the exchange into the chunk of stride 2^4
// (lanes 32 consecutive elements apart on one side, two groups of 16 2^(LOGG-lam0) apart on the
// other).  Padding every 32 elements and letting that chunk take the sub-transforms in the order
// of s with bits 4 and 5 swapped (the two 16-lane groups of a half wave then sit two blocks apart,
// 16 banks apart) makes every exchange conflict-free; any per-chunk order of the sub-transforms
// computes the same pass.
__host__ __device__ constexpr int row_pad(int logg, int loge) { return (loge == 4 && logg >= 10) ? 5 : loge; }
__host__ __device__ constexpr bool row_swap45(int logg, int loge, int yl) { return loge == 4 && logg >= 10 && yl == 4; }
__device__ __forceinline__ uint32_t swap_bits45(uint32_t s) {
    const uint32_t d = ((s >> 4) ^ (s >> 5)) & 1u;
    return s ^ (d * 0x30u);
}

// With MODE == ROW_SQUARE, `mul` (uniform over the workgroup) turns the pointwise square into the
// product by the multiplicand zp: a multiplying round of a batched power (r0 = r0 rm).
template <class M, int LOGE, int LOGG, int MODE>
__device__ __forceinline__ void row_pass(const M &md, uint32_t *__restrict__ data, const uint32_t *__restrict__ zp,
                                         const uint32_t *__restrict__ tw, int a, int NG, unsigned blk, uint32_t *lds,
                                         bool mul = false) {
    using GM = Geo<LOGG, LOGE>;
    constexpr int E = 1 << LOGE, G = 1 << LOGG, SPR = G >> LOGE, PAD = row_pad(LOGG, LOGE);
    const uint32_t t = threadIdx.x, q = t / SPR, s = t % SPR;
    const uint32_t row = blk * (uint32_t)NG + q;
    const uint32_t rowb = row * (uint32_t)G * 4u;  // byte offset of the row
    uint32_t *L = lds + q * (G + (G >> PAD));
    uint32_t x[E], w[E];
    static_for<0, GM::NCH>([&](auto cc) {
        constexpr int ci = decltype(cc)::value;
        constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
        const uint32_t sc = row_swap45(LOGG, LOGE, YL) ? swap_bits45(s) : s;
        const uint32_t hi = sc >> YL, base = (hi << (LOGG - lam0)) + (sc & ((1u << YL) - 1u));
        load_fwd_tw<LOGE>(w, tw, (1u << (a + lam0)) + (row << lam0) + hi, NE);
        if constexpr (ci == 0) {
            const uint32_t *p = at_bytes(data, rowb + base * 4u);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = p[rr << YL];
        } else {
            __syncthreads();
            const uint32_t *p = L + base + (base >> PAD);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = p[chunk_off<YL, PAD>(rr)];
        }
        fwd_chunk<M, LOGE, NE>(md, x, w);
        if constexpr (ci < GM::NCH - 1) {
            if constexpr (ci > 0) __syncthreads();  // chunk 0 follows no LDS read
            uint32_t *p = L + base + (base >> PAD);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) p[chunk_off<YL, PAD>(rr)] = x[rr];
        }
    });
    // the last forward chunk (YL = 0: E consecutive elements from s E) is the first inverse chunk
    constexpr int LAST = GM::NCH - 1;
    const uint32_t base_last = s << LOGE;
    if constexpr (MODE == ROW_FWD) {
        uint32_t *p = at_bytes(data, rowb + base_last * 4u);
#pragma unroll
        for (int rr = 0; rr < E; ++rr) p[rr] = x[rr];
        return;
    } else {
        load_inv_tw<LOGE>(w, tw, (1u << (a + GM::lam0(LAST))) + (row << GM::lam0(LAST)) + s, a + GM::lam0(LAST),
                          GM::ne(LAST));
        if (MODE == ROW_SQUARE && !mul) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = md.mont(x[rr], x[rr]);
        } else {
            const uint32_t *z = at_bytes(zp, rowb + base_last * 4u);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = md.mont(x[rr], z[rr]);
        }
        static_for_down<0, GM::NCH>([&](auto cc) {
            constexpr int ci = decltype(cc)::value;
            constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
            const uint32_t sc = row_swap45(LOGG, LOGE, YL) ? swap_bits45(s) : s;
            const uint32_t hi = sc >> YL, base = (hi << (LOGG - lam0)) + (sc & ((1u << YL) - 1u));
            if constexpr (ci != LAST) {
                load_inv_tw<LOGE>(w, tw, (1u << (a + lam0)) + (row << lam0) + hi, a + lam0, NE);
                __syncthreads();
                const uint32_t *p = L + base + (base >> PAD);
#pragma unroll
                for (int rr = 0; rr < E; ++rr) x[rr] = p[chunk_off<YL, PAD>(rr)];
            }
            inv_chunk<M, LOGE, NE>(md, x, w);
            if constexpr (ci == 0) {
                uint32_t *p = at_bytes(data, rowb + base * 4u);
#pragma unroll
                for (int rr = 0; rr < E; ++rr) p[rr << YL] = x[rr];
            } else {
                __syncthreads();
                uint32_t *p = L + base + (base >> PAD);
#pragma unroll
                for (int rr = 0; rr < E; ++rr) p[chunk_off<YL, PAD>(rr)] = x[rr];
            }
        });
    }
}

// LDS words of a row pass workgroup.
__host__ __device__ inline size_t row_pass_lds_words(int logg, int loge, int ng) {
    return (size_t)ng * (((size_t)1 << logg) + ((size_t)1 << (logg - row_pad(logg, loge))));
}

// ---------------------------------------------------------------------------------------------
// Column passes: levels 0..a-1 (a = LOGG) down NG adjacent columns c = blk NG + q of the
// 2^a x C matrix (element (k, c) at k C + c).  Thread t -> column q = t mod NG, sub-transform
// s = t / NG.  LDS: column q at q SQ, element k at k + (k >> LOGE), with the column stride SQ
// = 2^a + 2^a / E + pad = 32 / NG (mod 32; odd for NG >= 32).  gfx950 serves a wave's 32-bit LDS
// accesses half a wave at a time over 32 banks (the row pass's pattern, conflict-free in that
// model and not in a 64-bank one, measures no conflicts; a stride of 64 / NG mod 64 measured
// conflicts in the column passes): the 32 lanes of each half (NG consecutive columns by 32 / NG
// consecutive sub-transforms) then fall in distinct banks in every chunk.

__host__ __device__ inline uint32_t col_pass_stride(int logg, int loge, int ng) {
    const uint32_t w = (1u << logg) + (1u << (logg - loge));
    const uint32_t want = (ng >= 32) ? 1u : 32u / (uint32_t)ng;
    return w + ((want - (w & 31u)) & 31u);
}

__host__ __device__ inline size_t col_pass_lds_words(int logg, int loge, int ng) {
    return (size_t)ng * col_pass_stride(logg, loge, ng);
}

// Digits per workgroup of the normalisation (256 threads x 8): its carry-outs wc[] enter the next
// range's first 8 digits.
constexpr uint32_t NORM_RANGE = 2048;

// Adds the pending normalisation carries (carry_wg_body's work) into the digits a column workgroup
// has just loaded, so that the separate carry kernel can be skipped before a squaring.  Range
// g = i / 2048 starts at digit i = 2048 g, row i / C, column i mod C: with C >= 2048 at the columns
// k 2048 of every row, with C < 2048 at column 0 of every (2048 / C)-th row; its first 8 digits are
// columns c0..c0+7 of that row, held by lanes q < 8 of the workgroup whose first column c0 is the
// range start.  The workgroup's threads take one such row each (rows r = thread, thread + T, ...):
// re-read its 8 digits, add the carry of range g - 1 (-carry of the last range for g = 0:
// x^N = -1) exactly as carry_wg_body does, and leave them in LDS for lanes q < 8.  The patch area
// (2^LOGG x 8 words) is released by the barrier that precedes the pass's first LDS exchange.
template <int LOGE, int LOGG>
__device__ __forceinline__ bool col_absorb_carries(int32_t (&dg)[1 << LOGE], const int32_t *__restrict__ src,
                                                   const int64_t *__restrict__ wc, uint32_t b, uint32_t C,
                                                   uint32_t c0, uint32_t q, uint32_t base, int YL, uint32_t nrange,
                                                   uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    const bool starts = (C >= NORM_RANGE) ? (c0 % NORM_RANGE) == 0 : c0 == 0;
    if (!starts) return false;  // uniform over the workgroup
    const uint32_t every = (C >= NORM_RANGE) ? 1u : NORM_RANGE / C;
    int32_t *P = reinterpret_cast<int32_t *>(lds);
    const double binv = 1.0 / (double)b;
    for (uint32_t r = threadIdx.x; r < (1u << LOGG); r += blockDim.x) {
        if (r % every) continue;
        const uint32_t g = (uint32_t)(((size_t)r * C + c0) / NORM_RANGE);
        const int64_t cin = (g == 0) ? -wc[nrange - 1] : wc[g - 1];
        const int32_t *p = src + (size_t)r * C + c0;
        const int4 v0 = *reinterpret_cast<const int4 *>(p), v1 = *reinterpret_cast<const int4 *>(p + 4);
        int32_t e8[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        if (cin != 0) absorb_carry<8>(e8, cin, b, binv);
        *reinterpret_cast<int4 *>(P + r * 8) = make_int4(e8[0], e8[1], e8[2], e8[3]);
        *reinterpret_cast<int4 *>(P + r * 8 + 4) = make_int4(e8[4], e8[5], e8[6], e8[7]);
    }
    __syncthreads();
    if (q < 8) {
#pragma unroll
        for (int rr = 0; rr < E; ++rr) {
            const uint32_t r = base + ((uint32_t)rr << YL);
            if (r % every == 0) dg[rr] = P[r * 8 + q];
        }
    }
    return true;
}

// As col_absorb_carries for the two-digit-limb transform (gfe_limb.h): range g = i / 1024 of the
// limbs starts at limb 1024 g, and the boundary carries of range g - 1 (wl[2 (g-1)], wl[2 (g-1) + 1],
// the last range's negated for g = 0: B^M = -1) enter its first 8 digits (carry_limb_body's work),
// i.e. the limbs of columns c0..c0+3 of that row, held by lanes q < 4 as digit pairs d0, d1.
template <int LOGE, int LOGG>
__device__ __forceinline__ bool col_absorb_limb_carries(int32_t (&d0)[1 << LOGE], int32_t (&d1)[1 << LOGE],
                                                        const int32_t *__restrict__ src, const int64_t *__restrict__ wl,
                                                        uint32_t b, uint32_t C, uint32_t c0, uint32_t q, uint32_t base,
                                                        int YL, uint32_t nrange, uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    const bool starts = (C >= LIMB_RANGE) ? (c0 % LIMB_RANGE) == 0 : c0 == 0;
    if (!starts) return false;  // uniform over the workgroup
    const uint32_t every = (C >= LIMB_RANGE) ? 1u : LIMB_RANGE / C;
    int32_t *P = reinterpret_cast<int32_t *>(lds);
    const double binv = 1.0 / (double)b;
    for (uint32_t r = threadIdx.x; r < (1u << LOGG); r += blockDim.x) {
        if (r % every) continue;
        const uint32_t g = (uint32_t)(((size_t)r * C + c0) / LIMB_RANGE);
        const size_t w = 2 * (size_t)((g == 0) ? nrange - 1 : g - 1);
        const int64_t e0 = (g == 0) ? -wl[w] : wl[w], e1 = (g == 0) ? -wl[w + 1] : wl[w + 1];
        const int32_t *p = src + 2 * ((size_t)r * C + c0);
        const int4 v0 = *reinterpret_cast<const int4 *>(p), v1 = *reinterpret_cast<const int4 *>(p + 4);
        int32_t e8[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        if (e0 != 0 || e1 != 0) absorb_carry2<8>(e8, e0, e1, b, binv);
        *reinterpret_cast<int4 *>(P + r * 8) = make_int4(e8[0], e8[1], e8[2], e8[3]);
        *reinterpret_cast<int4 *>(P + r * 8 + 4) = make_int4(e8[4], e8[5], e8[6], e8[7]);
    }
    __syncthreads();
    if (q < 4) {
#pragma unroll
        for (int rr = 0; rr < E; ++rr) {
            const uint32_t r = base + ((uint32_t)rr << YL);
            if (r % every == 0) {
                d0[rr] = P[r * 8 + 2 * q];
                d1[rr] = P[r * 8 + 2 * q + 1];
            }
        }
    }
    return true;
}

// The range carries of the fused limb normalisation (norm_limb_rows), for a column workgroup whose
// NG columns are one range of every row (NG = 2^rlog, the plan that wrote them): range k C / NG + blk
// of row k starts at the group's first column, its carry ls[.] = (u0, u1, v, 0) goes to the digit
// pair of lane q = 0 as (u0, u1) and to lane q = 1 as (v, 0).  The workgroup stages one carry per row
// in LDS (the loads go out with the digits'), and after one barrier the two lanes add theirs, read
// from LDS, to every row they hold.  The staging area (2^LOGG x 16 bytes) is released by the barrier
// that precedes the pass's first LDS exchange.
template <int LOGE, int LOGG>
__device__ __forceinline__ bool col_add_range_carries(int32_t (&d0)[1 << LOGE], int32_t (&d1)[1 << LOGE],
                                                      const int4 *__restrict__ ls, int rlog, uint32_t C, uint32_t c0,
                                                      uint32_t q, uint32_t base, int YL, uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    int4 *P = reinterpret_cast<int4 *>(lds);
    const uint32_t G = C >> rlog, g = c0 >> rlog;
    for (uint32_t r = threadIdx.x; r < (1u << LOGG); r += blockDim.x) P[r] = ls[(size_t)r * G + g];
    __syncthreads();
    if (q < 2) {
        const int2 *P2 = reinterpret_cast<const int2 *>(lds) + q;
#pragma unroll
        for (int rr = 0; rr < E; ++rr) {
            const int2 v = P2[2 * (base + ((uint32_t)rr << YL))];
            d0[rr] += v.x;
            d1[rr] += v.y;
        }
    }
    return true;
}

// Forward levels 0..a-1 of one prime: signed digits `src` -> residues `dst` (both N values).  With
// `wc`, the carries of the last normalisation are added on the way in (col_absorb_carries).  LIMB:
// element i is the limb d_2i + b d_2i+1 of two digits (the two-digit-limb transform, gfe_limb.h),
// formed mod p with bm = b 2^32 mod p; `wc` then holds the limb normalisation's boundary carries
// (col_absorb_limb_carries), or `ls` the carries of the fused limb normalisation into its ranges of
// 2^rlog limbs (norm_limb_rows): the first two limbs of range r take (ls[r].x, ls[r].y) and
// (ls[r].z, 0) on their digit pairs, unnormalised (the transform's exactness bound for these
// plans counts those limbs at twice the size, limb_fused_ok on the host).
template <class M, int LOGE, int LOGG, bool LIMB = false>
__device__ __forceinline__ void col_fwd_pass(const M &md, const int32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                             const uint32_t *__restrict__ tw, uint32_t C, int NG, unsigned blk,
                                             uint32_t *lds, const int64_t *__restrict__ wc = nullptr,
                                             uint32_t b = 0, uint32_t bm = 0, const int4 *__restrict__ ls = nullptr,
                                             int rlog = 0) {
    using GM = Geo<LOGG, LOGE>;
    constexpr int E = 1 << LOGE;
    const uint32_t t = threadIdx.x, q = t & (uint32_t)(NG - 1), s = t / (uint32_t)NG;
    const uint32_t c = blk * (uint32_t)NG + q;
    uint32_t *L = lds + q * col_pass_stride(LOGG, LOGE, NG);
    uint32_t x[E], w[E];
    bool patched = false;  // the carry patch was read from LDS (uniform over the workgroup)
    static_for<0, GM::NCH>([&](auto cc) {
        constexpr int ci = decltype(cc)::value;
        constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
        const uint32_t hi = s >> YL, base = (hi << (LOGG - lam0)) + (s & ((1u << YL) - 1u));
        load_fwd_tw<LOGE>(w, tw, (1u << lam0) + hi, NE);
        if constexpr (ci == 0 && LIMB) {
            const uint32_t off = (base * C + c) * 8u;
            const int2 *s2 = reinterpret_cast<const int2 *>(src);
            int32_t d0[E], d1[E];
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const int2 v = *at_bytes(s2 + (size_t)((uint32_t)rr << YL) * C, off);
                d0[rr] = v.x;
                d1[rr] = v.y;
            }
            if (wc)
                patched = col_absorb_limb_carries<LOGE, LOGG>(d0, d1, src, wc, b, C, blk * (uint32_t)NG, q, base, YL,
                                                              (uint32_t)(((size_t)C << LOGG) / LIMB_RANGE), lds);
            if (ls) patched = col_add_range_carries<LOGE, LOGG>(d0, d1, ls, rlog, C, blk * (uint32_t)NG, q, base, YL, lds);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = md.add(to_res(d0[rr], md.p), md.mulw(to_res(d1[rr], md.p), bm));
        } else if constexpr (ci == 0) {
            const uint32_t off = (base * C + c) * 4u;
            int32_t dg[E];
#pragma unroll
            for (int rr = 0; rr < E; ++rr) dg[rr] = *at_bytes(src + (size_t)((uint32_t)rr << YL) * C, off);
            if (wc)
                patched = col_absorb_carries<LOGE, LOGG>(dg, src, wc, b, C, blk * (uint32_t)NG, q, base, YL,
                                                         (uint32_t)(((size_t)C << LOGG) / NORM_RANGE), lds);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = to_res(dg[rr], md.p);
        } else {
            __syncthreads();
            const uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = p[chunk_off<YL, LOGE>(rr)];
        }
        fwd_chunk<M, LOGE, NE>(md, x, w);
        if constexpr (ci == GM::NCH - 1) {
            const uint32_t off = (base * C + c) * 4u;
#pragma unroll
            for (int rr = 0; rr < E; ++rr) *at_bytes(dst + (size_t)((uint32_t)rr << YL) * C, off) = x[rr];
        } else {
            if (ci > 0 || patched) __syncthreads();  // chunk 0 follows no LDS read but the carry patch's
            uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) p[chunk_off<YL, LOGE>(rr)] = x[rr];
        }
    });
}

// The limb transform's forward column pass for the four primes in one workgroup (ks_col_fwd_limb_all):
// the group's 2^LOGG x NG digit pairs are loaded once, with wide loads along the rows, into LDS as
// D[k NG + j] = limb k C + blk NG + j, the range carries of the fused limb normalisation added to each
// row's first two limbs on the way (`ls`, when not null: the carry into range k C / NG + blk is
// ls[.] = (u0, u1, v, 0), (u0, u1) onto limb 0's digit pair and v onto limb 1's first digit); each
// prime's threads then read their column elements from LDS.  S = NG 2^LOGG / blockDim.x limbs per
// thread, S even.  Load i of thread t moves 16-byte word f = i blockDim.x + t of D (row f / (NG/2)):
// each wave's loads read whole runs of rows and its LDS stores are 16 consecutive bytes per lane.
template <int S>
__device__ __forceinline__ void stage_limb_digits(const int32_t *__restrict__ src, const int4 *__restrict__ ls,
                                                  uint32_t C, uint32_t NG, unsigned blk, int2 *D) {
    static_assert(S >= 2 && S % 2 == 0, "whole int4 loads of two limbs");
    const uint32_t NT = blockDim.x, tid = threadIdx.x;
    const uint32_t hl = 31u - (uint32_t)__clz((int)(NG / 2));  // NG / 2 int4 per row, a power of two
    const int4 *g = reinterpret_cast<const int4 *>(src) + (size_t)blk * (NG / 2);
    int4 v[S / 2], cy[S / 2];
#pragma unroll
    for (int i = 0; i < S / 2; ++i) {
        const uint32_t f = (uint32_t)i * NT + tid, k = f >> hl, m = f & ((1u << hl) - 1u);
        v[i] = g[(size_t)k * (C / 2) + m];
        cy[i] = (ls && m == 0) ? ls[(size_t)k * (C / NG) + blk] : make_int4(0, 0, 0, 0);
    }
    int4 *o = reinterpret_cast<int4 *>(D);
#pragma unroll
    for (int i = 0; i < S / 2; ++i) {
        v[i].x += cy[i].x;
        v[i].y += cy[i].y;
        v[i].z += cy[i].z;
        o[(uint32_t)i * NT + tid] = v[i];
    }
}

// As col_fwd_pass (LIMB) for prime-group thread t on its own exchange region L, the first chunk's
// digit pairs read from the staged D.  D overlaps the exchange regions: the first chunk's LDS
// writes wait for a workgroup barrier after every group's reads of D.
template <int LOGE, int LOGG>
__device__ __forceinline__ void col_fwd_limb_staged(const ModS &md, const int2 *D, uint32_t *__restrict__ dst,
                                                    const uint32_t *__restrict__ tw, uint32_t C, int NG, unsigned blk,
                                                    uint32_t *L, uint32_t t, uint32_t bm) {
    using GM = Geo<LOGG, LOGE>;
    constexpr int E = 1 << LOGE;
    const uint32_t q = t & (uint32_t)(NG - 1), s = t / (uint32_t)NG;
    const uint32_t c = blk * (uint32_t)NG + q;
    L += q * col_pass_stride(LOGG, LOGE, NG);
    uint32_t x[E], w[E];
    static_for<0, GM::NCH>([&](auto cc) {
        constexpr int ci = decltype(cc)::value;
        constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
        const uint32_t hi = s >> YL, base = (hi << (LOGG - lam0)) + (s & ((1u << YL) - 1u));
        load_fwd_tw<LOGE>(w, tw, (1u << lam0) + hi, NE);
        if constexpr (ci == 0) {
            const int2 *p = D + (size_t)base * NG + q;
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const int2 v = p[(size_t)((uint32_t)rr << YL) * NG];
                x[rr] = md.add(to_res(v.x, md.p), md.mulw(to_res(v.y, md.p), bm));
            }
        } else {
            __syncthreads();
            const uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = p[chunk_off<YL, LOGE>(rr)];
        }
        fwd_chunk<ModS, LOGE, NE>(md, x, w);
        if constexpr (ci == GM::NCH - 1) {
            const uint32_t off = (base * C + c) * 4u;
#pragma unroll
            for (int rr = 0; rr < E; ++rr) *at_bytes(dst + (size_t)((uint32_t)rr << YL) * C, off) = x[rr];
        } else {
            __syncthreads();  // chunk 0: after every group's reads of D
            uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) p[chunk_off<YL, LOGE>(rr)] = x[rr];
        }
    });
}

// Inverse levels a-1..0 of one prime, in place on the residues `data`.
template <class M, int LOGE, int LOGG>
__device__ __forceinline__ void col_inv_pass(const M &md, uint32_t *__restrict__ data, const uint32_t *__restrict__ tw,
                                             uint32_t C, int NG, unsigned blk, uint32_t *lds) {
    using GM = Geo<LOGG, LOGE>;
    constexpr int E = 1 << LOGE;
    const uint32_t t = threadIdx.x, q = t & (uint32_t)(NG - 1), s = t / (uint32_t)NG;
    const uint32_t c = blk * (uint32_t)NG + q;
    uint32_t *L = lds + q * col_pass_stride(LOGG, LOGE, NG);
    uint32_t x[E], w[E];
    static_for_down<0, GM::NCH>([&](auto cc) {
        constexpr int ci = decltype(cc)::value;
        constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
        const uint32_t hi = s >> YL, base = (hi << (LOGG - lam0)) + (s & ((1u << YL) - 1u));
        load_inv_tw<LOGE>(w, tw, (1u << lam0) + hi, lam0, NE);
        if constexpr (ci == GM::NCH - 1) {
            const uint32_t off = (base * C + c) * 4u;
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = *at_bytes(data + (size_t)((uint32_t)rr << YL) * C, off);
        } else {
            __syncthreads();
            const uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = p[chunk_off<YL, LOGE>(rr)];
        }
        inv_chunk<M, LOGE, NE>(md, x, w);
        if constexpr (ci == 0) {
            const uint32_t off = (base * C + c) * 4u;
#pragma unroll
            for (int rr = 0; rr < E; ++rr) *at_bytes(data + (size_t)((uint32_t)rr << YL) * C, off) = x[rr];
        } else {
            if constexpr (ci < GM::NCH - 1) __syncthreads();  // the first chunk follows no LDS read
            uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) p[chunk_off<YL, LOGE>(rr)] = x[rr];
        }
    });
}

// Inverse levels a-1..0 of one prime as col_inv_pass, run by thread t of a group of threads that
// has its own exchange region L (a workgroup holding several primes' column groups side by side),
// with the output left in LDS: element (k, q) of the NG columns at out[k NG + q], row-major, for a
// normalisation that walks the rows.  `out` may overlap any group's exchange region: a workgroup
// barrier separates every group's last exchange reads from these writes, so all the groups of the
// workgroup must call this with the same LOGE, LOGG and NG.
template <class M, int LOGE, int LOGG>
__device__ __forceinline__ void col_inv_to_lds(const M &md, const uint32_t *__restrict__ data, const uint32_t *__restrict__ tw,
                                               uint32_t C, int NG, unsigned blk, uint32_t *L, uint32_t t, uint32_t *out) {
    using GM = Geo<LOGG, LOGE>;
    constexpr int E = 1 << LOGE;
    const uint32_t q = t & (uint32_t)(NG - 1), s = t / (uint32_t)NG;
    const uint32_t c = blk * (uint32_t)NG + q;
    L += q * col_pass_stride(LOGG, LOGE, NG);
    uint32_t x[E], w[E];
    static_for_down<0, GM::NCH>([&](auto cc) {
        constexpr int ci = decltype(cc)::value;
        constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
        const uint32_t hi = s >> YL, base = (hi << (LOGG - lam0)) + (s & ((1u << YL) - 1u));
        load_inv_tw<LOGE>(w, tw, (1u << lam0) + hi, lam0, NE);
        if constexpr (ci == GM::NCH - 1) {
            const uint32_t off = (base * C + c) * 4u;
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = *at_bytes(data + (size_t)((uint32_t)rr << YL) * C, off);
        } else {
            __syncthreads();
            const uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = p[chunk_off<YL, LOGE>(rr)];
        }
        inv_chunk<M, LOGE, NE>(md, x, w);
        if constexpr (ci == 0) {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) out[(base + ((uint32_t)rr << YL)) * (uint32_t)NG + q] = x[rr];
        } else {
            if constexpr (ci < GM::NCH - 1) __syncthreads();  // the first chunk follows no LDS read
            uint32_t *p = L + base + (base >> LOGE);
#pragma unroll
            for (int rr = 0; rr < E; ++rr) p[chunk_off<YL, LOGE>(rr)] = x[rr];
        }
    });
}

// Instantiated shapes of the compile-time passes (host and device agree on these ranges).
constexpr int SPEC_ROW_MIN = 6, SPEC_ROW_MAX = 14;   // row levels
constexpr int SPEC_COL_MIN = 5, SPEC_COL_MAX = 11;   // column levels

__host__ __device__ inline bool spec_row_ok(int logg, int loge) {
    return logg >= SPEC_ROW_MIN && logg <= SPEC_ROW_MAX && (loge == 4 || loge == 5) && logg >= loge;
}
__host__ __device__ inline bool spec_col_ok(int logg, int loge, int ng) {
    return logg >= SPEC_COL_MIN && logg <= SPEC_COL_MAX && loge >= 3 && loge <= 5 && logg >= loge && ng >= 8 &&
           ng <= 64;
}

// The limb transform's fused inverse column pass and normalisation (ks_col_inv_norm_limb) runs the
// four primes' groups of T = ng 2^logg / 2^loge threads side by side: 4 T threads, whole waves per
// prime, within the workgroup limit (512 at 32 points per thread), so logg - loge <= 5.  The
// normalisation's threads of one row (ng / (2^loge / 4) of them) must share a 16-lane DPP row.
// LDS: the four exchange regions, reused for the rows (4 ng 2^logg words).
constexpr int LIMB_FUSED_MAX_SPAN = 5;

__host__ __device__ inline size_t limb_col_norm_lds_bytes(int logg, int loge, int ng) {
    const size_t ex = 4 * col_pass_lds_words(logg, loge, ng);
    const size_t rows = 4 * ((size_t)ng << logg);
    return 4 * (ex > rows ? ex : rows);
}

// LDS of ks_col_fwd_limb_all: the staged digit pairs (8 bytes per limb) or the four exchange regions.
__host__ __device__ inline size_t limb_col_fwd_all_lds_bytes(int logg, int loge, int ng) {
    const size_t ex = 16 * col_pass_lds_words(logg, loge, ng);
    const size_t st = 8 * ((size_t)ng << logg);
    return ex > st ? ex : st;
}

__host__ __device__ inline bool limb_col_norm_ok(int logg, int loge, int ng) {
    if (!spec_col_ok(logg, loge, ng) || logg - loge > LIMB_FUSED_MAX_SPAN) return false;
    const int T = (ng << logg) >> loge;
    return T % 64 == 0 && 4 * T <= (loge >= 5 ? 512 : 1024) && ng / ((1 << loge) / 4) <= 16 &&
           limb_col_norm_lds_bytes(logg, loge, ng) <= 160 * 1024;
}

}  // namespace gfe
