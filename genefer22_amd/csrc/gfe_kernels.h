// gfe_kernels.h -- gfx950 kernels of the squaring engine.
//
// Transform (per prime p, length N = 2^n, psi a primitive 2N-th root of unity mod p):
// negacyclic number-theoretic transform, Cooley-Tukey forward with roots
// Psi[m + j] = psi^bitrev_n(m + j) at level l (m = 2^l, block j), pointwise product in the
// bit-reversed domain, Gentleman-Sande inverse (u, v) -> (u + v, (u - v) Psi[m + j]^-1), computed
// as (u + v, (v - u) Psi[2m - 1 - j]) because Psi[m + j]^-1 = -Psi[2m - 1 - j] (psi^N = -1).
// The reference factors x^N + 1 the same way down to degree-1 residues (ocl/kernel.cl:235-273,
// "z-transform"); here the last level is also split and the product is pointwise.
//
// Decomposition per squaring (levels 0..a-1 = "columns", a..n-1 = "rows", C = N / 2^a):
//   col_fwd  : digits -> residues, levels 0..a-1 on each column c (elements c + i C), NG columns
//              per workgroup;
//   row      : levels a..n-1 on each row (C contiguous elements), pointwise square / product,
//              inverse levels n-1..a;
//   col_inv  : inverse levels a-1..0;
//   norm     : scale, Garner CRT, balanced carry over S consecutive digits per thread; carries
//              cross threads through LDS;
//   carry_wg : adds each workgroup's carry-out into the next one (negated across x^N = -1).
// Inside a pass, a thread holds E = 2^LOGE elements of one sub-transform in registers and runs
// LOGE levels per chunk; chunks exchange through LDS.
//
// Each pass is a __device__ body; the single-engine kernels below (grid.y = prime) and the batch
// kernels of gfe_batch.hip (grid.z = candidate slot) are thin wrappers around the bodies.
#pragma once

#include "gfe_device.h"

namespace gfe {

constexpr int MAX_RNS = 3;

// Primes of one transform: the two or three of the prime set, or the four of the two-digit-limb
// transform (gfe_limb.h).
struct PrimeArgs {
    uint32_t p[4];
    uint32_t q[4];
};

// Forward CT levels mu = 0..ne-1 of a 2^LOGE-point sub-transform held in x[].  twf(mu, bi) is
// the twiddle (Montgomery form) of level mu, register block bi; its Montgomery quotient factor
// w q is formed once per twiddle and shared by the block's butterflies (ModS::mulp).
template <class M, int LOGE, class TWF>
__device__ __forceinline__ void fwd_levels_t(const M &md, uint32_t (&x)[1 << LOGE], const TWF &twf, int ne) {
    constexpr int E = 1 << LOGE;
#pragma unroll
    for (int mu = 0; mu < LOGE; ++mu) {
        if (mu < ne) {
            const int h = E >> (mu + 1);
#pragma unroll
            for (int bi = 0; bi < (1 << mu); ++bi) {
                const uint32_t w = twf(mu, bi), wq = w * md.q;
#pragma unroll
                for (int i = 0; i < h; ++i) {
                    const int i0 = bi * 2 * h + i, i1 = i0 + h;
                    const uint32_t u = x[i0], v = md.mulp(x[i1], w, wq);
                    x[i0] = md.add(u, v);
                    x[i1] = md.sub(u, v);
                }
            }
        }
    }
}

// A chunk's twiddles in registers: level mu (< ne) occupies w[2^mu .. 2^(mu+1)), one aligned run
// of 2^mu table entries starting at `start(mu)`, loaded as 16-byte vectors where the run allows.
// Issued before a chunk's LDS exchange, the loads overlap the barrier.
template <int LOGE, class START>
__device__ __forceinline__ void load_chunk_tw(uint32_t (&w)[1 << LOGE], const uint32_t *__restrict__ tw, const START &start,
                                              int ne) {
#pragma unroll
    for (int mu = 0; mu < LOGE; ++mu) {
        if (mu < ne) {
            const uint32_t *p = tw + start(mu);
            const int c = 1 << mu;
            if (c >= 4) {
#pragma unroll
                for (int i = 0; i < c; i += 4) {
                    const uint4 v = *reinterpret_cast<const uint4 *>(p + i);
                    w[c + i] = v.x; w[c + i + 1] = v.y; w[c + i + 2] = v.z; w[c + i + 3] = v.w;
                }
            } else if (c == 2) {
                const uint2 v = *reinterpret_cast<const uint2 *>(p);
                w[2] = v.x; w[3] = v.y;
            } else {
                w[1] = p[0];
            }
        }
    }
}

// Forward twiddles of a chunk whose first-level block index is t0: level mu, block bi uses
// Psi[(t0 << mu) + bi].
template <int LOGE>
__device__ __forceinline__ void load_fwd_tw(uint32_t (&w)[1 << LOGE], const uint32_t *__restrict__ tw, uint32_t t0, int ne) {
    load_chunk_tw<LOGE>(w, tw, [&](int mu) { return t0 << mu; }, ne);
}

// Inverse twiddles: level mu, block bi uses Psi[3 2^(lvl0+mu) - 1 - (t0 << mu) - bi], i.e. entry
// 2^mu - 1 - bi of the aligned run from 3 2^(lvl0+mu) - (t0 + 1) 2^mu.
template <int LOGE>
__device__ __forceinline__ void load_inv_tw(uint32_t (&w)[1 << LOGE], const uint32_t *__restrict__ tw, uint32_t t0, int lvl0,
                                            int ne) {
    load_chunk_tw<LOGE>(w, tw, [&](int mu) { return (3u << (lvl0 + mu)) - ((t0 + 1u) << mu); }, ne);
}

template <class M, int LOGE>
__device__ __forceinline__ void fwd_levels_r(const M &md, uint32_t (&x)[1 << LOGE], const uint32_t (&w)[1 << LOGE], int ne) {
    fwd_levels_t<M, LOGE>(md, x, [&](int mu, int bi) { return w[(1 << mu) + bi]; }, ne);
}

// Inverse GS levels mu = ne-1..0 (the exact inverse of fwd_levels_t without the 1/2 factors);
// twf(mu, bi) is the inverse twiddle of level mu, block bi.
template <class M, int LOGE, class TWF>
__device__ __forceinline__ void inv_levels_t(const M &md, uint32_t (&x)[1 << LOGE], const TWF &twf, int ne) {
    constexpr int E = 1 << LOGE;
#pragma unroll
    for (int mu = LOGE - 1; mu >= 0; --mu) {
        if (mu < ne) {
            const int h = E >> (mu + 1);
#pragma unroll
            for (int bi = 0; bi < (1 << mu); ++bi) {
                const uint32_t w = twf(mu, bi), wq = w * md.q;
#pragma unroll
                for (int i = 0; i < h; ++i) {
                    const int i0 = bi * 2 * h + i, i1 = i0 + h;
                    const uint32_t u = x[i0], v = x[i1];
                    x[i0] = md.add(u, v);
                    x[i1] = md.mulp(md.sub(v, u), w, wq);
                }
            }
        }
    }
}

template <class M, int LOGE>
__device__ __forceinline__ void inv_levels_r(const M &md, uint32_t (&x)[1 << LOGE], const uint32_t (&w)[1 << LOGE], int ne) {
    inv_levels_t<M, LOGE>(md, x, [&](int mu, int bi) { return w[(2 << mu) - 1 - bi]; }, ne);
}

// Chunk schedule of a g-level pass: a partial first chunk of r = g mod LOGE levels (if r > 0),
// then full chunks.  Every chunk's element set spans LOGE levels starting at lam0.
struct Chunking {
    int g, r, nch;
    __device__ __forceinline__ Chunking(int g_, int loge) : g(g_), r(g_ % loge), nch((g_ + loge - 1) / loge) {}
    __device__ __forceinline__ int lam0(int ci, int loge) const { return (r == 0) ? ci * loge : (ci == 0 ? 0 : r + (ci - 1) * loge); }
    __device__ __forceinline__ int ne(int ci, int loge) const { return (r != 0 && ci == 0) ? r : loge; }
};

// ---------------------------------------------------------------------------------------------
// Column passes.  Thread t -> column q = t % NG, sub-transform s = t / NG of the column block
// starting at column blk * NG.  LDS element (q, k) at ((k + (k >> LOGE)) * NG + q).

// Forward levels 0..a-1 of one prime; x[] holds the thread's chunk-0 elements (column c0 + q,
// rows s + rr 2^(g - LOGE)); the residues are stored to `dst` (N uint32).
template <class M, int LOGE>
__device__ __forceinline__ void col_fwd_chunks(const M &md, uint32_t (&x)[1 << LOGE], uint32_t *__restrict__ dst,
                                               const uint32_t *__restrict__ tw, int n, int g, int NG, unsigned blk,
                                               uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    const uint32_t C = 1u << (n - g);
    const uint32_t c0 = blk * NG;
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const Chunking ch(g, LOGE);
    for (int ci = 0; ci < ch.nch; ++ci) {
        const int lam0 = ch.lam0(ci, LOGE), ne = ch.ne(ci, LOGE);
        const int ylog = g - lam0 - LOGE;
        const int hi = s >> ylog, lo = s & ((1 << ylog) - 1);
        const int base = (hi << (g - lam0)) + lo;
        uint32_t w[E];
        load_fwd_tw<LOGE>(w, tw, (1u << lam0) + hi, ne);
        if (ci > 0) {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = lds[(k + (k >> LOGE)) * NG + q];
            }
        }
        fwd_levels_r<M, LOGE>(md, x, w, ne);
        if (ci == ch.nch - 1) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                dst[(size_t)k * C + c0 + q] = x[rr];
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                lds[(k + (k >> LOGE)) * NG + q] = x[rr];
            }
        }
    }
}

// The thread's chunk-0 digits of column c0 + q (rows s + rr 2^(g - LOGE)).
template <int LOGE>
__device__ __forceinline__ void col_load_digits(int32_t (&dg)[1 << LOGE], const int32_t *__restrict__ src, int n, int g,
                                                int NG, unsigned blk) {
    constexpr int E = 1 << LOGE;
    const uint32_t C = 1u << (n - g);
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const int ylog = g - LOGE;
#pragma unroll
    for (int rr = 0; rr < E; ++rr) dg[rr] = src[(size_t)(s + (rr << ylog)) * C + blk * NG + q];
}

// Forward levels 0..a-1 from signed digits `src` (N int32) to residues `dst` (N uint32).
template <class M, int LOGE>
__device__ __forceinline__ void col_fwd_body(const M &md, const int32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                             const uint32_t *__restrict__ tw, int n, int g, int NG, unsigned blk,
                                             uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    int32_t dg[E];
    col_load_digits<LOGE>(dg, src, n, g, NG, blk);
    uint32_t x[E];
#pragma unroll
    for (int rr = 0; rr < E; ++rr) x[rr] = to_res(dg[rr], md.p);
    col_fwd_chunks<M, LOGE>(md, x, dst, tw, n, g, NG, blk, lds);
}

// Inverse levels a-1..0 of one prime from the residues `src`; x[] receives the thread's final
// elements (column c0 + q, rows s + rr 2^(g - LOGE)).
template <class M, int LOGE>
__device__ __forceinline__ void col_inv_chunks(const M &md, uint32_t (&x)[1 << LOGE], const uint32_t *__restrict__ src,
                                               const uint32_t *__restrict__ tw, int n, int g, int NG, unsigned blk,
                                               uint32_t *lds, bool preloaded = false) {
    constexpr int E = 1 << LOGE;
    const uint32_t C = 1u << (n - g);
    const uint32_t c0 = blk * NG;
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const Chunking ch(g, LOGE);
    for (int ci = ch.nch - 1; ci >= 0; --ci) {
        const int lam0 = ch.lam0(ci, LOGE), ne = ch.ne(ci, LOGE);
        const int ylog = g - lam0 - LOGE;
        const int hi = s >> ylog, lo = s & ((1 << ylog) - 1);
        const int base = (hi << (g - lam0)) + lo;
        uint32_t w[E];
        load_inv_tw<LOGE>(w, tw, (1u << lam0) + hi, lam0, ne);
        if (ci == ch.nch - 1) {
            if (!preloaded) {
#pragma unroll
                for (int rr = 0; rr < E; ++rr) {
                    const uint32_t k = base + (rr << ylog);
                    x[rr] = src[(size_t)k * C + c0 + q];
                }
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = lds[(k + (k >> LOGE)) * NG + q];
            }
        }
        inv_levels_r<M, LOGE>(md, x, w, ne);
        if (ci > 0) {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                lds[(k + (k >> LOGE)) * NG + q] = x[rr];
            }
        }
    }
}

// Inverse levels a-1..0, in place on the residues `dst`.
template <class M, int LOGE>
__device__ __forceinline__ void col_inv_body(const M &md, uint32_t *__restrict__ dst, const uint32_t *__restrict__ tw,
                                             int n, int g, int NG, unsigned blk, uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    uint32_t x[E];
    col_inv_chunks<M, LOGE>(md, x, dst, tw, n, g, NG, blk, lds);
    const uint32_t C = 1u << (n - g);
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const int ylog = g - LOGE;
#pragma unroll
    for (int rr = 0; rr < E; ++rr) dst[(size_t)(s + (rr << ylog)) * C + blk * NG + q] = x[rr];
}

// ---------------------------------------------------------------------------------------------
// Pipelined column passes.  A persistent workgroup transforms the column blocks blk = first,
// first + stride, ... (< C / NG) of one prime.  The twiddles of levels 0..a-1 -- the table's first
// 2^a entries, shared by every column -- are copied into LDS once, and the next block's elements
// are loaded while the current block is transformed, so that memory traffic overlaps the
// butterflies.  (Vector-memory loads complete in issue order: a twiddle load inside the loop
// would wait for the prefetch, hence the LDS copy.)  LDS: [0, 2^a) twiddles, then the exchange
// buffer of col_fwd_chunks / col_inv_chunks.

// The thread's elements of the first inverse chunk (the last forward chunk) of column block blk.
template <int LOGE>
__device__ __forceinline__ void col_load_inv_first(uint32_t (&x)[1 << LOGE], const uint32_t *__restrict__ src, int n,
                                                   int g, int NG, unsigned blk) {
    constexpr int E = 1 << LOGE;
    const uint32_t C = 1u << (n - g);
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const Chunking ch(g, LOGE);
    const int lam0 = ch.lam0(ch.nch - 1, LOGE);
    const int ylog = g - lam0 - LOGE;
    const int hi = s >> ylog, lo = s & ((1 << ylog) - 1);
    const int base = (hi << (g - lam0)) + lo;
#pragma unroll
    for (int rr = 0; rr < E; ++rr) x[rr] = src[(size_t)(base + (rr << ylog)) * C + blk * NG + q];
}

__device__ __forceinline__ uint32_t *col_twiddles_to_lds(const uint32_t *__restrict__ tw, int g, uint32_t *lds) {
    for (unsigned i = threadIdx.x; i < (1u << g); i += blockDim.x) lds[i] = tw[i];
    __syncthreads();
    return lds + (1u << g);
}

template <class M, int LOGE>
__device__ __forceinline__ void col_fwd_loop(const M &md, const int32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                             const uint32_t *__restrict__ tw, int n, int g, int NG, unsigned first,
                                             unsigned stride, uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    uint32_t *lx = col_twiddles_to_lds(tw, g, lds);
    const unsigned nblk = (1u << (n - g)) / NG;
    if (first >= nblk) return;
    int32_t dg[E];
    col_load_digits<LOGE>(dg, src, n, g, NG, first);
    for (unsigned blk = first; blk < nblk; blk += stride) {
        const bool more = blk + stride < nblk;
        int32_t dn[E];
        if (more) col_load_digits<LOGE>(dn, src, n, g, NG, blk + stride);
        uint32_t x[E];
#pragma unroll
        for (int rr = 0; rr < E; ++rr) x[rr] = to_res(dg[rr], md.p);
        col_fwd_chunks<M, LOGE>(md, x, dst, lds, n, g, NG, blk, lx);
        if (more) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) dg[rr] = dn[rr];
        }
    }
}

template <class M, int LOGE>
__device__ __forceinline__ void col_inv_loop(const M &md, uint32_t *__restrict__ data, const uint32_t *__restrict__ tw,
                                             int n, int g, int NG, unsigned first, unsigned stride, uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    uint32_t *lx = col_twiddles_to_lds(tw, g, lds);
    const unsigned nblk = (1u << (n - g)) / NG;
    if (first >= nblk) return;
    const uint32_t C = 1u << (n - g);
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const int ylog = g - LOGE;
    uint32_t xc[E];
    col_load_inv_first<LOGE>(xc, data, n, g, NG, first);
    for (unsigned blk = first; blk < nblk; blk += stride) {
        const bool more = blk + stride < nblk;
        uint32_t xn[E];
        if (more) col_load_inv_first<LOGE>(xn, data, n, g, NG, blk + stride);
        col_inv_chunks<M, LOGE>(md, xc, data, lds, n, g, NG, blk, lx, true);
#pragma unroll
        for (int rr = 0; rr < E; ++rr) data[(size_t)(s + (rr << ylog)) * C + blk * NG + q] = xc[rr];
        if (more) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) xc[rr] = xn[rr];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Column passes over every prime of a candidate.  One workgroup holds NG columns (all 2^a rows)
// and runs the primes one after another through the same LDS: the forward pass reads each digit
// once, and the inverse pass ends with the scale, the Garner CRT and the carry in registers, so
// that one squaring moves 56 N bytes instead of 80 N (no separate normalisation pass).
//
// Carry.  Digit i = k C + c of row k is held by lane q = c - c0, so a row segment
// (k, c0 .. c0 + NG - 1) lies in NG consecutive lanes of one wavefront (8 <= NG <= 64).  The exact
// coefficient v is divided by b (balanced) and every quotient moves one lane up, `rounds` times,
// all lanes at once (cross-lane shuffles); the quotients leaving the segment's last lane are summed
// into the segment's carry-out, which fold_body adds into the next segment after the kernel (the
// reference hands carries over in the same two steps, within a work-item group and then across
// groups: ocl/kernel.cl:2366-2459).  `rounds` is the smallest count after which every quotient is
// bounded by 1 (host: carry_rounds), so the digits end with |d| <= b/2 + 1.

template <int RNS, class M, int LOGE>
__device__ __forceinline__ void col_fwd_all_body(const PrimeArgs &pr, const int32_t *__restrict__ src,
                                                 uint32_t *__restrict__ ws, const uint32_t *__restrict__ tw, int n,
                                                 int g, int NG, unsigned blk, uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    const size_t N = (size_t)1 << n;
    int32_t dg[E];
    col_load_digits<LOGE>(dg, src, n, g, NG, blk);
#pragma unroll
    for (int k = 0; k < RNS; ++k) {
        const M md{pr.p[k], pr.q[k]};
        uint32_t x[E];
#pragma unroll
        for (int rr = 0; rr < E; ++rr) x[rr] = to_res(dg[rr], md.p);
        col_fwd_chunks<M, LOGE>(md, x, ws + k * N, tw + k * N, n, g, NG, blk, lds);
    }
}

template <int RNS, class M, int LOGE>
__device__ __forceinline__ void col_inv_norm_body(const CrtConst &c, const uint32_t *__restrict__ ws,
                                                  const uint32_t *__restrict__ tw, int32_t *__restrict__ dst,
                                                  int64_t *__restrict__ seg, uint32_t b, bool dup, int rounds, int n,
                                                  int g, int NG, unsigned blk, uint32_t *lds) {
    constexpr int E = 1 << LOGE;
    const size_t N = (size_t)1 << n;
    const uint32_t C = 1u << (n - g);
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const int ylog = g - LOGE;
    const M m1{c.p[0], c.q[0]}, m2{c.p[1], c.q[1]}, m3{c.p[2], c.q[2]};
    uint32_t xs[RNS][E];
    col_inv_chunks<M, LOGE>(m1, xs[0], ws, tw, n, g, NG, blk, lds);
    col_inv_chunks<M, LOGE>(m2, xs[1], ws + N, tw + N, n, g, NG, blk, lds);
    if (RNS == 3) col_inv_chunks<M, LOGE>(m3, xs[RNS - 1], ws + 2 * N, tw + 2 * N, n, g, NG, blk, lds);
    const double binv = 1.0 / (double)b;
    const uint32_t segs_per_row = C / NG;
#pragma unroll
    for (int rr = 0; rr < E; ++rr) {
        const uint32_t k = s + (rr << ylog);
        const uint32_t r1 = m1.mulw(xs[0][rr], c.scale[0]);
        const uint32_t r2 = m2.mulw(xs[1][rr], c.scale[1]);
        const uint32_t r3 = (RNS == 3) ? m3.mulw(xs[RNS - 1][rr], c.scale[2]) : 0u;
        uint64_t x64;
        double xd;
        garner<RNS>(c, m1, m2, m3, r1, r2, r3, &x64, &xd);
        if (dup) { x64 <<= 1; xd *= 2.0; }
        int64_t qc;
        int32_t r = div_bal(x64, xd, b, binv, &qc);
        int64_t acc = 0;
        for (int j = 1; j < rounds; ++j) {
            int64_t qin = __shfl_up(qc, 1u, NG);
            if (q == 0) qin = 0;
            acc += qc;
            const int64_t w = (int64_t)r + qin;
            r = div_bal((uint64_t)w, (double)w, b, binv, &qc);
        }
        int64_t qin = __shfl_up(qc, 1u, NG);
        if (q == 0) qin = 0;
        acc += qc;
        dst[(size_t)k * C + blk * NG + q] = r + (int32_t)qin;
        if (q == NG - 1) seg[(size_t)k * segs_per_row + blk] = acc;
    }
}

// Adds segment j - 1's carry-out (the last segment's, negated, for j = 0: x^N = -1) into the
// digits of segment j (NG digits from j NG), digit by digit while a carry remains.
__device__ __forceinline__ void fold_body(int32_t *dc, const int64_t *seg, size_t nseg, int NG, uint32_t b, size_t j) {
    int64_t cin = (j == 0) ? -seg[nseg - 1] : seg[j - 1];
    if (cin == 0) return;
    int32_t *p = dc + j * (size_t)NG;
    const double binv = 1.0 / (double)b;
    for (int i = 0; i < NG - 1 && cin != 0; ++i) {
        const int64_t f = (int64_t)p[i] + cin;
        p[i] = div_bal((uint64_t)f, (double)f, b, binv, &cin);
    }
    p[NG - 1] += (int32_t)cin;
}

// ---------------------------------------------------------------------------------------------
// Row pass: levels a..n-1 on rows of C = 2^(n-a) contiguous elements, the pointwise square or
// product (Montgomery, factor 2^-32), then inverse levels n-1..a.  ROW_FWD stops after the
// forward levels (multiplicand).  Thread t -> row q = t / (C/E) of the row block starting at row
// blk * NG, sub-transform s = t % (C/E).  LDS element (q, k) at (q * (C + C/E) + k + (k >> LOGE)).

enum RowMode { ROW_SQUARE = 0, ROW_MUL = 1, ROW_FWD = 2 };

// Row twiddles.  The bit fields of m = 2^l + row 2^(l-a) + j (j < 2^(l-a)) are disjoint, so
// bitrev_n(m) is the sum of their bit-reversals and Psi[m] = Psi[2^l + row 2^(l-a)] Psi[j]
// (Psi[0] = 1).  With TWF = 1 a row's twiddles are formed from one per-row factor per level and
// the table's first 2^(n-a-1) entries, which every row shares (cache-resident), instead of
// reading a distinct 2^(n-a)-entry slice of the table per row.  The inverse's index
// 3 2^l - 1 - m = 2^l + (R-1-row) 2^(l-a) + (2^(l-a) - 1 - j) factors the same way.
//
// `data` and `zp` point at the prime's N residues.
// With MODE == ROW_SQUARE, `mul` (uniform over the workgroup) takes the product by zp instead of the
// square (a multiplying round of a batched power).
template <class M, int LOGE, int MODE, int TWF>
__device__ __forceinline__ void row_body(const M &md, uint32_t *__restrict__ data, const uint32_t *__restrict__ zp,
                                         const uint32_t *__restrict__ tw, int n, int a, int NG, unsigned blk,
                                         uint32_t *lds, bool mul = false) {
    constexpr int E = 1 << LOGE;
    const int g = n - a, G = 1 << g;
    const int spr = G >> LOGE;  // sub-transforms per row
    const int t = threadIdx.x, q = t / spr, s = t % spr;
    const uint32_t row = blk * NG + q;
    uint32_t *dst = data + (size_t)row * G;
    uint32_t *L = lds + q * (G + spr);
    const Chunking ch(g, LOGE);
    // chunk geometry: chunk ci spans levels lam0 .. lam0 + LOGE - 1 of the row transform; the
    // thread's elements are base + rr 2^ylog, its first-level block index within the chunk is hi
    auto geom = [&](int ci, int &lam0, int &ne, int &ylog, int &hi, int &base) {
        lam0 = ch.lam0(ci, LOGE);
        ne = ch.ne(ci, LOGE);
        ylog = g - lam0 - LOGE;
        hi = s >> ylog;
        base = (hi << (g - lam0)) + (s & ((1 << ylog) - 1));
    };
    // table twiddles of the chunk, loaded before the chunk's LDS exchange (TWF = 0)
    auto fwd_tw = [&](uint32_t (&w)[E], int lam0, int hi, int ne) {
        load_fwd_tw<LOGE>(w, tw, (1u << (a + lam0)) + (row << lam0) + (uint32_t)hi, ne);
    };
    auto inv_tw = [&](uint32_t (&w)[E], int lam0, int hi, int ne) {
        load_inv_tw<LOGE>(w, tw, (1u << (a + lam0)) + (row << lam0) + (uint32_t)hi, a + lam0, ne);
    };
    uint32_t x[E], w[E];
    int lam0 = 0, ne = 0, hi = 0, base = 0, ylog = 0;
    for (int ci = 0; ci < ch.nch; ++ci) {
        geom(ci, lam0, ne, ylog, hi, base);
        if (!TWF) fwd_tw(w, lam0, hi, ne);
        if (ci == 0) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) x[rr] = dst[base + (rr << ylog)];
        } else {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = L[k + (k >> LOGE)];
            }
        }
        if (TWF) {
            uint32_t rb[LOGE];
#pragma unroll
            for (int mu = 0; mu < LOGE; ++mu) rb[mu] = (mu < ne) ? tw[(1u << (a + lam0 + mu)) + (row << (lam0 + mu))] : 0u;
            fwd_levels_t<M, LOGE>(md, x, [&](int mu, int bi) { return md.mont(rb[mu], tw[((uint32_t)hi << mu) + bi]); }, ne);
        } else {
            fwd_levels_r<M, LOGE>(md, x, w, ne);
        }
        if (ci < ch.nch - 1) {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                L[k + (k >> LOGE)] = x[rr];
            }
        }
    }
    if (MODE == ROW_FWD) {
#pragma unroll
        for (int rr = 0; rr < E; ++rr) dst[base + (rr << ylog)] = x[rr];
        return;
    }
    // the last forward chunk's element set is the first inverse chunk's: its twiddles load while
    // the pointwise square runs
    if (!TWF) inv_tw(w, lam0, hi, ne);
    if (MODE == ROW_SQUARE && !mul) {
#pragma unroll
        for (int rr = 0; rr < E; ++rr) x[rr] = md.mont(x[rr], x[rr]);
    } else {
        const uint32_t *z = zp + (size_t)row * G;
#pragma unroll
        for (int rr = 0; rr < E; ++rr) x[rr] = md.mont(x[rr], z[base + (rr << ylog)]);
    }
    for (int ci = ch.nch - 1; ci >= 0; --ci) {
        geom(ci, lam0, ne, ylog, hi, base);
        if (ci != ch.nch - 1) {
            if (!TWF) inv_tw(w, lam0, hi, ne);
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = L[k + (k >> LOGE)];
            }
        }
        if (TWF) {
            const uint32_t rrow = (1u << a) - 1u - row;
            uint32_t rb[LOGE];
#pragma unroll
            for (int mu = 0; mu < LOGE; ++mu) rb[mu] = (mu < ne) ? tw[(1u << (a + lam0 + mu)) + (rrow << (lam0 + mu))] : 0u;
            inv_levels_t<M, LOGE>(md, x, [&](int mu, int bi) {
                return md.mont(rb[mu], tw[(1u << (lam0 + mu)) - 1u - (((uint32_t)hi << mu) + bi)]);
            }, ne);
        } else {
            inv_levels_r<M, LOGE>(md, x, w, ne);
        }
        if (ci == 0) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) dst[base + (rr << ylog)] = x[rr];
        } else {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                L[k + (k >> LOGE)] = x[rr];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Carries

// Adds carry c into d[0..S-1] digit by digit while a carry remains; the last digit absorbs the rest
// (digits stay bounded: |d| <= b/2 + 4 for carries of the sizes a squaring produces).
template <int S>
__device__ __forceinline__ void absorb_carry(int32_t (&d)[S], int64_t c, uint32_t b, double binv) {
#pragma unroll
    for (int j = 0; j < S - 1; ++j) {
        if (c != 0) {
            const int64_t f = (int64_t)d[j] + c;
            d[j] = div_bal((uint64_t)f, (double)f, b, binv, &c);
        }
    }
    d[S - 1] += (int32_t)c;
}

// Second half of a carry pass: thread t owns the S digits d (computed with carry-in 0; its carry-out
// is cr) starting at digit i0 of register dc.  The carry-out of thread t is added into thread t + 1's
// digits through LDS; the workgroup's carry-out is stored to wc[blk] for carry_wg_body, which adds
// it into the next workgroup's first digits after this kernel (cross-workgroup data goes through a
// kernel boundary: gfx950's per-XCD L2s make an in-kernel device-scope fence an L2 write-back).
// blockDim.x must be 256.
template <int S>
__device__ __forceinline__ void carry_tail(int32_t (&d)[S], int64_t cr, int32_t *dc, size_t i0, int64_t *wc,
                                           unsigned blk, uint32_t b, double binv) {
    constexpr int T = 256;
    __shared__ int64_t cs[T];
    const int tid = threadIdx.x;
    cs[tid] = cr;
    __syncthreads();
    if (tid > 0) absorb_carry<S>(d, cs[tid - 1], b, binv);
    int32_t *o = dc + i0;
#pragma unroll
    for (int j = 0; j < S; j += 4) *reinterpret_cast<int4 *>(o + j) = make_int4(d[j], d[j + 1], d[j + 2], d[j + 3]);
    if (tid == T - 1) wc[blk] = cr;
}

// Scale, Garner CRT and balanced carry of one candidate: `ws` = its rns x N residues, `dst` = its
// register 0, `wc` = its workgroup carries; workgroup blk covers digits [blk 256 S, (blk+1) 256 S).
// Scale, Garner CRT, x2 (dup) and the balanced carry across one thread's S consecutive digits of
// residues a[k][j]: d receives the digits (carry-in 0), cr the thread's carry-out.
template <int RNS, class M, int S>
__device__ __forceinline__ void norm_digits(const CrtConst &c, const M &m1, const M &m2, const M &m3,
                                            const uint32_t (&a)[MAX_RNS][S], bool dup, uint32_t b, double binv,
                                            int32_t (&d)[S], int64_t &cr) {
    cr = 0;
    if (c.fast) {
        double cd = 0.0;
        uint32_t cl = 0;
#pragma unroll
        for (int j = 0; j < S; ++j)
            d[j] = crt_digit_fast<RNS>(c, m1, m2, m3, a[0][j], a[1][j], (RNS == 3) ? a[RNS - 1][j] : 0u, dup, b, binv,
                                       cd, cl);
        cr = carry_of(cd);
    } else {
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const uint32_t r1 = m1.mulw(a[0][j], c.scale[0]);
            const uint32_t r2 = m2.mulw(a[1][j], c.scale[1]);
            const uint32_t r3 = (RNS == 3) ? m3.mulw(a[RNS - 1][j], c.scale[2]) : 0u;
            uint64_t x64; double xd;
            garner<RNS>(c, m1, m2, m3, r1, r2, r3, &x64, &xd);
            if (dup) { x64 <<= 1; xd *= 2.0; }
            x64 += (uint64_t)cr;
            xd += (double)cr;
            d[j] = div_bal(x64, xd, b, binv, &cr);
        }
    }
}

// Scale, Garner CRT and balanced carry of one candidate: `ws` = its rns x N residues, `dst` = its
// register 0, `wc` = its workgroup carries; workgroup blk covers digits [blk 256 S, (blk+1) 256 S).
template <int RNS, class M, int S>
__device__ __forceinline__ void norm_body(const CrtConst &c, const uint32_t *__restrict__ ws, int32_t *dst, int64_t *wc,
                                          uint32_t b, bool dup, int n, unsigned blk) {
    const size_t N = (size_t)1 << n;
    const size_t i0 = ((size_t)blk * 256 + threadIdx.x) * S;
    const double binv = 1.0 / (double)b;
    const M m1{c.p[0], c.q[0]}, m2{c.p[1], c.q[1]}, m3{c.p[2], c.q[2]};
    const uint32_t *w0 = ws + i0;
    uint32_t a[MAX_RNS][S];
#pragma unroll
    for (int k = 0; k < RNS; ++k) {
#pragma unroll
        for (int j = 0; j < S; j += 4) {
            const uint4 v = *reinterpret_cast<const uint4 *>(w0 + (size_t)k * N + j);
            a[k][j] = v.x; a[k][j + 1] = v.y; a[k][j + 2] = v.z; a[k][j + 3] = v.w;
        }
    }
    int32_t d[S];
    int64_t cr;
    norm_digits<RNS, M, S>(c, m1, m2, m3, a, dup, b, binv, d, cr);
    carry_tail<S>(d, cr, dst, i0, wc, blk, b, binv);
}

// Normalisation of several ranges per workgroup: blocks blk = first, first + stride, ...
// (< N / (256 S)), the next block's residues loaded while the current block is converted.
template <int RNS, class M, int S>
__device__ __forceinline__ void norm_loop(const CrtConst &c, const uint32_t *__restrict__ ws, int32_t *dst, int64_t *wc,
                                          uint32_t b, bool dup, int n, unsigned first, unsigned stride) {
    const size_t N = (size_t)1 << n;
    const unsigned nblk = (unsigned)(N / (256 * S));
    if (first >= nblk) return;
    const double binv = 1.0 / (double)b;
    const M m1{c.p[0], c.q[0]}, m2{c.p[1], c.q[1]}, m3{c.p[2], c.q[2]};
    uint32_t a[MAX_RNS][S], an[MAX_RNS][S];
    auto load = [&](uint32_t (&v)[MAX_RNS][S], unsigned blk) {
        const uint32_t *w0 = ws + ((size_t)blk * 256 + threadIdx.x) * S;
#pragma unroll
        for (int k = 0; k < RNS; ++k) {
#pragma unroll
            for (int j = 0; j < S; j += 4) {
                const uint4 u = *reinterpret_cast<const uint4 *>(w0 + (size_t)k * N + j);
                v[k][j] = u.x; v[k][j + 1] = u.y; v[k][j + 2] = u.z; v[k][j + 3] = u.w;
            }
        }
    };
    load(a, first);
    for (unsigned blk = first; blk < nblk; blk += stride) {
        const bool more = blk + stride < nblk;
        if (more) load(an, blk + stride);
        int32_t d[S];
        int64_t cr;
        norm_digits<RNS, M, S>(c, m1, m2, m3, a, dup, b, binv, d, cr);
        carry_tail<S>(d, cr, dst, ((size_t)blk * 256 + threadIdx.x) * S, wc, blk, b, binv);
        __syncthreads();  // carry_tail's LDS is rewritten by the next block
        if (more) {
#pragma unroll
            for (int k = 0; k < RNS; ++k)
#pragma unroll
                for (int j = 0; j < S; ++j) a[k][j] = an[k][j];
        }
    }
}

// r = a r on the balanced digits (the reference's mulscalar, ocl/kernel.cl:2461-2492, run after
// the normalisation as baseModMul does, src/transformGPU.h:828-837), 8 digits per thread.
__device__ __forceinline__ void mulscalar_body(int32_t *dc, int64_t *wc, uint32_t b, int32_t a, unsigned blk) {
    constexpr int S = 8;
    const size_t i0 = ((size_t)blk * 256 + threadIdx.x) * S;
    const double binv = 1.0 / (double)b;
    int32_t d[S];
    const int4 v0 = *reinterpret_cast<const int4 *>(dc + i0), v1 = *reinterpret_cast<const int4 *>(dc + i0 + 4);
    d[0] = v0.x; d[1] = v0.y; d[2] = v0.z; d[3] = v0.w; d[4] = v1.x; d[5] = v1.y; d[6] = v1.z; d[7] = v1.w;
    int64_t c = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int64_t f = (int64_t)d[j] * a + c;
        d[j] = div_bal((uint64_t)f, (double)f, b, binv, &c);
    }
    carry_tail<S>(d, c, dc, i0, wc, blk, b, binv);
}

// Adds the carry-out of workgroup w - 1 (of the last one, negated, for w = 0: x^N = -1) into the
// first 8 digits of workgroup w's range (256 x 8 digits per workgroup), and clears it: a carry
// array holds only carries not yet in the digits (or zeros), so a later absorption of it is exact.
// Each entry is read by one thread only.
__device__ __forceinline__ void carry_wg_body(int32_t *dc, int64_t *wc, unsigned nwg, uint32_t b, unsigned w) {
    constexpr int S = 8;
    const unsigned src = (w == 0) ? nwg - 1 : w - 1;
    const int64_t c = wc[src];
    if (c == 0) return;
    wc[src] = 0;
    const int64_t cin = (w == 0) ? -c : c;
    int32_t *p = dc + (size_t)w * 256 * S;
    int32_t e[S];
    const int4 v0 = *reinterpret_cast<const int4 *>(p), v1 = *reinterpret_cast<const int4 *>(p + 4);
    e[0] = v0.x; e[1] = v0.y; e[2] = v0.z; e[3] = v0.w; e[4] = v1.x; e[5] = v1.y; e[6] = v1.z; e[7] = v1.w;
    absorb_carry<S>(e, cin, b, 1.0 / (double)b);
    *reinterpret_cast<int4 *>(p) = make_int4(e[0], e[1], e[2], e[3]);
    *reinterpret_cast<int4 *>(p + 4) = make_int4(e[4], e[5], e[6], e[7]);
}

// ---------------------------------------------------------------------------------------------
// Single-engine kernels (grid.y = prime)

struct ColArgs {
    const int32_t *src;   // digits (forward)
    uint32_t *ws;         // residues [prime][N]
    const uint32_t *tw;   // [prime][N], Montgomery form
    PrimeArgs pr;
    int n, a, ng;
    const int64_t *wc = nullptr;  // forward: workgroup carries still to add into src (carry_wg_body), or null
    uint32_t b = 0;
    uint32_t bm[4] = {0, 0, 0, 0};  // limb forward pass: b 2^32 mod p of each prime
    const int4 *ls = nullptr;       // limb forward pass: carries into ranges of 2^rlog limbs, or null
    int rlog = 0;
};

// 32 points per thread need more than the 128 registers a 1024-thread workgroup leaves: such plans
// run at most 512 threads.
template <class M, int LOGE>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) k_col_fwd(ColArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    col_fwd_body<M, LOGE>(md, A.src, A.ws + prime * N, A.tw + prime * N, A.n, A.a, A.ng, xcd_block(blockIdx.x, gridDim.x), lds);
}

template <class M, int LOGE>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) k_col_inv(ColArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    col_inv_body<M, LOGE>(md, A.ws + prime * N, A.tw + prime * N, A.n, A.a, A.ng, xcd_block(blockIdx.x, gridDim.x), lds);
}

// Pipelined (persistent) column kernels: grid (workgroups per prime, rns).
template <class M, int LOGE>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) k_col_fwd_p(ColArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    col_fwd_loop<M, LOGE>(md, A.src, A.ws + prime * N, A.tw + prime * N, A.n, A.a, A.ng, blockIdx.x, gridDim.x, lds);
}

template <class M, int LOGE>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) k_col_inv_p(ColArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    col_inv_loop<M, LOGE>(md, A.ws + prime * N, A.tw + prime * N, A.n, A.a, A.ng, blockIdx.x, gridDim.x, lds);
}

struct ColAllArgs {
    const int32_t *src;   // digits (forward)
    uint32_t *ws;         // residues [prime][N]
    const uint32_t *tw;   // [prime][N]
    int32_t *dst;         // digits (inverse)
    int64_t *seg;         // segment carries
    const CrtConst *crt;
    PrimeArgs pr;
    uint32_t b;
    int dup, rounds, n, a, ng;
};

// At most 512 threads (fused plans): the inverse holds every prime's elements in registers.
template <int RNS, class M, int LOGE>
__global__ void __launch_bounds__(512) k_col_fwd_all(ColAllArgs A) {
    extern __shared__ uint32_t lds[];
    col_fwd_all_body<RNS, M, LOGE>(A.pr, A.src, A.ws, A.tw, A.n, A.a, A.ng, xcd_block(blockIdx.x, gridDim.x), lds);
}

template <int RNS, class M, int LOGE>
__global__ void __launch_bounds__(512) k_col_inv_norm(ColAllArgs A) {
    extern __shared__ uint32_t lds[];
    col_inv_norm_body<RNS, M, LOGE>(*A.crt, A.ws, A.tw, A.dst, A.seg, A.b, A.dup != 0, A.rounds, A.n, A.a, A.ng,
                                    xcd_block(blockIdx.x, gridDim.x), lds);
}

struct RowArgs {
    uint32_t *ws;          // [prime][N]
    const uint32_t *zp;    // multiplicand [prime][N] (ROW_MUL)
    const uint32_t *tw;
    PrimeArgs pr;
    int n, a, ng;
};

template <class M, int LOGE, int MODE, int TWF>
__global__ void __launch_bounds__(1024) k_row(RowArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    row_body<M, LOGE, MODE, TWF>(md, A.ws + prime * N, A.zp + prime * N, A.tw + prime * N, A.n, A.a, A.ng, blockIdx.x, lds);
}

struct NormArgs {
    const uint32_t *ws;    // [prime][N]
    int32_t *dst;          // register 0
    int64_t *carry;        // [gridDim.x]
    const CrtConst *crt;
    uint32_t b;
    int n, dup;
};

template <int RNS, class M, int S>
__global__ void __launch_bounds__(256) k_norm_p(NormArgs A) {
    norm_loop<RNS, M, S>(*A.crt, A.ws, A.dst, A.carry, A.b, A.dup != 0, A.n, blockIdx.x, gridDim.x);
}

template <int RNS, class M, int S>
__global__ void __launch_bounds__(256) k_norm(NormArgs A) {
    norm_body<RNS, M, S>(*A.crt, A.ws, A.dst, A.carry, A.b, A.dup != 0, A.n, blockIdx.x);
}

}  // namespace gfe
