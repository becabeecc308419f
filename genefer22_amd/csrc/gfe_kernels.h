// gfe_kernels.h -- gfx950 kernels of the squaring engine (one candidate or a batch of candidates
// of the same n and prime set; grid.z indexes the candidate).
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
//   k_col_fwd : digits -> residues, levels 0..a-1 on each column c (elements c + i C), NG columns
//               per workgroup;
//   k_row     : levels a..n-1 on each row (C contiguous elements), pointwise square / product,
//               inverse levels n-1..a;
//   k_col_inv : inverse levels a-1..0;
//   k_norm    : scale, Garner CRT, balanced carry over S consecutive digits per thread; carries
//               cross threads through LDS;
//   k_carry_wg: adds each workgroup's carry-out into the next one (negated across x^N = -1).
// Inside a pass, a thread holds E = 2^LOGE elements of one sub-transform in registers and runs
// LOGE levels per chunk; chunks exchange through LDS.
#pragma once

#include "gfe_device.h"

namespace gfe {

constexpr int MAX_RNS = 3;

struct PrimeArgs {
    uint32_t p[MAX_RNS];
    uint32_t q[MAX_RNS];
};

// Forward CT levels mu = 0..ne-1 of a 2^LOGE-point sub-transform held in x[].  The twiddle of
// level mu, register block bi is tw[(t0 << mu) + bi].
template <class M, int LOGE>
__device__ __forceinline__ void fwd_levels(const M &md, uint32_t (&x)[1 << LOGE], const uint32_t *__restrict__ tw,
                                           uint32_t t0, int ne) {
    constexpr int E = 1 << LOGE;
#pragma unroll
    for (int mu = 0; mu < LOGE; ++mu) {
        if (mu < ne) {
            const int h = E >> (mu + 1);
#pragma unroll
            for (int bi = 0; bi < (1 << mu); ++bi) {
                const uint32_t w = tw[(t0 << mu) + bi];
#pragma unroll
                for (int i = 0; i < h; ++i) {
                    const int i0 = bi * 2 * h + i, i1 = i0 + h;
                    const uint32_t u = x[i0], v = md.mulw(x[i1], w);
                    x[i0] = md.add(u, v);
                    x[i1] = md.sub(u, v);
                }
            }
        }
    }
}

// Inverse GS levels mu = ne-1..0 (the exact inverse of fwd_levels without the 1/2 factors).
// lvl0 = global level of mu = 0.
template <class M, int LOGE>
__device__ __forceinline__ void inv_levels(const M &md, uint32_t (&x)[1 << LOGE], const uint32_t *__restrict__ tw,
                                           uint32_t t0, int lvl0, int ne) {
    constexpr int E = 1 << LOGE;
#pragma unroll
    for (int mu = LOGE - 1; mu >= 0; --mu) {
        if (mu < ne) {
            const int h = E >> (mu + 1);
            const uint32_t m3 = 3u << (lvl0 + mu);
#pragma unroll
            for (int bi = 0; bi < (1 << mu); ++bi) {
                const uint32_t w = tw[m3 - 1u - ((t0 << mu) + bi)];
#pragma unroll
                for (int i = 0; i < h; ++i) {
                    const int i0 = bi * 2 * h + i, i1 = i0 + h;
                    const uint32_t u = x[i0], v = x[i1];
                    x[i0] = md.add(u, v);
                    x[i1] = md.mulw(md.sub(v, u), w);
                }
            }
        }
    }
}

// Chunk schedule of a g-level pass: a partial first chunk of r = g mod LOGE levels (if r > 0),
// then full chunks.  Every chunk's element set spans LOGE levels starting at lam0.
struct Chunking {
    int g, r, nch;
    __device__ __forceinline__ Chunking(int g_, int loge) : g(g_), r(g_ % loge), nch((g_ + loge - 1) / loge) {}
    __device__ __forceinline__ int lam0(int ci, int loge) const { return (r == 0) ? ci * loge : (ci == 0 ? 0 : r + (ci - 1) * loge); }
    __device__ __forceinline__ int ne(int ci, int loge) const { return (r != 0 && ci == 0) ? r : loge; }
};

struct ColArgs {
    const int32_t *src;   // digits (candidate stride src_stride) -- forward only
    uint32_t *ws;         // residues, [cand][prime][N]
    const uint32_t *tw;   // [prime][N], Montgomery form
    PrimeArgs pr;
    long long src_stride; // in int32
    long long ws_stride;  // in uint32, per candidate
    int n, a, ng, rns;
};

// Column pass, forward: levels 0..a-1.  Thread t -> column q = t % NG, sub-transform s = t / NG.
// LDS element (q, k) at ((k + (k >> LOGE)) * NG + q).
template <class M, int LOGE, bool FROM_DIGITS>
__global__ void __launch_bounds__(1024) k_col_fwd(ColArgs A) {
    constexpr int E = 1 << LOGE;
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y, cand = blockIdx.z;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const int g = A.a, G = 1 << g, NG = A.ng;
    const uint32_t C = 1u << (A.n - g);
    const uint32_t c0 = blockIdx.x * NG;
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const size_t N = (size_t)1 << A.n;
    uint32_t *dst = A.ws + cand * A.ws_stride + (size_t)prime * N;
    const uint32_t *tw = A.tw + (size_t)prime * N;
    const Chunking ch(g, LOGE);
    uint32_t x[E];
    for (int ci = 0; ci < ch.nch; ++ci) {
        const int lam0 = ch.lam0(ci, LOGE), ne = ch.ne(ci, LOGE);
        const int ylog = g - lam0 - LOGE;
        const int hi = s >> ylog, lo = s & ((1 << ylog) - 1);
        const int base = (hi << (g - lam0)) + lo;
        if (ci == 0) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                const size_t gi = (size_t)k * C + c0 + q;
                if (FROM_DIGITS) x[rr] = to_res(A.src[cand * A.src_stride + gi], md.p);
                else x[rr] = dst[gi];
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = lds[(k + (k >> LOGE)) * NG + q];
            }
        }
        fwd_levels<M, LOGE>(md, x, tw, (1u << lam0) + hi, ne);
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

// Column pass, inverse: levels a-1..0, in place on ws.
template <class M, int LOGE>
__global__ void __launch_bounds__(1024) k_col_inv(ColArgs A) {
    constexpr int E = 1 << LOGE;
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y, cand = blockIdx.z;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const int g = A.a, NG = A.ng;
    const uint32_t C = 1u << (A.n - g);
    const uint32_t c0 = blockIdx.x * NG;
    const int t = threadIdx.x, q = t & (NG - 1), s = t / NG;
    const size_t N = (size_t)1 << A.n;
    uint32_t *dst = A.ws + cand * A.ws_stride + (size_t)prime * N;
    const uint32_t *tw = A.tw + (size_t)prime * N;
    const Chunking ch(g, LOGE);
    uint32_t x[E];
    for (int ci = ch.nch - 1; ci >= 0; --ci) {
        const int lam0 = ch.lam0(ci, LOGE), ne = ch.ne(ci, LOGE);
        const int ylog = g - lam0 - LOGE;
        const int hi = s >> ylog, lo = s & ((1 << ylog) - 1);
        const int base = (hi << (g - lam0)) + lo;
        if (ci == ch.nch - 1) {
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = dst[(size_t)k * C + c0 + q];
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = lds[(k + (k >> LOGE)) * NG + q];
            }
        }
        inv_levels<M, LOGE>(md, x, tw, (1u << lam0) + hi, lam0, ne);
        if (ci == 0) {
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

struct RowArgs {
    uint32_t *ws;          // [cand][prime][N]
    const uint32_t *zp;    // multiplicand [cand][prime][N] (MODE 1)
    const uint32_t *tw;
    PrimeArgs pr;
    long long ws_stride, zp_stride;
    int n, a, ng;
};

enum RowMode { ROW_SQUARE = 0, ROW_MUL = 1, ROW_FWD = 2 };

// Row pass: levels a..n-1 on rows of C = 2^(n-a) contiguous elements, the pointwise square or
// product (Montgomery, factor 2^-32), then inverse levels n-1..a.  ROW_FWD stops after the
// forward levels (multiplicand).  Thread t -> row q = t / (C/E), sub-transform s = t % (C/E).
// LDS element (q, k) at (q * (C + C/E) + k + (k >> LOGE)).
template <class M, int LOGE, int MODE>
__global__ void __launch_bounds__(1024) k_row(RowArgs A) {
    constexpr int E = 1 << LOGE;
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y, cand = blockIdx.z;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const int g = A.n - A.a, G = 1 << g;
    const int spr = G >> LOGE;  // sub-transforms per row
    const int t = threadIdx.x, q = t / spr, s = t % spr;
    const uint32_t row = blockIdx.x * A.ng + q;
    const size_t N = (size_t)1 << A.n;
    uint32_t *dst = A.ws + cand * A.ws_stride + (size_t)prime * N + (size_t)row * G;
    const uint32_t *tw = A.tw + (size_t)prime * N;
    uint32_t *L = lds + q * (G + spr);
    const Chunking ch(g, LOGE);
    uint32_t x[E];
    int hi = 0, base = 0, ylog = 0;
    for (int ci = 0; ci < ch.nch; ++ci) {
        const int lam0 = ch.lam0(ci, LOGE), ne = ch.ne(ci, LOGE);
        ylog = g - lam0 - LOGE;
        hi = s >> ylog;
        const int lo = s & ((1 << ylog) - 1);
        base = (hi << (g - lam0)) + lo;
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
        fwd_levels<M, LOGE>(md, x, tw, ((1u << (A.a + lam0)) + (row << lam0) + hi), ne);
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
    if (MODE == ROW_SQUARE) {
#pragma unroll
        for (int rr = 0; rr < E; ++rr) x[rr] = md.mont(x[rr], x[rr]);
    } else {
        const uint32_t *zp = A.zp + cand * A.zp_stride + (size_t)prime * N + (size_t)row * G;
#pragma unroll
        for (int rr = 0; rr < E; ++rr) x[rr] = md.mont(x[rr], zp[base + (rr << ylog)]);
    }
    for (int ci = ch.nch - 1; ci >= 0; --ci) {
        const int lam0 = ch.lam0(ci, LOGE), ne = ch.ne(ci, LOGE);
        ylog = g - lam0 - LOGE;
        hi = s >> ylog;
        const int lo = s & ((1 << ylog) - 1);
        base = (hi << (g - lam0)) + lo;
        if (ci != ch.nch - 1) {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < E; ++rr) {
                const uint32_t k = base + (rr << ylog);
                x[rr] = L[k + (k >> LOGE)];
            }
        }
        inv_levels<M, LOGE>(md, x, tw, ((1u << (A.a + lam0)) + (row << lam0) + hi), A.a + lam0, ne);
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

struct NormArgs {
    const uint32_t *ws;    // [cand][prime][N]
    int32_t *dst;          // digits [cand][N] (register 0 of each candidate)
    int64_t *carry;        // [cand][gridDim.x]: carry-out of each workgroup
    const CrtConst *crt;   // [cand] (or one shared entry when crt_stride == 0)
    const uint32_t *b;     // [cand]
    const uint8_t *dup;    // [cand] or null (then dup_all)
    long long ws_stride, dst_stride, carry_stride;
    int crt_stride;
    int n, dup_all;
};

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
// digits through LDS; the workgroup's carry-out is stored to wc[blockIdx.x] for k_carry_wg, which
// adds it into the next workgroup's first digits after this kernel (cross-workgroup data goes
// through a kernel boundary: gfx950's per-XCD L2s make an in-kernel device-scope fence an L2
// write-back).  blockDim.x must be 256.
template <int S>
__device__ __forceinline__ void carry_tail(int32_t (&d)[S], int64_t cr, int32_t *dc, size_t i0, int64_t *wc,
                                           uint32_t b, double binv) {
    constexpr int T = 256;
    __shared__ int64_t cs[T];
    const int tid = threadIdx.x;
    cs[tid] = cr;
    __syncthreads();
    if (tid > 0) absorb_carry<S>(d, cs[tid - 1], b, binv);
    int32_t *o = dc + i0;
#pragma unroll
    for (int j = 0; j < S; j += 4) *reinterpret_cast<int4 *>(o + j) = make_int4(d[j], d[j + 1], d[j + 2], d[j + 3]);
    if (tid == T - 1) wc[blockIdx.x] = cr;
}

struct CarryArgs {
    int32_t *dst;
    const int64_t *carry;
    const uint32_t *b;
    long long dst_stride, carry_stride;
    unsigned nwg;   // workgroups of the carry pass (each 256 x 8 digits)
};

// Adds the carry-out of workgroup w - 1 (of the last one, negated, for w = 0: x^N = -1) into the
// first digits of workgroup w's range.
__global__ void __launch_bounds__(256) k_carry_wg(CarryArgs A) {
    constexpr int S = 8;
    const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= A.nwg) return;
    const int cand = blockIdx.z;
    const int64_t *wc = A.carry + cand * A.carry_stride;
    const int64_t cin = (w == 0) ? -wc[A.nwg - 1] : wc[w - 1];
    if (cin == 0) return;
    const uint32_t b = A.b[cand];
    int32_t *p = A.dst + cand * A.dst_stride + (size_t)w * 256 * S;
    int32_t e[S];
    const int4 v0 = *reinterpret_cast<const int4 *>(p), v1 = *reinterpret_cast<const int4 *>(p + 4);
    e[0] = v0.x; e[1] = v0.y; e[2] = v0.z; e[3] = v0.w; e[4] = v1.x; e[5] = v1.y; e[6] = v1.z; e[7] = v1.w;
    absorb_carry<S>(e, cin, b, 1.0 / (double)b);
    *reinterpret_cast<int4 *>(p) = make_int4(e[0], e[1], e[2], e[3]);
    *reinterpret_cast<int4 *>(p + 4) = make_int4(e[4], e[5], e[6], e[7]);
}

// Scale, Garner CRT and balanced carry.  Thread t owns S consecutive digits; carries cross threads
// in carry_tail and workgroups in k_carry_wg.
template <int RNS, class M, int S>
__global__ void __launch_bounds__(256) k_norm(NormArgs A) {
    constexpr int T = 256;
    const int cand = blockIdx.z, tid = threadIdx.x;
    const size_t N = (size_t)1 << A.n;
    const size_t i0 = ((size_t)blockIdx.x * T + tid) * S;
    const CrtConst &c = A.crt[cand * A.crt_stride];
    const uint32_t b = A.b[cand];
    const double binv = 1.0 / (double)b;
    const bool dup = A.dup ? (A.dup[cand] != 0) : (A.dup_all != 0);
    const M m1{c.p[0], c.q[0]}, m2{c.p[1], c.q[1]}, m3{c.p[2], c.q[2]};
    const uint32_t *w0 = A.ws + cand * A.ws_stride + i0;
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
    int64_t cr = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const uint32_t r1 = m1.mulw(a[0][j], c.scale[0]);
        const uint32_t r2 = m2.mulw(a[1][j], c.scale[1]);
        const uint32_t r3 = (RNS == 3) ? m3.mulw(a[2][j], c.scale[2]) : 0u;
        uint64_t x64; double xd;
        garner<RNS>(c, m1, m2, m3, r1, r2, r3, &x64, &xd);
        if (dup) { x64 <<= 1; xd *= 2.0; }
        x64 += (uint64_t)cr;
        xd += (double)cr;
        d[j] = div_bal(x64, xd, b, binv, &cr);
    }
    carry_tail<S>(d, cr, A.dst + cand * A.dst_stride, i0, A.carry + cand * A.carry_stride, b, binv);
}

struct MulScalarArgs {
    int32_t *dst;
    int64_t *carry;
    const uint32_t *b;
    long long dst_stride, carry_stride;
    int n;
    int32_t a;
};

// r = a r on the balanced digits (the reference's mulscalar, ocl/kernel.cl:2461-2492, run after
// the normalisation as baseModMul does, src/transformGPU.h:828-837), 8 digits per thread.
__global__ void __launch_bounds__(256) k_mulscalar(MulScalarArgs A) {
    constexpr int S = 8;
    const int cand = blockIdx.z, tid = threadIdx.x;
    const size_t i0 = ((size_t)blockIdx.x * 256 + tid) * S;
    const uint32_t b = A.b[cand];
    const double binv = 1.0 / (double)b;
    int32_t *dc = A.dst + cand * A.dst_stride;
    int32_t d[S];
    const int4 v0 = *reinterpret_cast<const int4 *>(dc + i0), v1 = *reinterpret_cast<const int4 *>(dc + i0 + 4);
    d[0] = v0.x; d[1] = v0.y; d[2] = v0.z; d[3] = v0.w; d[4] = v1.x; d[5] = v1.y; d[6] = v1.z; d[7] = v1.w;
    int64_t c = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int64_t f = (int64_t)d[j] * A.a + c;
        d[j] = div_bal((uint64_t)f, (double)f, b, binv, &c);
    }
    carry_tail<S>(d, c, dc, i0, A.carry + cand * A.carry_stride, b, binv);
}

// Register copy, 16 bytes per thread (the reference's copy kernel, ocl/kernel.cl:2501-2506).
__global__ void __launch_bounds__(256) k_copy(int4 *__restrict__ dst, const int4 *__restrict__ src, size_t n4,
                                               long long dst_stride, long long src_stride) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) dst[blockIdx.z * dst_stride + i] = src[blockIdx.z * src_stride + i];
}

// r = a (digit 0 = a, others 0).
__global__ void __launch_bounds__(256) k_set(int32_t *dst, long long stride, size_t N, int32_t a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[blockIdx.z * stride + i] = (i == 0) ? a : 0;
}

}  // namespace gfe
