// gfe_fused.hip -- the whole squaring of one small candidate in one workgroup (n <= 13).
//
// For N = 2^n <= 8192 digits, the rns x N residues of a candidate (3 x 32 KiB at n = 13), its N
// digits and the carry hand-off fit in one workgroup's LDS.  One launch then runs a block of up to
// 64 squaring rounds for every candidate slot (one workgroup per slot): digits -> residues,
// forward NTT of length N (LDS-staged chunks of 8 or 16 points per thread, as gfe_pass.h), pointwise
// square, inverse NTT, then the CRT and the balanced carry with the negacyclic wrap -- with no
// global-memory round trip between rounds and no cross-workgroup carry.  The transform is the one
// of the four-step passes (the same twiddle table, Psi[2^l + j] at level l), so the residues and
// the digit bound |d| <= b/2 + 4 are the same; the reference runs the same squaring as 5 kernels
// per iteration (src/transformGPU.h:665-693).
//
// LDS (words): 3 regions of N + N/E residues (padded, element k at k + k/E; E = 8 points per
// thread and prime, 16 at n = 13), N + N/E digits (same padding), T = N/E int64 carries.
#include <mutex>

#include "gfe_fused.h"
#include "gfe_pass.h"

namespace gfe {

// Points per thread per prime: 8 up to n = 12, 16 at n = 13 (at most 512 threads, so that the
// primes' interleaved chunks fit the registers).
__host__ __device__ constexpr int fused_loge(int logn) { return logn >= 13 ? 4 : 3; }

__host__ __device__ inline size_t fused_lds_bytes_n(int logn) {
    const size_t N = (size_t)1 << logn, E = (size_t)1 << fused_loge(logn), T = N / E;
    return 3 * (N + N / E) * 4 + (N + N / E) * 4 + T * 8;
}

// Forward NTT, square, inverse NTT of primes K0 .. K0 + RNS - 1 on the LDS digits `dig`, the primes'
// chunks interleaved so that they share the LDS exchanges' barriers (and give each thread RNS times
// the independent butterflies).  Prime k's residues end in the padded region res3 + k RS (element
// j at j + (j >> FLOGE)).
// tw3: the primes' twiddle tables, `tws` entries apart (a table of a larger N serves: its first N
// entries are the table of N).
template <class M, int LOGN, int RNS, int K0>
__device__ __forceinline__ void fused_ntt(const M (&md3)[3], uint32_t *res3, const int32_t *dig,
                                          const uint32_t *__restrict__ tw3, uint32_t tws) {
    constexpr int FLOGE = fused_loge(LOGN), FE = 1 << FLOGE;
    using GM = Geo<LOGN, FLOGE>;
    constexpr uint32_t N = 1u << LOGN, RS = N + N / FE;
    // primes K0 .. K0 + RNS - 1
    M md[RNS];
#pragma unroll
    for (int k = 0; k < RNS; ++k) md[k] = md3[K0 + k];
    uint32_t *res = res3 + K0 * RS;
    const uint32_t *tw = tw3 + K0 * tws;
    const uint32_t s = threadIdx.x;
    uint32_t x[RNS][FE], w[RNS][FE];
    static_for<0, GM::NCH>([&](auto cc) {
        constexpr int ci = decltype(cc)::value;
        constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
        const uint32_t hi = s >> YL, base = (hi << (LOGN - lam0)) + (s & ((1u << YL) - 1u));
        const uint32_t lb = base + (base >> FLOGE);
#pragma unroll
        for (int k = 0; k < RNS; ++k) load_fwd_tw<FLOGE>(w[k], tw + k * tws, (1u << lam0) + hi, NE);
        if constexpr (ci == 0) {
#pragma unroll
            for (int rr = 0; rr < FE; ++rr) {
                const int32_t d = dig[lb + chunk_off<YL, FLOGE>(rr)];
#pragma unroll
                for (int k = 0; k < RNS; ++k) x[k][rr] = to_res(d, md[k].p);
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RNS; ++k)
#pragma unroll
                for (int rr = 0; rr < FE; ++rr) x[k][rr] = res[k * RS + lb + chunk_off<YL, FLOGE>(rr)];
        }
#pragma unroll
        for (int k = 0; k < RNS; ++k) fwd_chunk<M, FLOGE, NE>(md[k], x[k], w[k]);
        if constexpr (ci < GM::NCH - 1) {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RNS; ++k)
#pragma unroll
                for (int rr = 0; rr < FE; ++rr) res[k * RS + lb + chunk_off<YL, FLOGE>(rr)] = x[k][rr];
        }
    });
    constexpr int LAST = GM::NCH - 1;
#pragma unroll
    for (int k = 0; k < RNS; ++k) {
        load_inv_tw<FLOGE>(w[k], tw + k * tws, (1u << GM::lam0(LAST)) + s, GM::lam0(LAST), GM::ne(LAST));
#pragma unroll
        for (int rr = 0; rr < FE; ++rr) x[k][rr] = md[k].mont(x[k][rr], x[k][rr]);
    }
    static_for_down<0, GM::NCH>([&](auto cc) {
        constexpr int ci = decltype(cc)::value;
        constexpr int lam0 = GM::lam0(ci), NE = GM::ne(ci), YL = GM::ylog(ci);
        const uint32_t hi = s >> YL, base = (hi << (LOGN - lam0)) + (s & ((1u << YL) - 1u));
        const uint32_t lb = base + (base >> FLOGE);
        if constexpr (ci != LAST) {
#pragma unroll
            for (int k = 0; k < RNS; ++k) load_inv_tw<FLOGE>(w[k], tw + k * tws, (1u << lam0) + hi, lam0, NE);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RNS; ++k)
#pragma unroll
                for (int rr = 0; rr < FE; ++rr) x[k][rr] = res[k * RS + lb + chunk_off<YL, FLOGE>(rr)];
        }
#pragma unroll
        for (int k = 0; k < RNS; ++k) inv_chunk<M, FLOGE, NE>(md[k], x[k], w[k]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RNS; ++k)
#pragma unroll
            for (int rr = 0; rr < FE; ++rr) res[k * RS + lb + chunk_off<YL, FLOGE>(rr)] = x[k][rr];
    });
}

// CRT, doubling and balanced carry of the whole candidate: thread t owns digits t E .. t E + E - 1;
// its carry-out enters thread t + 1, the last thread's enters thread 0 negated (x^N = -1).
template <int RNS, class M, int LOGN>
__device__ __forceinline__ void fused_crt(const CrtConst &c, const uint32_t *res, int32_t *dig, int64_t *cs,
                                          uint32_t b, bool dup) {
    constexpr int FLOGE = fused_loge(LOGN), FE = 1 << FLOGE;
    constexpr uint32_t N = 1u << LOGN, T = N >> FLOGE, RS = N + N / FE;
    const uint32_t t = threadIdx.x;
    const M m1{c.p[0], c.q[0]}, m2{c.p[1], c.q[1]}, m3{c.p[2], c.q[2]};
    const double binv = 1.0 / (double)b;
    const uint32_t o = t * (FE + 1);
    int64_t cr = 0;
    // digits in runs of 8 (a rolled loop keeps one run's CRT values live at a time)
    if (c.fast) {
        double cd = 0.0;
        uint32_t cl = 0;
#pragma unroll 1
        for (int h = 0; h < FE; h += 8) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t k = o + h + j;
                dig[k] = crt_digit_fast<RNS>(c, m1, m2, m3, res[k], res[RS + k], (RNS == 3) ? res[2 * RS + k] : 0u, dup,
                                             b, binv, cd, cl);
            }
        }
        cr = carry_of(cd);
    } else {
#pragma unroll 1
        for (int h = 0; h < FE; h += 8) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t k = o + h + j;
                const uint32_t r1 = m1.mulw(res[k], c.scale[0]);
                const uint32_t r2 = m2.mulw(res[RS + k], c.scale[1]);
                const uint32_t r3 = (RNS == 3) ? m3.mulw(res[2 * RS + k], c.scale[2]) : 0u;
                uint64_t x64;
                double xd;
                garner<RNS>(c, m1, m2, m3, r1, r2, r3, &x64, &xd);
                if (dup) { x64 <<= 1; xd *= 2.0; }
                x64 += (uint64_t)cr;
                xd += (double)cr;
                dig[k] = div_bal(x64, xd, b, binv, &cr);
            }
        }
    }
    cs[t] = cr;
    __syncthreads();
    // the incoming carry runs into the thread's digits while it lasts; the last digit keeps the rest
    int64_t cin = (t == 0) ? -cs[T - 1] : cs[t - 1];
    for (int j = 0; j < FE - 1 && cin != 0; ++j) {
        const int64_t f = (int64_t)dig[o + j] + cin;
        dig[o + j] = div_bal((uint64_t)f, (double)f, b, binv, &cin);
    }
    dig[o + FE - 1] += (int32_t)cin;
}

// One block of rounds of one slot.  act / dupf: per-round flag rows (row r at r * stride), or
// null: every round active / not doubled (dup0 then gives the doubling of a single round).
template <class M, int LOGN>
__device__ __forceinline__ void fused_block(const FusedArgs &A, const M (&md)[3], int rns, uint32_t b,
                                            const CrtConst &c, int32_t *gdig, const uint32_t *tw, int slot,
                                            uint32_t *lds) {
    constexpr int FLOGE = fused_loge(LOGN), FE = 1 << FLOGE;
    constexpr uint32_t N = 1u << LOGN, RS = N + N / FE;
    uint32_t *res = lds;
    int32_t *dig = reinterpret_cast<int32_t *>(lds + 3 * RS);
    int64_t *cs = reinterpret_cast<int64_t *>(lds + 4 * RS);
    const uint32_t t = threadIdx.x, o = t * (FE + 1);
#pragma unroll
    for (int j = 0; j < FE; j += 4) {
        const int4 v = *reinterpret_cast<const int4 *>(gdig + t * FE + j);
        dig[o + j] = v.x; dig[o + j + 1] = v.y; dig[o + j + 2] = v.z; dig[o + j + 3] = v.w;
    }
    bool any = false;
    for (int r = 0; r < A.rounds; ++r) {
        if (A.act && !A.act[(size_t)r * A.stride + slot]) continue;
        const bool dup = A.dupf ? (A.dupf[(size_t)r * A.stride + slot] != 0) : (A.dup0 != 0);
        any = true;
        __syncthreads();  // digits of the previous round (or the load) complete
        if constexpr (LOGN >= 13) {  // one prime at a time: 16 points per thread x 3 primes exceed the registers
            fused_ntt<M, LOGN, 1, 0>(md, res, dig, tw, A.tws);
            fused_ntt<M, LOGN, 1, 1>(md, res, dig, tw, A.tws);
            if (rns == 3) fused_ntt<M, LOGN, 1, 2>(md, res, dig, tw, A.tws);
        } else {
            if (rns == 3) fused_ntt<M, LOGN, 3, 0>(md, res, dig, tw, A.tws);
            else fused_ntt<M, LOGN, 2, 0>(md, res, dig, tw, A.tws);
        }
        __syncthreads();
        if (rns == 3) fused_crt<3, M, LOGN>(c, res, dig, cs, b, dup);
        else fused_crt<2, M, LOGN>(c, res, dig, cs, b, dup);
    }
    if (!any) return;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < FE; j += 4)
        *reinterpret_cast<int4 *>(gdig + t * FE + j) = make_int4(dig[o + j], dig[o + j + 1], dig[o + j + 2], dig[o + j + 3]);
}

// The block of rounds of one workgroup's slot (batch) or of the engine's candidate at N = 2^LOGN.
template <int LOGN>
__device__ __forceinline__ void fused_slot(const FusedArgs &A, int slot, uint32_t *lds) {
    uint32_t b;
    int rns, is32;
    const CrtConst *c;
    int32_t *gdig;
    if (A.slots) {
        const SlotInfo si = A.slots[slot];
        b = si.b; rns = si.rns; is32 = si.is32;
        c = A.crt + slot;
        gdig = A.regs + (size_t)slot * A.reg_stride;
    } else {
        b = A.b; rns = A.rns; is32 = A.is32;
        c = A.crt;
        gdig = A.regs;
    }
    const uint32_t *tw = A.tw + (A.slots ? (size_t)is32 * 3 * A.tws : 0);
    if (is32) {
        const ModU md[3] = {{A.prU.p[0], A.prU.q[0]}, {A.prU.p[1], A.prU.q[1]}, {A.prU.p[2], A.prU.q[2]}};
        fused_block<ModU, LOGN>(A, md, rns, b, *c, gdig, tw, slot, lds);
    } else {
        const ModS md[3] = {{A.prS.p[0], A.prS.q[0]}, {A.prS.p[1], A.prS.q[1]}, {A.prS.p[2], A.prS.q[2]}};
        fused_block<ModS, LOGN>(A, md, rns, b, *c, gdig, tw, slot, lds);
    }
}

// grid.x = slots (batch) or 1 (single engine); blockDim.x = fused_threads(n).  Batch slots of
// another n (or empty, rns = 0) exit at once.
template <int LOGN>
__global__ void __launch_bounds__((1 << LOGN) >> fused_loge(LOGN)) kf_square(FusedArgs A) {
    extern __shared__ uint32_t lds[];
    const int slot = (int)blockIdx.x;
    if (A.slots && (A.slots[slot].rns == 0 || A.slots[slot].n != LOGN)) return;
    fused_slot<LOGN>(A, slot, lds);
}

// Mixed batches: one launch runs the block for slots of n = 12 and n = 13 alike (both take 512
// threads; the LDS is sized for n = 13), each workgroup at its slot's n; with only_n != 0 the
// slots of that n alone.
__global__ void __launch_bounds__(512) kf_square_mixed(FusedArgs A) {
    extern __shared__ uint32_t lds[];
    const int slot = (int)blockIdx.x;
    const SlotInfo &si = A.slots[slot];
    if (si.rns == 0 || (A.only_n && si.n != A.only_n)) return;
    if (si.n == 12) fused_slot<12>(A, slot, lds);
    else if (si.n == 13) fused_slot<13>(A, slot, lds);
}

static_assert(((1 << 12) >> fused_loge(12)) == 512 && ((1 << 13) >> fused_loge(13)) == 512,
              "the mixed kernel's slots of n = 12 and 13 take the same workgroup");

FusedFn fused_square_mixed_fn() { return kf_square_mixed; }

FusedFn fused_square_fn(int logn) {
    switch (logn) {
        case 10: return kf_square<10>;
        case 11: return kf_square<11>;
        case 12: return kf_square<12>;
        case 13: return kf_square<13>;
        default: return nullptr;
    }
}

size_t fused_lds_bytes(int logn) { return fused_lds_bytes_n(logn); }
unsigned fused_threads(int logn) { return (1u << logn) >> fused_loge(logn); }

void fused_set_lds_attributes() {
    static std::once_flag once;
    std::call_once(once, []() {
        for (int n = FUSED_MIN_N; n <= FUSED_MAX_N; ++n)
            (void)hipFuncSetAttribute((const void *)fused_square_fn(n), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024);
        (void)hipFuncSetAttribute((const void *)kf_square_mixed, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipGetLastError();
    });
}

}  // namespace gfe
