// gfe_spec_engine.hip -- single-candidate kernels of the compile-time passes (gfe_pass.h):
// grid (blocks, primes), one prime's residues per workgroup.
#include <mutex>

#include "gfe_pass.h"
#include "gfe_spec.h"

namespace gfe {

template <class M, int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) ks_col_fwd(ColArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    col_fwd_pass<M, LOGE, LOGG>(md, A.src, A.ws + prime * N, A.tw + prime * N, (uint32_t)(N >> LOGG), A.ng,
                                xcd_block(blockIdx.x, gridDim.x), lds, A.wc, A.b);
}

// Forward column pass of the two-digit-limb transform (gfe_limb.h): element i = limb of digits
// 2i, 2i + 1; grid (blocks, 4 primes), A.n = log2 of the limb count.
template <int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) ks_col_fwd_limb(ColArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const ModS md{A.pr.p[prime], A.pr.q[prime]};
    const size_t M = (size_t)1 << A.n;
    col_fwd_pass<ModS, LOGE, LOGG, true>(md, A.src, A.ws + prime * M, A.tw + prime * M, (uint32_t)(M >> LOGG), A.ng,
                                         xcd_block(blockIdx.x, gridDim.x), lds, A.wc, A.b, A.bm[prime], A.ls, A.rlog);
}

// Forward column pass of the two-digit-limb transform for the four primes in one workgroup (the
// fused plans' shape: threads [p T, (p + 1) T) for prime p): the group's digit pairs are loaded once
// into LDS with the range carries of the last normalisation (A.ls) added, then each prime's threads
// run its forward column levels from there.
template <int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) ks_col_fwd_limb_all(ColArgs A) {
    extern __shared__ uint32_t lds[];
    constexpr int S = (1 << LOGE) / 4;
    const uint32_t T = ((uint32_t)A.ng << LOGG) >> LOGE;
    const uint32_t prime = threadIdx.x / T, t = threadIdx.x - prime * T;
    const ModS md{A.pr.p[prime], A.pr.q[prime]};
    const size_t M = (size_t)1 << A.n;
    const uint32_t C = (uint32_t)(M >> LOGG);
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    int2 *D = reinterpret_cast<int2 *>(lds);
    stage_limb_digits<S>(A.src, A.ls, C, (uint32_t)A.ng, blk, D);
    __syncthreads();
    const uint32_t exw = (uint32_t)col_pass_lds_words(LOGG, LOGE, A.ng);
    col_fwd_limb_staged<LOGE, LOGG>(md, D, A.ws + prime * M, A.tw + prime * M, C, A.ng, blk, lds + prime * exw, t,
                                    A.bm[prime]);
}

// Inverse column pass of the two-digit-limb transform fused with its normalisation (gfe_limb.h):
// one workgroup per group of ng columns runs the four primes' inverse column transforms side by side
// (threads [p T, (p + 1) T) for prime p, T = ng 2^LOGG / 2^LOGE, each prime on its own exchange
// region), leaves them in LDS and normalises the group's 2^LOGG rows of ng limbs, 2^LOGE / 4 limbs
// per thread.  The residues are read once and never written back; register 0 and the range carries
// are the only outputs.
template <int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) ks_col_inv_norm_limb(LimbColNormArgs A) {
    extern __shared__ uint32_t lds[];
    constexpr int S = (1 << LOGE) / 4;
    const uint32_t T = ((uint32_t)A.ng << LOGG) >> LOGE;
    const uint32_t prime = threadIdx.x / T, t = threadIdx.x - prime * T;
    const ModS md{A.pr.p[prime], A.pr.q[prime]};
    const size_t M = (size_t)1 << A.n;
    const uint32_t C = (uint32_t)(M >> LOGG);
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t exw = (uint32_t)col_pass_lds_words(LOGG, LOGE, A.ng), cells = (uint32_t)A.ng << LOGG;
    col_inv_to_lds<ModS, LOGE, LOGG>(md, A.ws + prime * M, A.tw + prime * M, C, A.ng, blk, lds + prime * exw, t,
                                     lds + prime * cells);
    __syncthreads();
    norm_limb_rows<S>(*A.crt, lds, (uint32_t)A.ng, 1u << LOGG, C, blk, A.dst, A.slots, (uint32_t)(M / (uint32_t)A.ng),
                      A.dup != 0);
}

template <class M, int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) ks_col_inv(ColArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    col_inv_pass<M, LOGE, LOGG>(md, A.ws + prime * N, A.tw + prime * N, (uint32_t)(N >> LOGG), A.ng,
                                xcd_block(blockIdx.x, gridDim.x), lds);
}

// 32 points per thread: rows of 2^11 and 2^13 points need 130-147 registers and spill when held to
// 128 (4 waves per SIMD), so the bound is 3 waves; the other shapes stay within 128 registers
// (2^14 points: 119) and keep 4 waves per SIMD.
template <class M, int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024, LOGE >= 5 ? 3 : 1) ks_row_sq(RowArgs A) {
    extern __shared__ uint32_t lds[];
    const int prime = blockIdx.y;
    const M md{A.pr.p[prime], A.pr.q[prime]};
    const size_t N = (size_t)1 << A.n;
    row_pass<M, LOGE, LOGG, ROW_SQUARE>(md, A.ws + prime * N, A.zp + prime * N, A.tw + prime * N, A.a, A.ng,
                                        blockIdx.x, lds);
}

template <int KIND, class M, int LOGE, int LOGG = SPEC_COL_MIN>
static SpecColFn col_lookup(int logg) {
    if constexpr (LOGG > SPEC_COL_MAX) {
        return nullptr;
    } else {
        if (logg == LOGG) {
            if constexpr (LOGG >= LOGE) return KIND == 0 ? ks_col_fwd<M, LOGE, LOGG> : ks_col_inv<M, LOGE, LOGG>;
            else return nullptr;
        }
        return col_lookup<KIND, M, LOGE, LOGG + 1>(logg);
    }
}

template <int LOGE, int LOGG = SPEC_COL_MIN>
static SpecColFn col_limb_lookup(int logg) {
    if constexpr (LOGG > SPEC_COL_MAX) {
        return nullptr;
    } else {
        if (logg == LOGG) {
            if constexpr (LOGG >= LOGE) return ks_col_fwd_limb<LOGE, LOGG>;
            else return nullptr;
        }
        return col_limb_lookup<LOGE, LOGG + 1>(logg);
    }
}

template <int LOGE, int LOGG = SPEC_COL_MIN>
static SpecLimbNormFn col_inv_norm_limb_lookup(int logg) {
    if constexpr (LOGG > SPEC_COL_MAX || LOGG > LOGE + LIMB_FUSED_MAX_SPAN) {
        return nullptr;
    } else {
        if (logg == LOGG) {
            if constexpr (LOGG >= LOGE) return ks_col_inv_norm_limb<LOGE, LOGG>;
            else return nullptr;
        }
        return col_inv_norm_limb_lookup<LOGE, LOGG + 1>(logg);
    }
}

template <int LOGE, int LOGG = SPEC_COL_MIN>
static SpecColFn col_fwd_limb_all_lookup(int logg) {
    if constexpr (LOGG > SPEC_COL_MAX || LOGG > LOGE + LIMB_FUSED_MAX_SPAN) {
        return nullptr;
    } else {
        if (logg == LOGG) {
            if constexpr (LOGG >= LOGE) return ks_col_fwd_limb_all<LOGE, LOGG>;
            else return nullptr;
        }
        return col_fwd_limb_all_lookup<LOGE, LOGG + 1>(logg);
    }
}

template <class M, int LOGE, int LOGG = SPEC_ROW_MIN>
static SpecRowFn row_lookup(int logg) {
    if constexpr (LOGG > SPEC_ROW_MAX) {
        return nullptr;
    } else {
        if (logg == LOGG) return ks_row_sq<M, LOGE, LOGG>;
        return row_lookup<M, LOGE, LOGG + 1>(logg);
    }
}

template <int KIND, class M>
static SpecColFn col_fn(int loge, int logg) {
    switch (loge) {
        case 3: return col_lookup<KIND, M, 3>(logg);
        case 4: return col_lookup<KIND, M, 4>(logg);
        case 5: return col_lookup<KIND, M, 5>(logg);
        default: return nullptr;
    }
}

SpecColFn spec_col_fwd(int is32, int loge, int logg) {
    return is32 ? col_fn<0, ModU>(loge, logg) : col_fn<0, ModS>(loge, logg);
}
SpecColFn spec_col_inv(int is32, int loge, int logg) {
    return is32 ? col_fn<1, ModU>(loge, logg) : col_fn<1, ModS>(loge, logg);
}
SpecColFn spec_col_fwd_limb(int loge, int logg) {
    switch (loge) {
        case 3: return col_limb_lookup<3>(logg);
        case 4: return col_limb_lookup<4>(logg);
        case 5: return col_limb_lookup<5>(logg);
        default: return nullptr;
    }
}

SpecColFn spec_col_fwd_limb_all(int loge, int logg) {
    switch (loge) {
        case 3: return col_fwd_limb_all_lookup<3>(logg);
        case 4: return col_fwd_limb_all_lookup<4>(logg);
        case 5: return col_fwd_limb_all_lookup<5>(logg);
        default: return nullptr;
    }
}

SpecLimbNormFn spec_col_inv_norm_limb(int loge, int logg) {
    switch (loge) {
        case 3: return col_inv_norm_limb_lookup<3>(logg);
        case 4: return col_inv_norm_limb_lookup<4>(logg);
        case 5: return col_inv_norm_limb_lookup<5>(logg);
        default: return nullptr;
    }
}

SpecRowFn spec_row_square(int is32, int loge, int logg) {
    if (loge == 4) return is32 ? row_lookup<ModU, 4>(logg) : row_lookup<ModS, 4>(logg);
    if (loge == 5) return is32 ? row_lookup<ModU, 5>(logg) : row_lookup<ModS, 5>(logg);
    return nullptr;
}

void spec_set_lds_attributes() {
    static std::once_flag once;
    std::call_once(once, []() {
        const int lds = 160 * 1024;
        for (int is32 = 0; is32 < 2; ++is32)
            for (int loge = 3; loge <= 5; ++loge) {
                for (int g = SPEC_COL_MIN; g <= SPEC_COL_MAX; ++g) {
                    if (SpecColFn f = spec_col_fwd(is32, loge, g))
                        (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                    if (SpecColFn f = spec_col_inv(is32, loge, g))
                        (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                    if (SpecColFn f = is32 ? nullptr : spec_col_fwd_limb(loge, g))
                        (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                    if (SpecLimbNormFn f = is32 ? nullptr : spec_col_inv_norm_limb(loge, g))
                        (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                    if (SpecColFn f = is32 ? nullptr : spec_col_fwd_limb_all(loge, g))
                        (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                }
                for (int g = SPEC_ROW_MIN; g <= SPEC_ROW_MAX; ++g)
                    if (SpecRowFn f = spec_row_square(is32, loge, g))
                        (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            }
        (void)hipGetLastError();
    });
}

}  // namespace gfe
