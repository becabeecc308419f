// gfe_spec_batch.hip -- batched kernels of the compile-time passes (gfe_pass.h): grid
// (blocks, primes, slots); each workgroup runs one prime of one candidate slot and picks the
// 31-bit or 32-bit arithmetic from the slot's prime set.
#include <mutex>

#include "gfe_pass.h"
#include "gfe_spec.h"

namespace gfe {

template <int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) kbs_col_fwd(BArgs A, int to_zp) {
    extern __shared__ uint32_t lds[];
    const int s = bslot(A);
    if (s < 0) return;
    const SlotInfo si = A.slots[s];
    const int prime = blockIdx.y;
    if (prime >= si.rns) return;
    const size_t N = (size_t)1 << A.n, NS = (size_t)1 << A.sn;
    const int32_t *src = A.regs + s * A.reg_stride + A.src_reg * N;
    uint32_t *dst = (to_zp ? A.zp : A.ws) + (size_t)s * 3 * NS + prime * N;
    const uint32_t *tw = A.tw + ((size_t)si.is32 * 3 + prime) * NS;
    const uint32_t C = (uint32_t)(N >> LOGG);
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t *wc = A.absorb ? A.carry + (size_t)s * (NS >> 3) : nullptr;
    if (si.is32)
        col_fwd_pass<ModU, LOGE, LOGG>(ModU{A.prU.p[prime], A.prU.q[prime]}, src, dst, tw, C, A.ng, blk, lds, wc, si.b);
    else
        col_fwd_pass<ModS, LOGE, LOGG>(ModS{A.prS.p[prime], A.prS.q[prime]}, src, dst, tw, C, A.ng, blk, lds, wc, si.b);
}

template <int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024) kbs_col_inv(BArgs A, int) {
    extern __shared__ uint32_t lds[];
    const int s = bslot(A);
    if (s < 0) return;
    const SlotInfo si = A.slots[s];
    const int prime = blockIdx.y;
    if (prime >= si.rns) return;
    const size_t N = (size_t)1 << A.n, NS = (size_t)1 << A.sn;
    uint32_t *data = A.ws + (size_t)s * 3 * NS + prime * N;
    const uint32_t *tw = A.tw + ((size_t)si.is32 * 3 + prime) * NS;
    const uint32_t C = (uint32_t)(N >> LOGG);
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    if (si.is32) col_inv_pass<ModU, LOGE, LOGG>(ModU{A.prU.p[prime], A.prU.q[prime]}, data, tw, C, A.ng, blk, lds);
    else col_inv_pass<ModS, LOGE, LOGG>(ModS{A.prS.p[prime], A.prS.q[prime]}, data, tw, C, A.ng, blk, lds);
}

// 32 points per thread with both prime sets' arithmetic in one kernel need up to 145 registers:
// bounded at 128 (4 waves per SIMD) some shapes spill to scratch, so these run at 3 waves per SIMD
// (n = 20, one slot: 52.3 us per round against 52.9 us with the spills).
template <int LOGE, int LOGG>
__global__ void __launch_bounds__(LOGE >= 5 ? 512 : 1024, LOGE >= 5 ? 3 : 1) kbs_row_sq(BArgs A) {
    extern __shared__ uint32_t lds[];
    const int s = bslot(A);
    if (s < 0) return;
    const SlotInfo si = A.slots[s];
    const int prime = blockIdx.y;
    if (prime >= si.rns) return;
    const size_t N = (size_t)1 << A.n, NS = (size_t)1 << A.sn;
    uint32_t *data = A.ws + (size_t)s * 3 * NS + prime * N;
    const uint32_t *zp = A.zp + (size_t)s * 3 * NS + prime * N;
    const uint32_t *tw = A.tw + ((size_t)si.is32 * 3 + prime) * NS;
    const bool mul = A.mulf && A.mulf[s];
    if (si.is32) row_pass<ModU, LOGE, LOGG, ROW_SQUARE>(ModU{A.prU.p[prime], A.prU.q[prime]}, data, zp, tw, A.a, A.ng, blockIdx.x, lds, mul);
    else row_pass<ModS, LOGE, LOGG, ROW_SQUARE>(ModS{A.prS.p[prime], A.prS.q[prime]}, data, zp, tw, A.a, A.ng, blockIdx.x, lds, mul);
}

template <int KIND, int LOGE, int LOGG = SPEC_COL_MIN>
static SpecBatchColFn bcol_lookup(int logg) {
    if constexpr (LOGG > SPEC_COL_MAX) {
        return nullptr;
    } else {
        if (logg == LOGG) {
            if constexpr (LOGG >= LOGE) return KIND == 0 ? kbs_col_fwd<LOGE, LOGG> : kbs_col_inv<LOGE, LOGG>;
            else return nullptr;
        }
        return bcol_lookup<KIND, LOGE, LOGG + 1>(logg);
    }
}

template <int LOGE, int LOGG = SPEC_ROW_MIN>
static SpecBatchRowFn brow_lookup(int logg) {
    if constexpr (LOGG > SPEC_ROW_MAX) {
        return nullptr;
    } else {
        if (logg == LOGG) return kbs_row_sq<LOGE, LOGG>;
        return brow_lookup<LOGE, LOGG + 1>(logg);
    }
}

template <int KIND>
static SpecBatchColFn bcol_fn(int loge, int logg) {
    switch (loge) {
        case 3: return bcol_lookup<KIND, 3>(logg);
        case 4: return bcol_lookup<KIND, 4>(logg);
        case 5: return bcol_lookup<KIND, 5>(logg);
        default: return nullptr;
    }
}

SpecBatchColFn spec_batch_col_fwd(int loge, int logg) { return bcol_fn<0>(loge, logg); }
SpecBatchColFn spec_batch_col_inv(int loge, int logg) { return bcol_fn<1>(loge, logg); }
SpecBatchRowFn spec_batch_row_square(int loge, int logg) {
    if (loge == 4) return brow_lookup<4>(logg);
    if (loge == 5) return brow_lookup<5>(logg);
    return nullptr;
}

void spec_batch_set_lds_attributes() {
    static std::once_flag once;
    std::call_once(once, []() {
        const int lds = 160 * 1024;
        for (int loge = 3; loge <= 5; ++loge) {
            for (int g = SPEC_COL_MIN; g <= SPEC_COL_MAX; ++g) {
                if (SpecBatchColFn f = spec_batch_col_fwd(loge, g))
                    (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                if (SpecBatchColFn f = spec_batch_col_inv(loge, g))
                    (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            }
            for (int g = SPEC_ROW_MIN; g <= SPEC_ROW_MAX; ++g)
                if (SpecBatchRowFn f = spec_batch_row_square(loge, g))
                    (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        }
        (void)hipGetLastError();
    });
}

}  // namespace gfe
