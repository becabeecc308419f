// gfe_batch_dev.h -- device-side argument block of the batched engine (gfe_batch.hip and the
// compile-time pass kernels of gfe_spec_batch.hip).
#pragma once

#include "gfe_kernels.h"

namespace gfe {

struct SlotInfo {
    uint32_t b;
    int32_t rns, is32, rounds;   // rounds: carry rounds of the fused inverse column pass
    int32_t n;                   // the candidate's n (a mixed batch holds several)
};

struct BArgs {
    int32_t *regs;
    uint32_t *ws, *zp;
    int64_t *carry;
    const CrtConst *crt;
    const SlotInfo *slots;
    const uint32_t *tw;      // [2][3][N]
    const uint8_t *active;   // [cap] or null (all)
    const uint8_t *dup;      // [cap] or null
    const uint8_t *mulf;     // [cap] or null: squaring rounds of these slots multiply by zp instead
    PrimeArgs prS, prU;
    long long reg_stride;    // 3 N
    int n, a, ng;            // n: the launch's transform size; slots of another n exit at once
    int sn;                  // log2 of the per-slot allocation: slot s's residues at s 3 2^sn
    uint32_t src_reg, dst_reg;
    int32_t value;
    int32_t absorb;          // squaring rounds: the column forward pass adds the pending carries
};

__device__ __forceinline__ int bslot(const BArgs &A) {
    const int s = (int)blockIdx.z;
    if (A.active && !A.active[s]) return -1;
    if (A.slots[s].n != A.n) return -1;
    return s;
}

}  // namespace gfe
