// gfe_fused.h -- host-side interface of the single-workgroup squaring kernels (gfe_fused.hip).
#pragma once

#include "gfe_batch_dev.h"
#include "gfe_kernels.h"

namespace gfe {

constexpr int FUSED_MIN_N = 10, FUSED_MAX_N = 13;

struct FusedArgs {
    int32_t *regs;            // batch: slot s's register 0 at regs + s reg_stride; engine: register 0
    const uint32_t *tw;       // batch: [2][3][N] (31-bit set, 32-bit set); engine: [rns][N]
    const CrtConst *crt;      // batch: per slot; engine: one
    const SlotInfo *slots;    // batch: per slot (rns == 0: empty); engine: null
    const uint8_t *act;       // per-round activity rows (row r at r stride), or null: every round
    const uint8_t *dupf;      // per-round doubling rows, or null: dup0
    long long reg_stride;
    uint32_t stride;          // flag row length (slots)
    int rounds;
    int dup0;
    uint32_t b;               // engine only
    int rns, is32;            // engine only
    PrimeArgs prS, prU;
    uint32_t tws;             // entries between the primes' twiddle tables
    int only_n;               // mixed kernel: run the slots of this n only (0: all)
};

typedef void (*FusedFn)(FusedArgs);

// Kernel for N = 2^logn digits (null outside FUSED_MIN_N..FUSED_MAX_N); launch with grid.x =
// slots, blockDim.x = fused_threads(logn) and fused_lds_bytes(logn) bytes of dynamic LDS.
FusedFn fused_square_fn(int logn);
// One launch for the slots of n = 12 and 13 of a mixed batch (512 threads, fused_lds_bytes(13)).
FusedFn fused_square_mixed_fn();
size_t fused_lds_bytes(int logn);
unsigned fused_threads(int logn);
void fused_set_lds_attributes();

}  // namespace gfe
