// gfe_spec.h -- host-side lookup of the compile-time pass kernels (gfe_pass.h), instantiated in
// gfe_spec_engine.hip (single candidate, grid.y = prime) and gfe_spec_batch.hip (grid.z = slot).
#pragma once

#include "gfe_batch_dev.h"
#include "gfe_kernels.h"
#include "gfe_limb.h"

namespace gfe {

typedef void (*SpecColFn)(ColArgs);
typedef void (*SpecLimbNormFn)(LimbColNormArgs);
typedef void (*SpecRowFn)(RowArgs);
typedef void (*SpecBatchColFn)(BArgs, int);
typedef void (*SpecBatchRowFn)(BArgs);

// Null when the shape is not instantiated (spec_col_ok / spec_row_ok).
SpecColFn spec_col_fwd(int is32, int loge, int logg);
SpecColFn spec_col_inv(int is32, int loge, int logg);
SpecColFn spec_col_fwd_limb(int loge, int logg);  // two-digit-limb transform, 31-bit primes
// Its four primes in one workgroup (the fused plans' shape: grid column groups, 4 ng 2^logg / 2^loge
// threads, limb_col_fwd_all_lds_bytes of LDS; null outside limb_col_norm_ok).
SpecColFn spec_col_fwd_limb_all(int loge, int logg);
// The limb transform's inverse column pass fused with its normalisation (grid: column groups,
// 4 ng 2^logg / 2^loge threads, limb_col_norm_lds_bytes of LDS; null outside limb_col_norm_ok).
SpecLimbNormFn spec_col_inv_norm_limb(int loge, int logg);
SpecRowFn spec_row_square(int is32, int loge, int logg);
SpecBatchColFn spec_batch_col_fwd(int loge, int logg);
SpecBatchColFn spec_batch_col_inv(int loge, int logg);  // second argument unused
SpecBatchRowFn spec_batch_row_square(int loge, int logg);

// Raises the dynamic-LDS limit of every instantiated kernel to 160 KiB (once per process).
void spec_set_lds_attributes();
void spec_batch_set_lds_attributes();

}  // namespace gfe
