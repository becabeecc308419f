/* gfe.h -- C ABI of the gfx950 modular-squaring engine for b^(2^n) + 1.
 *
 * Drop-in boundary for genefer's GPU transform (the reference's `transform` interface,
 * src/transform.h:42-256, implemented on the GPU by transformGPUs, src/transformGPU.h:931-1195).
 * Plain C types only; every entry point returns GFE_OK (0) or a negative error code and never
 * throws; gfe_last_error() describes the last failure of the calling thread.
 *
 * Register model (src/transform.h:54-73): num_regs registers of N = 2^n digits; register 0 is the
 * working value r0; rm is the multiplicand set by gfe_init_multiplicand.  Registers hold signed
 * digits in base b with |digit| <= b/2 + 4 (a bounded, non-canonical representation, as the
 * reference's GPU registers are); gfe_get_digits returns them as they are and the caller
 * canonicalises (gint::unbalance, src/gint.h:40-102).
 *
 * All work is queued on the engine's own HIP stream; gfe_get_digits / gfe_save / gfe_sync wait
 * for it.  One engine must not be used from two threads at once (as the reference's engines,
 * src/transformGPU.h:205-208).
 */
#ifndef GFE_H
#define GFE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFE_OK 0
#define GFE_EINVAL (-1)     /* bad argument (n, b, register index, size) */
#define GFE_EHIP (-2)       /* HIP runtime error */
#define GFE_ENOMEM (-3)     /* device allocation failed */
#define GFE_EKIND (-4)      /* context of another engine kind */
#define GFE_ENODEV (-5)     /* no such device / no HIP device */

typedef struct gfe_engine gfe_engine;

/* transform::create_gpu (src/transform.h:114-120) -> create_ocl (src/transform_ocl.cpp:20-52):
 * the prime set (2 or 3 primes of 31 or 32 bits) is chosen from (b, n) by the reference's rule;
 * it names the engine kind and the context payload's residues.  Three-prime transforms compute in
 * the 31-bit primes for every b (balanced digits, |d| <= b/2 + 4, keep them exact).
 * 12 <= n <= 23 as the reference's command line allows (src/main.cpp:258-259), 1024 <= b <= 2e9. */
int gfe_create(uint32_t b, uint32_t n, uint32_t num_regs, int device, gfe_engine **out);
int gfe_destroy(gfe_engine *e);

/* transform::set, src/transform.h:57 -- r0 = a.  Registers hold balanced digits, so a must not
 * exceed b/2 + 4 (GFE_EINVAL otherwise); the reference's callers set 1 or small bases
 * (src/genefer.h:445-458, 489). */
int gfe_set(gfe_engine *e, uint32_t a);
/* transform::squareDup, src/transform.h:58 -- r0 = r0^2, or 2 r0^2 when dup != 0. */
int gfe_square_dup(gfe_engine *e, int dup);
/* transform::squareMul, src/transform.h:59 -- r0 = a r0^2. */
int gfe_square_mul(gfe_engine *e, int32_t a);
/* transform::initMultiplicand, src/transform.h:60 -- rm = transform(r_src). */
int gfe_init_multiplicand(gfe_engine *e, uint32_t src);
/* transform::mul, src/transform.h:61 -- r0 = r0 * rm. */
int gfe_mul(gfe_engine *e);
/* transform::copy, src/transform.h:63 -- r_dst = r_src. */
int gfe_copy(gfe_engine *e, uint32_t dst, uint32_t src);
/* transform::getZi / setZi, src/transform.h:53-54 (getInt/setInt 234-246 add the size/base checks
 * and gint canonicalisation on the caller side); any register. */
int gfe_get_digits(gfe_engine *e, uint32_t reg, int32_t *out);
int gfe_set_digits(gfe_engine *e, uint32_t reg, const int32_t *in);
/* transform::getMemSize / getCacheSize, src/transform.h:65-66. */
size_t gfe_mem_size(const gfe_engine *e);
size_t gfe_cache_size(const gfe_engine *e);
/* transform::saveContext / readContext, src/transform.h:68-69: the payload of the context file
 * after the driver's header, in the layout of the reference's GPU transform
 * (src/transformGPU.h:1127-1150): int kind (transform::EKind, src/transform.h:45: 6 = NTT2,
 * 7 = NTT3), then for each of the num_regs registers and each prime of the set, N uint32 plain
 * residues of the signed digits (ZPT::set_int, src/transformGPU.h:147).  A context written by the
 * reference's GPU build loads here and the other way round.  gfe_context_size gives the bytes
 * gfe_save writes.  gfe_load returns GFE_EKIND for a payload of another kind (the reference's
 * readContext returns false) and GFE_EINVAL when the residues of one digit disagree between primes.
 * Loaded digits of any size (the reference's registers hold digits in (-b, b)) are balanced to
 * |d| <= b/2 on the host, which keeps the value. */
size_t gfe_context_size(const gfe_engine *e, uint32_t num_regs);
int gfe_save(gfe_engine *e, uint32_t num_regs, void *buf, size_t size);
int gfe_load(gfe_engine *e, uint32_t num_regs, const void *buf, size_t size);
/* transform::getError, src/transform.h:71 -- always 0: the arithmetic is exact. */
double gfe_get_error(const gfe_engine *e);
/* transform::info, src/transform.h:72 -- prime set computed in (and the reference's set when it
 * differs) and pass plan as text. */
int gfe_info(const gfe_engine *e, char *buf, size_t len);
/* No counterpart in the reference (one candidate per process there): moves the engine to base b of
 * the same n, for servers that keep engines between candidates instead of building their twiddle
 * tables again.  b must select the prime set the engine computes in (GFE_EKIND otherwise; create
 * an engine for it); the CRT constants, the carry rounds and the plan become b's -- the two-digit-
 * limb transform where it is exact for b and the engine was created with its tables (for a b that
 * has it), else the digit transform's default.  The registers keep their digits (carries pending
 * from the last squaring are added first, in the old base): the caller sets them (gfe_set,
 * gfe_set_digits) before squaring. */
int gfe_rebase(gfe_engine *e, uint32_t b);
/* Waits for the engine's queued work. */
int gfe_sync(gfe_engine *e);
/* The reference's prime-set selection without creating an engine: *rns = 2 or 3, *is32 = 0 (31-bit)
 * or 1 (32-bit). */
int gfe_select(uint32_t b, uint32_t n, int *rns, int *is32);
const char *gfe_last_error(void);

/* Times `count` squarings (dup pattern: dup_bits bit i for squaring i mod 64) with HIP events on
 * the engine's stream, after the queued work.  *ms = total.  When kernel_ms is not null,
 * kernel_ms[0..4] = average duration of one launch of each pass (column forward, row, column
 * inverse, normalisation; [4] = 0; a single-workgroup plan reports its one kernel in [0]), each
 * from `count` back-to-back launches of that pass alone between one pair of events; that run
 * squares register 0 once more (dup = 0), so register 0 ends as the value squared count (+ 1)
 * times. */
int gfe_time_squarings(gfe_engine *e, uint32_t count, uint64_t dup_bits, double *ms, double *kernel_ms);

/* Autotuner: times every admissible pass plan for this engine's n on the device and keeps the
 * fastest (the reference's splitter + engines::tune, src/ocl.h:31-90, src/transformGPU.h:851-900).
 * `reps` squarings per plan.  The chosen plan is reported by gfe_info. */
int gfe_tune(gfe_engine *e, uint32_t reps);

/* The pass plan as {a (column levels), column E log2, columns per workgroup, row E log2, rows per
 * workgroup, flags: bit 0 factored row twiddles, bit 1 all-prime column passes with the CRT and
 * carry fused into the inverse one, bit 2 persistent column passes, bit 3 compile-time passes,
 * bit 4 the squaring in one workgroup (n <= 13), bit 5 the two-digit-limb transform (see
 * gfe_engine_plan; refused where it is not exact for the engine's b); with bit 5, bits 1 and 3
 * together fuse the limb transform's inverse column pass with its normalisation (ranges of
 * columns-per-workgroup limbs whose carries the next squaring adds; refused where the larger
 * range-start limbs this leaves would break exactness for the engine's b)}.  No reference counterpart (the reference's tuned split is
 * internal to its engines): fixes a plan for reproducible benchmarks and for tests that cover
 * every kernel variant.  gfe_set_plan returns GFE_EINVAL for a plan not admissible at this n. */
int gfe_get_plan(const gfe_engine *e, int32_t plan[6]);
int gfe_set_plan(gfe_engine *e, const int32_t plan[6]);

/* Plan queries that need no device (no reference counterpart, for tooling and tests):
 * gfe_default_plan writes the digit transform's default plan at n (capacity 0) or the plan a batch
 * of `capacity` slots starts with; gfe_plan_admissible returns GFE_OK for a plan the kernels accept at n, else
 * GFE_EINVAL; gfe_candidate_plans writes how many plans the autotuner times at n to *count and,
 * when plans is not null, the first min(max, count) of them (6 int32 each). */
int gfe_default_plan(uint32_t n, uint32_t capacity, int32_t plan[6]);
/* The plan gfe_create gives an engine of (b, n): gfe_default_plan(n, 0), or for the (b, n) the
 * two-digit-limb transform covers (flags bit 5: limbs d_2j + b d_2j+1 of base b^2, N/2-point
 * transforms over four 31-bit primes; a and the row fields then describe the N/2-point transform)
 * its default plan, with the inverse column pass fused to the normalisation (flags 42) where that
 * stays exact.  No device needed. */
int gfe_engine_plan(uint32_t b, uint32_t n, int32_t plan[6]);
int gfe_plan_admissible(uint32_t n, const int32_t plan[6]);
int gfe_candidate_plans(uint32_t n, int32_t *plans, uint32_t max, uint32_t *count);

/* The twiddle tables the engine uploads (host only, no device; the reference's root tables,
 * src/transformGPU.h:1046-1078, in this engine's layout): for each of the three primes of the
 * 31-bit (is32 = 0) or 32-bit (is32 = 1) set, N = 2^n entries Psi[j] = psi^bitrev_n(j) 2^32 mod p,
 * psi = g^((p-1) / 2N), written to out + k N.  The tables are built on up to max_threads host
 * threads (0: on the calling thread only); the result does not depend on it. */
int gfe_host_twiddles(uint32_t n, int is32, int max_threads, uint32_t *out);

/* No counterpart in the reference: squarings that allocate nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * Squares the N = 2^n balanced digits at `digits` -- device memory of `device`, read and written in
 * place, digits as gfe_get_digits returns them (|d| <= b/2 + 4) -- `count` times, squaring i
 * doubled when bit i mod 64 of dup_bits is set, queued on `stream` (a hipStream_t; NULL = the null
 * stream), and returns without waiting.  The squarings run the digit transform's default plan
 * (gfe_default_plan(n, 0)): the digits equal those of `count` gfe_square_dup calls on an engine of
 * the same (b, n) on that plan followed by gfe_get_digits, and their value equals the engine's on
 * any plan (where an engine starts on the two-digit-limb transform, its balanced digits may differ
 * from these while the residue is the same).  The transform work space, twiddles,
 * carries and CRT constants live in one static device array of GFE_NOALLOC_BYTES that the library
 * holds on every device it runs on until the process ends, which limits n to
 * 12 <= n <= GFE_NOALLOC_MAX_N (GFE_EINVAL beyond); no device or host memory is allocated or freed.
 * Every call rebuilds the twiddles and constants of its (b, n) in that array, so a call does not
 * depend on an earlier one; calls that overlap on one device race on the array. */
#define GFE_NOALLOC_MAX_N 16
#define GFE_NOALLOC_BYTES 1638656 /* 2 x 3 x 2^16 x 4 (work space, twiddles) + 2^16 (carries) + 256 */
int gfe_square_noalloc(uint32_t b, uint32_t n, int device, int32_t *digits, uint32_t count, uint64_t dup_bits,
                       void *stream);

#ifdef __cplusplus
}
#endif
#endif
