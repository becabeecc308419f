/* gfe_batch.h -- C ABI of the batched squaring engine: many candidates b of the same n share one
 * set of launches (grid.z = candidate slot), with every candidate's registers resident in HBM.
 *
 * The reference runs one candidate per process and per device (src/genefer.h:1363-1573 with
 * transformGPUs, src/transformGPU.h:931-1195).  A batch runs the same per-candidate work -- the
 * Fermat PRP loop with Gerbicz-Li updates and the final Gerbicz-Li check (src/genefer.h:479-615,
 * 736-748) -- for up to `capacity` candidates at once; each candidate's residues are bit-exact with
 * its own single-candidate run.  Prime sets (2/3 primes, 31/32-bit) may differ between slots.
 *
 * Gerbicz-Li cadence: the d(t) product is updated every B iterations with B = the reference's
 * B_GerbiczLi(esize) rounded up to a multiple of 64, so that updates fall on the 64-iteration
 * block boundaries of the batch; r0 (and so RES64) does not depend on B, and the check is the
 * reference's with that B.
 *
 * Conventions of gfe.h: GFE_OK or a negative code, gfe_last_error() for the message; one batch
 * must not be used from two threads at once.
 */
#ifndef GFE_BATCH_H
#define GFE_BATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFE_EFULL (-6) /* no free slot */
#define GFE_BATCH_MAX_DEPTH 10  /* largest Pietrzak-Li depth of a proof batch */

typedef struct gfe_batch gfe_batch;

typedef struct gfe_batch_result {
    uint64_t tag;      /* caller's tag of the request */
    uint32_t b, n;
    uint64_t res64, old64, hash64;
    uint64_t esize;    /* bits of E = b^(2^n) */
    uint64_t iters;    /* PRP squarings run (esize for a full test) */
    int32_t is_prp;    /* residue == 1 (full tests) */
    int32_t gl_ok;     /* Gerbicz-Li check: 1 pass, 0 fail, -1 not run (truncated request) */
    double t_submit, t_done;  /* seconds on the batch's clock (steady clock since creation) */
    int32_t depth;     /* proof batches: the Pietrzak-Li depth of the proof below, else 0 */
    int32_t pad;
    uint64_t pkey;     /* proof key: gethash64 of v1 = prod_k mu_k^(w_k) (src/genefer.h:620-734) */
    uint64_t mu_hash64[GFE_BATCH_MAX_DEPTH + 1];  /* gethash64 of mu_0 .. mu_depth */
} gfe_batch_result;

/* A batch of `capacity` slots for candidates b^(2^n) + 1, 12 <= n <= 23. */
int gfe_batch_create(uint32_t n, uint32_t capacity, int device, gfe_batch **out);
/* A proof batch: every submission runs the reference's proof mode (src/genefer.h:750-766) -- the
 * full PRP loop with Pietrzak-Li checkpoints every B_PL = B_PietrzakLi(esize, depth) iterations
 * kept in the slot's registers 3 .. 3 + 2^depth - 1 in HBM (the reference's fast checkpoints,
 * src/genefer.h:1377-1389), the Gerbicz-Li check, then the proof (PL, src/genefer.h:620-734) with
 * its powers run as squaring and multiplying rounds of the batch's blocks.  Results carry the proof
 * key and the hashes of mu_0 .. mu_depth, equal to gfe_run_proof's.  1 <= depth <=
 * GFE_BATCH_MAX_DEPTH (0: a quick-test batch). */
int gfe_batch_create_proof(uint32_t n, uint32_t capacity, uint32_t depth, int device, gfe_batch **out);
/* A mixed batch: candidates of n = 12 and n = 13 in the same slots (submit with gfe_batch_submit_n;
 * gfe_batch_submit means n = 13).  Every squaring block is one launch of the single-workgroup
 * squaring in which each workgroup works at its slot's n; while fewer than 128 slots of n = 13 are
 * active (where the single-workgroup squaring at n = 13 is slower than the four-step passes, one
 * candidate: 166 ms against 76 ms for 4096 squarings on MI355X) that launch takes the slots of
 * n = 12 and the slots of n = 13 run the four-step passes after it.  Register operations run on the
 * four-step passes of each slot's n.  A candidate's residues are those of its own n's batch. */
int gfe_batch_create_mixed(uint32_t capacity, int device, gfe_batch **out);
int gfe_batch_destroy(gfe_batch *bt);
/* Device bytes one slot holds (registers, work space, multiplicand, carries): the admission unit. */
size_t gfe_batch_slot_bytes(uint32_t n);
/* As gfe_batch_slot_bytes for a proof batch of `depth` (3 + 2^depth registers per slot). */
size_t gfe_batch_proof_slot_bytes(uint32_t n, uint32_t depth);
/* Chooses the pass plan by timing the admissible ones on `reps` rounds of the batch's busy slots,
 * the fastest eight again on 8 x reps and, when the batch runs its blocks as graphs, the fastest
 * three on captured 64-round blocks (their registers are saved and restored, so their programs
 * continue unchanged); returns
 * GFE_EINVAL when no slot is busy: submit the candidates first.  The reference's tune,
 * src/transformGPU.h:851-900. */
int gfe_batch_tune(gfe_batch *bt, uint32_t reps);
int gfe_batch_info(const gfe_batch *bt, char *buf, size_t len);
/* The batch's pass plan, as gfe_get_plan / gfe_set_plan of gfe.h. */
int gfe_batch_get_plan(const gfe_batch *bt, int32_t plan[6]);
int gfe_batch_set_plan(gfe_batch *bt, const int32_t plan[6]);

/* Submits the quick test of b^(2^n) + 1 (iters == 0: the whole PRP loop, then the Gerbicz-Li
 * check) or the first `iters` squarings of its PRP loop (no check); in a proof batch always the
 * whole test and the proof.  *slot receives the slot; returns
 * GFE_EFULL when every slot is busy. */
int gfe_batch_submit(gfe_batch *bt, uint32_t b, uint64_t iters, uint64_t tag, int32_t *slot);
/* As gfe_batch_submit for a candidate of n (the batch's n, or 12 or 13 in a mixed batch; GFE_EINVAL
 * otherwise). */
int gfe_batch_submit_n(gfe_batch *bt, uint32_t n, uint32_t b, uint64_t iters, uint64_t tag, int32_t *slot);
/* Runs the pending register operations of every slot, then one block of up to 64 squaring rounds
 * for every slot in a squaring run.  *active receives the number of busy slots after the step. */
int gfe_batch_step(gfe_batch *bt, int32_t *active);
/* Moves up to `max` finished results to out (their slots become free); *count receives how many. */
int gfe_batch_poll(gfe_batch *bt, gfe_batch_result *out, uint32_t max, uint32_t *count);
/* Number of busy slots. */
int gfe_batch_active(const gfe_batch *bt);
/* Squaring blocks enqueued on the device and not finished yet (>= 0), or a negative error code. */
int gfe_batch_backlog(gfe_batch *bt);
/* Device bytes a batch of `capacity` slots allocates (slots, shared twiddles, flags). */
size_t gfe_batch_device_bytes(uint32_t n, uint32_t capacity);
size_t gfe_batch_proof_device_bytes(uint32_t n, uint32_t capacity, uint32_t depth);
/* Canonical digits of register `reg` (< 3, or < 3 + 2^depth in a proof batch) of a busy slot (for
 * checks; waits for the queued work). */
int gfe_batch_get_digits(gfe_batch *bt, int32_t slot, uint32_t reg, int32_t *out);

/* Benchmark: times `rounds` squaring rounds (one squaring of every busy slot per round) with HIP
 * events, without advancing the slots' programs (their register 0 is squared in place).  Rounds
 * are enqueued as gfe_batch_step enqueues them (graph launches of up to 64 rounds when on). */
int gfe_batch_time_rounds(gfe_batch *bt, uint32_t rounds, double *ms);

/* Squaring blocks as HIP graphs (on by default): each block of up to 64 rounds is captured once
 * per (rounds, slot range, plan) and replayed with one launch; off, its kernels are launched one by
 * one.  Results are identical either way.  Replaces the reference's per-kernel enqueue loop
 * (src/transformGPU.h:665-693, src/ocl.h:626-651) for batches. */
int gfe_batch_set_graphs(gfe_batch *bt, int on);

/* Runs quick tests (or their first `iters` squarings) of `count` candidates of one n to completion
 * in one batch of min(count, capacity) slots, refilling slots as candidates finish.
 * results[k] is the result of b[k]. */
int gfe_batch_run(uint32_t n, const uint32_t *b, uint32_t count, uint64_t iters, uint32_t capacity, int device,
                  gfe_batch_result *results);

#ifdef __cplusplus
}
#endif
#endif
