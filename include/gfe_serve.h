/* gfe_serve.h -- C ABI of the serving loop: continuous batching of candidate requests of mixed n on
 * one device, with every admitted candidate's state resident in HBM.
 *
 * No counterpart in the reference, which tests one candidate per process
 * (src/genefer.h:1363-1573).  Each request runs the reference's PRP work for one candidate
 * (src/genefer.h:479-615): its full quick test, or the first `iters` squarings of its PRP loop.
 *
 * Requests are queued in arrival order.  A request is admitted into a pool (a gfe_batch of its n;
 * for the quick tests of n = 12 and 13 one mixed batch, gfe_batch_create_mixed, that squares both
 * in one launch per block; for proof requests a proof batch of its n and depth) that has a free slot; when no pool of its n has one, a new pool is created if its device bytes
 * (gfe_batch_device_bytes) fit in the HBM budget, after releasing idle pools of other n if needed.
 * Otherwise the request waits.  Every pool runs on its own stream; a pump enqueues one block of
 * squaring rounds on each pool whose device backlog is below two blocks, so the pools' work
 * overlaps on the device.  A candidate's residues do not depend on what else is admitted.
 *
 * Lone requests: at light load (at most `max_rate` arrivals in the last second, gfe_server_set_solo;
 * 8 by default) a quick or truncated test that is admitted while no pool has work runs alone on a
 * warm single-candidate engine of its kind (one per n and prime set, moved between bases with
 * gfe_rebase; gfe_server_prepare warms the one for n's NTT3 candidates), on a host thread, exactly
 * as gfe_run_quick / gfe_run_prp run it: the request then takes its time alone on the plain
 * engine instead of a batch's at one active slot.  Two such runs may go side by side (a single
 * candidate below n = 18 leaves most of the device idle); a request that finds both busy waits for
 * them when one is short (at most 65,536 squarings), else it goes to the pools;
 * under heavier load every request does, since a lone run would hold the device while the pools
 * fill.  Warm engines count in the HBM budget and are freed when a pool needs the room.
 *
 * Conventions of gfe.h: GFE_OK or a negative code, gfe_last_error() for the message; one server
 * must not be used from two threads at once.
 */
#ifndef GFE_SERVE_H
#define GFE_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "gfe_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gfe_server gfe_server;

typedef struct gfe_serve_result {
    uint64_t tag;
    uint32_t n, b;
    uint64_t res64, old64, hash64;
    uint64_t esize, iters;
    int32_t is_prp, gl_ok;     /* as gfe_batch_result */
    double t_arrival;          /* seconds on the server clock: submitted */
    double t_admit;            /* admitted into a pool slot */
    double t_done;             /* result available to the caller */
    int32_t depth;             /* proof requests: the proof's depth, else 0 */
    int32_t pad;
    uint64_t pkey;             /* proof requests: the proof key (as gfe_batch_result) */
    uint64_t mu_hash64[GFE_BATCH_MAX_DEPTH + 1];  /* gethash64 of mu_0 .. mu_depth */
} gfe_serve_result;

typedef struct gfe_server_stats {
    uint64_t submitted, admitted, completed, queued;
    uint64_t bytes_budget, bytes_reserved, bytes_peak;
    uint32_t pools, pools_created, pools_released, busy_slots;
    uint64_t pumps, blocks;    /* pump calls; squaring blocks enqueued */
    uint64_t solo;             /* requests run alone on a warm single-candidate engine */
} gfe_server_stats;

/* A server on `device` whose pools may hold at most `hbm_budget` device bytes (0: 3/4 of the
 * device's free memory at creation). */
int gfe_server_create(int device, uint64_t hbm_budget, gfe_server **out);
int gfe_server_destroy(gfe_server *s);
/* Seconds since the server was created (steady clock). */
double gfe_server_clock(const gfe_server *s);
/* Creates a pool for n ahead of its first request (`capacity` slots, 0 = default for n); for
 * n <= 13 the mixed pool, unless one exists; and, with the lone-request path on, a warm engine for
 * n's NTT3 candidates. */
int gfe_server_prepare(gfe_server *s, uint32_t n, uint32_t capacity);
/* The lone-request path while at most `max_rate` requests arrived in the last second (default 8,
 * about four fifths of one plain engine's capacity for a mix of n = 12..20; a saturated mix sees
 * two to three times that); 0: every request through a pool. */
int gfe_server_set_solo(gfe_server *s, int max_rate);
/* Queues a request (arrival time = now). */
int gfe_server_submit(gfe_server *s, uint32_t n, uint32_t b, uint64_t iters, uint64_t tag);
/* Queues a proof request: the full test and the Pietrzak-Li proof of `depth` (1 ..
 * GFE_BATCH_MAX_DEPTH; 0 = gfe_server_submit), run in a proof pool of its (n, depth) whose slots
 * keep the checkpoints in HBM; admitted by that pool's larger footprint
 * (gfe_batch_proof_device_bytes).  `iters` is ignored for depth > 0. */
int gfe_server_submit_proof(gfe_server *s, uint32_t n, uint32_t b, uint64_t iters, uint32_t depth, uint64_t tag);
/* Admits queued requests, then enqueues one block on every pool that has work and a short
 * backlog.  *busy receives the number of admitted, unfinished requests. */
int gfe_server_pump(gfe_server *s, int32_t *busy);
/* Moves up to `max` finished results to out; *count receives how many. */
int gfe_server_poll(gfe_server *s, gfe_serve_result *out, uint32_t max, uint32_t *count);
int gfe_server_get_stats(const gfe_server *s, gfe_server_stats *out);

/* Open-loop replay: request k (n[k], b[k], iters[k], proof depth depth[k] -- depth may be null:
 * no proofs) is submitted when the server clock reaches arrival[k] seconds (non-decreasing); runs
 * until every request has finished.  prepare_n lists the n whose pools (and warm engines) are made
 * before the clock starts; solo_rate is gfe_server_set_solo's max_rate (-1: the default, 0: every
 * request through a pool).  results[k] is the result of request k; stats (optional) receives the
 * final statistics. */
int gfe_serve_replay(const uint32_t *n, const uint32_t *b, const uint64_t *iters, const uint32_t *depth,
                     const double *arrival, uint32_t count, int device, uint64_t hbm_budget, const uint32_t *prepare_n,
                     uint32_t prepare_count, int solo_rate, gfe_serve_result *results, gfe_server_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
