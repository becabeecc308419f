/* gfe_driver.h -- C ABI of the PRP driver that runs on the squaring engine (genefer's caller side,
 * src/genefer.h:441-766).  Used by tests and bench.py; a genefer-style C++ driver uses the
 * transform interface (genefer22_amd/csrc/transform_hip.h) directly instead.
 * Same conventions as gfe.h: GFE_OK or a negative code, gfe_last_error() for the message. */
#ifndef GFE_DRIVER_H
#define GFE_DRIVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gfe_result {
    uint64_t res64;    /* canonical residue mod 2^64 (gint::isOne, src/gint.h:169-202) */
    uint64_t old64;    /* low bytes of the 8 top canonical digits */
    uint64_t hash64;   /* gint::gethash64 of the canonical residue, 0 when it is 1 */
    uint64_t pkey;     /* Pietrzak-Li proof key (gfe_run_proof), else 0 */
    uint64_t esize;    /* bits of E = b^(2^n) */
    uint64_t iters;    /* squarings of the PRP loop that ran */
    int32_t is_prp;    /* residue == 1 */
    int32_t gl_ok;     /* Gerbicz-Li check: 1 pass, 0 fail, -1 not run */
    double seconds;    /* wall time of the run */
} gfe_result;

/* Host only (no device): the top bits of the PRP exponent E = b^(2^n) that a run of the PRP loop's
 * first k squarings reads (src/genefer.h:479-552), without computing E -- T = floor(E / 2^shift)
 * with at least k bits, exact, as little-endian 64-bit words (*nwords of them, at most max_words),
 * and *esize = bits of E.  The batch and the serving loop use it for truncated runs. */
int gfe_exponent_top(uint32_t b, uint32_t n, uint64_t k, uint64_t *words, uint32_t max_words, uint32_t *nwords,
                     uint64_t *shift, uint64_t *esize);

/* genefer's quick test (quick(), src/genefer.h:736-748): PRP of b^(2^n)+1 in base 2 with the
 * Gerbicz-Li check at the reference's cadence B_GL = B_GerbiczLi(esize). */
int gfe_run_quick(uint32_t b, uint32_t n, int device, gfe_result *out);

/* The first `iters` squarings of the PRP loop (all when iters == 0 or >= esize), with the d(t)
 * update every b_gl iterations (0 -> the reference's cadence, < 0 -> none).  When digits_out is not
 * null it receives the canonical digits of r0 (N int32). */
int gfe_run_prp(uint32_t b, uint32_t n, int device, uint64_t iters, int32_t b_gl, int32_t *digits_out,
                gfe_result *out);

/* As gfe_run_prp, recording RES64 / old64 of r0 after every `every`-th squaring (counted from 1)
 * into res64[] / old64[] (iters / every entries); the d(t) update runs at the reference's cadence. */
int gfe_run_prp_trace(uint32_t b, uint32_t n, int device, uint64_t iters, uint64_t every, uint64_t *res64,
                      uint64_t *old64, gfe_result *out);

/* proof(), src/genefer.h:750-766: PRP with Pietrzak-Li checkpoints, the Gerbicz-Li check, then the
 * proof.  The checkpoints stay in device registers (the reference's fast_checkpoints) whenever the
 * 3 + 2^depth registers fit in 64 GiB of HBM (n = 22 at depth 8: 4.1 GiB), host copies otherwise.
 * mu_hash64 (may be null) receives gethash64 of mu_0..mu_depth (depth + 1 values). */
int gfe_run_proof(uint32_t b, uint32_t n, int device, int depth, uint64_t *mu_hash64, gfe_result *out);

/* As gfe_run_proof, writing the reference's files: `<main_path>.proof` (int version = 1, int depth,
 * then mu_0..mu_depth as residue records -- uint32 size, uint32 base, size int32 canonical digits --
 * and the CRC-32 trailer ~crc ^ 0xa23777ac; src/genefer.h:620-734, src/gint.h:160-167,
 * src/file.h:150-161).  flags: GFE_PROOF_CKPT_FILES also writes the checkpoint files
 * `<main_path>_<i>.ckpt` (one residue record and the trailer each; src/genefer.h:140-144, 536-546),
 * which the reference's proof and server modes read; GFE_PROOF_HOST_CKPTS keeps the checkpoints in
 * host memory even when they fit in HBM. */
#define GFE_PROOF_CKPT_FILES 1
#define GFE_PROOF_HOST_CKPTS 2
int gfe_write_proof(uint32_t b, uint32_t n, int device, int depth, const char *main_path, int flags,
                    uint64_t *mu_hash64, gfe_result *out);

/* The server role of the proof protocol (server(), src/genefer.h:771-880): reads
 * `<main_path>.proof`, recomputes the proof key from its records (out->pkey; out->res64 / old64 /
 * is_prp are mu_0's, the PRP residue) and writes the certificate `<main_path>.cert` (int version 1,
 * int B = B_PietrzakLi(esize, depth), the residue record of v2^rnd2, p2 rnd2 + rnd3 in GMP's raw
 * integer format outside the CRC, the CRC-32 trailer).  *ckey receives the check key that a check
 * of this certificate must reproduce.  The reference draws rnd1 in [2, 64), rnd2 in [16, 256) and
 * rnd3 in [4, 65536) from std::rand seeded with the time; here the caller passes them (outside
 * those ranges: GFE_EINVAL), so a certificate can be reproduced.  A missing, malformed or corrupted
 * proof file is GFE_EINVAL. */
int gfe_run_server(uint32_t b, uint32_t n, int device, const char *main_path, uint32_t rnd1, uint32_t rnd2,
                   uint32_t rnd3, uint64_t *ckey, gfe_result *out);

/* The check role (check(), src/genefer.h:882-1062, without its context file): reads the certificate
 * `<main_path>.cert`, raises its residue to 2^B under the Gerbicz test and computes 2^p2 under the
 * Gerbicz-Li test.  out->gl_ok = 1 when both tests pass (*ckey then holds the check key, equal to
 * the server's for a correct proof), 0 when one fails (*ckey = 0).  A missing or malformed
 * certificate is GFE_EINVAL. */
int gfe_run_check(uint32_t b, uint32_t n, int device, const char *main_path, uint64_t *ckey, gfe_result *out);

/* The PRP loop with the reference's context file at ctx_path (src/genefer.h:198-276, 479-552):
 * resumes from it (or from ctx_path.old) when it holds a PRP-loop context of this engine kind,
 * else starts from r0 = 1.  Writes a context every `save_every` squarings of this call (0: none)
 * and, once `stop_after` squarings of this call have run (0: no limit), writes one and returns, as
 * the reference does when asked to quit.  The file is the reference GPU build's: int -1, int 0
 * (the PRP loop), int i (next loop index, -1 at the end), double elapsed seconds, the transform
 * payload of gfe_save for registers 0 and 1, the CRC-32 trailer; each save goes through
 * ctx_path.new and keeps the previous file as ctx_path.old.  out->iters counts the squarings of
 * the loop done over all calls; once the loop has ended the Gerbicz-Li check runs at the
 * reference's cadence (out->gl_ok 1/0) and out->is_prp is set; before that out->gl_ok = -1 and
 * out->is_prp = -1.  out->seconds is the elapsed time the context records.  *done (may be null)
 * is 1 when the loop has ended. */
int gfe_run_prp_resumable(uint32_t b, uint32_t n, int device, const char *ctx_path, uint64_t save_every,
                          uint64_t stop_after, int32_t *done, gfe_result *out);

/* Canonical residue of a digit vector with the product's gint (src/gint.h:40-225):
 * canon (may be null) receives the canonical digits. */
int gfe_residue(const int32_t *digits, uint32_t n, uint32_t b, int32_t *canon, gfe_result *out);

#ifdef __cplusplus
}
#endif
#endif
