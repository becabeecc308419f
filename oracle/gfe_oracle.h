/* gfe_oracle.h -- CPU restatement of genefer's exact-integer squaring engine.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg (CPU timings and the RES64 check beside them) may load
 * liboracle.so; the product (genefer22_amd/) never does.
 *
 * Every function states the reference file:line it restates.  "parity unpinned" marks a
 * function that no golden vector, reference build or closed form checks (none at present:
 * see the pinning table in tests/test_oracle.py).
 */
#ifndef GFE_ORACLE_H
#define GFE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Prime index 0..2 = S set {127*2^24+1, 63*2^25+1, 15*2^27+1} (generators 3, 5, 31),
 * 3..5 = U set {125*2^25+1, 243*2^24+1, 235*2^24+1} (generators 3, 7, 3).
 * src/transformGPU.h:37-115; src/transformCPUi32.h:19-22,84-86. */
uint32_t gfo_prime(int idx);
uint32_t gfo_generator(int idx);

/* The reference's per-prime Montgomery constants, derived from p and g by closed forms
 * (src/transformGPU.h:37-105): out[0..8] = P, Q (p^-1 mod 2^32), R (2^32 mod p),
 * RSQ (2^64 mod p), H (g*R), IM (I*R, I = g^((p-1)/4)), MFIM (I*R^2), SQRTI (g^((p-1)/8)*R),
 * ISQRTI (I*g^((p-1)/8)*R). */
void gfo_prime_constants(int idx, uint32_t out[9]);

/* Garner constants of a 3-prime set (src/transformGPU.h:67-75,107-115):
 * out32[0..2] = MF of 1/P2 mod P1, 1/P3 mod P1, 1/P3 mod P2; out32[3] = (P1P2P3) mod 2^32;
 * out32[4] = (P1P2P3/2) mod 2^32; out64[0] = (P1P2P3) >> 32; out64[1] = (P1P2P3/2) >> 32. */
void gfo_garner_constants(int is32, uint32_t out32[5], uint64_t out64[2]);

/* NORM = (N/2)^-1 mod p as ZPT::norm computes it (src/transformGPU.h:163, 1009-1011). */
uint32_t gfo_norm(int idx, uint32_t n);

/* Prime-set selection of create_ocl (src/transform_ocl.cpp:35-50): *rns = 2 or 3,
 * *is32 = 0 (S primes) or 1 (U primes). */
void gfo_select(uint32_t b, uint32_t n, int *rns, int *is32);

/* Twiddle table of transformGPUs (src/transformGPU.h:1046-1078), plain (non-Montgomery)
 * values: w[s + j] = r_{4s}^{bitrev(j, 2s) + 1} for s = 1, 2, ..., N/4, j < s
 * (entries 1 .. N/2-1); w[0] is unused and set to 0.  `w` has N/2 entries. */
void gfo_twiddles(int idx, uint32_t n, uint32_t *w);

/* Engine: registers of N int32 digits, a multiplicand in transform domain, prime set chosen
 * by gfo_select unless rns/is32 are forced (rns = 0 -> select). */
typedef struct gfo_engine gfo_engine;
gfo_engine *gfo_create(uint32_t b, uint32_t n, int num_regs, int rns, int is32);
void gfo_destroy(gfo_engine *e);
int gfo_rns(const gfo_engine *e);
int gfo_is32(const gfo_engine *e);

void gfo_set(gfo_engine *e, uint32_t a);                 /* r0 = a          src/transform.h:57 */
void gfo_square_dup(gfo_engine *e, int dup);             /* r0 = r0^2 (x2)  src/transform.h:58 */
void gfo_square_mul(gfo_engine *e, int32_t a);           /* r0 = a*r0^2     src/transform.h:59 */
void gfo_init_multiplicand(gfo_engine *e, int src);      /* rm = T(r_src)   src/transform.h:60 */
void gfo_mul(gfo_engine *e);                             /* r0 *= rm        src/transform.h:61 */
void gfo_copy(gfo_engine *e, int dst, int src);          /*                 src/transform.h:63 */
void gfo_get_digits(const gfo_engine *e, int reg, int32_t *out);
void gfo_set_digits(gfo_engine *e, int reg, const int32_t *in);

/* One squaring (dup) or product of two digit vectors mod b^N + 1 by plain schoolbook
 * negacyclic convolution with exact 128-bit accumulation and a full carry; O(N^2), for tiny N.
 * out may alias a.  Independent of the transform: used to pin it. */
void gfo_schoolbook_mul(const int32_t *a, const int32_t *c, int32_t *out, uint32_t n, uint32_t b);

/* gint (src/gint.h:40-225). */
void gfo_unbalance(int32_t *d, uint32_t n, uint32_t b);
void gfo_balance(int32_t *d, uint32_t n, uint32_t b);
int gfo_is_one(const int32_t *canon, uint32_t n, uint32_t b, uint64_t *res64, uint64_t *old64);
uint64_t gfo_hash64(const int32_t *canon, uint32_t n);
uint32_t gfo_hash32(const int32_t *canon, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
