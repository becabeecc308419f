// gfe_engine.h -- host-side engine object behind the C ABI of include/gfe.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "gfe_kernels.h"
#include "gfe_pass.h"
#include "gfe_spec.h"
#include "gfe_fused.h"
#include "gfe_limb.h"

namespace gfe {

// Pass plan of one n: a column levels (2^a-point column transforms, col_ng columns per
// workgroup, 2^col_loge elements per thread) and n - a row levels (row_ng rows per workgroup,
// 2^row_loge elements per thread); norm_s digits per thread in the normalisation.
struct Plan {
    int a = 0, col_loge = 4, col_ng = 16, row_loge = 4, row_ng = 1, norm_s = 8;
    int row_tw = 0;  // 1: row twiddles formed from per-row factors (row_body TWF)
    int fused = 0;   // 1: all-prime column passes with the CRT and carry in the inverse one
    int pipe = 0;    // 1: persistent column and normalisation passes that prefetch the next block
    int spec = 0;    // 1: column passes and the squaring row pass with compile-time geometry (gfe_pass.h)
    int fuse_sq = 0; // 1: the squaring in one workgroup per candidate (gfe_fused.hip, n <= 13); the
                     //    other fields still shape the multiplicand passes
    int limb = 0;    // 1: the two-digit-limb transform (gfe_limb.h): length N/2 over four 31-bit primes,
                     //    compile-time passes; a and the row fields then describe the N/2-point transform
};

Plan default_plan(uint32_t n);
Plan default_batch_plan(uint32_t n, uint32_t capacity);
bool plan_valid(const Plan &P, uint32_t n);
std::string plan_str(const Plan &P, uint32_t n);
void plan_to_array(const Plan &P, int32_t out[6]);
std::vector<Plan> candidate_plans(uint32_t n);
unsigned persistent_grid(unsigned nblk, unsigned groups, unsigned resident);
bool plan_from_array(const int32_t in[6], uint32_t n, Plan *out);
void select_primes(uint32_t b, uint32_t n, int *rns, int *is32);
// The prime set the transform computes in for a set chosen by select_primes.
int transform_is32(int rns, int is32);
uint32_t prime_of(int is32, int k);
uint32_t mulw_const(uint32_t w, uint32_t p);
// psi 2^32 mod p for prime k of a set: psi = g^((p-1) / 2^(n+1)), the primitive 2^(n+1)-th root the
// twiddle tables are powers of; and p^-1 mod 2^32.
uint32_t psi_mont(int is32, int k, uint32_t n);
uint32_t prime_qinv(uint32_t p);
// Sets the kernels' LDS limits once per process (engine, compile-time and single-workgroup passes).
void engine_set_lds_attributes();
CrtConst make_crt(int rns, int is32, uint32_t n, uint32_t b = 0);
LimbCrt make_limb_crt(uint32_t n, uint32_t b);
int carry_rounds(uint32_t b, uint32_t n);
// The two-digit-limb transform is exact for (b, n): an NTT3 candidate (the reference's rule),
// b >= 2^23 (B = b^2 >= 2^46) and N l_max^2 < P/4 with l_max = (b/2 + 4)(b + 1), P the product of
// the four primes.
bool limb_ok(uint32_t b, uint32_t n);
// The limb transform with its inverse column pass fused to the normalisation (flags 2 | 8 | 32):
// the range-start limbs enter the next transform unnormalised, so its bound counts them larger.
bool limb_fused_ok(uint32_t b, uint32_t n, int ng);
// The plan an engine of (b, n) starts with (gfe_engine_plan).
Plan engine_plan_for(uint32_t b, uint32_t n);
size_t engine_bytes(int rns, size_t N, uint32_t num_regs, bool limb_tables);
// Plans of the limb transform the autotuner times (n: digits exponent).
std::vector<Plan> candidate_limb_plans(uint32_t n);
Plan default_limb_plan(uint32_t n);
void make_twiddle_tables(const int *sets, const int *primes, int count, uint32_t n, uint32_t *out, int max_threads = 8);
int set_error(int code, const std::string &msg);
extern thread_local std::string g_last_error;

struct Engine {
    uint32_t b = 0, n = 0, num_regs = 0;
    int device = 0, rns = 0, is32 = 0;
    int ref_is32 = 0;  // the reference's set (select_primes): the context payload's residues
    size_t N = 0;
    Plan plan;
    PrimeArgs pr;
    hipStream_t stream = nullptr;
    int32_t *regs = nullptr;
    uint32_t *ws = nullptr, *zp = nullptr;
    int64_t *carry = nullptr;
    uint32_t *tw = nullptr;
    CrtConst *crt = nullptr;
    // two-digit-limb transform (limb_ok): four twiddle tables of N/2 entries, its CRT constants,
    // its primes and b 2^32 mod p of each
    uint32_t *tw_limb = nullptr;
    void *limb_crt = nullptr;
    PrimeArgs pr4;
    uint32_t bm4[4] = {0, 0, 0, 0};
    bool can_limb = false;
    int tn() const { return plan.limb ? (int)n - 1 : (int)n; }   // log2 of the transform length
    int tp() const { return plan.limb ? 4 : rns; }               // primes of the transform
    const uint32_t *cur_tw() const { return plan.limb ? tw_limb : tw; }
    const PrimeArgs &cur_pr() const { return plan.limb ? pr4 : pr; }
    int rounds = 0;   // carry rounds of the fused inverse column pass
    int cus = 0;      // compute units of the device
    std::map<std::tuple<const void *, int, size_t>, unsigned> resident;  // per (kernel, threads, LDS)
    unsigned resident_blocks(const void *fn, int threads, size_t lds);
    std::vector<hipEvent_t> events;
    bool dev_ok = false;
    // The carries of the last normalisation are not yet in register 0: the next squaring's column
    // forward pass adds them on the way in (plans with compile-time passes); every other use of the
    // registers runs the carry kernel first (flush_carry).
    bool carry_pending = false;
    // what the pending carries are: 0 k_norm's workgroup carries, 1 k_norm_limb's boundary pairs,
    // 2 the range slots of ks_col_inv_norm_limb (ranges of 2^pend_rlog limbs)
    int pend_kind = 0, pend_rlog = 0;
    bool defer_carry() const { return plan.spec && !plan.pipe && (!plan.fused || plan.limb) && !plan.fuse_sq; }
    void flush_carry();

    ~Engine();
    int init(uint32_t b, uint32_t n, uint32_t num_regs, int device);
    // Binds the engine to device memory it does not own (no allocation, nothing freed on
    // destruction): register 0 = `digits`, the squaring's work space, twiddles, carries and CRT
    // constants at the given addresses, work queued on `stream`, pass plan `plan`.  Only squarings
    // (square, flush_carry) may run on a bound engine.
    void bind(uint32_t b, uint32_t n, int device, hipStream_t stream, int32_t *digits, uint32_t *ws, uint32_t *tw,
              int64_t *carry, CrtConst *crt, const Plan &plan);
    size_t lds_col() const;
    size_t lds_row() const;
    void launch_col_fwd(uint32_t src, uint32_t *dst);
    void launch_col_inv(int dup);
    void launch_row(int mode, uint32_t *data);
    void launch_norm(int dup);
    void launch_carry();          // the pending carries of the current plan's normalisation
    void launch_carry_digits();   // workgroup carries of k_norm / k_mulscalar
    void launch_carry_limbs();    // boundary carries of k_norm_limb
    void launch_carry_slots();    // range carries of ks_col_inv_norm_limb
    void square(int dup);
    void launch_fused(int dup);
    void mul();
    void init_multiplicand(uint32_t src);
    void mul_scalar(int32_t a);
    void set(uint32_t a);
    void copy(uint32_t dst, uint32_t src);
    int check(const char *what);
};

}  // namespace gfe
