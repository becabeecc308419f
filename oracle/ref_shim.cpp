// ref_shim.cpp -- C-ABI shim over the reference genefer CPU transforms.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the reference's own
// sources (src/transform_{sse2,sse4,avx,fma,512,i32}.cpp, read in place from the
// reference checkout) into oracle/_ref/libgfref.so by `make -C oracle ref`.  Nothing under
// genefer22_amd/ links or loads it: it is the checker that the HIP engine is compared
// with and the CPU baseline that bench.py times beside it.
//
// It drives the reference's abstract transform (src/transform.h:42-256) exactly the way
// the reference driver does (src/genefer.h:479-552 prp, 557-615 GL) and returns the
// canonical residue through the reference's own gint (src/gint.h:40-225).

#include <cstdint>
#include <cstring>
#include <string>
#include <stdexcept>
#include <gmp.h>

#include "transform.h"
#include "gint.h"

namespace {

struct ref_engine {
    transform *t = nullptr;
    gint *g = nullptr;
    uint32_t b = 0, n = 0;
    std::string ttype;
    std::string err;
};

thread_local std::string g_last_error;

int ilog2_32(uint32_t x) { return 31 - __builtin_clz(x); }

}  // namespace

extern "C" {

const char *gfref_last_error() { return g_last_error.c_str(); }

// create_cpu: src/transform.h:122-192.  impl "" = best for this CPU, or one of
// 512 fma avx sse4 sse2 i32.
void *gfref_create(uint32_t b, uint32_t n, uint32_t num_threads, const char *impl, uint32_t num_regs, int check_error) {
    try {
        auto *e = new ref_engine;
        e->b = b; e->n = n;
        e->t = transform::create_cpu(b, n, num_threads, std::string(impl ? impl : ""), num_regs, check_error != 0, e->ttype);
        e->g = new gint(size_t(1) << n, b);
        return e;
    } catch (const std::exception &ex) {
        g_last_error = ex.what();
        return nullptr;
    }
}

void gfref_destroy(void *h) {
    auto *e = static_cast<ref_engine *>(h);
    if (!e) return;
    delete e->t; delete e->g; delete e;
}

const char *gfref_type(void *h) { return static_cast<ref_engine *>(h)->ttype.c_str(); }

void gfref_set(void *h, uint32_t a) { static_cast<ref_engine *>(h)->t->set(a); }
void gfref_square_dup(void *h, int dup) { static_cast<ref_engine *>(h)->t->squareDup(dup != 0); }
void gfref_square_mul(void *h, int32_t a) { static_cast<ref_engine *>(h)->t->squareMul(a); }
void gfref_init_multiplicand(void *h, uint32_t src) { static_cast<ref_engine *>(h)->t->initMultiplicand(src); }
void gfref_mul(void *h) { static_cast<ref_engine *>(h)->t->mul(); }
void gfref_copy(void *h, uint32_t dst, uint32_t src) { static_cast<ref_engine *>(h)->t->copy(dst, src); }
double gfref_get_error(void *h) { return static_cast<ref_engine *>(h)->t->getError(); }

// Register 0 as digits: raw (as getZi returns them) when canonical == 0, else
// canonical digits in [0, b) (gint::unbalance, src/gint.h:40-102).
void gfref_get_digits(void *h, int32_t *out, int canonical) {
    auto *e = static_cast<ref_engine *>(h);
    e->t->getInt(*e->g);
    if (canonical) e->g->unbalance();
    std::memcpy(out, e->g->data(), (size_t(1) << e->n) * sizeof(int32_t));
}

// setInt: src/transform.h:241-246 (balances first).
void gfref_set_digits(void *h, const int32_t *in) {
    auto *e = static_cast<ref_engine *>(h);
    std::memcpy(e->g->data(), in, (size_t(1) << e->n) * sizeof(int32_t));
    e->g->reset();
    e->t->setInt(*e->g);
}

// RES64 / old64 / isOne (src/gint.h:169-202) and hash64 (204-219) of register 0.
int gfref_res64(void *h, uint64_t *res64, uint64_t *old64, uint64_t *hash64) {
    auto *e = static_cast<ref_engine *>(h);
    e->t->getInt(*e->g);
    const bool one = e->g->isOne(*res64, *old64);
    *hash64 = one ? 0 : e->g->gethash64();
    return one ? 1 : 0;
}

// Runs the reference's Fermat PRP loop (src/genefer.h:479-552) for the first `iters`
// exponent bits of E = b^(2^n) (all of them when iters == 0), with the Gerbicz-Li
// d(t) update every B_GL iterations (B_GL = 0 -> the reference's B_GerbiczLi(esize),
// src/genefer.h:467-471; B_GL < 0 -> no d(t) update).  When the whole exponent is run
// and `do_gl` is set, the Gerbicz-Li check (src/genefer.h:557-615) runs afterwards and
// its outcome is returned in *gl_ok (register 0 then holds d(t+1), as in the reference;
// the residue outputs are read before the check, as quick() does, src/genefer.h:736-748).
// Returns esize (number of exponent bits), or -1 on error.
long gfref_prp(void *h, long iters, int b_gl, int do_gl, uint64_t *res64, uint64_t *old64,
               int *is_one, int *gl_ok) {
    auto *e = static_cast<ref_engine *>(h);
    transform *t = e->t;
    mpz_t exponent; mpz_init(exponent);
    mpz_ui_pow_ui(exponent, e->b, 1ul << e->n);
    const long esize = (long)mpz_sizeinbase(exponent, 2);
    const int B_GL = (b_gl == 0) ? (int)((esize - 1) / (2l << (ilog2_32((uint32_t)esize) / 2)) + 1) : b_gl;
    t->set(1);
    t->copy(1, 0);
    const long i0 = esize - 1;
    const long i_end = (iters <= 0 || iters >= esize) ? 0 : esize - iters;
    for (long i = i0; i >= i_end; --i) {
        t->squareDup(mpz_tstbit(exponent, (mp_bitcnt_t)i) != 0);
        if (B_GL > 0 && (i % B_GL == 0) && (i / B_GL != 0)) {
            t->copy(2, 0);
            t->mul(1);
            t->copy(1, 0);
            t->copy(0, 2);
        }
    }
    uint64_t h64 = 0;
    *is_one = gfref_res64(h, res64, old64, &h64);
    *gl_ok = -1;
    if (i_end == 0 && do_gl && B_GL > 0) {
        // GL: src/genefer.h:557-615
        t->mul(1);
        t->copy(2, 0);
        t->copy(0, 1);
        for (int i = B_GL - 1; i >= 0; --i) t->squareDup(false);
        t->copy(1, 0);
        mpz_t res, ee, tt; mpz_init_set_ui(res, 0); mpz_init_set(ee, exponent); mpz_init(tt);
        while (mpz_sgn(ee) != 0) {
            mpz_mod_2exp(tt, ee, (unsigned long)B_GL);
            mpz_add(res, res, tt);
            mpz_div_2exp(ee, ee, (unsigned long)B_GL);
        }
        t->set(1);
        for (int i = (int)mpz_sizeinbase(res, 2) - 1; i >= 0; --i) t->squareDup(mpz_tstbit(res, (mp_bitcnt_t)i) != 0);
        t->mul(1);
        t->getInt(*e->g);
        const uint64_t h1 = e->g->gethash64();
        t->copy(0, 2);
        t->getInt(*e->g);
        const uint64_t h2 = e->g->gethash64();
        *gl_ok = (h1 == h2) ? 1 : 0;
        mpz_clear(res); mpz_clear(ee); mpz_clear(tt);
    }
    mpz_clear(exponent);
    return esize;
}

// The reference's gint on a digit vector: canonical digits (into canon, may be null), RES64 /
// old64 / isOne (src/gint.h:169-202), hash64 / hash32 (204-225; 0 when the value is 1).
int gfref_gint(const int32_t *digits, uint32_t n, uint32_t b, int32_t *canon, uint64_t *res64, uint64_t *old64,
               uint64_t *hash64, uint32_t *hash32) {
    gint g(size_t(1) << n, b);
    std::memcpy(g.data(), digits, (size_t(1) << n) * sizeof(int32_t));
    g.reset();
    const bool one = g.isOne(*res64, *old64);
    *hash64 = one ? 0 : g.gethash64();
    *hash32 = one ? 0 : g.gethash32();
    if (canon) std::memcpy(canon, g.data(), (size_t(1) << n) * sizeof(int32_t));
    return one ? 1 : 0;
}

// The PRP loop as gfref_prp, recording RES64 / old64 of register 0 after every `every`-th
// squaring (squarings counted from 1) up to `iters`; returns the number of records.
long gfref_prp_trace(void *h, long iters, long every, uint64_t *res64, uint64_t *old64) {
    auto *e = static_cast<ref_engine *>(h);
    transform *t = e->t;
    mpz_t exponent; mpz_init(exponent);
    mpz_ui_pow_ui(exponent, e->b, 1ul << e->n);
    const long esize = (long)mpz_sizeinbase(exponent, 2);
    const int B_GL = (int)((esize - 1) / (2l << (ilog2_32((uint32_t)esize) / 2)) + 1);
    t->set(1);
    t->copy(1, 0);
    const long i0 = esize - 1;
    const long i_end = (iters <= 0 || iters >= esize) ? 0 : esize - iters;
    long rec = 0;
    for (long i = i0; i >= i_end; --i) {
        t->squareDup(mpz_tstbit(exponent, (mp_bitcnt_t)i) != 0);
        if ((i % B_GL == 0) && (i / B_GL != 0)) {
            t->copy(2, 0);
            t->mul(1);
            t->copy(1, 0);
            t->copy(0, 2);
        }
        const long k = i0 - i + 1;
        if (every > 0 && k % every == 0) {
            uint64_t hh = 0;
            gfref_res64(h, &res64[rec], &old64[rec], &hh);
            ++rec;
        }
    }
    mpz_clear(exponent);
    return rec;
}

}  // extern "C"
