// genefer_driver.cpp -- C ABI of the PRP driver (include/gfe_driver.h).
#include <chrono>
#include <cstring>
#include <exception>
#include <string>

#include "../../include/gfe_driver.h"
#include "genefer.h"

namespace gfe {
int set_error(int code, const std::string &msg);
}

namespace {

// Pietrzak-Li checkpoints stay in device registers (3 + i) whenever the 3 + 2^depth registers fit in
// 64 GiB of HBM: n = 22 at proof power 8 takes 4.1 GiB.  The reference keeps them on the GPU only
// for n <= 17 and writes files otherwise (src/genefer.h:1377-1385).
bool checkpoints_in_hbm(uint32_t n, int depth) {
    const double bytes = (3.0 + (double)(size_t(1) << depth)) * (double)(size_t(1) << n) * 4.0;
    return depth <= 11 && bytes <= 64.0 * 1024 * 1024 * 1024;  // gfe_create takes at most 4096 registers
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The candidate range of gfe_create (include/gfe.h): argument errors are GFE_EINVAL, reported
// before any device work.
int check_candidate(uint32_t b, uint32_t n) {
    if (n < 12 || n > 23) return gfe::set_error(GFE_EINVAL, "n must be in [12, 23]");
    if (b < 1024 || b > 2000000000u) return gfe::set_error(GFE_EINVAL, "b must be in [1024, 2000000000]");
    return GFE_OK;
}

void fill_residue(gfe::gint &g, gfe_result *out) {
    uint64_t r = 0, o = 0;
    const bool one = g.isOne(r, o);
    out->res64 = r;
    out->old64 = o;
    out->is_prp = one ? 1 : 0;
    out->hash64 = one ? 0 : g.gethash64();
}

}  // namespace

extern "C" {

int gfe_exponent_top(uint32_t b, uint32_t n, uint64_t k, uint64_t *words, uint32_t max_words, uint32_t *nwords,
                     uint64_t *shift, uint64_t *esize) {
    if (!words || !nwords || !shift || !esize) return gfe::set_error(GFE_EINVAL, "exponent_top arguments");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    mpz_t T;
    mpz_init(T);
    gfe::exponent_top(b, n, k, T, shift, esize);
    const size_t need = (mpz_sizeinbase(T, 2) + 63) / 64;
    int rc = GFE_OK;
    if (need > max_words) {
        rc = gfe::set_error(GFE_EINVAL, "exponent_top: output too small");
    } else {
        size_t cnt = 0;
        mpz_export(words, &cnt, -1, 8, 0, 0, T);
        *nwords = (uint32_t)cnt;
    }
    mpz_clear(T);
    return rc;
}

int gfe_run_quick(uint32_t b, uint32_t n, int device, gfe_result *out) {
    if (!out) return gfe::set_error(GFE_EINVAL, "null output");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    try {
        const double t0 = now_s();
        gfe::genefer G(b, n, device, 3);
        const size_t es = G.esize();
        const int B_GL = gfe::genefer::B_GerbiczLi(es);
        G.prp(B_GL, 0, false);
        G.t().getInt(G.gi());
        fill_residue(G.gi(), out);
        out->esize = es;
        out->iters = es;
        out->gl_ok = G.GL(B_GL) ? 1 : 0;
        out->seconds = now_s() - t0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EHIP, ex.what());
    }
}

int gfe_run_prp(uint32_t b, uint32_t n, int device, uint64_t iters, int32_t b_gl, int32_t *digits_out,
                gfe_result *out) {
    if (!out) return gfe::set_error(GFE_EINVAL, "null output");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    try {
        const double t0 = now_s();
        gfe::genefer G(b, n, device, 3);
        const size_t es = G.esize();
        const int B_GL = (b_gl == 0) ? gfe::genefer::B_GerbiczLi(es) : (b_gl < 0 ? 0 : b_gl);
        const long i_end = (iters == 0 || iters >= es) ? 0 : (long)(es - iters);
        G.prp(B_GL, 0, false, i_end);
        G.t().getInt(G.gi());
        fill_residue(G.gi(), out);
        if (digits_out) memcpy(digits_out, G.gi().data(), (size_t(1) << n) * sizeof(int32_t));
        out->esize = es;
        out->iters = es - (size_t)i_end;
        out->gl_ok = -1;
        out->seconds = now_s() - t0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EHIP, ex.what());
    }
}

int gfe_run_prp_resumable(uint32_t b, uint32_t n, int device, const char *ctx_path, uint64_t save_every,
                          uint64_t stop_after, int32_t *done, gfe_result *out) {
    if (!out || !ctx_path || !*ctx_path) return gfe::set_error(GFE_EINVAL, "resumable prp arguments");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    if (done) *done = 0;
    try {
        gfe::genefer G(b, n, device, 3);
        const size_t es = G.esize();
        const int B_GL = gfe::genefer::B_GerbiczLi(es);
        int resumed_at = -2, i_next = 0;
        double elapsed = 0;
        const bool ended = G.prp_ctx(B_GL, ctx_path, (long)save_every, (long)stop_after, resumed_at, i_next, elapsed);
        G.t().getInt(G.gi());
        fill_residue(G.gi(), out);
        out->esize = es;
        out->iters = (uint64_t)((long)es - 1 - (long)i_next);
        out->seconds = elapsed;
        if (ended) {
            out->gl_ok = G.GL(B_GL) ? 1 : 0;
        } else {
            out->gl_ok = -1;
            out->is_prp = -1;
        }
        if (done) *done = ended ? 1 : 0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EHIP, ex.what());
    }
}

int gfe_run_prp_trace(uint32_t b, uint32_t n, int device, uint64_t iters, uint64_t every, uint64_t *res64,
                      uint64_t *old64, gfe_result *out) {
    if (!out || !res64 || !old64 || every == 0) return gfe::set_error(GFE_EINVAL, "trace arguments");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    try {
        const double t0 = now_s();
        gfe::genefer G(b, n, device, 3);
        const size_t es = G.esize();
        const int B_GL = gfe::genefer::B_GerbiczLi(es);
        const long i_end = (iters == 0 || iters >= es) ? 0 : (long)(es - iters);
        size_t rec = 0;
        G.on_square = [&](long k) {
            if ((uint64_t)k % every == 0) {
                G.t().getInt(G.gi());
                G.gi().isOne(res64[rec], old64[rec]);
                ++rec;
            }
        };
        G.prp(B_GL, 0, false, i_end);
        G.on_square = nullptr;
        G.t().getInt(G.gi());
        fill_residue(G.gi(), out);
        out->esize = es;
        out->iters = es - (size_t)i_end;
        out->gl_ok = -1;
        out->seconds = now_s() - t0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EHIP, ex.what());
    }
}

int gfe_run_proof(uint32_t b, uint32_t n, int device, int depth, uint64_t *mu_hash64, gfe_result *out) {
    if (!out || depth < 1 || depth > 12) return gfe::set_error(GFE_EINVAL, "proof arguments");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    try {
        const double t0 = now_s();
        const bool fast = checkpoints_in_hbm(n, depth);
        gfe::genefer G(b, n, device, fast ? 3 + (size_t(1) << depth) : 3);
        const size_t es = G.esize();
        const int B_GL = gfe::genefer::B_GerbiczLi(es), B_PL = gfe::genefer::B_PietrzakLi(es, depth);
        G.prp(B_GL, B_PL, fast);
        G.t().getInt(G.gi());
        fill_residue(G.gi(), out);
        out->esize = es;
        out->iters = es;
        out->gl_ok = G.GL(B_GL) ? 1 : 0;
        std::vector<std::vector<int32_t>> mus;
        out->pkey = G.PL(depth, fast, mus);
        if (mu_hash64) {
            gfe::gint g(size_t(1) << n, b);
            for (size_t k = 0; k < mus.size(); ++k) {
                std::copy(mus[k].begin(), mus[k].end(), g.data());
                g.reset();
                mu_hash64[k] = g.gethash64();
            }
        }
        out->seconds = now_s() - t0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EHIP, ex.what());
    }
}

static bool write_gint_record(gfe::file &f, const std::vector<int32_t> &canon, uint32_t b) {
    return gfe::genefer::write_record(f, canon.data(), canon.size(), b);
}

int gfe_write_proof(uint32_t b, uint32_t n, int device, int depth, const char *main_path, int flags,
                    uint64_t *mu_hash64, gfe_result *out) {
    if (!out || !main_path || depth < 1 || depth > 12 || (flags & ~3)) return gfe::set_error(GFE_EINVAL, "proof arguments");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    const bool write_ckpts = (flags & GFE_PROOF_CKPT_FILES) != 0;
    try {
        const double t0 = now_s();
        const bool fast = checkpoints_in_hbm(n, depth) && !(flags & GFE_PROOF_HOST_CKPTS);
        gfe::genefer G(b, n, device, fast ? 3 + (size_t(1) << depth) : 3);
        const size_t es = G.esize();
        const int B_GL = gfe::genefer::B_GerbiczLi(es), B_PL = gfe::genefer::B_PietrzakLi(es, depth);
        G.prp(B_GL, B_PL, fast);
        G.t().getInt(G.gi());
        fill_residue(G.gi(), out);
        out->esize = es;
        out->iters = es;
        out->gl_ok = G.GL(B_GL) ? 1 : 0;
        const std::string main(main_path);
        if (write_ckpts) {  // <main>_<i>.ckpt: one residue record and the CRC (src/genefer.h:140-144, 536-546)
            for (size_t i = 0; i <= (es - 1) / (size_t)B_PL; ++i) {
                gfe::file f(main + "_" + std::to_string(i) + ".ckpt", "wb");
                if (!f.exists() || !write_gint_record(f, G.ckpt_digits(i, fast), b))
                    return gfe::set_error(GFE_EINVAL, "cannot write checkpoint file");
                f.write_crc32();
            }
        }
        std::vector<std::vector<int32_t>> mus;
        out->pkey = G.PL(depth, fast, mus);
        // <main>.proof: int version = 1, int depth, mu_0 .. mu_depth, CRC (src/genefer.h:629-717)
        gfe::file pf(main + ".proof", "wb");
        const int version = 1;
        if (!pf.exists() || !pf.write(reinterpret_cast<const char *>(&version), sizeof(version)) ||
            !pf.write(reinterpret_cast<const char *>(&depth), sizeof(depth)))
            return gfe::set_error(GFE_EINVAL, "cannot write proof file");
        gfe::gint g(size_t(1) << n, b);
        for (size_t k = 0; k < mus.size(); ++k) {
            if (!write_gint_record(pf, mus[k], b)) return gfe::set_error(GFE_EINVAL, "cannot write proof file");
            if (mu_hash64) {
                std::copy(mus[k].begin(), mus[k].end(), g.data());
                g.reset();
                mu_hash64[k] = g.gethash64();
            }
        }
        pf.write_crc32();
        out->seconds = now_s() - t0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EHIP, ex.what());
    }
}

int gfe_run_server(uint32_t b, uint32_t n, int device, const char *main_path, uint32_t rnd1, uint32_t rnd2,
                   uint32_t rnd3, uint64_t *ckey, gfe_result *out) {
    if (!out || !ckey || !main_path || !*main_path) return gfe::set_error(GFE_EINVAL, "server arguments");
    if (rnd1 < 2 || rnd1 >= 64 || rnd2 < 16 || rnd2 >= 256 || rnd3 < 4 || rnd3 >= 65536)
        return gfe::set_error(GFE_EINVAL, "server random values out of range (rnd1 in [2, 64), rnd2 in [16, 256), rnd3 in [4, 65536))");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    try {
        const double t0 = now_s();
        gfe::genefer G(b, n, device, 4);
        const std::string main(main_path);
        bool one = false;
        uint64_t r = 0, o = 0, pk = 0, ck = 0;
        G.server(main + ".proof", main + ".cert", rnd1, rnd2, rnd3, one, r, o, pk, ck);
        out->res64 = r;
        out->old64 = o;
        out->is_prp = one ? 1 : 0;
        out->pkey = pk;
        out->esize = G.esize();
        out->gl_ok = -1;
        *ckey = ck;
        out->seconds = now_s() - t0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EINVAL, ex.what());
    }
}

int gfe_run_check(uint32_t b, uint32_t n, int device, const char *main_path, uint64_t *ckey, gfe_result *out) {
    if (!out || !ckey || !main_path || !*main_path) return gfe::set_error(GFE_EINVAL, "check arguments");
    if (check_candidate(b, n) != GFE_OK) return GFE_EINVAL;
    memset(out, 0, sizeof(*out));
    try {
        const double t0 = now_s();
        gfe::genefer G(b, n, device, 4);
        uint64_t ck = 0;
        const bool ok = G.check(std::string(main_path) + ".cert", ck);
        out->gl_ok = ok ? 1 : 0;
        out->esize = G.esize();
        *ckey = ok ? ck : 0;
        out->seconds = now_s() - t0;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EINVAL, ex.what());
    }
}

int gfe_residue(const int32_t *digits, uint32_t n, uint32_t b, int32_t *canon, gfe_result *out) {
    if (!digits || !out || n > 24) return gfe::set_error(GFE_EINVAL, "residue arguments");
    memset(out, 0, sizeof(*out));
    try {
        gfe::gint g(size_t(1) << n, b);
        memcpy(g.data(), digits, (size_t(1) << n) * sizeof(int32_t));
        fill_residue(g, out);
        if (canon) memcpy(canon, g.data(), (size_t(1) << n) * sizeof(int32_t));
        out->gl_ok = -1;
        return GFE_OK;
    } catch (const std::exception &ex) {
        return gfe::set_error(GFE_EINVAL, ex.what());
    }
}

}  // extern "C"
