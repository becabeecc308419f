// genefer.h -- the Fermat PRP driver of the squaring engine: the caller side of the transform
// interface, restating the reference driver's sequencing (src/genefer.h:441-766):
//   prp   : r0 = 2^E, E = b^(2^n), one squareDup(bit) per exponent bit, with the Gerbicz-Li
//           product d(t) updated every B_GL iterations and Pietrzak-Li checkpoints every B_PL;
//   GL    : the Gerbicz-Li check d(t)^(2^B) * 2^res == d(t+1);
//   PL    : the Pietrzak-Li proof (mu_k residues and the proof key) from the checkpoints;
//   power / powerz : r0 = r_reg^e by square-and-multiply.
// GMP provides E and its bits as in the reference.
#pragma once

#include <gmp.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "exponent.h"
#include "transform_hip.h"

namespace gfe {

inline int ilog2_32(uint32_t x) { return 31 - __builtin_clz(x); }

// res = sum of the B-bit chunks of E (the exponent of the Gerbicz-Li check, src/genefer.h:580-588),
// in one linear pass over E's words (the reference's repeated mpz_div_2exp is quadratic in |E|).
inline void gl_chunk_sum(mpz_t res, mpz_srcptr E, unsigned long B) {
    const size_t nbits = mpz_sizeinbase(E, 2);
    std::vector<uint64_t> w((nbits + 63) / 64 + 1, 0);
    size_t cnt = 0;
    mpz_export(w.data(), &cnt, -1, 8, 0, 0, E);
    mpz_set_ui(res, 0);
    mpz_t t;
    mpz_init(t);
    for (size_t s0 = 0; s0 < nbits; s0 += B) {
        const size_t wb = s0 / 64, we = std::min(w.size(), (s0 + B + 63) / 64 + 1);
        mpz_import(t, we - wb, -1, 8, 0, 0, w.data() + wb);
        mpz_tdiv_q_2exp(t, t, (mp_bitcnt_t)(s0 % 64));
        mpz_tdiv_r_2exp(t, t, (mp_bitcnt_t)B);
        mpz_add(res, res, t);
    }
    mpz_clear(t);
}

class genefer {
public:
    genefer(uint32_t b, uint32_t n, int device, size_t num_regs)
        : _t(new transformHIP(b, n, device, num_regs)), _g(size_t(1) << n, b) {
        mpz_init(_E);
        mpz_ui_pow_ui(_E, b, 1ul << n);
    }
    // On an engine of (b, n) with num_regs registers that the caller keeps (a server's warm engine,
    // gfe_rebase).
    genefer(gfe_engine *engine, uint32_t b, uint32_t n, size_t num_regs)
        : _t(new transformHIP(engine, b, n, num_regs)), _g(size_t(1) << n, b) {
        mpz_init(_E);
        mpz_ui_pow_ui(_E, b, 1ul << n);
    }
    // As above for a run of at most the PRP loop's first `top_bits` squarings: E holds its top bits
    // (exponent_top) in place, zeros below them, which such a run never reads; esize() is exact.
    genefer(gfe_engine *engine, uint32_t b, uint32_t n, size_t num_regs, uint64_t top_bits)
        : _t(new transformHIP(engine, b, n, num_regs)), _g(size_t(1) << n, b) {
        mpz_init(_E);
        uint64_t S = 0, es = 0;
        exponent_top(b, n, top_bits, _E, &S, &es);
        mpz_mul_2exp(_E, _E, S);
    }
    ~genefer() { mpz_clear(_E); }

    transformHIP &t() { return *_t; }
    gint &gi() { return _g; }
    const mpz_t &exponent() const { return _E; }
    size_t esize() const { return mpz_sizeinbase(_E, 2); }

    // src/genefer.h:467-476
    static int B_GerbiczLi(size_t esize) {
        const size_t L = size_t(2) << (ilog2_32((uint32_t)esize) / 2);
        return (int)((esize - 1) / L) + 1;
    }
    static int B_PietrzakLi(size_t esize, int depth) { return (int)((esize - 1) >> depth) + 1; }

    // src/genefer.h:441-465
    void power(size_t reg, uint32_t e) {
        transformHIP &T = *_t;
        T.initMultiplicand(reg);
        T.set(1);
        if (e == 0) return;
        for (int i = ilog2_32(e); i >= 0; --i) {
            T.squareDup(false);
            if ((e >> i) & 1u) T.mul();
        }
    }
    void powerz(size_t reg, mpz_srcptr e) {
        transformHIP &T = *_t;
        T.initMultiplicand(reg);
        T.set(1);
        if (mpz_sgn(e) == 0) return;
        for (int i = (int)mpz_sizeinbase(e, 2) - 1; i >= 0; --i) {
            T.squareDup(false);
            if (mpz_tstbit(e, (mp_bitcnt_t)i)) T.mul();
        }
    }

    // The PRP loop, src/genefer.h:479-552.  Runs exponent bits esize-1 .. i_end (i_end = 0: all).
    // Pietrzak-Li checkpoints go to registers 3 + i/B_PL when ckpt_in_regs (the reference's
    // fast_checkpoints), else to host digit vectors.
    // Called after each squaring k = 1, 2, ... of prp() when set.
    std::function<void(long)> on_square;

    void prp(int B_GL, int B_PL, bool ckpt_in_regs, long i_end = 0) {
        transformHIP &T = *_t;
        T.set(1);
        T.copy(1, 0);
        const long i0 = (long)esize() - 1;
        _ckpts.clear();
        for (long i = i0; i >= i_end; --i) {
            T.squareDup(mpz_tstbit(_E, (mp_bitcnt_t)i) != 0);
            if (B_GL > 0 && (i % B_GL == 0) && (i / B_GL != 0)) {
                T.copy(2, 0);
                T.mul(1);
                T.copy(1, 0);
                T.copy(0, 2);
            }
            if (B_PL != 0 && (i % B_PL == 0)) {
                if (ckpt_in_regs) T.copy(3 + size_t(i / B_PL), 0);
                else {
                    T.getInt(_g);
                    _ckpts[size_t(i / B_PL)].assign(_g.data(), _g.data() + _g.getSize());
                }
            }
            if (on_square) on_square(i0 - i + 1);
        }
    }

    // Context files, src/genefer.h:198-276: int version (-1, the reference's GPU builds), int where
    // (0: the PRP loop), int i (the loop index to run next, -1 once the loop has ended), double
    // elapsed seconds, the transform payload of registers 0..1 (0..2 for where != 0; all of them
    // with fast checkpoints), the CRC trailer.  A save writes <path>.new, then moves <path> to
    // <path>.old and <path>.new to <path>; a read that finds <path> missing or invalid tries
    // <path>.old.
    bool saveContext(const std::string &path, int where, bool fast, int i, double elapsed) const {
        const std::string old_path = path + ".old", new_path = path + ".new";
        {
            file f(new_path, "wb");
            const int version = -1;
            if (!f.write(reinterpret_cast<const char *>(&version), sizeof(version))) return false;
            if (!f.write(reinterpret_cast<const char *>(&where), sizeof(where))) return false;
            if (!f.write(reinterpret_cast<const char *>(&i), sizeof(i))) return false;
            if (!f.write(reinterpret_cast<const char *>(&elapsed), sizeof(elapsed))) return false;
            _t->saveContext(f, fast ? 0 : context_regs(where));
            f.write_crc32();
            if (!f.close()) return false;
        }
        std::remove(old_path.c_str());
        if (file_exists(path) && std::rename(path.c_str(), old_path.c_str()) != 0) return false;
        return std::rename(new_path.c_str(), path.c_str()) == 0;
    }

    // 0 on success, -1 when the file is missing, -2 when it is not a context of this kind
    // (src/genefer.h:198-218).
    int readContextFile(const std::string &path, int where, bool fast, int &i, double &elapsed) {
        file f(path, "rb");
        if (!f.exists()) return -1;
        int version = 0, rwhere = 0;
        if (!f.read(reinterpret_cast<char *>(&version), sizeof(version)) || -version != 1) return -2;
        if (!f.read(reinterpret_cast<char *>(&rwhere), sizeof(rwhere)) || rwhere != where) return -2;
        if (!f.read(reinterpret_cast<char *>(&i), sizeof(i))) return -2;
        if (!f.read(reinterpret_cast<char *>(&elapsed), sizeof(elapsed))) return -2;
        if (!_t->readContext(f, fast ? 0 : context_regs(where))) return -2;
        if (!f.check_crc32()) return -2;
        return 0;
    }

    bool readContext(const std::string &path, int where, bool fast, int &i, double &elapsed) {
        if (readContextFile(path, where, fast, i, elapsed) == 0) return true;
        return readContextFile(path + ".old", where, fast, i, elapsed) == 0;
    }

    // The PRP loop of src/genefer.h:479-552 with its context file at `path`: resumes from a context
    // of this kind when there is one, else starts from r0 = 1.  Saves a context every `save_every`
    // squarings of this call (the reference saves every 600 s of wall time; 0: none) and, once
    // `stop_after` squarings of this call have run (0: no limit), saves one and stops, as the
    // reference does when asked to quit.  Returns true when the loop has run to its end (the
    // context then holds i = -1).  `resumed_at` receives the loop index of the resumed context
    // (-2 for a fresh start), `i_next` the index a later call starts at, `elapsed` the accumulated
    // wall time.
    bool prp_ctx(int B_GL, const std::string &path, long save_every, long stop_after, int &resumed_at, int &i_next,
                 double &elapsed) {
        transformHIP &T = *_t;
        int ri = 0;
        double rt = 0;
        const bool found = readContext(path, 0, false, ri, rt);
        if (!found) {
            ri = 0;
            rt = 0;
            T.set(1);
            T.copy(1, 0);
        }
        resumed_at = found ? ri : -2;
        const auto t0 = std::chrono::steady_clock::now();
        auto clock = [&] { return rt + std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
        auto save = [&](int i) {
            if (!saveContext(path, 0, false, i, clock())) throw std::runtime_error("cannot save context " + path);
        };
        const int i0 = (int)esize() - 1, i_start = found ? ri : i0;
        if (found && ri == -1) {
            elapsed = rt;
            i_next = -1;
            return true;
        }
        long done = 0;
        for (int i = i_start; i >= 0; --i) {
            if (stop_after > 0 && done == stop_after) {
                save(i);
                elapsed = clock();
                i_next = i;
                return false;
            }
            if (save_every > 0 && done > 0 && done % save_every == 0) save(i);
            T.squareDup(mpz_tstbit(_E, (mp_bitcnt_t)i) != 0);
            if (B_GL > 0 && (i % B_GL == 0) && (i / B_GL != 0)) {
                T.copy(2, 0);
                T.mul(1);
                T.copy(1, 0);
                T.copy(0, 2);
            }
            ++done;
        }
        save(-1);
        elapsed = clock();
        i_next = -1;
        return true;
    }

    // The Gerbicz-Li check, src/genefer.h:557-615.  In: r0 = 2^E, r1 = d(t).
    bool GL(int B_GL) {
        transformHIP &T = *_t;
        T.mul(1);
        T.copy(2, 0);
        T.copy(0, 1);
        for (int i = B_GL - 1; i >= 0; --i) T.squareDup(false);
        T.copy(1, 0);
        mpz_t res;
        mpz_init(res);
        gl_chunk_sum(res, _E, (unsigned long)B_GL);
        T.set(1);
        for (int i = (int)mpz_sizeinbase(res, 2) - 1; i >= 0; --i) T.squareDup(mpz_tstbit(res, (mp_bitcnt_t)i) != 0);
        mpz_clear(res);
        T.mul(1);
        T.getInt(_g);
        const uint64_t h1 = _g.gethash64();
        T.copy(0, 2);
        T.getInt(_g);
        const uint64_t h2 = _g.gethash64();
        return h1 == h2;
    }

    // The Pietrzak-Li proof, src/genefer.h:620-734, returning the mu_k residues (canonical
    // digits) and the proof key hash64(v1).  Needs the checkpoints of prp(.., B_PL, ..).
    uint64_t PL(int depth, bool ckpt_in_regs, std::vector<std::vector<int32_t>> &mus) {
        transformHIP &T = *_t;
        mus.clear();
        load_ckpt(0, ckpt_in_regs);  // mu[0] = ckpt[0]
        _g.unbalance();
        mus.emplace_back(_g.data(), _g.data() + _g.getSize());
        const uint32_t q0 = _g.gethash32();
        T.setInt(_g);
        power(0, q0);
        T.copy(2, 0);
        const size_t L = size_t(1) << depth;
        std::vector<__mpz_struct> w(L / 2);
        for (auto &x : w) mpz_init(&x);
        mpz_set_ui(&w[0], q0);
        for (int k = 1; k <= depth; ++k) {
            const size_t i = size_t(1) << (depth - k);
            load_ckpt(i, ckpt_in_regs);
            if (!ckpt_in_regs) T.setInt(_g);
            powerz(0, &w[0]);
            T.copy(1, 0);
            for (size_t j = i; j < L / 2; j += i) {
                load_ckpt(i + 2 * j, ckpt_in_regs);
                if (!ckpt_in_regs) T.setInt(_g);
                powerz(0, &w[j]);
                T.mul(1);
                T.copy(1, 0);
            }
            T.getInt(_g);
            _g.unbalance();
            mus.emplace_back(_g.data(), _g.data() + _g.getSize());
            const uint32_t q = _g.gethash32();
            power(0, q);
            T.mul(2);
            T.copy(2, 0);
            if (i > 1)
                for (size_t j = 0; j < L / 2; j += i) mpz_mul_ui(&w[i / 2 + j], &w[j], q);
        }
        for (auto &x : w) mpz_clear(&x);
        T.copy(0, 2);
        T.getInt(_g);
        return _g.gethash64();
    }

    // Canonical digits of Pietrzak-Li checkpoint i (for the checkpoint files of the reference's
    // proof mode, src/genefer.h:536-546).
    std::vector<int32_t> ckpt_digits(size_t i, bool in_regs) {
        load_ckpt(i, in_regs);
        _g.unbalance();
        return std::vector<int32_t>(_g.data(), _g.data() + _g.getSize());
    }

    // The server's side of the proof protocol (src/genefer.h:771-880): verifies nothing by itself
    // but turns a proof file into the proof key and a certificate that only a correct proof lets a
    // checker close.  From the records mu_0..mu_depth of `proof_path` (q_k = hash32(mu_k)):
    // v1 = prod_k mu_k^{w_k} as the prover formed it (proof key = hash64(v1)), v2 folded as
    // v2 <- v2^{q_k} mu_k, and the exponent p2 = sum_i w_i e_i over the B-bit chunks e_i of E
    // (B = B_PietrzakLi), with the weights w grown as the prover grew them.  The certificate
    // (`cert_path`) holds int version 1, int B, the residue record of v2^rnd2 (v2 first multiplied
    // by 2^rnd1 when p2 > rnd1 2^B, p2 then losing rnd1 2^B), p2 rnd2 + rnd3 in GMP's raw format
    // (outside the CRC, as the reference's file writes it) and the CRC trailer; the check key is
    // hash64(2^rnd3 v1^rnd2).  The reference draws rnd1 in [2, 64), rnd2 in [16, 256) and rnd3 in
    // [4, 65536) from std::rand; here they are arguments.  mu_0 is the PRP residue: is_prp, res64
    // and old64 come from it.  Needs 4 registers; throws on a malformed or corrupted file.
    void server(const std::string &proof_path, const std::string &cert_path, uint32_t rnd1, uint32_t rnd2,
                uint32_t rnd3, bool &is_prp, uint64_t &res64, uint64_t &old64, uint64_t &pkey, uint64_t &ckey) {
        transformHIP &T = *_t;
        file pf(proof_path, "rb");
        if (!pf.exists()) throw std::runtime_error("cannot open " + proof_path);
        int version = 0, depth = 0;
        if (!pf.read(reinterpret_cast<char *>(&version), sizeof(version)) ||
            !pf.read(reinterpret_cast<char *>(&depth), sizeof(depth)) || version != 1 || depth < 1 || depth > 20)
            throw std::runtime_error("bad proof file header " + proof_path);
        const size_t L = size_t(1) << depth;
        const int B = B_PietrzakLi(esize(), depth);
        mpz_vec w(L);
        read_record(pf, proof_path);  // v1 = mu_0^w_0
        is_prp = _g.isOne(res64, old64);
        uint32_t q = _g.gethash32();
        mpz_set_ui(w[0], q);
        T.setInt(_g);
        power(0, q);
        T.copy(1, 0);
        T.set(1);  // v2 = 1
        T.copy(2, 0);
        for (int k = 1; k <= depth; ++k) {
            read_record(pf, proof_path);  // mu_k -> r3
            q = _g.gethash32();
            T.setInt(_g);
            T.copy(3, 0);
            power(0, q);  // v1 = v1 mu_k^q_k
            T.mul(1);
            T.copy(1, 0);
            power(2, q);  // v2 = v2^q_k mu_k
            T.mul(3);
            T.copy(2, 0);
            const size_t i = size_t(1) << (depth - k);
            for (size_t j = 0; j < L; j += 2 * i) mpz_mul_ui(w[i + j], w[j], q);
        }
        if (!pf.check_crc32()) throw std::runtime_error("bad proof file (crc32) " + proof_path);
        T.copy(0, 1);
        T.getInt(_g);
        pkey = _g.gethash64();
        // p2 = sum_i w_i (E >> (i B) mod 2^B)
        mpz_t p2, t, e;
        mpz_init_set_ui(p2, 0);
        mpz_init(t);
        mpz_init_set(e, _E);
        for (size_t i = 0; i < L; ++i) {
            mpz_tdiv_r_2exp(t, e, (mp_bitcnt_t)B);
            mpz_addmul(p2, t, w[i]);
            mpz_tdiv_q_2exp(e, e, (mp_bitcnt_t)B);
        }
        mpz_clear(e);
        mpz_set_ui(t, rnd1);
        mpz_mul_2exp(t, t, (mp_bitcnt_t)B);
        if (mpz_cmp(p2, t) > 0) {
            mpz_sub(p2, p2, t);
            T.set(2);
            power(0, rnd1);
            T.mul(2);
            T.copy(2, 0);
        }
        power(1, rnd2);
        T.copy(1, 0);
        T.set(2);
        power(0, rnd3);
        T.mul(1);
        T.getInt(_g);
        ckey = _g.gethash64();
        power(2, rnd2);
        T.getInt(_g);
        mpz_mul_ui(p2, p2, rnd2);
        mpz_add_ui(p2, p2, rnd3);
        {
            file cf(cert_path, "wb");
            const int cversion = 1;
            bool ok = cf.exists() && cf.write(reinterpret_cast<const char *>(&cversion), sizeof(cversion)) &&
                      cf.write(reinterpret_cast<const char *>(&B), sizeof(B));
            _g.unbalance();
            ok = ok && write_record(cf, _g.data(), _g.getSize(), _g.getBase());
            ok = ok && mpz_out_raw(cf.handle(), p2) != 0;
            if (ok) cf.write_crc32();
            ok = ok && cf.close();
            mpz_clear(p2);
            mpz_clear(t);
            if (!ok) throw std::runtime_error("cannot write " + cert_path);
        }
    }

    // The checker's side (src/genefer.h:882-1062, without its context file): reads the certificate
    // (int version, int B, residue record u, p2 in GMP's raw format, CRC) and computes
    // v2' = u^(2^B) under the Gerbicz test of block L = B_GerbiczLi(B) (d(t) = prod of u(jL),
    // checked as u(0) d(t)^(2^L) == d(t+1)), then 2^p2 under the Gerbicz-Li test of block
    // GL = B_GerbiczLi(bits of p2); the check key is hash64(v2' 2^p2), which equals the server's
    // when the proof was correct.  Returns false when a test fails.  Needs 4 registers.
    bool check(const std::string &cert_path, uint64_t &ckey) {
        transformHIP &T = *_t;
        int B = 0;
        mpz_t p2;
        mpz_init(p2);
        {
            file cf(cert_path, "rb");
            if (!cf.exists()) {
                mpz_clear(p2);
                throw std::runtime_error("cannot open " + cert_path);
            }
            int version = 0;
            bool ok = cf.read(reinterpret_cast<char *>(&version), sizeof(version)) &&
                      cf.read(reinterpret_cast<char *>(&B), sizeof(B)) && version == 1 && B > 0;
            if (ok) read_record(cf, cert_path);
            ok = ok && mpz_inp_raw(p2, cf.handle()) != 0 && cf.check_crc32();
            if (!ok) {
                mpz_clear(p2);
                throw std::runtime_error("bad certificate file " + cert_path);
            }
        }
        const int p2size = (int)mpz_sizeinbase(p2, 2);
        const int L = B_GerbiczLi((size_t)B), GLb = B_GerbiczLi((size_t)p2size);
        // v2' = u^(2^B), d(t) = prod u(jL)
        const int i0 = p2size + B - 1;
        T.setInt(_g);
        T.copy(1, 0);
        for (int i = i0; i >= p2size; --i) {
            const int j = i0 - i;
            if (j % L == 0 && j != 0) {
                T.copy(2, 0);
                T.mul(1);
                T.copy(1, 0);
                T.copy(0, 2);
            }
            T.squareDup(false);
        }
        T.copy(2, 0);
        if (B % L != 0)
            for (int i = L - (B % L); i > 0; --i) T.squareDup(false);  // u((t + 1) L)
        T.mul(1);  // d(t + 1)
        T.copy(3, 0);
        T.copy(0, 1);
        for (int i = L; i > 0; --i) T.squareDup(false);
        T.copy(1, 0);
        T.setInt(_g);  // u(0) d(t)^(2^L) == d(t + 1)
        T.mul(1);
        T.getInt(_g);
        uint64_t h1 = _g.gethash64();
        T.copy(0, 3);
        T.getInt(_g);
        uint64_t h2 = _g.gethash64();
        if (h1 != h2) {
            mpz_clear(p2);
            return false;
        }
        // 2^p2 with the Gerbicz-Li product d(t) every GL bits
        T.set(1);
        T.copy(1, 0);
        for (int i = p2size - 1; i >= 0; --i) {
            T.squareDup(mpz_tstbit(p2, (mp_bitcnt_t)i) != 0);
            if (i % GLb == 0 && i / GLb != 0) {
                T.copy(3, 0);
                T.mul(1);
                T.copy(1, 0);
                T.copy(0, 3);
            }
        }
        T.copy(3, 0);
        T.mul(2);  // v2' 2^p2
        T.getInt(_g);
        ckey = _g.gethash64();
        T.copy(0, 3);
        T.mul(1);  // d(t + 1) = d(t) 2^p2
        T.copy(2, 0);
        T.copy(0, 1);
        for (int i = GLb - 1; i >= 0; --i) T.squareDup(false);
        T.copy(1, 0);
        mpz_t res;
        mpz_init(res);
        gl_chunk_sum(res, p2, (unsigned long)GLb);
        mpz_clear(p2);
        T.set(1);
        for (int i = (int)mpz_sizeinbase(res, 2) - 1; i >= 0; --i) T.squareDup(mpz_tstbit(res, (mp_bitcnt_t)i) != 0);
        mpz_clear(res);
        T.mul(1);
        T.getInt(_g);
        h1 = _g.gethash64();
        T.copy(0, 2);
        T.getInt(_g);
        h2 = _g.gethash64();
        return h1 == h2;
    }

    // A residue record of the reference's files (gint::write, src/gint.h:160-167): uint32 size,
    // uint32 base, size int32 canonical digits.
    static bool write_record(file &f, const int32_t *canon, size_t size, uint32_t base) {
        const uint32_t sz = (uint32_t)size;
        return f.write(reinterpret_cast<const char *>(&sz), sizeof(sz)) &&
               f.write(reinterpret_cast<const char *>(&base), sizeof(base)) &&
               f.write(reinterpret_cast<const char *>(canon), size * sizeof(int32_t));
    }

private:
    std::unique_ptr<transformHIP> _t;

    static size_t context_regs(int where) { return (where == 0) ? 2 : 3; }
    static bool file_exists(const std::string &path) {
        std::FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) return false;
        std::fclose(f);
        return true;
    }

    gint _g;
    mpz_t _E;
    std::map<size_t, std::vector<int32_t>> _ckpts;

    // mpz_t array with its own clean-up
    struct mpz_vec {
        std::vector<__mpz_struct> v;
        explicit mpz_vec(size_t n) : v(n) { for (auto &x : v) mpz_init(&x); }
        ~mpz_vec() { for (auto &x : v) mpz_clear(&x); }
        mpz_ptr operator[](size_t i) { return &v[i]; }
    };

    // One residue record into the gint (gint::read, src/gint.h:151-158: size and base must be the
    // candidate's).
    void read_record(file &f, const std::string &path) {
        uint32_t size = 0, base = 0;
        if (!f.read(reinterpret_cast<char *>(&size), sizeof(size)) ||
            !f.read(reinterpret_cast<char *>(&base), sizeof(base)) || size != _g.getSize() || base != _g.getBase() ||
            !f.read(reinterpret_cast<char *>(_g.data()), _g.getSize() * sizeof(int32_t)))
            throw std::runtime_error("bad residue record in " + path);
        _g.reset();
    }

    // Checkpoint i into r0 (registers) or into the gint (host copies).
    void load_ckpt(size_t i, bool in_regs) {
        if (in_regs) {
            _t->copy(0, 3 + i);
            _t->getInt(_g);
        } else {
            const std::vector<int32_t> &c = _ckpts.at(i);
            std::copy(c.begin(), c.end(), _g.data());
            _g.reset();
        }
    }
};

}  // namespace gfe
