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
