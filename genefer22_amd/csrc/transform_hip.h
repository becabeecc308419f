// transform_hip.h -- C++ host side above the C ABI: the reference's transform interface and its
// HIP implementation.
//
// `transform` mirrors the abstract class of src/transform.h:42-256 (same virtuals, same argument
// meaning, same getInt/setInt checks and exceptions).  `transformHIP` implements it by forwarding
// each virtual to one gfe_* entry point of include/gfe.h, so a genefer-style driver (genefer.h in
// this directory) uses it exactly as the reference driver uses transformGPUs.  Errors become
// std::runtime_error, as the reference's oclFatal does (src/ocl.h:166-174).
#pragma once

#include <stdint.h>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gfe.h"
#include "gint.h"

namespace gfe {

// File with a running CRC-32 (reflected polynomial 0xEDB88320), the container of context,
// checkpoint and proof files; write_crc32 / check_crc32 append / verify ~crc ^ 0xa23777ac
// (src/file.h:70-161).
class file {
public:
    file(const std::string &name, const char *mode) : _name(name), _f(std::fopen(name.c_str(), mode)) {}
    ~file() { if (_f) std::fclose(_f); }
    file(const file &) = delete;
    file &operator=(const file &) = delete;
    bool exists() const { return _f != nullptr; }
    // Flushes and closes; false when a write did not reach the file.
    bool close() {
        if (!_f) return false;
        const bool ok = std::fclose(_f) == 0;
        _f = nullptr;
        return ok;
    }
    uint32_t crc32() const { return _crc; }
    // The stream itself, for records written outside the running CRC (GMP's raw integers).
    std::FILE *handle() const { return _f; }
    bool read(char *p, size_t n) {
        if (!_f) return false;
        const size_t r = std::fread(p, 1, n, _f);
        _crc = update(_crc, p, n);
        return r == n;
    }
    bool write(const char *p, size_t n) {
        if (!_f) return false;
        const size_t r = std::fwrite(p, 1, n, _f);
        _crc = update(_crc, p, n);
        return r == n;
    }
    void write_crc32() {
        const uint32_t c = ~_crc ^ 0xa23777acu;
        write(reinterpret_cast<const char *>(&c), sizeof(c));
    }
    bool check_crc32() {
        const uint32_t expect = ~_crc ^ 0xa23777acu;
        uint32_t c = 0;
        if (!read(reinterpret_cast<char *>(&c), sizeof(c))) return false;
        return c == expect;
    }
    static uint32_t update(uint32_t crc, const char *p, size_t n) {
        static const std::vector<uint32_t> table = [] {
            std::vector<uint32_t> t(256);
            for (uint32_t i = 0; i < 256; ++i) {
                uint32_t r = i;
                for (int j = 0; j < 8; ++j) r = (r & 1) ? (r >> 1) ^ 0xedb88320u : (r >> 1);
                t[i] = r;
            }
            return t;
        }();
        uint32_t c = ~crc;
        for (size_t i = 0; i < n; ++i) c = (c >> 8) ^ table[(c ^ (uint8_t)p[i]) & 0xff];
        return ~c;
    }

private:
    std::string _name;
    std::FILE *_f;
    uint32_t _crc = 0;
};

class transform {
public:
    transform(size_t size, uint32_t n, uint32_t b) : _size(size), _n(n), _b(b) {}
    virtual ~transform() {}

    virtual void set(uint32_t a) = 0;                  // r_0 = a
    virtual void squareDup(bool dup) = 0;              // r_0 = r_0^2 or 2 r_0^2
    virtual void squareMul(int32_t a) = 0;             // r_0 = a r_0^2
    virtual void initMultiplicand(size_t src) = 0;     // r_m = transform(r_src)
    virtual void mul() = 0;                            // r_0 *= r_m
    virtual void copy(size_t dst, size_t src) const = 0;
    virtual size_t getMemSize() const = 0;
    virtual size_t getCacheSize() const = 0;
    virtual bool readContext(file &cFile, size_t num_regs) = 0;
    virtual void saveContext(file &cFile, size_t num_regs) const = 0;
    virtual double getError() const { return 0; }
    virtual void info() const {}

    void mul(size_t src) {
        initMultiplicand(src);
        mul();
    }
    void getInt(gint &g) const {
        if (g.getSize() != _size || g.getBase() != _b) throw std::runtime_error("getInt");
        getZi(g.data());
        g.reset();
    }
    void setInt(gint &g) {
        if (g.getSize() != _size || g.getBase() != _b) throw std::runtime_error("setInt");
        g.balance();
        setZi(g.data());
    }
    size_t getSize() const { return _size; }
    uint32_t getB() const { return _b; }
    uint32_t getN() const { return _n; }

protected:
    virtual void getZi(int32_t *zi) const = 0;
    virtual void setZi(const int32_t *zi) = 0;

private:
    const size_t _size;
    const uint32_t _n, _b;
};

class transformHIP : public transform {
public:
    transformHIP(uint32_t b, uint32_t n, int device, size_t num_regs)
        : transform(size_t(1) << n, n, b), _num_regs(num_regs) {
        ck(gfe_create(b, n, (uint32_t)num_regs, device, &_e));
    }
    // On an engine the caller keeps (num_regs registers of n): not destroyed with the transform.
    transformHIP(gfe_engine *borrowed, uint32_t b, uint32_t n, size_t num_regs)
        : transform(size_t(1) << n, n, b), _e(borrowed), _num_regs(num_regs), _own(false) {}
    ~transformHIP() override {
        if (_own) gfe_destroy(_e);
    }
    transformHIP(const transformHIP &) = delete;
    transformHIP &operator=(const transformHIP &) = delete;

    void set(uint32_t a) override { ck(gfe_set(_e, a)); }
    void squareDup(bool dup) override { ck(gfe_square_dup(_e, dup ? 1 : 0)); }
    void squareMul(int32_t a) override { ck(gfe_square_mul(_e, a)); }
    void initMultiplicand(size_t src) override { ck(gfe_init_multiplicand(_e, (uint32_t)src)); }
    void mul() override { ck(gfe_mul(_e)); }
    using transform::mul;
    void copy(size_t dst, size_t src) const override { ck(gfe_copy(_e, (uint32_t)dst, (uint32_t)src)); }
    size_t getMemSize() const override { return gfe_mem_size(_e); }
    size_t getCacheSize() const override { return gfe_cache_size(_e); }
    double getError() const override { return gfe_get_error(_e); }
    void info() const override {
        char buf[256];
        if (gfe_info(_e, buf, sizeof(buf)) == GFE_OK) std::printf("%s\n", buf);
    }
    bool readContext(file &cFile, size_t nregs) override {
        const uint32_t nr = (uint32_t)(nregs ? nregs : _num_regs);
        std::vector<char> buf(gfe_context_size(_e, nr));
        if (!cFile.read(buf.data(), buf.size())) return false;
        return gfe_load(_e, nr, buf.data(), buf.size()) == GFE_OK;
    }
    void saveContext(file &cFile, size_t nregs) const override {
        const uint32_t nr = (uint32_t)(nregs ? nregs : _num_regs);
        std::vector<char> buf(gfe_context_size(_e, nr));
        ck(gfe_save(_e, nr, buf.data(), buf.size()));
        cFile.write(buf.data(), buf.size());
    }
    void getRegister(size_t reg, int32_t *out) const { ck(gfe_get_digits(_e, (uint32_t)reg, out)); }
    void setRegister(size_t reg, const int32_t *in) { ck(gfe_set_digits(_e, (uint32_t)reg, in)); }
    gfe_engine *engine() const { return _e; }

protected:
    void getZi(int32_t *zi) const override { ck(gfe_get_digits(_e, 0, zi)); }
    void setZi(const int32_t *zi) override { ck(gfe_set_digits(_e, 0, zi)); }

private:
    gfe_engine *_e = nullptr;
    size_t _num_regs;
    bool _own = true;
    static void ck(int rc) {
        if (rc != GFE_OK) throw std::runtime_error(std::string("gfe: ") + gfe_last_error());
    }
};

}  // namespace gfe
