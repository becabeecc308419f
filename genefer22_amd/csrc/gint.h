// gint.h -- digit vector of a residue mod b^N + 1 and its canonical form and hashes.
//
// Same contract as the reference's gint (src/gint.h:17-226): unbalance() gives canonical digits
// in [0, b) with the negacyclic wrap (the value -1 = b^N is kept as digit 0 = -1), balance() gives
// digits in (-b/2, b/2], isOne() returns RES64 (canonical value mod 2^64) and old64 (low bytes of
// the 8 top digits), gethash64/gethash32 are the reference's residue hashes.
#pragma once

#include <stdint.h>
#include <stddef.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace gfe {

class gint {
public:
    gint(size_t size, uint32_t base) : _size(size), _base(base), _d(size, 0), _state(State::Unknown) {}

    size_t getSize() const { return _size; }
    uint32_t getBase() const { return _base; }
    int32_t *data() { return _d.data(); }
    const int32_t *data() const { return _d.data(); }
    void reset() { _state = State::Unknown; }

    // Canonical digits in [0, b): full carry, then the wrap f b^N = -f until no carry is left.
    void unbalance() {
        if (_state == State::Unbalanced) return;
        _state = State::Unbalanced;
        const int64_t B = _base;
        int64_t f = carry_pass(0, [B](int64_t v) { int64_t r = v % B; return r < 0 ? r + B : r; }, false);
        while (f != 0) {
            f = carry_pass(-f, [B](int64_t v) { int64_t r = v % B; return r < 0 ? r + B : r; }, true);
            if (f == 1 && all_zero()) {  // -1 cannot be written with digits in [0, b)
                _d[0] = -1;
                break;
            }
        }
    }

    // Balanced digits in (-b/2, b/2].
    void balance() {
        if (_state == State::Balanced) return;
        _state = State::Balanced;
        const int64_t B = _base;
        auto red = [B](int64_t v) {
            int64_t r = v % B;
            if (r > B / 2) r -= B;
            if (r <= -B / 2) r += B;
            return r;
        };
        int64_t f = carry_pass(0, red, false);
        while (f != 0) f = carry_pass(-f, red, true);
    }

    bool isOne(uint64_t &res64, uint64_t &old64) {
        unbalance();
        bool one = (_d[0] == 1);
        for (size_t i = 1; one && i < _size; ++i) one = (_d[i] == 0);
        uint64_t r = 0, bp = 1;
        for (size_t i = 0; i < _size; ++i) {
            r += (uint64_t)(uint32_t)_d[i] * bp;
            bp *= _base;
        }
        res64 = r;
        uint64_t o = 0;
        for (size_t i = 8; i != 0; --i) o = (o << 8) | (uint8_t)_d[_size - i];
        old64 = o;
        return one;
    }

    uint64_t gethash64() {
        unbalance();
        uint64_t h = 0;
        bool zero = true;
        for (size_t i = 0; i < _size; ++i) {
            const uint32_t a = (uint32_t)_d[i];
            h += a;
            const unsigned s = (unsigned)((17 * (uint64_t)a + 5) % 64);
            const uint64_t v = a + 0xc39d8a0552b073e8ull;
            h ^= (v << s) | (v >> ((64 - s) & 63));
            zero &= (a == 0);
        }
        if (zero) throw std::runtime_error("value is zero");
        return h;
    }

    uint32_t gethash32() {
        const uint64_t h = gethash64();
        return std::max<uint32_t>(2u, (uint32_t)h ^ (uint32_t)(h >> 32));
    }

private:
    enum class State { Unknown, Balanced, Unbalanced };
    const size_t _size;
    const uint32_t _base;
    std::vector<int32_t> _d;
    State _state;

    bool all_zero() const {
        for (size_t i = 0; i < _size; ++i)
            if (_d[i] != 0) return false;
        return true;
    }

    // One carry sweep from digit 0 with incoming carry f; `early` stops as soon as the carry is 0.
    template <class Red> int64_t carry_pass(int64_t f, Red red, bool early) {
        const int64_t B = _base;
        for (size_t i = 0; i < _size; ++i) {
            f += _d[i];
            const int64_t r = red(f);
            _d[i] = (int32_t)r;
            f = (f - r) / B;
            if (early && f == 0) break;
        }
        return f;
    }
};

}  // namespace gfe
