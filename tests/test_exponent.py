"""gfe_exponent_top (genefer22_amd/csrc/exponent.h): the top bits of E = b^(2^n) that a truncated
PRP run reads, from bounded truncated squarings, against E computed exactly with Python integers --
T = floor(E / 2^shift) with at least k bits and the exact bit length of E, for random bases over
n = 12..18, the largest b, b a power of two times 3 (long runs of equal bits), and k from 1 to
beyond esize (E itself then)."""
import ctypes
import random

import numpy as np
import pytest


def top(G, b, n, k):
    L = G.lib()
    words = np.zeros((1 << (n - 1)) + 8, dtype=np.uint64)  # E has fewer than 32 2^n bits
    nw, shift, esize = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = L.gfe_exponent_top(b, n, k, words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(words),
                            ctypes.byref(nw), ctypes.byref(shift), ctypes.byref(esize))
    assert rc == 0, L.gfe_last_error()
    T = int.from_bytes(words[:nw.value].tobytes(), "little")
    return T, shift.value, esize.value


@pytest.fixture(scope="module")
def G():
    import genefer22_amd
    genefer22_amd.lib()
    return genefer22_amd


CASES = [(1999999990, 12), (1999999990, 16), (3 << 29, 16), (1026, 18), (146950, 14), (4098, 13)]
rs = random.Random(11)
CASES += [(2 * rs.randint(512, 999999999), rs.randint(12, 17)) for _ in range(10)]


@pytest.mark.parametrize("b,n", CASES)
def test_exponent_top_bits(G, b, n):
    E = b ** (1 << n)
    es = E.bit_length()
    for k in (1, 64, 4096, 65536, es - 10, es + 5):
        if k <= 0:
            continue
        T, s, esize = top(G, b, n, k)
        assert esize == es, (b, n, k)
        assert T == E >> s, (b, n, k, s)
        assert T.bit_length() >= min(k, es), (b, n, k)
