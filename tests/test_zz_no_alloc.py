"""gfe_square_noalloc (squarings that allocate nothing) against an engine's gfe_square_dup on the same
digits: bit-exact digit vectors, two calls in a row on one device (the second with a smaller input,
n = 16 then n = 12), each checked against its own engine run.  Device memory for the caller's digits
comes from the HIP runtime directly (the entry point itself allocates nothing)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipSetDevice.argtypes = [ctypes.c_int]

    def ck(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ck(self.h.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p

    def to_device(self, dst, arr):
        self.ck(self.h.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")

    def to_host(self, arr, src):
        self.ck(self.h.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2), "hipMemcpy D2H")


def start_digits(gfe, b, n, warm):
    """A non-trivial balanced residue: the PRP loop's value after `warm` squarings of 1."""
    with gfe.Transform(b, n) as t:
        t.set(1)
        E = b ** (1 << n)
        es = E.bit_length()
        for i in range(es - 1, es - 1 - warm, -1):
            t.square_dup((E >> i) & 1)
        return t.get_digits(0)


def engine_squarings(gfe, b, n, d0, count, dup_bits, plan=None):
    with gfe.Transform(b, n) as t:
        if plan is not None:
            t.set_plan(plan)
        t.set_digits(d0, 0)
        for i in range(count):
            t.square_dup((dup_bits >> (i % 64)) & 1)
        return t.get_digits(0)


# (b, n, squarings, dup pattern): C2's first seeded b at n = 16 (three primes), then a smaller
# input, C1's first b at n = 12 (two primes)
CASES = [(1433042190, 16, 300, 0x9E3779B97F4A7C15), (146950, 12, 200, 0x00F0F0F00FF00F0F)]


def test_noalloc_equals_engine_two_calls_in_a_row(gfe):
    L = gfe.lib()
    hip = Hip()
    hip.ck(hip.h.hipSetDevice(0), "hipSetDevice")
    N0 = 1 << CASES[0][1]
    buf = hip.malloc(N0 * 4)  # the caller's digits; the larger case's size serves both calls
    try:
        for b, n, count, dup in CASES:
            print(f"noalloc case b={b} n={n} count={count}", flush=True)
            d0 = start_digits(gfe, b, n, 40)
            want = engine_squarings(gfe, b, n, d0, count, dup)
            hip.to_device(buf, np.ascontiguousarray(d0, dtype=np.int32))
            rc = L.gfe_square_noalloc(b, n, 0, buf, count, dup, None)
            assert rc == 0, L.gfe_last_error().decode()
            hip.ck(hip.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
            got = np.empty(1 << n, dtype=np.int32)
            hip.to_host(got, buf)
            assert np.array_equal(got, want), (b, n, int(np.count_nonzero(got != want)))
            r_got, r_want = gfe.residue(got, n, b), gfe.residue(want, n, b)
            assert r_got["res64"] == r_want["res64"]
    finally:
        hip.h.hipFree(buf)


def test_noalloc_where_the_engine_starts_on_limbs(gfe):
    """b = 5e7 at n = 12: an engine starts on the two-digit-limb transform, the no-alloc squarings
    run the digit transform's default plan.  Digits equal the engine's on that plan; the residue
    equals the engine's on its own (limb) plan."""
    L = gfe.lib()
    hip = Hip()
    b, n, count, dup = 50000000, 12, 150, 0x0123456789ABCDEF
    with gfe.Transform(b, n) as t:
        assert t.get_plan()[5] & 32
    d0 = start_digits(gfe, b, n, 30)
    want_digits = engine_squarings(gfe, b, n, d0, count, dup, plan=gfe.default_plan(n))
    want_limb = engine_squarings(gfe, b, n, d0, count, dup)
    buf = hip.malloc(4 << n)
    try:
        hip.to_device(buf, np.ascontiguousarray(d0, dtype=np.int32))
        rc = L.gfe_square_noalloc(b, n, 0, buf, count, dup, None)
        assert rc == 0, L.gfe_last_error().decode()
        hip.ck(hip.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        got = np.empty(1 << n, dtype=np.int32)
        hip.to_host(got, buf)
    finally:
        hip.h.hipFree(buf)
    assert np.array_equal(got, want_digits)
    assert gfe.residue(got, n, b)["res64"] == gfe.residue(want_limb, n, b)["res64"]


def test_noalloc_rejects_what_it_cannot_hold(gfe):
    L = gfe.lib()
    hip = Hip()
    buf = hip.malloc(4 << 17)
    try:
        assert L.gfe_square_noalloc(1000000, 17, 0, buf, 1, 0, None) == -1  # beyond the array's n
        assert b"12, 16" in L.gfe_last_error()
        host = np.zeros(1 << 12, dtype=np.int32)
        assert L.gfe_square_noalloc(146950, 12, 0, host.ctypes.data, 1, 0, None) == -1  # host memory
        assert L.gfe_square_noalloc(146950, 12, 0, None, 1, 0, None) == -1
    finally:
        hip.h.hipFree(buf)
