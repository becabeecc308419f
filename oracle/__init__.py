"""The oracle: CPU restatement of genefer's exact squaring engine and of its driver.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg (the reference CPU timings and the RES64 check beside them) may import this package.  The product (genefer22_amd/) never imports, loads or calls it.

Layers:
  * liboracle.so (gfe_oracle.c): the transform engine, CRT, carry and gint functions, each
    citing the reference file:line it restates;
  * this module: ctypes bindings, the exponent E = b^(2^n), the PRP loop with the Gerbicz-Li
    d(t) cadence (src/genefer.h:479-552), the Gerbicz-Li check (557-615), the Pietrzak-Li proof
    (620-734) with power/powerz (441-465), and closed forms used to pin the engine;
  * RefEngine: the reference's own CPU transforms compiled from the reference sources
    (oracle/_ref/libgfref.so, built by `make -C oracle ref`), when present.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

I32P = ctypes.POINTER(ctypes.c_int32)
U32P = ctypes.POINTER(ctypes.c_uint32)
U64P = ctypes.POINTER(ctypes.c_uint64)


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and, when the reference checkout is present, _ref/libgfref.so and the
    reference's own program _ref/genefer)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True, stdout=out)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-C", _HERE, "-j8", "ref", "cli"], check=True, stdout=out)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.gfo_prime.restype = ctypes.c_uint32
        L.gfo_generator.restype = ctypes.c_uint32
        L.gfo_norm.restype = ctypes.c_uint32
        L.gfo_create.restype = ctypes.c_void_p
        L.gfo_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for name in ("gfo_destroy", "gfo_rns", "gfo_is32"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.gfo_set.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.gfo_square_dup.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.gfo_square_mul.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.gfo_init_multiplicand.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.gfo_mul.argtypes = [ctypes.c_void_p]
        L.gfo_copy.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.gfo_get_digits.argtypes = [ctypes.c_void_p, ctypes.c_int, I32P]
        L.gfo_set_digits.argtypes = [ctypes.c_void_p, ctypes.c_int, I32P]
        L.gfo_schoolbook_mul.argtypes = [I32P, I32P, I32P, ctypes.c_uint32, ctypes.c_uint32]
        L.gfo_unbalance.argtypes = [I32P, ctypes.c_uint32, ctypes.c_uint32]
        L.gfo_balance.argtypes = [I32P, ctypes.c_uint32, ctypes.c_uint32]
        L.gfo_is_one.argtypes = [I32P, ctypes.c_uint32, ctypes.c_uint32, U64P, U64P]
        L.gfo_hash64.restype = ctypes.c_uint64
        L.gfo_hash64.argtypes = [I32P, ctypes.c_uint32]
        L.gfo_hash32.restype = ctypes.c_uint32
        L.gfo_hash32.argtypes = [I32P, ctypes.c_uint32]
        L.gfo_twiddles.argtypes = [ctypes.c_int, ctypes.c_uint32, U32P]
        L.gfo_prime_constants.argtypes = [ctypes.c_int, U32P]
        L.gfo_garner_constants.argtypes = [ctypes.c_int, U32P, U64P]
        L.gfo_select.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _LIB = L
    return _LIB


def _i32(a: np.ndarray):
    return a.ctypes.data_as(I32P)


# ----------------------------------------------------------------------------- constants

PRIMES_S = (127 * 2**24 + 1, 63 * 2**25 + 1, 15 * 2**27 + 1)
PRIMES_U = (125 * 2**25 + 1, 243 * 2**24 + 1, 235 * 2**24 + 1)


def prime_constants(idx: int) -> dict:
    out = (ctypes.c_uint32 * 9)()
    lib().gfo_prime_constants(idx, out)
    keys = ("P", "Q", "R", "RSQ", "H", "IM", "MFIM", "SQRTI", "ISQRTI")
    return dict(zip(keys, list(out)))


def garner_constants(is32: bool) -> dict:
    o32 = (ctypes.c_uint32 * 5)()
    o64 = (ctypes.c_uint64 * 2)()
    lib().gfo_garner_constants(int(is32), o32, o64)
    return {"INVP2_P1": o32[0], "INVP3_P1": o32[1], "INVP3_P2": o32[2], "P1P2P3L": o32[3],
            "P1P2P3_2L": o32[4], "P1P2P3H": o64[0], "P1P2P3_2H": o64[1]}


def select(b: int, n: int) -> tuple[int, bool]:
    """Prime-set selection of create_ocl (src/transform_ocl.cpp:35-50)."""
    rns, is32 = ctypes.c_int(), ctypes.c_int()
    lib().gfo_select(b, n, ctypes.byref(rns), ctypes.byref(is32))
    return rns.value, bool(is32.value)


def twiddles(idx: int, n: int) -> np.ndarray:
    w = np.zeros(2 ** (n - 1), dtype=np.uint32)
    lib().gfo_twiddles(idx, n, w.ctypes.data_as(U32P))
    return w


def norm(idx: int, n: int) -> int:
    return lib().gfo_norm(idx, n)


# ----------------------------------------------------------------------------- gint

def unbalance(d: np.ndarray, n: int, b: int) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.int32).copy()
    lib().gfo_unbalance(_i32(d), n, b)
    return d


def balance(d: np.ndarray, n: int, b: int) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.int32).copy()
    lib().gfo_balance(_i32(d), n, b)
    return d


@dataclass
class Residue:
    is_one: bool
    res64: int
    old64: int
    hash64: int


def residue(d: np.ndarray, n: int, b: int) -> Residue:
    """RES64 / old64 / isOne / hash64 of a digit vector (src/gint.h:169-219)."""
    c = unbalance(d, n, b)
    r, o = ctypes.c_uint64(), ctypes.c_uint64()
    one = lib().gfo_is_one(_i32(c), n, b, ctypes.byref(r), ctypes.byref(o))
    h = 0 if one else lib().gfo_hash64(_i32(c), n)
    return Residue(bool(one), r.value, o.value, h)


def hash64(d: np.ndarray, n: int, b: int) -> int:
    c = unbalance(d, n, b)
    return lib().gfo_hash64(_i32(c), n)


def hash32(d: np.ndarray, n: int, b: int) -> int:
    c = unbalance(d, n, b)
    return lib().gfo_hash32(_i32(c), n)


def digits_to_int(d: np.ndarray, b: int) -> int:
    """Value of sum d_i b^i (any representation, any sign)."""
    v = 0
    for x in reversed([int(t) for t in d]):
        v = v * b + x
    return v


def int_to_digits(v: int, n: int, b: int) -> np.ndarray:
    """Canonical digits of v mod b^N + 1 (v = b^N, i.e. -1, is digit 0 = -1 as gint keeps it)."""
    N = 1 << n
    F = b ** N + 1
    v %= F
    if v == F - 1:
        d = np.zeros(N, dtype=np.int32)
        d[0] = -1
        return d
    out = np.zeros(N, dtype=np.int32)
    for i in range(N):
        v, out[i] = divmod(v, b)
    return out


# ----------------------------------------------------------------------------- engine

class Engine:
    """The restated transform engine (the reference's transform interface, src/transform.h:54-73)."""

    def __init__(self, b: int, n: int, num_regs: int = 3, rns: int = 0, is32: bool = False):
        self.b, self.n, self.N = b, n, 1 << n
        self._h = lib().gfo_create(b, n, num_regs, rns, int(is32))
        if not self._h:
            raise ValueError("gfo_create failed")
        self.num_regs = num_regs

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gfo_destroy(self._h)
            self._h = None

    @property
    def rns(self) -> int:
        return lib().gfo_rns(self._h)

    @property
    def is32(self) -> bool:
        return bool(lib().gfo_is32(self._h))

    def set(self, a: int): lib().gfo_set(self._h, a)
    def square_dup(self, dup: bool): lib().gfo_square_dup(self._h, int(dup))
    def square_mul(self, a: int): lib().gfo_square_mul(self._h, a)
    def init_multiplicand(self, src: int): lib().gfo_init_multiplicand(self._h, src)
    def mul(self, src: int | None = None):
        if src is not None:
            self.init_multiplicand(src)
        lib().gfo_mul(self._h)
    def copy(self, dst: int, src: int): lib().gfo_copy(self._h, dst, src)

    def get_digits(self, reg: int = 0) -> np.ndarray:
        d = np.zeros(self.N, dtype=np.int32)
        lib().gfo_get_digits(self._h, reg, _i32(d))
        return d

    def set_digits(self, d: np.ndarray, reg: int = 0):
        d = np.ascontiguousarray(d, dtype=np.int32)
        lib().gfo_set_digits(self._h, reg, _i32(d))

    def residue(self, reg: int = 0) -> Residue:
        return residue(self.get_digits(reg), self.n, self.b)


def schoolbook_mul(a: np.ndarray, c: np.ndarray, n: int, b: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    c = np.ascontiguousarray(c, dtype=np.int32)
    out = np.zeros(1 << n, dtype=np.int32)
    lib().gfo_schoolbook_mul(_i32(a), _i32(c), _i32(out), n, b)
    return out


# ----------------------------------------------------------------------------- driver

def exponent(b: int, n: int) -> int:
    """E = b^(2^n) (src/genefer.h:1476, mpz_ui_pow_ui)."""
    return b ** (1 << n)


def b_gerbicz_li(esize: int) -> int:
    """src/genefer.h:467-471."""
    L = 2 << ((esize.bit_length() - 1) // 2) if esize > 0 else 2
    return (esize - 1) // L + 1


def b_pietrzak_li(esize: int, depth: int) -> int:
    """src/genefer.h:473-476."""
    return ((esize - 1) >> depth) + 1


def closed_form_residue(b: int, n: int, iters: int) -> int:
    """Register 0 after `iters` iterations of the PRP loop: 2^(E >> (esize - iters)) mod b^N+1."""
    E = exponent(b, n)
    esize = E.bit_length()
    k = min(iters, esize)
    return pow(2, E >> (esize - k), b ** (1 << n) + 1)


def prp(eng, E: int, iters: int | None = None, B_GL: int | None = None, B_PL: int = 0,
        checkpoints: dict | None = None, resume_at: int | None = None):
    """The Fermat PRP loop, src/genefer.h:479-552, on any engine with the transform interface.

    Registers: 0 = 2^(partial exponent), 1 = d(t) (Gerbicz-Li product), 2 = scratch.
    Runs the first `iters` exponent bits (all when None).  With B_PL, register 0 is recorded into
    `checkpoints[i // B_PL]` at every i % B_PL == 0 (the proof checkpoints, 536-546).  With
    `resume_at`, registers 0 and 1 already hold a resumed context (484-503) and the loop runs from
    index resume_at for `iters` bits (to the end when None).
    Returns the index i of the last processed bit."""
    esize = E.bit_length()
    if B_GL is None:
        B_GL = b_gerbicz_li(esize)
    i_start = esize - 1
    if resume_at is None:
        eng.set(1)
        eng.copy(1, 0)
    else:
        i_start = resume_at
    i_end = 0 if iters is None or iters > i_start else i_start + 1 - iters
    for i in range(i_start, i_end - 1, -1):
        eng.square_dup(((E >> i) & 1) != 0)
        if B_GL > 0 and i % B_GL == 0 and i // B_GL != 0:
            eng.copy(2, 0)
            eng.mul(1)
            eng.copy(1, 0)
            eng.copy(0, 2)
        if B_PL and checkpoints is not None and i % B_PL == 0:
            checkpoints[i // B_PL] = eng.get_digits(0)
    return i_end


def gerbicz_li(eng, E: int, B_GL: int) -> bool:
    """The Gerbicz-Li check, src/genefer.h:557-615.  In: r0 = 2^E, r1 = d(t)."""
    n, b = eng.n, eng.b
    eng.mul(1)
    eng.copy(2, 0)
    eng.copy(0, 1)
    for _ in range(B_GL):
        eng.square_dup(False)
    eng.copy(1, 0)
    res, e = 0, E
    while e:
        res += e & ((1 << B_GL) - 1)
        e >>= B_GL
    eng.set(1)
    for i in range(res.bit_length() - 1, -1, -1):
        eng.square_dup(((res >> i) & 1) != 0)
    eng.mul(1)
    h1 = hash64(eng.get_digits(0), n, b)
    eng.copy(0, 2)
    h2 = hash64(eng.get_digits(0), n, b)
    return h1 == h2


def quick(eng, E: int):
    """quick(), src/genefer.h:736-748: PRP then GL.  Returns (Residue, gl_ok)."""
    B_GL = b_gerbicz_li(E.bit_length())
    prp(eng, E, B_GL=B_GL)
    r = eng.residue(0)
    return r, gerbicz_li(eng, E, B_GL)


def power(eng, reg: int, e: int):
    """power / powerz, src/genefer.h:441-465: r0 = r_reg^e."""
    eng.init_multiplicand(reg)
    eng.set(1)
    if e == 0:
        return
    for i in range(e.bit_length() - 1, -1, -1):
        eng.square_dup(False)
        if (e >> i) & 1:
            eng.mul()


def pietrzak_li(eng, ckpts: dict, depth: int):
    """Pietrzak-Li proof generation, src/genefer.h:620-734, with checkpoints held in memory
    (the reference's fast_checkpoints registers 3 + i).  Returns (list of mu digit vectors,
    pkey = hash64(v1)).  Needs num_regs >= 3."""
    n, b = eng.n, eng.b
    mus = []
    gi = ckpts[0]
    mus.append(unbalance(gi, n, b))
    q = hash32(gi, n, b)
    eng.set_digits(balance(gi, n, b), 0)
    power(eng, 0, q)
    eng.copy(2, 0)
    L = 1 << depth
    w = [0] * (L // 2)
    w[0] = q
    for k in range(1, depth + 1):
        i = 1 << (depth - k)
        eng.set_digits(balance(ckpts[i], n, b), 0)
        power(eng, 0, w[0])
        eng.copy(1, 0)
        for j in range(i, L // 2, i):
            eng.set_digits(balance(ckpts[i + 2 * j], n, b), 0)
            power(eng, 0, w[j])
            eng.mul(1)
            eng.copy(1, 0)
        gi = eng.get_digits(0)
        mus.append(unbalance(gi, n, b))
        q = hash32(gi, n, b)
        power(eng, 0, q)
        eng.mul(2)
        eng.copy(2, 0)
        if i > 1:
            for j in range(0, L // 2, i):
                w[i // 2 + j] = w[j] * q
    eng.copy(0, 2)
    pkey = hash64(eng.get_digits(0), n, b)
    return mus, pkey


# ----------------------------------------------------------------------------- reference build

class RefEngine:
    """The reference's own CPU transform (src/transform.h create_cpu), compiled from the reference
    sources by `make -C oracle ref` into oracle/_ref/libgfref.so."""

    PATH = os.path.join(_HERE, "_ref", "libgfref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    @classmethod
    def gint(cls, digits: np.ndarray, n: int, b: int):
        """The reference's gint on a digit vector: (canonical digits, res64, old64, hash64, hash32, is_one)."""
        d = np.ascontiguousarray(digits, dtype=np.int32)
        c = np.zeros(1 << n, dtype=np.int32)
        r, o, h = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        h32 = ctypes.c_uint32()
        one = cls._lib().gfref_gint(_i32(d), n, b, _i32(c), ctypes.byref(r), ctypes.byref(o), ctypes.byref(h),
                                    ctypes.byref(h32))
        return c, r.value, o.value, h.value, h32.value, bool(one)

    @classmethod
    def _lib(cls):
        global _REF
        if _REF is None:
            L = ctypes.CDLL(cls.PATH)
            L.gfref_create.restype = ctypes.c_void_p
            L.gfref_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p,
                                       ctypes.c_uint32, ctypes.c_int]
            L.gfref_destroy.argtypes = [ctypes.c_void_p]
            L.gfref_type.restype = ctypes.c_char_p
            L.gfref_type.argtypes = [ctypes.c_void_p]
            L.gfref_set.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
            L.gfref_square_dup.argtypes = [ctypes.c_void_p, ctypes.c_int]
            L.gfref_square_mul.argtypes = [ctypes.c_void_p, ctypes.c_int32]
            L.gfref_init_multiplicand.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
            L.gfref_mul.argtypes = [ctypes.c_void_p]
            L.gfref_copy.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
            L.gfref_get_error.restype = ctypes.c_double
            L.gfref_get_error.argtypes = [ctypes.c_void_p]
            L.gfref_get_digits.argtypes = [ctypes.c_void_p, I32P, ctypes.c_int]
            L.gfref_set_digits.argtypes = [ctypes.c_void_p, I32P]
            L.gfref_gint.restype = ctypes.c_int
            L.gfref_gint.argtypes = [I32P, ctypes.c_uint32, ctypes.c_uint32, I32P, U64P, U64P, U64P, U32P]
            L.gfref_prp_trace.restype = ctypes.c_long
            L.gfref_prp_trace.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_long, U64P, U64P]
            L.gfref_prp.restype = ctypes.c_long
            L.gfref_prp.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_int, U64P, U64P,
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
            _REF = L
        return _REF

    def __init__(self, b: int, n: int, threads: int = 1, impl: str = "", num_regs: int = 3, check_error: bool = False):
        L = self._lib()
        self.b, self.n, self.N = b, n, 1 << n
        self._h = L.gfref_create(b, n, threads, impl.encode(), num_regs, int(check_error))
        if not self._h:
            raise RuntimeError("reference create_cpu failed")
        self.impl = L.gfref_type(self._h).decode()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib().gfref_destroy(self._h)
            self._h = None

    def set(self, a): self._lib().gfref_set(self._h, a)
    def square_dup(self, dup): self._lib().gfref_square_dup(self._h, int(dup))
    def square_mul(self, a): self._lib().gfref_square_mul(self._h, a)
    def init_multiplicand(self, src): self._lib().gfref_init_multiplicand(self._h, src)
    def mul(self, src=None):
        if src is not None:
            self.init_multiplicand(src)
        self._lib().gfref_mul(self._h)
    def copy(self, dst, src): self._lib().gfref_copy(self._h, dst, src)
    def error(self) -> float: return self._lib().gfref_get_error(self._h)

    def get_digits(self, reg: int = 0, canonical: bool = False) -> np.ndarray:
        if reg != 0:
            raise ValueError("the reference exposes register 0 only")
        d = np.zeros(self.N, dtype=np.int32)
        self._lib().gfref_get_digits(self._h, _i32(d), int(canonical))
        return d

    def set_digits(self, d: np.ndarray, reg: int = 0):
        if reg != 0:
            raise ValueError("the reference exposes register 0 only")
        d = np.ascontiguousarray(d, dtype=np.int32).copy()
        self._lib().gfref_set_digits(self._h, _i32(d))

    def residue(self, reg: int = 0) -> Residue:
        return residue(self.get_digits(reg), self.n, self.b)

    def prp_trace(self, iters: int, every: int):
        """RES64 / old64 after every `every`-th squaring of the reference PRP loop, up to iters."""
        cnt = iters // every
        r = (ctypes.c_uint64 * max(cnt, 1))()
        o = (ctypes.c_uint64 * max(cnt, 1))()
        got = self._lib().gfref_prp_trace(self._h, iters, every, r, o)
        return [(r[i], o[i]) for i in range(got)]

    def prp(self, iters: int = 0, b_gl: int = 0, do_gl: bool = False):
        """Reference PRP loop in native code (ref_shim.cpp gfref_prp).  Returns
        (esize, Residue-without-hash, gl_ok or None)."""
        r, o = ctypes.c_uint64(), ctypes.c_uint64()
        one, gl = ctypes.c_int(), ctypes.c_int()
        es = self._lib().gfref_prp(self._h, iters, b_gl, int(do_gl), ctypes.byref(r), ctypes.byref(o),
                                   ctypes.byref(one), ctypes.byref(gl))
        return es, Residue(bool(one.value), r.value, o.value, 0), (None if gl.value < 0 else bool(gl.value))
