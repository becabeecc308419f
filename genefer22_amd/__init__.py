"""genefer22_amd -- gfx950 modular-squaring engine for b^(2^n) + 1 (genefer's transform path).

Python mirror of the reference's transform interface (src/transform.h:42-256) over the C ABI of
include/gfe.h and include/gfe_driver.h (libgfe.so, built by genefer22_amd.build).  Every method
maps to one gfe_* call; errors raise GfeError with the library's message.  There is no CPU fallback:
importing works without a GPU, but creating an engine without libgfe.so or a HIP device raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgfe.so")
_LIB = None

I32P = ctypes.POINTER(ctypes.c_int32)
U32P = ctypes.POINTER(ctypes.c_uint32)
U64P = ctypes.POINTER(ctypes.c_uint64)


class GfeError(RuntimeError):
    pass


class Result(ctypes.Structure):
    _fields_ = [("res64", ctypes.c_uint64), ("old64", ctypes.c_uint64), ("hash64", ctypes.c_uint64),
                ("pkey", ctypes.c_uint64), ("esize", ctypes.c_uint64), ("iters", ctypes.c_uint64),
                ("is_prp", ctypes.c_int32), ("gl_ok", ctypes.c_int32), ("seconds", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


MAX_PROOF_DEPTH = 10  # GFE_BATCH_MAX_DEPTH


def _proof_dict(d: dict) -> dict:
    """Turns the mu_hash64 array of a result into the list of its depth + 1 hashes."""
    d["mu_hash64"] = list(d["mu_hash64"])[:d["depth"] + 1] if d["depth"] else []
    d.pop("pad", None)
    return d


class ServeResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("n", ctypes.c_uint32), ("b", ctypes.c_uint32),
                ("res64", ctypes.c_uint64), ("old64", ctypes.c_uint64), ("hash64", ctypes.c_uint64),
                ("esize", ctypes.c_uint64), ("iters", ctypes.c_uint64), ("is_prp", ctypes.c_int32),
                ("gl_ok", ctypes.c_int32), ("t_arrival", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("depth", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("pkey", ctypes.c_uint64), ("mu_hash64", ctypes.c_uint64 * (MAX_PROOF_DEPTH + 1))]

    def as_dict(self) -> dict:
        return _proof_dict({k: getattr(self, k) for k, _ in self._fields_})


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("admitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("queued", ctypes.c_uint64), ("bytes_budget", ctypes.c_uint64), ("bytes_reserved", ctypes.c_uint64),
                ("bytes_peak", ctypes.c_uint64), ("pools", ctypes.c_uint32), ("pools_created", ctypes.c_uint32),
                ("pools_released", ctypes.c_uint32), ("busy_slots", ctypes.c_uint32), ("pumps", ctypes.c_uint64),
                ("blocks", ctypes.c_uint64), ("solo", ctypes.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every exported entry point; tests check each is present.
ABI = [
    ("gfe_create", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("gfe_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("gfe_set", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gfe_square_dup", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("gfe_square_mul", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("gfe_init_multiplicand", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gfe_mul", ctypes.c_int, [ctypes.c_void_p]),
    ("gfe_copy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    ("gfe_get_digits", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, I32P]),
    ("gfe_set_digits", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, I32P]),
    ("gfe_mem_size", ctypes.c_size_t, [ctypes.c_void_p]),
    ("gfe_cache_size", ctypes.c_size_t, [ctypes.c_void_p]),
    ("gfe_context_size", ctypes.c_size_t, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gfe_save", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t]),
    ("gfe_load", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t]),
    ("gfe_get_error", ctypes.c_double, [ctypes.c_void_p]),
    ("gfe_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    ("gfe_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("gfe_select", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    ("gfe_last_error", ctypes.c_char_p, []),
    ("gfe_time_squarings", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    ("gfe_tune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gfe_run_quick", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(Result)]),
    ("gfe_run_prp", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, ctypes.c_int32, I32P, ctypes.POINTER(Result)]),
    ("gfe_run_prp_trace", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, U64P, U64P, ctypes.POINTER(Result)]),
    ("gfe_default_plan", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, I32P]),
    ("gfe_plan_admissible", ctypes.c_int, [ctypes.c_uint32, I32P]),
    ("gfe_engine_plan", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, I32P]),
    ("gfe_candidate_plans", ctypes.c_int, [ctypes.c_uint32, I32P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    ("gfe_host_twiddles", ctypes.c_int, [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, U32P]),
    ("gfe_rebase", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gfe_exponent_top", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, U64P, ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_uint32), U64P, U64P]),
    ("gfe_square_noalloc", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                          ctypes.c_uint64, ctypes.c_void_p]),
    ("gfe_run_prp_resumable", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64,
                                             ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(Result)]),
    ("gfe_run_proof", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, U64P, ctypes.POINTER(Result)]),
    ("gfe_residue", ctypes.c_int, [I32P, ctypes.c_uint32, ctypes.c_uint32, I32P, ctypes.POINTER(Result)]),
    ("gfe_write_proof", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                       ctypes.c_int, U64P, ctypes.POINTER(Result)]),
    ("gfe_run_server", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, U64P, ctypes.POINTER(Result)]),
    ("gfe_run_check", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_char_p, U64P,
                                     ctypes.POINTER(Result)]),
    ("gfe_batch_create", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("gfe_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("gfe_batch_create_mixed", ctypes.c_int, [ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("gfe_batch_submit_n", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_int32)]),
    ("gfe_batch_create_proof", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_void_p)]),
    ("gfe_batch_proof_slot_bytes", ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32]),
    ("gfe_batch_proof_device_bytes", ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("gfe_batch_slot_bytes", ctypes.c_size_t, [ctypes.c_uint32]),
    ("gfe_batch_tune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gfe_batch_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    ("gfe_batch_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32)]),
    ("gfe_batch_step", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    ("gfe_batch_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    ("gfe_batch_active", ctypes.c_int, [ctypes.c_void_p]),
    ("gfe_batch_get_digits", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, I32P]),
    ("gfe_batch_time_rounds", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    ("gfe_batch_run", ctypes.c_int, [ctypes.c_uint32, U32P, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p]),
    ("gfe_batch_backlog", ctypes.c_int, [ctypes.c_void_p]),
    ("gfe_batch_set_graphs", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("gfe_get_plan", ctypes.c_int, [ctypes.c_void_p, I32P]),
    ("gfe_set_plan", ctypes.c_int, [ctypes.c_void_p, I32P]),
    ("gfe_batch_get_plan", ctypes.c_int, [ctypes.c_void_p, I32P]),
    ("gfe_batch_set_plan", ctypes.c_int, [ctypes.c_void_p, I32P]),
    ("gfe_batch_device_bytes", ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32]),
    ("gfe_server_create", ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    ("gfe_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("gfe_server_clock", ctypes.c_double, [ctypes.c_void_p]),
    ("gfe_server_prepare", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    ("gfe_server_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]),
    ("gfe_server_submit_proof", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                               ctypes.c_uint32, ctypes.c_uint64]),
    ("gfe_server_pump", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    ("gfe_server_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    ("gfe_server_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServerStats)]),
    ("gfe_server_set_solo", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("gfe_serve_replay", ctypes.c_int, [U32P, U32P, U64P, U32P, ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.c_int,
                                        ctypes.c_uint64, U32P, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.POINTER(ServerStats)]),
]


class BatchResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("b", ctypes.c_uint32), ("n", ctypes.c_uint32),
                ("res64", ctypes.c_uint64), ("old64", ctypes.c_uint64), ("hash64", ctypes.c_uint64),
                ("esize", ctypes.c_uint64), ("iters", ctypes.c_uint64), ("is_prp", ctypes.c_int32),
                ("gl_ok", ctypes.c_int32), ("t_submit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("depth", ctypes.c_int32), ("pad", ctypes.c_int32), ("pkey", ctypes.c_uint64),
                ("mu_hash64", ctypes.c_uint64 * (MAX_PROOF_DEPTH + 1))]

    def as_dict(self) -> dict:
        return _proof_dict({k: getattr(self, k) for k, _ in self._fields_})


def lib():
    """Loads libgfe.so; raises GfeError when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise GfeError(f"{LIB_PATH} is missing: build it with `python -m genefer22_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in ABI:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ck(rc: int) -> None:
    if rc != 0:
        raise GfeError(f"gfe error {rc}: {lib().gfe_last_error().decode(errors='replace')}")


def _i32(a: np.ndarray):
    return a.ctypes.data_as(I32P)


def select(b: int, n: int) -> tuple[int, bool]:
    """Prime set of create_ocl (src/transform_ocl.cpp:35-50): (2 or 3 primes, 32-bit primes?)."""
    rns, is32 = ctypes.c_int(), ctypes.c_int()
    _ck(lib().gfe_select(b, n, ctypes.byref(rns), ctypes.byref(is32)))
    return rns.value, bool(is32.value)


class Transform:
    """The reference's transform interface on the HIP engine (one candidate, one device)."""

    def __init__(self, b: int, n: int, num_regs: int = 3, device: int = 0):
        self.b, self.n, self.N, self.num_regs = b, n, 1 << n, num_regs
        h = ctypes.c_void_p()
        _ck(lib().gfe_create(b, n, num_regs, device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().gfe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # src/transform.h:57-73
    def set(self, a: int): _ck(lib().gfe_set(self._h, a))
    def square_dup(self, dup: bool = False): _ck(lib().gfe_square_dup(self._h, int(bool(dup))))
    def square_mul(self, a: int): _ck(lib().gfe_square_mul(self._h, a))
    def init_multiplicand(self, src: int): _ck(lib().gfe_init_multiplicand(self._h, src))

    def mul(self, src: int | None = None):
        """mul() (r0 *= rm) or, with src, transform::mul(src) = initMultiplicand(src) + mul()."""
        if src is not None:
            self.init_multiplicand(src)
        _ck(lib().gfe_mul(self._h))

    def copy(self, dst: int, src: int): _ck(lib().gfe_copy(self._h, dst, src))
    def sync(self): _ck(lib().gfe_sync(self._h))
    def mem_size(self) -> int: return lib().gfe_mem_size(self._h)
    def cache_size(self) -> int: return lib().gfe_cache_size(self._h)
    def error(self) -> float: return lib().gfe_get_error(self._h)

    def info(self) -> str:
        buf = ctypes.create_string_buffer(512)
        _ck(lib().gfe_info(self._h, buf, len(buf)))
        return buf.value.decode()

    def get_digits(self, reg: int = 0) -> np.ndarray:
        d = np.zeros(self.N, dtype=np.int32)
        _ck(lib().gfe_get_digits(self._h, reg, _i32(d)))
        return d

    def set_digits(self, d, reg: int = 0):
        d = np.ascontiguousarray(d, dtype=np.int32)
        if d.shape != (self.N,):
            raise ValueError("setInt")
        _ck(lib().gfe_set_digits(self._h, reg, _i32(d)))

    def save_context(self, num_regs: int = 0) -> bytes:
        size = lib().gfe_context_size(self._h, num_regs or self.num_regs)
        buf = ctypes.create_string_buffer(size)
        _ck(lib().gfe_save(self._h, num_regs, buf, size))
        return buf.raw

    def read_context(self, data: bytes, num_regs: int = 0) -> bool:
        """readContext: False for a context of another engine kind (src/transformGPU.h:1127-1139)."""
        buf = ctypes.create_string_buffer(data, len(data))
        rc = lib().gfe_load(self._h, num_regs, buf, len(data))
        if rc == -4:
            return False
        _ck(rc)
        return True

    def time_squarings(self, count: int, dup_bits: int = 0, per_kernel: bool = False):
        ms = ctypes.c_double()
        km = (ctypes.c_double * 5)()
        _ck(lib().gfe_time_squarings(self._h, count, dup_bits & (2**64 - 1), ctypes.byref(ms), km if per_kernel else None))
        return (ms.value, list(km)) if per_kernel else ms.value

    def tune(self, reps: int = 8) -> str:
        _ck(lib().gfe_tune(self._h, reps))
        return self.info()

    def get_plan(self) -> tuple:
        """(a, column E log2, columns/workgroup, row E log2, rows/workgroup, factored row twiddles)."""
        a = np.zeros(6, dtype=np.int32)
        _ck(lib().gfe_get_plan(self._h, _i32(a)))
        return tuple(int(v) for v in a)

    def set_plan(self, plan) -> None:
        a = np.ascontiguousarray(plan, dtype=np.int32)
        _ck(lib().gfe_set_plan(self._h, _i32(a)))

    def rebase(self, b: int) -> None:
        """Moves the engine to base b of the same n and prime set (gfe_rebase); set the registers
        before squaring."""
        _ck(lib().gfe_rebase(self._h, b))
        self.b = b

    def residue(self, reg: int = 0) -> dict:
        return residue(self.get_digits(reg), self.n, self.b)


def residue(digits, n: int, b: int, canonical: bool = False):
    """RES64 / old64 / hash64 / isOne of a digit vector by the library's gint (src/gint.h:40-225)."""
    d = np.ascontiguousarray(digits, dtype=np.int32)
    r = Result()
    canon = np.zeros(1 << n, dtype=np.int32) if canonical else None
    _ck(lib().gfe_residue(_i32(d), n, b, _i32(canon) if canonical else None, ctypes.byref(r)))
    out = r.as_dict()
    if canonical:
        out["digits"] = canon
    return out


def run_quick(b: int, n: int, device: int = 0) -> dict:
    """genefer -q: PRP + Gerbicz-Li check (quick(), src/genefer.h:736-748)."""
    r = Result()
    _ck(lib().gfe_run_quick(b, n, device, ctypes.byref(r)))
    return r.as_dict()


def run_prp(b: int, n: int, iters: int = 0, b_gl: int = 0, device: int = 0, digits: bool = False) -> dict:
    """The first `iters` squarings of the PRP loop (src/genefer.h:479-552)."""
    r = Result()
    d = np.zeros(1 << n, dtype=np.int32) if digits else None
    _ck(lib().gfe_run_prp(b, n, device, iters, b_gl, _i32(d) if digits else None, ctypes.byref(r)))
    out = r.as_dict()
    if digits:
        out["digits"] = d
    return out


def default_plan(n: int, capacity: int = 0) -> tuple:
    """The plan an engine (capacity 0) or a batch of `capacity` slots of n starts with (no device)."""
    p = np.zeros(6, dtype=np.int32)
    _ck(lib().gfe_default_plan(n, capacity, _i32(p)))
    return tuple(int(x) for x in p)


def plan_admissible(n: int, plan) -> bool:
    """Whether the kernels accept `plan` at n (no device)."""
    p = np.ascontiguousarray(plan, dtype=np.int32)
    if p.shape != (6,):
        raise ValueError("a plan has 6 fields")
    return lib().gfe_plan_admissible(n, _i32(p)) == 0


def engine_plan(b: int, n: int) -> tuple:
    """The plan an engine of (b, n) starts with (no device): the limb transform's where it applies."""
    p = np.zeros(6, dtype=np.int32)
    _ck(lib().gfe_engine_plan(b, n, _i32(p)))
    return tuple(int(x) for x in p)


def candidate_plans(n: int) -> list[tuple]:
    """The plans the autotuner times at n (no device)."""
    cnt = ctypes.c_uint32(0)
    _ck(lib().gfe_candidate_plans(n, None, 0, ctypes.byref(cnt)))
    buf = np.zeros(6 * cnt.value, dtype=np.int32)
    _ck(lib().gfe_candidate_plans(n, _i32(buf), cnt.value, ctypes.byref(cnt)))
    return [tuple(int(x) for x in buf[6 * i:6 * i + 6]) for i in range(cnt.value)]


def host_twiddles(n: int, is32: int, max_threads: int = 8) -> np.ndarray:
    """The engine's twiddle tables of a prime set (no device): 3 x 2^n uint32, entry j of prime k is
    psi_k^bitrev_n(j) 2^32 mod p_k; `max_threads` host threads build them (0: the calling thread)."""
    out = np.zeros((3, 1 << n), dtype=np.uint32)
    _ck(lib().gfe_host_twiddles(n, is32, max_threads, out.ctypes.data_as(U32P)))
    return out


def run_prp_resumable(b: int, n: int, ctx_path: str, save_every: int = 0, stop_after: int = 0,
                      device: int = 0) -> dict:
    """The PRP loop with the reference's context file (src/genefer.h:198-276, 479-552): resumes from
    `ctx_path` when it holds a context of this kind, saves every `save_every` squarings and stops
    (saving) after `stop_after` squarings of this call; 'done' tells whether the loop has ended, in
    which case the Gerbicz-Li check has run ('gl_ok')."""
    r = Result()
    done = ctypes.c_int32(0)
    _ck(lib().gfe_run_prp_resumable(b, n, device, os.fsencode(ctx_path), save_every, stop_after,
                                    ctypes.byref(done), ctypes.byref(r)))
    out = r.as_dict()
    out["done"] = bool(done.value)
    return out


def run_prp_trace(b: int, n: int, iters: int, every: int, device: int = 0) -> list[tuple[int, int]]:
    """(RES64, old64) after every `every`-th squaring of the PRP loop, up to `iters`."""
    cnt = iters // every
    r = (ctypes.c_uint64 * max(cnt, 1))()
    o = (ctypes.c_uint64 * max(cnt, 1))()
    res = Result()
    _ck(lib().gfe_run_prp_trace(b, n, device, iters, every, r, o, ctypes.byref(res)))
    return [(r[i], o[i]) for i in range(cnt)]


def run_proof(b: int, n: int, depth: int, device: int = 0) -> dict:
    """proof(): PRP with Pietrzak-Li checkpoints, GL, proof key (src/genefer.h:750-766)."""
    r = Result()
    mu = (ctypes.c_uint64 * (depth + 1))()
    _ck(lib().gfe_run_proof(b, n, device, depth, mu, ctypes.byref(r)))
    out = r.as_dict()
    out["mu_hash64"] = list(mu)
    return out


def write_proof(b: int, n: int, depth: int, main_path: str, checkpoints: bool = False, host_checkpoints: bool = False,
                device: int = 0) -> dict:
    """proof() writing the reference's files: `<main_path>.proof` and, with `checkpoints`, the
    checkpoint files `<main_path>_<i>.ckpt` (src/genefer.h:536-546, 620-734).  `host_checkpoints`
    keeps the checkpoints in host memory instead of HBM registers."""
    r = Result()
    mu = (ctypes.c_uint64 * (depth + 1))()
    flags = (1 if checkpoints else 0) | (2 if host_checkpoints else 0)
    _ck(lib().gfe_write_proof(b, n, device, depth, main_path.encode(), flags, mu, ctypes.byref(r)))
    out = r.as_dict()
    out["mu_hash64"] = list(mu)
    return out


def run_server(b: int, n: int, main_path: str, rnd: tuple[int, int, int], device: int = 0) -> dict:
    """server(): reads `<main_path>.proof`, recomputes the proof key and writes the certificate
    `<main_path>.cert` with the random values rnd = (rnd1, rnd2, rnd3) (src/genefer.h:771-880);
    returns the result fields (pkey, and res64 / old64 / is_prp of mu_0) and the check key."""
    r = Result()
    ck = ctypes.c_uint64()
    _ck(lib().gfe_run_server(b, n, device, main_path.encode(), rnd[0], rnd[1], rnd[2], ctypes.byref(ck), ctypes.byref(r)))
    out = r.as_dict()
    out["ckey"] = ck.value
    return out


def run_check(b: int, n: int, main_path: str, device: int = 0) -> dict:
    """check(): reads the certificate `<main_path>.cert` and returns gl_ok (1: both Gerbicz tests
    passed) and the check key (src/genefer.h:882-1062)."""
    r = Result()
    ck = ctypes.c_uint64()
    _ck(lib().gfe_run_check(b, n, device, main_path.encode(), ctypes.byref(ck), ctypes.byref(r)))
    out = r.as_dict()
    out["ckey"] = ck.value
    return out


def read_proof_file(path: str) -> dict:
    """Parses a proof file of the reference's layout (src/genefer.h:629-717): version, depth, the
    mu_k residue records (uint32 size, uint32 base, size int32 digits) and the CRC-32 trailer,
    which is checked (~crc32 ^ 0xa23777ac over the preceding bytes, src/file.h:150-161)."""
    import zlib
    data = open(path, "rb").read()
    body, trailer = data[:-4], int.from_bytes(data[-4:], "little")
    if trailer != ((~zlib.crc32(body)) & 0xFFFFFFFF) ^ 0xA23777AC:
        raise GfeError(f"{path}: bad file (crc32)")
    version, depth = np.frombuffer(body[:8], dtype=np.int32)
    mus, off = [], 8
    for _ in range(int(depth) + 1):
        size, base = np.frombuffer(body[off:off + 8], dtype=np.uint32)
        off += 8
        mus.append((int(base), np.frombuffer(body[off:off + 4 * int(size)], dtype=np.int32).copy()))
        off += 4 * int(size)
    if off != len(body):
        raise GfeError(f"{path}: trailing bytes")
    return {"version": int(version), "depth": int(depth), "mus": mus}


class Batch:
    """Many candidates of one n squared in shared launches (include/gfe_batch.h); with depth > 0 a
    proof batch (every submission: full test and Pietrzak-Li proof, checkpoints in HBM)."""

    def __init__(self, n: int, capacity: int, device: int = 0, depth: int = 0, mixed: bool = False):
        """mixed: candidates of n = 12 and 13 in one batch (n must then be 13), one launch per block."""
        self.n, self.capacity, self.depth, self.mixed = n, capacity, depth, mixed
        h = ctypes.c_void_p()
        if mixed:
            if n != 13 or depth:
                raise GfeError("a mixed batch holds quick tests of n = 12 and 13 (n = 13, depth 0)")
            _ck(lib().gfe_batch_create_mixed(capacity, device, ctypes.byref(h)))
        else:
            _ck(lib().gfe_batch_create_proof(n, capacity, depth, device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().gfe_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def slot_bytes(n: int) -> int:
        return lib().gfe_batch_slot_bytes(n)

    def info(self) -> str:
        buf = ctypes.create_string_buffer(512)
        _ck(lib().gfe_batch_info(self._h, buf, len(buf)))
        return buf.value.decode()

    def tune(self, reps: int = 4) -> str:
        _ck(lib().gfe_batch_tune(self._h, reps))
        return self.info()

    def get_plan(self) -> tuple:
        a = np.zeros(6, dtype=np.int32)
        _ck(lib().gfe_batch_get_plan(self._h, _i32(a)))
        return tuple(int(v) for v in a)

    def set_plan(self, plan) -> None:
        a = np.ascontiguousarray(plan, dtype=np.int32)
        _ck(lib().gfe_batch_set_plan(self._h, _i32(a)))

    def set_graphs(self, on: bool) -> None:
        """Squaring blocks as captured HIP graphs (default) or launched kernel by kernel."""
        _ck(lib().gfe_batch_set_graphs(self._h, 1 if on else 0))

    def submit(self, b: int, iters: int = 0, tag: int = 0, n: int | None = None) -> int | None:
        """Slot of the new candidate (of n, a mixed batch's 12 or 13; default the batch's n), or None
        when the batch is full."""
        slot = ctypes.c_int32(-1)
        rc = lib().gfe_batch_submit_n(self._h, self.n if n is None else n, b, iters, tag, ctypes.byref(slot))
        if rc == -6:
            return None
        _ck(rc)
        return slot.value

    def step(self) -> int:
        a = ctypes.c_int32(0)
        _ck(lib().gfe_batch_step(self._h, ctypes.byref(a)))
        return a.value

    def poll(self, max_results: int = 4096) -> list[dict]:
        arr = (BatchResult * max_results)()
        cnt = ctypes.c_uint32(0)
        _ck(lib().gfe_batch_poll(self._h, arr, max_results, ctypes.byref(cnt)))
        return [arr[i].as_dict() for i in range(cnt.value)]

    def active(self) -> int:
        return lib().gfe_batch_active(self._h)

    def backlog(self) -> int:
        r = lib().gfe_batch_backlog(self._h)
        _ck(min(r, 0))
        return r

    def get_digits(self, slot: int, reg: int = 0, n: int | None = None) -> np.ndarray:
        d = np.zeros(1 << (self.n if n is None else n), dtype=np.int32)
        _ck(lib().gfe_batch_get_digits(self._h, slot, reg, _i32(d)))
        return d

    def time_rounds(self, rounds: int) -> float:
        """ms for `rounds` squarings of every busy slot (benchmark: squares r0 in place)."""
        ms = ctypes.c_double()
        _ck(lib().gfe_batch_time_rounds(self._h, rounds, ctypes.byref(ms)))
        return ms.value


def run_batch(n: int, bases, iters: int = 0, capacity: int = 0, device: int = 0) -> list[dict]:
    """Quick tests (iters = 0) or truncated PRP runs of many b of one n in one batch."""
    b = np.ascontiguousarray(bases, dtype=np.uint32)
    res = (BatchResult * len(b))()
    _ck(lib().gfe_batch_run(n, b.ctypes.data_as(U32P), len(b), iters, capacity or len(b), device, res))
    return [res[i].as_dict() for i in range(len(b))]


class Server:
    """The serving loop (include/gfe_serve.h): requests of mixed n, continuous batching."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = ctypes.c_void_p()
        _ck(lib().gfe_server_create(device, hbm_budget, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().gfe_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return lib().gfe_server_clock(self._h)

    def prepare(self, n: int, capacity: int = 0):
        _ck(lib().gfe_server_prepare(self._h, n, capacity))

    def set_solo(self, max_rate: int):
        """The lone-request path (a request admitted while no pool has work runs alone on a warm
        single-candidate engine) while at most `max_rate` requests arrived in the last second
        (default 8); 0 turns it off."""
        _ck(lib().gfe_server_set_solo(self._h, int(max_rate)))

    def submit(self, n: int, b: int, iters: int = 0, tag: int = 0, depth: int = 0):
        """A quick test (or its first `iters` squarings), or with depth > 0 a proof request."""
        _ck(lib().gfe_server_submit_proof(self._h, n, b, iters, depth, tag))

    def pump(self) -> int:
        busy = ctypes.c_int32(0)
        _ck(lib().gfe_server_pump(self._h, ctypes.byref(busy)))
        return busy.value

    def poll(self, max_results: int = 4096) -> list[dict]:
        arr = (ServeResult * max_results)()
        cnt = ctypes.c_uint32(0)
        _ck(lib().gfe_server_poll(self._h, arr, max_results, ctypes.byref(cnt)))
        return [arr[i].as_dict() for i in range(cnt.value)]

    def stats(self) -> dict:
        st = ServerStats()
        _ck(lib().gfe_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()


def serve_replay(requests, device: int = 0, hbm_budget: int = 0, prepare=(), solo_rate: int = -1) -> tuple[list[dict], dict]:
    """Open-loop replay of (n, b, iters, arrival_seconds[, proof depth]) requests through the
    serving loop (solo_rate: the lone-request path's arrival-rate limit, -1 the default, 0 off).
    Returns (results in request order, server statistics)."""
    k = len(requests)
    n = np.ascontiguousarray([r[0] for r in requests], dtype=np.uint32)
    b = np.ascontiguousarray([r[1] for r in requests], dtype=np.uint32)
    it = np.ascontiguousarray([r[2] for r in requests], dtype=np.uint64)
    t = np.ascontiguousarray([r[3] for r in requests], dtype=np.float64)
    dp = np.ascontiguousarray([r[4] if len(r) > 4 else 0 for r in requests], dtype=np.uint32)
    pn = np.ascontiguousarray(list(prepare) or [0], dtype=np.uint32)
    res = (ServeResult * k)()
    st = ServerStats()
    _ck(lib().gfe_serve_replay(n.ctypes.data_as(U32P), b.ctypes.data_as(U32P), it.ctypes.data_as(U64P),
                               dp.ctypes.data_as(U32P), t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), k, device,
                               hbm_budget,
                               pn.ctypes.data_as(U32P), len(prepare), int(solo_rate), res, ctypes.byref(st)))
    return [res[i].as_dict() for i in range(k)], st.as_dict()
