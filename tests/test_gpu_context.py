"""Context files in the layout of the reference's GPU build (src/genefer.h:198-276 header and CRC,
src/transformGPU.h:1127-1150 payload): the engine's payload against its own registers, the
resumable PRP loop against one uninterrupted run and the reference's KATs, and files that pass
between the engine and the oracle in both directions.  The file layout is built and parsed here
from the reference's description, independently of the library's writer."""
import os
import struct
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRIMES_S = [127 * 2**24 + 1, 63 * 2**25 + 1, 15 * 2**27 + 1]
PRIMES_U = [125 * 2**25 + 1, 243 * 2**24 + 1, 235 * 2**24 + 1]


def primes_of(oracle, b, n):
    rns, is32 = oracle.select(b, n)
    return (PRIMES_U if is32 else PRIMES_S)[:rns]


def residues(d, p):
    d = np.asarray(d, dtype=np.int64)
    return np.where(d < 0, d + p, d).astype(np.uint32)


def digits_of(z0, p0):
    """ZPT::get_int of the first prime's residues (src/transformGPU.h:146)."""
    z0 = z0.astype(np.int64)
    return np.where(z0 >= p0 // 2, z0 - p0, z0).astype(np.int32)


def crc_trailer(body: bytes) -> bytes:
    return struct.pack("<I", (~zlib.crc32(body) & 0xFFFFFFFF) ^ 0xA23777AC)


def context_file(i, elapsed, regs, primes) -> bytes:
    """A PRP-loop context as the reference's GPU build writes it: int -1, int 0, int i, double
    elapsed, int kind (6 NTT2, 7 NTT3), residues [register][prime][N], CRC trailer."""
    body = struct.pack("<iiid", -1, 0, i, elapsed) + struct.pack("<i", 6 if len(primes) == 2 else 7)
    body += b"".join(residues(d, p).tobytes() for d in regs for p in primes)
    return body + crc_trailer(body)


def parse_context(data: bytes, N: int, primes):
    version, where, i, elapsed, kind = struct.unpack_from("<iiidi", data, 0)
    z = np.frombuffer(data, dtype=np.uint32, offset=24, count=(len(data) - 28) // 4)
    return {"version": version, "where": where, "i": i, "elapsed": elapsed, "kind": kind,
            "regs": z.reshape(-1, len(primes), N), "crc_ok": data[-4:] == crc_trailer(data[:-4])}


@pytest.mark.parametrize("b", [1026, 1000000000, 1999999266])
def test_payload_is_the_reference_gpu_layout(gfe, oracle, b):
    n = 12
    N = 1 << n
    primes = primes_of(oracle, b, n)
    print(f"payload layout b={b} n={n} primes={len(primes)}", flush=True)
    t = gfe.Transform(b, n, num_regs=3)
    t.set(1)
    for k in range(40):
        t.square_dup(k % 3 == 0)
    t.copy(1, 0)
    t.square_dup(True)
    t.copy(2, 0)
    t.square_dup(False)
    data = t.save_context()
    assert len(data) == 4 + 3 * len(primes) * N * 4
    kind = struct.unpack_from("<i", data)[0]
    assert kind == (6 if len(primes) == 2 else 7)
    z = np.frombuffer(data, dtype=np.uint32, offset=4).reshape(3, len(primes), N)
    regs = [t.get_digits(r) for r in range(3)]
    for r in range(3):
        for k, p in enumerate(primes):
            assert np.array_equal(z[r, k], residues(regs[r], p)), (r, k)
        assert np.array_equal(digits_of(z[r, 0], primes[0]), regs[r])
    u = gfe.Transform(b, n, num_regs=3)
    assert u.read_context(data)
    for r in range(3):
        assert np.array_equal(oracle.unbalance(u.get_digits(r), n, b), oracle.unbalance(regs[r], n, b))
    # digits in (-b, b), as the reference's normalisation leaves them, load as the same value
    rs = np.random.RandomState(b % 1000)
    wide = [rs.randint(-(b - 1), b, N).astype(np.int32) for _ in range(3)]
    wide[0][:4] = b - 1
    wide[1][-4:] = -(b - 1)
    wdata = struct.pack("<i", kind) + b"".join(residues(d, p).tobytes() for d in wide for p in primes)
    assert u.read_context(wdata)
    F = b ** N + 1
    for r in range(3):
        got = u.get_digits(r)
        assert np.abs(got.astype(np.int64)).max() <= b // 2 + 4
        assert oracle.digits_to_int(got, b) % F == oracle.digits_to_int(wide[r], b) % F, r
    # registers 0..1 only, as the PRP loop's context holds them
    two = t.save_context(2)
    assert two == data[:4 + 2 * len(primes) * N * 4]
    # the other kind is refused (readContext false); residues that disagree between primes or a
    # truncated payload are errors
    assert not u.read_context(struct.pack("<i", 13 - kind) + data[4:])
    bad = bytearray(data)
    bad[4 + (N + 5) * 4] ^= 1  # register 0, prime 1, digit 5
    with pytest.raises(gfe.GfeError):
        u.read_context(bytes(bad))
    with pytest.raises(gfe.GfeError):
        u.read_context(data[:-8])
    t.close()
    u.close()


def test_resumed_prp_matches_the_kat(gfe, tmp_path):
    """Three calls, stopped after 12345 and 20000 squarings, end at the RES64 of the reference's KAT
    (src/main.cpp:424) with the Gerbicz-Li check passing; a fourth call finds i = -1 and only
    reruns the check."""
    b, n, want = 1000000000, 12, 0x4E43A5B93273C649
    path = str(tmp_path / "prp.ctx")
    print(f"resume b={b} n={n}", flush=True)
    r1 = gfe.run_prp_resumable(b, n, path, save_every=5000, stop_after=12345)
    assert not r1["done"] and r1["iters"] == 12345 and r1["gl_ok"] == -1 and r1["is_prp"] == -1
    assert os.path.exists(path) and os.path.exists(path + ".old") and not os.path.exists(path + ".new")
    r2 = gfe.run_prp_resumable(b, n, path, stop_after=20000)
    assert not r2["done"] and r2["iters"] == 32345
    assert r2["seconds"] >= r1["seconds"]
    r3 = gfe.run_prp_resumable(b, n, path, save_every=50000)
    assert r3["done"] and r3["iters"] == r3["esize"]
    assert r3["res64"] == want, hex(r3["res64"])
    assert r3["gl_ok"] == 1
    r4 = gfe.run_prp_resumable(b, n, path)
    assert r4["done"] and r4["res64"] == want and r4["gl_ok"] == 1


def test_context_files_pass_between_oracle_and_engine(gfe, oracle, tmp_path):
    """The oracle runs 2000 iterations and its registers go into a context file written here in the
    reference GPU build's layout; the engine resumes from it for 1000 iterations and its own file
    holds the registers the oracle reaches after those 1000; the engine then finishes with the
    RES64 of an uninterrupted run and a passing Gerbicz-Li check.  A file from an engine of the
    other kind (another b with another prime set) is not taken: the run starts afresh."""
    b, n = 1026, 12
    N = 1 << n
    E = oracle.exponent(b, n)
    esize = E.bit_length()
    B_GL = oracle.b_gerbicz_li(esize)
    primes = primes_of(oracle, b, n)
    print(f"oracle <-> engine contexts b={b} n={n} esize={esize}", flush=True)
    o = oracle.Engine(b, n, num_regs=3)
    oracle.prp(o, E, iters=2000, B_GL=B_GL)
    path = str(tmp_path / "c.ctx")
    with open(path, "wb") as f:
        f.write(context_file(esize - 1 - 2000, 1.5, [o.get_digits(0), o.get_digits(1)], primes))
    r = gfe.run_prp_resumable(b, n, path, stop_after=1000)
    assert not r["done"] and r["iters"] == 3000 and r["seconds"] >= 1.5
    with open(path, "rb") as f:
        c = parse_context(f.read(), N, primes)
    assert c["version"] == -1 and c["where"] == 0 and c["crc_ok"]
    assert c["i"] == esize - 1 - 3000 and c["kind"] == (6 if len(primes) == 2 else 7)
    assert c["regs"].shape == (2, len(primes), N)
    oracle.prp(o, E, iters=1000, B_GL=B_GL, resume_at=esize - 1 - 2000)
    for reg in range(2):
        d = digits_of(c["regs"][reg, 0], primes[0])
        for k, p in enumerate(primes):
            assert np.array_equal(c["regs"][reg, k], residues(d, p))
        assert np.array_equal(oracle.unbalance(d, n, b), oracle.unbalance(o.get_digits(reg), n, b)), reg
    fin = gfe.run_prp_resumable(b, n, path)
    full = gfe.run_quick(b, n)
    assert fin["done"] and fin["res64"] == full["res64"] and fin["gl_ok"] == 1

    other = str(tmp_path / "other.ctx")
    with open(other, "wb") as f:
        f.write(context_file(esize - 1 - 2000, 0.0, [o.get_digits(0), o.get_digits(1)], PRIMES_S[:2]))
    b3 = 1000000000  # NTT3 at n = 12: the NTT2 file is of another kind
    fresh = gfe.run_prp_resumable(b3, n, other, stop_after=10)
    assert fresh["iters"] == 10
