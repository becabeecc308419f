"""The proof protocol on the GPU engine against the reference's own program.

tests/golden/ref_proof.json and the certificates ref_cert_<b>_<n>.cert were written by the
reference's command-line program (CPU build, tests/golden/make_proof_golden.py) in its proof,
server and check modes (src/genefer.h:620-1062, depth 7 as src/main.cpp:237 fixes it).  Checked here:

  * the proof file the engine writes is byte-for-byte the reference's (SHA-256), with the same
    RES64, old64, proof key and mu_k hashes, and its checkpoint files hold the PRP residues;
  * the engine's server role recomputes the reference's proof key from the proof file, and a
    certificate it writes is closed by its own check with the server's check key;
  * the engine's check role closes the certificates the reference wrote, with the check key the
    reference printed, and a certificate with one digit changed gives another check key;
  * where the reference's program was built into oracle/_ref (it travels with the tree), it reads
    the engine's files: its server mode prints the engine's proof key and its check mode closes the
    engine's certificate with the engine's check key.
"""
import hashlib
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(ROOT, "oracle", "_ref", "genefer")


def _fixture():
    return json.load(open(os.path.join(HERE, "golden", "ref_proof.json")))


CASES = [(c["b"], c["n"]) for c in json.load(open(os.path.join(HERE, "golden", "ref_proof.json")))["cases"]]


def _case(b, n):
    fx = _fixture()
    return fx, next(c for c in fx["cases"] if (c["b"], c["n"]) == (b, n))


def _cli(args, cwd):
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return open(os.path.join(cwd, "results.txt")).read().splitlines()[-1]


@pytest.mark.parametrize("b,n", CASES)
def test_proof_file_is_the_reference_s(gfe, tmp_path, b, n):
    fx, c = _case(b, n)
    depth = fx["depth"]
    main = str(tmp_path / "g")
    print(f"[proof] b={b} n={n} depth={depth}", flush=True)
    r = gfe.write_proof(b, n, depth, main, checkpoints=True)
    assert r["gl_ok"] == 1
    assert (f"{r['res64']:016X}", f"{r['old64']:016X}", f"{r['pkey']:016X}") == (c["res64"], c["old64"], c["pkey"])
    assert [f"{h:016X}" for h in r["mu_hash64"]] == c["mu_hash64"]
    data = open(main + ".proof", "rb").read()
    assert len(data) == c["proof_bytes"]
    assert hashlib.sha256(data).hexdigest() == c["proof_sha256"]
    # checkpoint i holds 2^(E >> (i B_PL)) mod b^N + 1: the earliest (largest i: an exponent of at
    # most B_PL bits) against pow() where the modulus is small enough for Python integers to take
    # seconds (esize < 50,000 bits); checkpoint 0 is mu_0, the proof file's first record, which the
    # SHA-256 above pins
    E = b ** (1 << n)
    es = E.bit_length()
    bpl = ((es - 1) >> depth) + 1
    F = b ** (1 << n) + 1
    last = (es - 1) // bpl
    for i in ((last,) if es < 50000 else ()):
        raw = open(f"{main}_{i}.ckpt", "rb").read()
        d = np.frombuffer(raw[8:-4], dtype=np.int32)
        v = 0
        for x in reversed(d.tolist()):
            v = v * b + x
        v %= F
        assert v == pow(2, E >> (i * bpl), F), i
        print(f"[proof] checkpoint {i} = 2^(E >> {i * bpl})", flush=True)
    d0 = np.frombuffer(open(f"{main}_0.ckpt", "rb").read()[8:-4], dtype=np.int32)
    assert np.array_equal(d0, np.frombuffer(data[16:16 + 4 * (1 << n)], dtype=np.int32))


@pytest.mark.parametrize("b,n", CASES[:2])
def test_server_and_check_roles(gfe, tmp_path, b, n):
    fx, c = _case(b, n)
    main = str(tmp_path / "g")
    print(f"[server] b={b} n={n}", flush=True)
    gfe.write_proof(b, n, fx["depth"], main)
    s = gfe.run_server(b, n, main, (37, 101, 4242))
    assert f"{s['pkey']:016X}" == c["pkey"]
    assert (f"{s['res64']:016X}", f"{s['old64']:016X}") == (c["res64"], c["old64"])
    k = gfe.run_check(b, n, main)
    assert k["gl_ok"] == 1 and k["ckey"] == s["ckey"]
    # other random values give another certificate and check key, closed all the same
    s2 = gfe.run_server(b, n, main, (2, 16, 4))
    k2 = gfe.run_check(b, n, main)
    assert k2["gl_ok"] == 1 and k2["ckey"] == s2["ckey"] != s["ckey"]


@pytest.mark.parametrize("b,n", CASES)
def test_check_of_the_reference_s_certificate(gfe, tmp_path, b, n):
    _, c = _case(b, n)
    main = str(tmp_path / "g")
    shutil.copyfile(os.path.join(HERE, "golden", c["cert"]), main + ".cert")
    print(f"[check] b={b} n={n} {c['cert']}", flush=True)
    k = gfe.run_check(b, n, main)
    assert k["gl_ok"] == 1
    assert f"{k['ckey']:016X}" == c["ckey"]


def test_changed_certificate_gives_another_check_key(gfe, tmp_path):
    b, n = CASES[0]
    _, c = _case(b, n)
    data = bytearray(open(os.path.join(HERE, "golden", c["cert"]), "rb").read())
    # digit 100 of the residue record (after int version, int B, uint32 size, uint32 base) moves by
    # one, and the CRC-32 trailer (it covers the header and the record) is recomputed, so that the
    # change reaches the arithmetic checks
    off = 16 + 4 * 100
    d = int.from_bytes(data[off:off + 4], "little", signed=True)
    data[off:off + 4] = int((d + 1) % b).to_bytes(4, "little", signed=True)
    import zlib
    rec_end = 16 + 4 * (1 << n)
    crc = ((~zlib.crc32(bytes(data[:rec_end]))) & 0xFFFFFFFF) ^ 0xA23777AC
    data[-4:] = crc.to_bytes(4, "little")
    main = str(tmp_path / "g")
    open(main + ".cert", "wb").write(bytes(data))
    k = gfe.run_check(b, n, main)
    # as the reference's check (src/genefer.h:887-1062): its Gerbicz tests guard the checker's own
    # arithmetic and pass; the check key, which the server compares with its own, no longer matches
    assert k["gl_ok"] == 1
    assert f"{k['ckey']:016X}" != c["ckey"]
    # a corrupted trailer is a malformed file
    data[-1] ^= 0xFF
    open(main + ".cert", "wb").write(bytes(data))
    with pytest.raises(gfe.GfeError):
        gfe.run_check(b, n, main)


@pytest.mark.skipif(not os.path.exists(CLI), reason="the reference's program is not built (make -C oracle cli)")
@pytest.mark.parametrize("b,n", CASES[:2])
def test_reference_program_reads_the_engine_s_files(gfe, tmp_path, b, n):
    fx, c = _case(b, n)
    main = str(tmp_path / "g")
    print(f"[interop] b={b} n={n}", flush=True)
    gfe.write_proof(b, n, fx["depth"], main)
    # the reference's server on the engine's proof prints the same proof key
    line = _cli(["-s", "-b", str(b), "-n", str(n), "-f", "g"], str(tmp_path))
    assert re.search(r"pkey = ([0-9A-F]{16})", line).group(1) == c["pkey"]
    # the engine's certificate is closed by the reference's check with the engine's check key
    s = gfe.run_server(b, n, main, (53, 200, 60000))
    line = _cli(["-c", "-b", str(b), "-n", str(n), "-f", "g"], str(tmp_path))
    assert "is checked" in line
    assert re.search(r"ckey = ([0-9A-F]{16})", line).group(1) == f"{s['ckey']:016X}"
