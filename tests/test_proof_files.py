"""The reference's proof-protocol files (tests/golden/ref_proof.json and the certificates the
reference's own program wrote) against mathematics, on the CPU.

  * certificate layout (src/genefer.h:871-879): int version 1, int B = B_PietrzakLi(esize, 7), the
    residue record (uint32 size, uint32 base, size canonical int32 digits), p2 in GMP's raw format
    (int32 big-endian byte count, big-endian magnitude) and the CRC-32 trailer, which covers the
    header and the record but not p2 (the reference's file updates its CRC for raw integers only
    when the write fails, src/file.h:124-138);
  * the check key the reference printed is hash64(u^(2^B) 2^p2) for the certificate's residue u,
    recomputed with Python integers;
  * the Pietrzak-Li proof computed with Python integers alone (tests/golden/pl_n12.json, depth 7)
    has the reference's proof key and mu_k hashes, so the restated protocol (src/genefer.h:620-734)
    is pinned against an execution of the reference.
"""
import hashlib
import json
import os
import struct
import zlib

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FX = json.load(open(os.path.join(GOLD, "ref_proof.json")))


def parse_cert(data: bytes, n: int):
    version, B = struct.unpack_from("<ii", data, 0)
    size, base = struct.unpack_from("<II", data, 8)
    assert size == 1 << n
    digits = np.frombuffer(data, dtype="<i4", count=size, offset=16)
    off = 16 + 4 * size
    (nbytes,) = struct.unpack_from(">i", data, off)
    mag = int.from_bytes(data[off + 4:off + 4 + abs(nbytes)], "big")
    p2 = -mag if nbytes < 0 else mag
    end = off + 4 + abs(nbytes)
    assert end + 4 == len(data)
    (crc,) = struct.unpack_from("<I", data, end)
    return version, B, base, digits, p2, crc, off


def to_int(d, b):
    v = 0
    for x in reversed([int(t) for t in d]):
        v = v * b + x
    return v


@pytest.mark.parametrize("case", FX["cases"], ids=lambda c: f"b{c['b']}_n{c['n']}")
def test_reference_certificate_layout(oracle, case):
    b, n = case["b"], case["n"]
    data = open(os.path.join(GOLD, case["cert"]), "rb").read()
    assert hashlib.sha256(data).hexdigest() == case["cert_sha256"]
    version, B, base, digits, p2, crc, rec_end = parse_cert(data, n)
    assert version == 1 and base == b and p2 > 0
    es = (b ** (1 << n)).bit_length()
    assert B == case["cert_B"] == oracle.b_pietrzak_li(es, FX["depth"])
    assert int(digits.min()) >= -1 and int(digits.max()) < b
    assert crc == ((~zlib.crc32(data[:rec_end])) & 0xFFFFFFFF) ^ 0xA23777AC


@pytest.mark.parametrize("case", FX["cases"][:2], ids=lambda c: f"b{c['b']}_n{c['n']}")
def test_check_key_of_reference_certificate_from_integers(oracle, case):
    b, n = case["b"], case["n"]
    data = open(os.path.join(GOLD, case["cert"]), "rb").read()
    _, B, _, digits, p2, _, _ = parse_cert(data, n)
    F = b ** (1 << n) + 1
    u = to_int(digits, b) % F
    v = pow(u, 1 << B, F) * pow(2, p2, F) % F
    assert f"{oracle.hash64(oracle.int_to_digits(v, n, b), n, b):016X}" == case["ckey"] == case["check_ckey"]


def test_integer_proof_equals_the_reference_s():
    pl = json.load(open(os.path.join(GOLD, "pl_n12.json")))
    if "7" not in pl["proofs"]:
        pytest.skip("pl_n12.json has no depth-7 proof")
    case = next(c for c in FX["cases"] if (c["b"], c["n"]) == (pl["b"], pl["n"]))
    assert (pl["res64"], pl["old64"]) == (case["res64"], case["old64"])
    assert pl["proofs"]["7"]["pkey"] == case["pkey"]
    assert pl["proofs"]["7"]["mu_hash64"] == case["mu_hash64"]
