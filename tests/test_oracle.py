"""CPU tests that pin the oracle (oracle/gfe_oracle.c + oracle/__init__.py).

Pinning table (what fixes each restated function, so that a dropped term, a wrong sign or index,
or a transposed operand fails at least one check):
  constants (src/transformGPU.h:37-115) ...... literal values in tests/golden/reference_values.json
  prime-set rule (src/transform_ocl.cpp:35-50) closed form and boundary values
  twiddles (src/transformGPU.h:1046-1078) ..... direct powers of the generator, root orders, mirror identity
  transform + pair square + CRT + carry ...... schoolbook negacyclic product; closed form 2^k mod b^N+1
  gint unbalance / RES64 / old64 ............. Python big integers
  gint hash64 / hash32 ....................... the reference's own gint (compiled, oracle/_ref)
  PRP loop + Gerbicz-Li ...................... the reference's KATs (src/main.cpp, design/rns.cpp);
                                                the compiled reference at intermediate iterations;
                                                a planted error makes the check fail
  Pietrzak-Li proof .......................... checkpoints against the closed form; the proof key
                                                against v1 = prod mu_k^hash32(mu_k) in Python integers;
                                                the mu_k themselves: parity unpinned (no reference
                                                execution of proof mode)
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))


def test_prime_constants_match_reference(oracle):
    for set_name, base in (("S", 0), ("U", 3)):
        for k in range(3):
            assert oracle.prime_constants(base + k) == REF["constants"][set_name][k], (set_name, k)


def test_garner_constants_match_reference(oracle):
    assert oracle.garner_constants(False) == REF["constants"]["garner_S"]
    assert oracle.garner_constants(True) == REF["constants"]["garner_U"]


def test_norm_matches_reference(oracle):
    want = REF["constants"]["norm_n16"]
    assert [oracle.norm(k, 16) for k in range(3)] == [want["NORM1"], want["NORM2"], want["NORM3"]]
    for k in range(6):
        p = oracle.lib().gfo_prime(k)
        for n in (5, 12, 23):
            assert oracle.norm(k, n) * (1 << (n - 1)) % p == 1


def test_prime_set_rule(oracle):
    S, U = oracle.PRIMES_S, oracle.PRIMES_U
    for n in (12, 15, 16, 20, 22, 23):
        N = 1 << n
        for b in (1024, 20000, 1_000_000, 8_000_000, 30_000_000, 999_999_998, 1_000_000_000, 1_000_000_002, 2_000_000_000):
            if b * b < S[0] * S[1] // 2 // N:
                want = (2, False)
            elif b * b < U[0] * U[1] // 2 // N:
                want = (2, True)
            elif b <= 1_000_000_000:
                want = (3, False)
            else:
                want = (3, True)
            assert oracle.select(b, n) == want, (b, n)
    # thresholds quoted for n = 12: NTT2-S below 2.34e7, NTT2-U below 4.57e7
    assert oracle.select(23_400_000, 12) == (2, False) and oracle.select(23_500_000, 12) == (2, True)
    assert oracle.select(45_600_000, 12) == (2, True) and oracle.select(45_800_000, 12) == (3, False)


@pytest.mark.parametrize("idx", range(6))
def test_twiddles(oracle, idx):
    n = 9
    p, g = oracle.lib().gfo_prime(idx), oracle.lib().gfo_generator(idx)
    w = [int(x) for x in oracle.twiddles(idx, n)]
    N = 1 << n
    assert pow(g, (p - 1) // 2, p) == p - 1  # g is a quadratic non-residue (a generator of the 2-part)
    s = 1
    while s < N // 2:
        r = pow(g, (p - 1) // (4 * s), p)
        assert pow(r, 4 * s, p) == 1 and pow(r, 2 * s, p) == p - 1  # primitive root of order 4s
        for j in range(s):
            e = int(bin(j)[2:].zfill(max(1, (2 * s).bit_length() - 1))[::-1], 2) + 1
            assert w[s + j] == pow(r, e, p)
            # inverse of the root of block j is minus the root of block s - 1 - j
            assert w[s + j] * ((p - w[s + (s - 1 - j)]) % p) % p == 1
        s *= 2


@pytest.mark.parametrize("n", [4, 5, 6, 7])
@pytest.mark.parametrize("rns,is32", [(2, False), (2, True), (3, False), (3, True)])
def test_transform_product_vs_schoolbook(oracle, n, rns, is32):
    rs = np.random.RandomState(100 * n + rns + is32)
    b = 1_000_000 if rns == 2 else 999_999_998
    N = 1 << n
    a = oracle.balance(rs.randint(0, b, N).astype(np.int32), n, b)
    c = oracle.balance(rs.randint(0, b, N).astype(np.int32), n, b)
    e = oracle.Engine(b, n, num_regs=2, rns=rns, is32=is32)
    e.set_digits(a, 0)
    e.square_dup(False)
    assert np.array_equal(oracle.unbalance(e.get_digits(0), n, b), oracle.unbalance(oracle.schoolbook_mul(a, a, n, b), n, b))
    e.set_digits(a, 0)
    e.square_dup(True)
    two_a2 = oracle.schoolbook_mul(oracle.schoolbook_mul(a, a, n, b), np.array([2] + [0] * (N - 1), np.int32), n, b)
    assert np.array_equal(oracle.unbalance(e.get_digits(0), n, b), oracle.unbalance(two_a2, n, b))
    e.set_digits(a, 0)
    e.set_digits(c, 1)
    e.mul(1)
    assert np.array_equal(oracle.unbalance(e.get_digits(0), n, b), oracle.unbalance(oracle.schoolbook_mul(a, c, n, b), n, b))
    # schoolbook itself against Python integers mod b^N + 1
    F = b ** N + 1
    v = oracle.digits_to_int(oracle.schoolbook_mul(a, c, n, b), b) % F
    assert v == oracle.digits_to_int(a, b) * oracle.digits_to_int(c, b) % F


@pytest.mark.parametrize("b,n,iters", [(123456, 5, 80), (1_999_999_266, 6, 150), (30_000_000, 7, 120), (1024 + 2, 8, 60)])
def test_prp_closed_form(oracle, b, n, iters):
    """r0 after k PRP iterations equals 2^(E >> (esize - k)) mod b^N + 1, for any GL cadence."""
    E = oracle.exponent(b, n)
    for B_GL in (None, 7, -1):
        e = oracle.Engine(b, n)
        oracle.prp(e, E, iters=iters, B_GL=B_GL)
        v = oracle.digits_to_int(e.get_digits(0), b) % (b ** (1 << n) + 1)
        assert v == oracle.closed_form_residue(b, n, iters)


def test_prp_resume_equals_one_run(oracle):
    """A loop stopped after 70 iterations and resumed from its registers 0 and 1 (a context,
    src/genefer.h:484-503) reaches the registers of one uninterrupted run of 150, and r0 the closed
    form; the Gerbicz-Li product crosses the stop (B_GL = 7)."""
    b, n, B_GL = 1_999_999_266, 6, 7
    E = oracle.exponent(b, n)
    esize = E.bit_length()
    one = oracle.Engine(b, n)
    oracle.prp(one, E, iters=150, B_GL=B_GL)
    first = oracle.Engine(b, n)
    oracle.prp(first, E, iters=70, B_GL=B_GL)
    second = oracle.Engine(b, n)
    second.set_digits(first.get_digits(0), 0)
    second.set_digits(first.get_digits(1), 1)
    oracle.prp(second, E, iters=80, B_GL=B_GL, resume_at=esize - 1 - 70)
    for reg in (0, 1):
        assert np.array_equal(oracle.unbalance(second.get_digits(reg), n, b), oracle.unbalance(one.get_digits(reg), n, b))
    v = oracle.digits_to_int(second.get_digits(0), b) % (b ** (1 << n) + 1)
    assert v == oracle.closed_form_residue(b, n, 150)


def test_square_mul(oracle):
    b, n = 777_778, 6
    rs = np.random.RandomState(3)
    a = oracle.balance(rs.randint(0, b, 1 << n).astype(np.int32), n, b)
    e = oracle.Engine(b, n)
    e.set_digits(a, 0)
    e.square_mul(12345)
    F = b ** (1 << n) + 1
    assert oracle.digits_to_int(e.get_digits(0), b) % F == 12345 * oracle.digits_to_int(a, b) ** 2 % F


def test_gint_canonical_and_res64(oracle):
    rs = np.random.RandomState(7)
    for n, b in ((4, 1026), (6, 999_999_998), (8, 2_000_000_000)):
        N = 1 << n
        F = b ** N + 1
        d = rs.randint(-b, b, N).astype(np.int32)
        v = oracle.digits_to_int(d, b) % F
        c = oracle.unbalance(d, n, b)
        assert np.array_equal(c, oracle.int_to_digits(v, n, b))
        r = oracle.residue(d, n, b)
        assert r.res64 == v % (1 << 64) or (v == F - 1)
        top = [int(x) & 0xFF for x in c[N - 8:]]
        assert r.old64 == int.from_bytes(bytes(top), "big")
        bal = oracle.balance(d, n, b)
        assert np.abs(bal).max() <= b // 2 and oracle.digits_to_int(bal, b) % F == v
    # -1 = b^N is kept as digit 0 = -1 (src/gint.h:86-100)
    d = np.zeros(16, np.int32)
    d[0] = -1
    assert oracle.unbalance(d, 4, 1026)[0] == -1


def test_gint_hashes_match_reference(oracle):
    if not oracle.RefEngine.available():
        pytest.skip("oracle/_ref not built")
    rs = np.random.RandomState(11)
    for n, b in ((12, 1026), (12, 999_999_998), (13, 2_000_000_000), (12, 45_687_570)):
        d = rs.randint(-(b // 2), b // 2, 1 << n).astype(np.int32)
        c, r64, o64, h64, h32, one = oracle.RefEngine.gint(d, n, b)
        assert np.array_equal(oracle.unbalance(d, n, b), c)
        mine = oracle.residue(d, n, b)
        assert (mine.res64, mine.old64, mine.hash64) == (r64, o64, h64)
        assert oracle.hash32(d, n, b) == h32


@pytest.mark.parametrize("kat", [k for k in REF["kats"] if k["n"] <= 10], ids=lambda k: f"b{k['b']}_n{k['n']}")
def test_kats_full_quick_test(oracle, kat):
    b, n = kat["b"], kat["n"]
    e = oracle.Engine(b, n)
    r, gl = oracle.quick(e, oracle.exponent(b, n))
    assert f"{r.res64:016X}" == kat["res64"]
    assert gl


def test_gerbicz_li_detects_a_planted_error(oracle):
    """add1() (src/transform.h:248-255) during the PRP loop must make the check fail."""
    b, n = 1_000_000_000, 6
    E = oracle.exponent(b, n)
    B = oracle.b_gerbicz_li(E.bit_length())

    class Faulty(oracle.Engine):
        calls = 0

        def square_dup(self, dup):
            super().square_dup(dup)
            Faulty.calls += 1
            if Faulty.calls == 100:
                d = self.get_digits(0)
                d[0] += 1
                self.set_digits(d, 0)

    e = Faulty(b, n)
    oracle.prp(e, E, B_GL=B)
    assert not oracle.gerbicz_li(e, E, B)
    e2 = oracle.Engine(b, n)
    oracle.prp(e2, E, B_GL=B)
    assert oracle.gerbicz_li(e2, E, B)


@pytest.mark.parametrize("b,n,iters", [(1_000_000_000, 12, 300), (23_448_336, 12, 300), (2_000_000_000, 13, 120), (1026, 12, 300)])
def test_oracle_vs_compiled_reference(oracle, b, n, iters):
    if not oracle.RefEngine.available():
        pytest.skip("oracle/_ref not built")
    ref = oracle.RefEngine(b, n)
    _, rr, _ = ref.prp(iters=iters, b_gl=0)
    e = oracle.Engine(b, n)
    oracle.prp(e, oracle.exponent(b, n), iters=iters)
    r = e.residue(0)
    assert (r.res64, r.old64) == (rr.res64, rr.old64)


def test_pietrzak_li_proof_identity(oracle):
    """The proof's mu_k follow the protocol of src/genefer.h:620-734: with the checkpoints of the
    PRP loop, mu_0 is the first checkpoint, and the proof key is hash64(v1) where
    v1 = prod_k mu_k^{q_k} (q_k = hash32(mu_k)) -- recomputed here from the mu's with Python
    integers (parity of the proof against an execution of the reference is unpinned)."""
    b, n, depth = 1_000_000_000, 5, 3
    E = oracle.exponent(b, n)
    es = E.bit_length()
    B_PL = oracle.b_pietrzak_li(es, depth)
    e = oracle.Engine(b, n)
    ck = {}
    oracle.prp(e, E, B_GL=oracle.b_gerbicz_li(es), B_PL=B_PL, checkpoints=ck)
    assert sorted(ck) == list(range(0, (es - 1) // B_PL + 1))
    F = b ** (1 << n) + 1
    for i, d in ck.items():  # checkpoint i = 2^(E >> (i B_PL)) mod F
        assert oracle.digits_to_int(d, b) % F == pow(2, E >> (i * B_PL), F)
    mus, pkey = oracle.pietrzak_li(e, ck, depth)
    assert len(mus) == depth + 1
    assert np.array_equal(mus[0], oracle.unbalance(ck[0], n, b))
    v1 = 1
    for mu in mus:
        q = oracle.hash32(mu, n, b)
        v1 = v1 * pow(oracle.digits_to_int(mu, b), q, F) % F
    assert pkey == oracle.hash64(oracle.int_to_digits(v1, n, b), n, b)


def test_pietrzak_li_proof_against_integer_fixture(oracle):
    """The oracle engine's proof at b = 1026, n = 12 equals tests/golden/pl_n12.json, which holds the
    proof computed with Python integers alone (tests/golden/make_pl_golden.py): pins the mu_k and
    the proof key of the restated protocol (src/genefer.h:620-734) against mathematics."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pl_n12.json")
    if not os.path.exists(path):
        pytest.skip("pl_n12.json not generated")
    fx = json.load(open(path))
    b, n = fx["b"], fx["n"]
    E = oracle.exponent(b, n)
    es = E.bit_length()
    assert es == fx["esize"]
    want = fx["proofs"]["3"]
    B_PL = oracle.b_pietrzak_li(es, 3)
    assert B_PL == want["B_PL"]
    e = oracle.Engine(b, n)
    ck = {}
    oracle.prp(e, E, B_GL=oracle.b_gerbicz_li(es), B_PL=B_PL, checkpoints=ck)
    mus, pkey = oracle.pietrzak_li(e, ck, 3)
    assert [f"{oracle.hash64(m, n, b):016X}" for m in mus] == want["mu_hash64"]
    assert f"{pkey:016X}" == want["pkey"]
