"""Pietrzak-Li proof on the GPU engine against integer arithmetic, and the reference's file layouts.

tests/golden/pl_n12.json holds the proof of b = 1026, n = 12 at depths 3 and 5 computed with Python
integers alone (tests/golden/make_pl_golden.py: the PRP residues from the exponent bits, then the
protocol of src/genefer.h:620-734), so it pins the mu_k and the proof key against mathematics."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture():
    path = os.path.join(HERE, "golden", "pl_n12.json")
    if not os.path.exists(path):
        pytest.skip("pl_n12.json not generated")
    return json.load(open(path))


@pytest.mark.parametrize("depth", [3, 5])
def test_proof_and_files_match_integer_arithmetic(gfe, tmp_path, depth):
    fx = _fixture()
    want = fx["proofs"][str(depth)]
    b, n = fx["b"], fx["n"]
    main = str(tmp_path / f"cand{depth}")
    r = gfe.write_proof(b, n, depth, main, checkpoints=True)
    assert r["gl_ok"] == 1
    assert (f"{r['res64']:016X}", f"{r['old64']:016X}") == (fx["res64"], fx["old64"])
    assert [f"{h:016X}" for h in r["mu_hash64"]] == want["mu_hash64"]
    assert f"{r['pkey']:016X}" == want["pkey"]
    # the proof file: version 1, depth, depth + 1 residue records of canonical digits, CRC trailer
    p = gfe.read_proof_file(main + ".proof")
    assert (p["version"], p["depth"]) == (1, depth)
    for (base, d), h in zip(p["mus"], want["mu_hash64"]):
        assert base == b and d.shape == (1 << n,)
        assert int(d.min()) >= -1 and int(d.max()) < b
        assert f"{gfe.residue(d, n, b)['hash64']:016X}" == h
    # checkpoint files: one residue record each; checkpoint 0 is mu_0 (the PRP's final residue)
    es, bpl = fx["esize"], want["B_PL"]
    count = (es - 1) // bpl + 1
    for i in range(count):
        data = open(f"{main}_{i}.ckpt", "rb").read()
        assert len(data) == 8 + 4 * (1 << n) + 4
        if i == 0:
            d0 = np.frombuffer(data[8:-4], dtype=np.int32)
            assert np.array_equal(d0, p["mus"][0][1])
    assert not os.path.exists(f"{main}_{count}.ckpt")


def test_proof_with_host_checkpoints(gfe, tmp_path):
    """Checkpoints in host memory (the reference's file mode) give the same proof as in HBM."""
    fx = _fixture()
    r = gfe.write_proof(fx["b"], fx["n"], 3, str(tmp_path / "host"), host_checkpoints=True)
    assert [f"{h:016X}" for h in r["mu_hash64"]] == fx["proofs"]["3"]["mu_hash64"]
    assert f"{r['pkey']:016X}" == fx["proofs"]["3"]["pkey"]


def test_proof_without_files_equals_run_proof(gfe):
    fx = _fixture()
    r = gfe.run_proof(fx["b"], fx["n"], 3)
    assert [f"{h:016X}" for h in r["mu_hash64"]] == fx["proofs"]["3"]["mu_hash64"]
    assert f"{r['pkey']:016X}" == fx["proofs"]["3"]["pkey"]
