"""bench.py's output contract: one JSON line with the headline metric, the roofline and CPU-baseline
objects, and the refusal of more than one GPU (checked without a device)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=timeout)


def test_bench_refuses_more_than_one_gpu():
    r = run_bench("--gpus", "2", timeout=60)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "one GPU" in r.stderr


@pytest.mark.gpu
def test_bench_json_line():
    r = run_bench("--steps", "50", "--warmup", "5", "--no-batch", "--no-serve", "--job-iters", "0",
                  "--cpu-seconds", "1")
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1, r.stdout
    line = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in line, k
    assert line["n_gpus"] == 1 and line["steps"] == 50 and line["warmup"] == 5
    assert line["unit"] == "ns/squaring" and line["higher_is_better"] is False
    assert line["value"] > 0 and abs(line["value"] - line["ms_per_step"] * 1e6) < 1.0
    assert line["config"]["workload"] and line["config"]["n"] == 22
    rf = line["roofline"]
    assert rf["bound"] in ("hbm", "mfma") and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["achieved"] > 0 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    par = line["config"]["parity"]
    cpu = line["cpu_baseline"]
    if par is not None:  # the reference CPU build is present
        assert par["match"], par
        assert cpu["kind"] == "reference" and cpu["cores"] >= 1 and cpu["value"] > 0 and cpu["sample"]
        small = line["config"]["small_configs"]
        assert set(small) == {"C1", "C2", "C3"}
        assert small["C3"]["cpu_candidates_per_hour"] > 0


def test_measured_counters_cover_the_default_plan():
    """bench.py's roofline.traffic and VALU figures are profile constants of the profiled plan: the
    n = 22 default plan must be one of them, the kernel that plan's row pass instantiates must be the
    one the entry names, and the named profile summary must hold that kernel and those figures."""
    import importlib.util
    import genefer22_amd
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    if not os.path.exists(genefer22_amd.LIB_PATH):
        from genefer22_amd import build
        build.build()
    genefer22_amd.lib()
    from genefer22_amd import workloads
    plan = genefer22_amd.engine_plan(workloads.config_bases(workloads.C4, 1)[0], 22)  # what bench_n22 runs
    m = bench.measured("row", plan)
    assert m.get("traffic") and m.get("valu_instr_per_butterfly"), m
    a, cl, cng, rl, rng, fl = plan
    assert fl & 8, plan
    tn = 21 if fl & 32 else 22  # the two-digit-limb transform has 2^21 points
    assert m["kernel"].endswith(f", {rl}, {tn - a}>"), (m["kernel"], plan)
    text = open(os.path.join(ROOT, m["profile"])).read()
    short = m["kernel"].replace("gfe::", "")
    assert m["kernel"] in text and short in text
    row = next(l for l in text.splitlines() if l.startswith(f"| `{short}`") and "MiB" not in l and "M |" in l)
    assert f"{m['fetch_mib']}" in row and f"{m['write_mib']}" in row, row
    assert f"{m['valu_instr_per_butterfly']} VALU instructions" in text


def test_serving_parity_finds_a_changed_residue(tmp_path):
    """bench.py's serving.parity: every served RES64 / old64 is looked up by request id in the
    reference fixture; requests the fixture does not hold (or holds for another n, b or length) are
    not counted, a changed residue is reported and fails the comparison, a missing fixture gives None."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    from genefer22_amd import workloads
    reqs = workloads.serving_requests(6, 24.0)
    assert [(q.n, q.b) for q in reqs] == [(q.n, q.b) for q in workloads.serving_requests(6, 2.0)]
    ref = [{"rid": q.rid, "n": q.n, "b": q.b, "iters": q.iters, "res64": f"{0x1000 + q.rid:016X}",
            "old64": f"{0x2000 + q.rid:016X}"} for q in reqs[:5]]
    ref[4]["b"] += 2  # another candidate under the same id: not compared
    path = tmp_path / "ref.json"
    path.write_text(json.dumps({"results": ref}))
    results = [{"res64": 0x1000 + q.rid, "old64": 0x2000 + q.rid} for q in reqs]
    p = bench.serving_parity(reqs, results, str(path))
    assert p["checked"] == 4 and p["of"] == 6 and p["match"] and p["mismatches"] == []
    results[2]["old64"] ^= 1
    p = bench.serving_parity(reqs, results, str(path))
    assert not p["match"] and p["mismatches"] == [reqs[2].rid]
    assert bench.serving_parity(reqs, results, str(tmp_path / "absent.json")) is None
