"""The reference CUDA CLI's command contract on the gfx950 transform
(reference src/cuda_cli.cpp:1-256; bench bodies src/cuda/ntt_kernels.cu:491-821;
vector-file format include/ntt/vector_io.hpp:1-104; timing protocol
include/ntt/measurement.hpp:13-95, benchmark_mode.hpp:18-22).

CPU: the vector-file parser and decimal reduction against Python big
integers, the percentile / JSON / operand-stream helpers against numpy,
json and a Python restatement, and the CLI's preflight errors (raised before
any device call). GPU: transform / identity / verify / bench and the three
JSONL modes, every transform output against the oracle; the resident
conversion benchmark (reference src/resident_conversion_bench.cpp:97-300) on
the HIP Session.
"""
import json
import os
import random
import subprocess

import numpy as np
import pytest

import oracle as o
from common import log_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ntt-gpu_amd", "bin", "ntt_hip_cli")
CONVERSION = os.path.join(ROOT, "ntt-gpu_amd", "bin", "session_conversion_bench")
TOOL = os.path.join(ROOT, "tests", "cpp", "build", "vector_io_tool")


def _run(*args, timeout=300):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)


def _params(n, count, bits=46):
    q = o.find_primes(n, count, bits)
    return q, [o.min_root(m, n) for m in q]


def _vector_file(path, n, q, psi, coeffs, comment="# test vector\n"):
    with open(path, "w") as f:
        f.write(comment)
        f.write(f"{n} {len(q)}\n")
        for m, r in zip(q, psi):
            f.write(f"{m} {r}\n")
        f.write("COEFFS\n")
        f.write("".join(f"{c}\n" for c in coeffs))
    return path


def _big_coeffs(n, q, seed):
    big_q = 1
    for m in q:
        big_q *= m
    rng = random.Random(seed)
    return [rng.randrange(big_q) for _ in range(n)]


# ---------------------------------------------------------------- CPU
def test_vector_file_staging_matches_python_big_integers(tmp_path):
    n = 16
    q, psi = _params(n, 4, 60)
    coeffs = _big_coeffs(n, q, 11) + []
    coeffs[0], coeffs[1] = 0, 10 ** 70  # zero and a value far beyond the product of the primes
    path = _vector_file(tmp_path / "v.txt", n, q, psi, coeffs, "# c1\n# c2\n")
    r = _run(TOOL, "stage", path)
    assert r.returncode == 0, r.stderr
    got = [int(x) for x in r.stdout.split()]
    assert got == [c % m for m in q for c in coeffs]


@pytest.mark.parametrize("body,message", [
    ("4 1\n17 2\nCOEFFS\n1\n2\n3\n", "truncated"),
    ("6 1\n17 2\nCOEFFS\n1\n2\n3\n4\n5\n6\n", "power of two"),
    ("4 0\nCOEFFS\n1\n2\n3\n4\n", "need >= 1 prime"),
    ("4 1\n17 2\nCOEFF\n1\n2\n3\n4\n", "expected COEFFS"),
    ("4 1\n17 2\nCOEFFS\n1\n-2\n3\n4\n", "bad coefficient"),
    ("4 1\n17 2\nCOEFFS\n1\n0x2\n3\n4\n", "bad coefficient"),
])
def test_vector_file_rejections(tmp_path, body, message):
    path = tmp_path / "bad.txt"
    path.write_text(body)
    r = _run(TOOL, "stage", path)
    assert r.returncode == 1 and message in r.stderr, r.stderr


def test_vector_file_small_known_answer(tmp_path):
    """The worked example N=4, q=17, psi=2 (reference docs/modular-arithmetic-contract.md:200-209)
    staged exactly."""
    path = _vector_file(tmp_path / "kat.txt", 4, [17], [2], [1, 15, 3, 15])
    r = _run(TOOL, "stage", path)
    assert [int(x) for x in r.stdout.split()] == [1, 15, 3, 15]


def test_percentiles_match_numpy():
    rng = np.random.default_rng(5)
    v = rng.exponential(2.0, 37)
    for f in (0.0, 0.05, 0.5, 0.95, 1.0):
        r = _run(TOOL, "percentile", f, *[repr(float(x)) for x in v])
        assert r.returncode == 0
        assert float(r.stdout) == pytest.approx(np.percentile(v, 100 * f), rel=1e-12, abs=1e-15)


def test_json_escaping_round_trips():
    text = 'MI355X "gfx950" \\ tab\tend\x01'
    r = _run(TOOL, "json", text)
    assert json.loads(r.stdout) == text


def test_sample_and_summary_lines():
    v = [1.5, 0.25, 3.0, 2.0]
    r = _run(TOOL, "samples", "gpu_forward", *v)
    lines = [json.loads(x) for x in r.stdout.splitlines()]
    assert [x["milliseconds"] for x in lines[:4]] == v
    assert [x["iteration"] for x in lines[:4]] == [0, 1, 2, 3]
    s = lines[4]
    assert s["type"] == "summary" and s["operation"] == "gpu_forward" and s["count"] == 4
    assert s["mean_ms"] == pytest.approx(np.mean(v)) and s["median_ms"] == pytest.approx(np.median(v))
    assert s["p5_ms"] == pytest.approx(np.percentile(v, 5)) and s["p95_ms"] == pytest.approx(np.percentile(v, 95))


def _operands(n, q, seed=0x7251D9):
    """Left and right operands drawn alternately from one MMIX LCG stream
    (reference include/ntt/benchmark_inputs.hpp:9-22)."""
    mask, s = (1 << 64) - 1, seed
    left, right = [], []
    for i in range(n * len(q)):
        s = (s * 6364136223846793005 + 1442695040888963407) & mask
        left.append(s % q[i // n])
        s = (s * 6364136223846793005 + 1442695040888963407) & mask
        right.append(s % q[i // n])
    return left, right


def test_benchmark_operand_stream():
    q, _ = _params(8, 3)
    r = _run(TOOL, "inputs", 8, *q)
    pairs = [tuple(int(x) for x in line.split()) for line in r.stdout.splitlines()]
    left, right = _operands(8, q)
    assert pairs == list(zip(left, right))


def test_cli_usage_and_preflight(tmp_path):
    """Usage errors exit 2; malformed counts and parameters are rejected before
    any device call (exit 1 with the contract message), so this runs anywhere."""
    assert _run(CLI).returncode == 2
    assert _run(CLI, "verify").returncode == 2
    n = 16
    q, psi = _params(n, 2)
    good = _vector_file(tmp_path / "p.txt", n, q, psi, [0] * n)
    r = _run(CLI, "bench-jsonl", good, 0, 10)
    assert r.returncode == 1 and "warmups" in r.stderr
    r = _run(CLI, "gpu-bench-jsonl", good, 1, 1)
    assert r.returncode == 1 and "samples" in r.stderr
    r = _run(CLI, "events-jsonl", good, 1)
    assert r.returncode == 1 and "usage" in r.stderr
    bad = _vector_file(tmp_path / "np.txt", n, [q[0] + 2 * n], [psi[0]], [0] * n)  # not prime or wrong class
    r = _run(CLI, bad, "transform")
    assert r.returncode == 1 and "requires" in r.stderr, r.stderr
    r = _run(CLI, tmp_path / "missing.txt", "identity")
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = _run(CLI, good, "frobnicate")
    assert r.returncode == 2


def test_conversion_bench_preflight(tmp_path):
    """The resident conversion benchmark rejects bad counts and moduli before any device call."""
    n = 16
    q, psi = _params(n, 2)
    good = _vector_file(tmp_path / "p.txt", n, q, psi, [0] * n)
    r = _run(CONVERSION, good, 0, 10)
    assert r.returncode == 1 and "warmups" in r.stderr
    q61, psi61 = _params(n, 1, 61)
    wide = _vector_file(tmp_path / "w.txt", n, q61, psi61, [0] * n)
    r = _run(CONVERSION, wide, 1, 2)
    assert r.returncode == 1 and "below 2^60" in r.stderr
    assert _run(CONVERSION, good).returncode == 1


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("n,count,bits", [(4, 1, 0), (1024, 3, 46), (65536, 2, 60)])
def test_cli_transform_matches_oracle(tmp_path, n, count, bits):
    if bits == 0:
        q, psi, coeffs = [17], [2], [1, 15, 3, 15]
    else:
        q, psi = _params(n, count, bits)
        coeffs = _big_coeffs(n, q, n + count)
    log_case(f"cli transform n={n} limbs={len(q)}")
    path = _vector_file(tmp_path / "v.txt", n, q, psi, coeffs)
    r = _run(CLI, path, "transform")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    for k, m in enumerate(q):
        block = lines[k * (n + 1):(k + 1) * (n + 1)]
        assert block[0] == f"LIMB {k}"
        want = o.ntt_forward(np.array([c % m for c in coeffs], dtype=np.uint64), m, psi[k])
        assert [int(x) for x in block[1:]] == [int(x) for x in want], f"limb {k}"
    if bits == 0:
        assert lines[1:5] == ["10", "16", "3", "9"]  # docs/modular-arithmetic-contract.md:200-209
    r = _run(CLI, path, "identity")
    assert r.returncode == 0 and r.stdout == "IDENTITY OK\n", r.stderr


@pytest.mark.gpu
def test_cli_verify(tmp_path):
    log_case("cli verify (files + generated n=65536)")
    files = []
    for n, count in ((64, 2), (2048, 3)):
        q, psi = _params(n, count)
        files.append(_vector_file(tmp_path / f"v{n}.txt", n, q, psi, _big_coeffs(n, q, n)))
    r = _run(CLI, "verify", *files)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines()[-1] == "VERIFIED OK"
    assert sum(line.startswith("VERIFIED: ") for line in r.stdout.splitlines()) == 3


@pytest.mark.gpu
@pytest.mark.parametrize("mode,operations", [("bench-jsonl", 6), ("gpu-bench-jsonl", 5), ("events-jsonl", 6)])
def test_cli_jsonl_modes(tmp_path, mode, operations):
    n = 4096
    q, psi = _params(n, 3, 60)
    log_case(f"cli {mode} n={n} limbs=3")
    path = _vector_file(tmp_path / "p.txt", n, q, psi, [0] * n)
    r = _run(CLI, mode, path, 1, 4)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(x) for x in r.stdout.splitlines()]
    head, tail = lines[0], lines[-1]
    assert head["type"] == "workload" and head["n"] == n and head["towers"] == 3 and head["moduli"] == q
    assert head["roots"] == psi and head["seed"] == 0x7251D9 and head["samples"] == 4
    assert tail == {"type": "complete", "operations": operations}
    summaries = [x for x in lines if x["type"] == "summary"]
    assert len(summaries) == operations and all(s["count"] == 4 and s["median_ms"] > 0 for s in summaries)
    assert sum(x["type"] == "sample" for x in lines) == 4 * operations
    names = [s["operation"] for s in summaries]
    assert names[-1] in ("gpu_fused_multiply", "d2h_pinned") and "gpu_forward" in names


@pytest.mark.gpu
def test_cli_bench_table():
    log_case("cli bench 10 11 2")
    r = _run(CLI, "bench", 10, 11, 2)
    assert r.returncode == 0, r.stderr
    rows = [line for line in r.stdout.splitlines() if line.startswith("| 2^")]
    assert [row.split()[1] for row in rows] == ["2^10", "2^11"]
    assert all(float(c) > 0 for row in rows for c in row.strip("|").split("|")[2:])


@pytest.mark.gpu
def test_conversion_bench_resident_pair(tmp_path):
    """The reference's resident conversion benchmark on the HIP Session: ModUp,
    ModDown, the pair and the host-boundary pair, every sample checked inside
    the tool against the CPU reference (exit status 1 on a mismatch)."""
    n = 4096
    q, psi = _params(n, 5, 60)
    log_case(f"session conversion bench n={n} towers=5")
    path = _vector_file(tmp_path / "p.txt", n, q, psi, [0] * n)
    r = _run(CONVERSION, path, 1, 3)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(x) for x in r.stdout.splitlines()]
    assert lines[0]["type"] == "workload" and lines[0]["towers"] == 5 and lines[0]["p_towers"] == 3
    names = [x["operation"] for x in lines if x["type"] == "summary"]
    assert names == ["resident_modup_coefficient", "resident_modup_evaluation", "resident_moddown_evaluation",
                     "resident_modup_moddown", "resident_conversion_host_boundary"]
    assert lines[-1] == {"type": "complete", "operations": 5}
