"""The C++ include/ntt runtime, as the reference gates it (ctest binaries):
session_tests runs the Session on the deferred CPU transport; the GPU
binary drives the same API over the gfx950 transport; arith_tests checks the
NTT kernels' wide-lazy modular arithmetic on the host against 128-bit math."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "cpp", "build")


def _run(name, timeout):
    exe = os.path.join(BUILD, name)
    assert os.path.exists(exe), f"{exe} not built (run __graft_entry__.build())"
    print(f"[case] {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_wide_lazy_arithmetic_ranges():
    assert "ARITH CHECKS PASSED" in _run("arith_tests", 120)


def test_resident_runtime_on_host_transport():
    assert "SESSION RUNTIME CHECKS PASSED" in _run("session_tests", 600)


@pytest.mark.gpu
def test_resident_runtime_on_hip_transport():
    assert "HIP RESIDENT CHECKS PASSED" in _run("hip_resident_tests", 600)


def test_openfhe_adapter_on_host_transport():
    """include/ntt/openfhe_hybrid_adapter.hpp with an OpenFHE test double and the
    replay transport: the five entry points against the oracle, commit order
    and the unchanged ciphertext on failure."""
    assert "ADAPTER CHECKS PASSED" in _run("adapter_tests", 600)
