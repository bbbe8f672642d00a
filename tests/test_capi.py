"""CPU checks of the C ABI boundary: the library loads, exports exactly what
include/ntt_hip.h declares, and rejects invalid parameters before any device
work (reference include/ntt/cuda_validation.hpp:17-48,
tests/resident_cuda_preflight.cpp:1-21). No compute call runs here."""
import ctypes
import os
import re
import subprocess

import pytest

from nttgpu import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ntt_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ntt_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(hip.LIB_PATH), "libntt_hip.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (ntt_hip_\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    extra = sorted(exported - set(declared()))
    assert not extra, f"exported but undeclared: {extra}"


def test_binding_covers_the_header():
    assert sorted(hip.SIGNATURES) == declared()
    lib = hip.lib()
    for name in declared():
        assert getattr(lib, name) is not None
    assert lib.ntt_hip_abi_version() == 1


SERVE_HEADER = os.path.join(ROOT, "include", "ntt_serve.h")


def test_serving_library_exports_and_binding_cover_its_header():
    from nttgpu import serve
    text = re.sub(r"/\*.*?\*/", "", open(SERVE_HEADER).read(), flags=re.S)
    want = sorted(set(re.findall(r"\b(ntt_serve_\w+)\s*\(", text)))
    out = subprocess.run(["nm", "-D", "--defined-only", serve.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (ntt_serve_\w+)", out))
    assert [s for s in want if s not in exported] == []
    assert sorted(exported - set(want)) == []
    assert sorted(serve.SIGNATURES) == want


def _create(n, q, psi, device=0):
    h = ctypes.c_void_p()
    qa = (ctypes.c_uint64 * len(q))(*q)
    pa = (ctypes.c_uint64 * len(psi))(*psi)
    return hip.lib().ntt_hip_context_create(ctypes.byref(h), device, n, qa, pa, len(q))


Q65 = 1152921504606584833
PSI65 = 18043022392882


@pytest.mark.parametrize("n,q,psi,why", [
    (3, [Q65], [PSI65], "power-of-two"),
    (131072, [Q65], [PSI65], "power-of-two"),
    (1, [Q65], [PSI65], "power-of-two"),
    (65536, [Q65 + 2 * 65536], [PSI65], "prime"),        # composite or wrong shape
    (65536, [1 << 63], [PSI65], "prime"),
    (65536, [Q65], [PSI65 + 1], "primitive"),
    (65536, [Q65], [0], "primitive"),
    (65536, [Q65], [Q65 + 5], "primitive"),
])
def test_context_preflight_rejects_before_device_work(n, q, psi, why):
    status = _create(n, q, psi)
    assert status == hip.EINVAL
    assert why in hip.lib().ntt_hip_last_error().decode()


def test_negative_device_rejected():
    assert _create(65536, [Q65], [PSI65], device=-1) == hip.EINVAL


def test_null_arguments_rejected():
    lib = hip.lib()
    assert lib.ntt_hip_transform(None, None, None, 0, 1, 0, None) == hip.EINVAL
    assert lib.ntt_hip_keyswitch_run(None, None, 1, None, None, None, None) == hip.EINVAL
    assert lib.ntt_hip_memcpy(0, None, None, 8, hip.H2D, None) == hip.EINVAL
    assert lib.ntt_hip_memcpy(0, ctypes.c_void_p(8), ctypes.c_void_p(8), 8, 7, None) == hip.EINVAL


def test_null_keys_rejected_before_any_use():
    """A null key pointer in the key array is a contract error, not a crash."""
    lib = hip.lib()
    keys = (ctypes.c_void_p * 1)(None)
    one = ctypes.c_void_p(8)
    higher = (ctypes.c_void_p * 1)(ctypes.c_void_p(24))
    assert lib.ntt_hip_relinearize(keys, 1, ctypes.c_void_p(8), ctypes.c_void_p(16), higher, 1, one,
                                   None) == hip.EINVAL
    assert "null" in lib.ntt_hip_last_error().decode()
    autos = (ctypes.c_uint32 * 1)(3)
    outs = (ctypes.c_void_p * 1)(ctypes.c_void_p(32))
    assert lib.ntt_hip_keyswitch_rotate(keys, 1, autos, ctypes.c_void_p(8), ctypes.c_void_p(16), 1, outs, outs,
                                        one, None) == hip.EINVAL
    assert "null" in lib.ntt_hip_last_error().decode()
