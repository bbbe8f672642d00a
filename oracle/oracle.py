"""ctypes view of the CPU oracle library (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker. Nothing in the product path imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libntt_oracle.so")
_lib = None

U64 = ctypes.c_uint64
U64P = ctypes.POINTER(ctypes.c_uint64)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        _lib = ctypes.CDLL(_LIB_PATH)
        for name in ("or_add_mod", "or_sub_mod", "or_mul_mod", "or_pow_mod", "or_shoup_precompute",
                     "or_log2", "or_bit_reverse", "or_find_psi", "or_min_root", "or_last_prime",
                     "or_previous_prime", "or_special_product", "or_mix", "or_barrett_quotient",
                     "or_barrett_multiply", "or_shoup_multiply", "or_automorphism_source",
                     "or_lcg_fill"):
            getattr(_lib, name).restype = U64
        _lib.or_inv_mod.restype = U64
        _lib.or_inv_mod.argtypes = [U64, U64, ctypes.POINTER(ctypes.c_int)]
        for name in ("or_add_mod", "or_sub_mod", "or_mul_mod", "or_pow_mod"):
            getattr(_lib, name).argtypes = [U64, U64, U64]
        _lib.or_shoup_precompute.argtypes = [U64, U64]
        _lib.or_shoup_multiply.argtypes = [U64, U64, U64, U64]
        _lib.or_barrett_precompute.argtypes = [U64, U64P]
        _lib.or_barrett_quotient.argtypes = [U64, U64, U64, U64]
        _lib.or_barrett_multiply.argtypes = [U64, U64, U64, U64, U64]
        _lib.or_log2.argtypes = [U64]
        _lib.or_bit_reverse.argtypes = [U64, U64]
        _lib.or_is_prime.argtypes = [U64]
        _lib.or_find_primes.argtypes = [U64, U64, U64, U64P]
        _lib.or_find_psi.argtypes = [U64, U64]
        _lib.or_min_root.argtypes = [U64, U64]
        _lib.or_last_prime.argtypes = [U64, U64]
        _lib.or_previous_prime.argtypes = [U64, U64]
        _lib.or_automorphism_source.argtypes = [U64, U64, ctypes.c_uint32]
        _lib.or_mix.argtypes = [U64]
        _lib.or_special_product.argtypes = [U64P, U64, U64]
        _lib.or_lcg_fill.argtypes = [U64, U64P, U64, U64, U64P]
        _lib.or_set_threads.argtypes = [ctypes.c_int]
        _lib.or_pointwise_towers.argtypes = [U64P, U64P, U64P, U64, U64, U64P]
        _lib.or_ntt_forward.argtypes = [U64P, U64, U64, U64]
        _lib.or_ntt_inverse.argtypes = [U64P, U64, U64, U64]
        _lib.or_ntt_table.argtypes = [U64P, U64, U64, U64, ctypes.c_int]
        _lib.or_ntt_towers.argtypes = [U64P, U64, U64, U64P, U64P, ctypes.c_int]
        _lib.or_ntt_naive.argtypes = [U64P, U64P, U64, U64, U64]
        _lib.or_twiddles.argtypes = [U64, U64, U64, U64P]
        _lib.or_basis_switch.argtypes = [U64P, U64, U64P, U64, U64P, U64, U64P]
        _lib.or_mod_up.argtypes = [U64P, ctypes.c_int, U64, U64P, U64P, U64, U64P, U64P, U64, U64P]
        _lib.or_mod_down.argtypes = [U64P, U64, U64P, U64P, U64, U64P, U64P, U64, U64P]
        _lib.or_hybrid_digits.argtypes = [U64P, U64, U64P, U64P, U64, U64P, U64P, U64, U64, U64P]
        _lib.or_hybrid_mac.argtypes = [U64P, U64, U64P, U64P, U64, U64P, U64, U64P, U64, U64, U64P, U64P]
        _lib.or_hybrid_rotate.argtypes = [U64P, U64P, U64P, U64, U64P, U64, U64P, U64, ctypes.c_uint32,
                                          U64P, U64P]
        _lib.or_key_switch.argtypes = [U64P, U64, U64P, U64P, U64, U64, U64P, U64P, U64, U64, U64P, U64P,
                                       U64P, U64P]
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U64P)


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def add_mod(a, b, q): return lib().or_add_mod(a, b, q)
def sub_mod(a, b, q): return lib().or_sub_mod(a, b, q)
def mul_mod(a, b, q): return lib().or_mul_mod(a, b, q)
def pow_mod(a, e, q): return lib().or_pow_mod(a, e, q)


def inv_mod(x, q):
    ok = ctypes.c_int(0)
    v = lib().or_inv_mod(x, q, ctypes.byref(ok))
    if not ok.value:
        raise ValueError("no inverse")
    return v


def shoup_precompute(w, q): return lib().or_shoup_precompute(w, q)
def shoup_multiply(x, w, wp, q): return lib().or_shoup_multiply(x, w, wp, q)


def barrett_precompute(q):
    out = np.zeros(2, dtype=np.uint64)
    lib().or_barrett_precompute(q, _p(out))
    return int(out[0]), int(out[1])


def barrett_multiply(a, b, q):
    lo, hi = barrett_precompute(q)
    return lib().or_barrett_multiply(a, b, q, lo, hi)


def barrett_quotient(low, high, q):
    lo, hi = barrett_precompute(q)
    return lib().or_barrett_quotient(low, high, lo, hi)


def bit_reverse(x, lg): return lib().or_bit_reverse(x, lg)
def log2(n): return lib().or_log2(n)
def is_prime(n): return bool(lib().or_is_prime(n))


def find_primes(n, count, bits=46):
    out = np.zeros(count, dtype=np.uint64)
    got = lib().or_find_primes(n, count, bits, _p(out))
    if got < count:
        raise RuntimeError("prime search exhausted the bit range")
    return [int(v) for v in out]


def find_psi(q, n): return lib().or_find_psi(q, n)
def min_root(q, n): return lib().or_min_root(q, n)
def last_prime(bits, m): return lib().or_last_prime(bits, m)
def previous_prime(q, m): return lib().or_previous_prime(q, m)
def mix(x): return lib().or_mix(x)
def automorphism_source(dest, n, k): return lib().or_automorphism_source(dest, n, k)


def special_product(p, q):
    pa = _arr(p)
    return lib().or_special_product(_p(pa), len(pa), q)


def ntt_forward(values, q, psi):
    """Reference DIF forward (include/ntt/ntt.hpp:96-121)."""
    a = _arr(values).copy()
    if lib().or_ntt_forward(_p(a), len(a), q, psi):
        raise ValueError("invalid transform shape")
    return a


def ntt_inverse(values, q, psi):
    a = _arr(values).copy()
    if lib().or_ntt_inverse(_p(a), len(a), q, psi):
        raise ValueError("invalid transform shape")
    return a


def ntt_table(values, q, psi, inverse=False):
    a = _arr(values).copy()
    if lib().or_ntt_table(_p(a), len(a), q, psi, int(inverse)):
        raise ValueError("invalid transform shape")
    return a


def ntt_towers(values, n, q, psi, inverse=False):
    """Tower-major batch transform (group-twiddle restatement, same outputs)."""
    a = _arr(values).copy()
    qa, pa = _arr(q), _arr(psi)
    assert a.size == n * len(qa)
    if lib().or_ntt_towers(_p(a), n, len(qa), _p(qa), _p(pa), int(inverse)):
        raise ValueError("invalid transform shape")
    return a


def ntt_naive(values, q, psi):
    a = _arr(values)
    out = np.zeros_like(a)
    lib().or_ntt_naive(_p(a), _p(out), len(a), q, psi)
    return out


def twiddles(n, root, q):
    out = np.zeros(n + 1, dtype=np.uint64)
    lib().or_twiddles(n, root, q, _p(out))
    return out


def basis_switch(values, n, src, dst):
    a, s, d = _arr(values), _arr(src), _arr(dst)
    out = np.zeros(n * len(d), dtype=np.uint64)
    err = lib().or_basis_switch(_p(a), n, _p(s), len(s), _p(d), len(d), _p(out))
    if err:
        raise ValueError(f"basis switch rejected ({err})")
    return out


def mod_up(values, n, q, qpsi, p, ppsi, evaluation=True):
    a, qa, qr, pa, pr = map(_arr, (values, q, qpsi, p, ppsi))
    out = np.zeros(n * (len(qa) + len(pa)), dtype=np.uint64)
    if lib().or_mod_up(_p(a), int(evaluation), n, _p(qa), _p(qr), len(qa), _p(pa), _p(pr), len(pa), _p(out)):
        raise ValueError("mod_up failed")
    return out


def mod_down(values, n, q, qpsi, p, ppsi):
    a, qa, qr, pa, pr = map(_arr, (values, q, qpsi, p, ppsi))
    out = np.zeros(n * len(qa), dtype=np.uint64)
    if lib().or_mod_down(_p(a), n, _p(qa), _p(qr), len(qa), _p(pa), _p(pr), len(pa), _p(out)):
        raise ValueError("mod_down failed")
    return out


def hybrid_digits(values, n, q, qpsi, p, ppsi, alpha):
    a, qa, qr, pa, pr = map(_arr, (values, q, qpsi, p, ppsi))
    L, K = len(qa), len(pa)
    beta = (L + alpha - 1) // alpha
    out = np.zeros(beta * (L + K) * n, dtype=np.uint64)
    if lib().or_hybrid_digits(_p(a), n, _p(qa), _p(qr), L, _p(pa), _p(pr), K, alpha, _p(out)):
        raise ValueError("hybrid_digits failed")
    return out


def hybrid_mac(digits, beta, key_b, key_a, n, q, p, M):
    d, kb, ka, qa, pa = map(_arr, (digits, key_b, key_a, q, p))
    L, K = len(qa), len(pa)
    ob = np.zeros((L + K) * n, dtype=np.uint64)
    oa = np.zeros_like(ob)
    lib().or_hybrid_mac(_p(d), beta, _p(kb), _p(ka), n, _p(qa), L, _p(pa), K, M, _p(ob), _p(oa))
    return ob, oa


def hybrid_rotate(ext_b, ext_a, c0, n, q, p, k):
    b, a, c, qa, pa = map(_arr, (ext_b, ext_a, c0, q, p))
    L, K = len(qa), len(pa)
    ob = np.zeros((L + K) * n, dtype=np.uint64)
    oa = np.zeros_like(ob)
    lib().or_hybrid_rotate(_p(b), _p(a), _p(c), n, _p(qa), L, _p(pa), K, ctypes.c_uint32(k), _p(ob), _p(oa))
    return ob, oa


def key_switch(values, n, q, qpsi, M, p, ppsi, alpha, key_b, key_a):
    """KeySwitchHybrid core: q/qpsi are the ACTIVE Q towers (L), M the full Q count."""
    a, qa, qr, pa, pr, kb, ka = map(_arr, (values, q, qpsi, p, ppsi, key_b, key_a))
    L, K = len(qa), len(pa)
    ob = np.zeros(L * n, dtype=np.uint64)
    oa = np.zeros_like(ob)
    err = lib().or_key_switch(_p(a), n, _p(qa), _p(qr), L, M, _p(pa), _p(pr), K, alpha,
                              _p(kb), _p(ka), _p(ob), _p(oa))
    if err:
        raise ValueError("key_switch failed")
    return ob, oa


def set_threads(t):
    lib().or_set_threads(int(t))


def max_threads():
    return lib().or_max_threads()


def pointwise_towers(a, b, n, q):
    a, b, qa = _arr(a), _arr(b), _arr(q)
    out = np.zeros_like(a)
    lib().or_pointwise_towers(_p(a), _p(b), _p(out), n, len(qa), _p(qa))
    return out


def lcg_fill(state, n, q):
    qa = _arr(q)
    out = np.zeros(n * len(qa), dtype=np.uint64)
    new_state = lib().or_lcg_fill(state, _p(out), n, len(qa), _p(qa))
    return out, new_state
