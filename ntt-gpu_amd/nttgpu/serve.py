"""ctypes binding of the continuous-batching server (include/ntt_serve.h,
implemented in libntt_host.so on top of the gfx950 kernels)."""
import ctypes
import os

import numpy as np

from . import hip

LIB_PATH = os.path.join(os.path.dirname(hip.LIB_PATH), "libntt_host.so")
KEYSWITCH, FORWARD, INVERSE, MODUP, MODDOWN = 0, 1, 2, 3, 4
OP_NAMES = {KEYSWITCH: "keyswitch", FORWARD: "ntt", INVERSE: "intt", MODUP: "modup", MODDOWN: "moddown"}

U64P, SZ, VP, INT = hip.U64P, hip.SZ, hip.VP, hip.INT
SIGNATURES = {
    "ntt_serve_last_error": [],
    "ntt_serve_create": [ctypes.POINTER(VP), INT, SZ, SZ, SZ],
    "ntt_serve_destroy": [VP],
    "ntt_serve_add_params": [VP, ctypes.c_uint64, U64P, U64P, SZ, U64P, U64P, SZ, SZ, U64P, U64P,
                             ctypes.POINTER(ctypes.c_uint32)],
    "ntt_serve_output_words": [VP, ctypes.c_uint32, INT, SZ, ctypes.POINTER(SZ)],
    "ntt_serve_submit": [VP, ctypes.c_uint32, INT, SZ, U64P, SZ, ctypes.POINTER(ctypes.c_uint64)],
    "ntt_serve_submit_device": [VP, ctypes.c_uint32, INT, SZ, VP, VP, VP, ctypes.POINTER(ctypes.c_uint64)],
    "ntt_serve_ready": [VP, ctypes.c_uint64],
    "ntt_serve_wait": [VP, ctypes.c_uint64, U64P, SZ, U64P, SZ, ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(SZ)],
    "ntt_serve_stats": [VP, ctypes.POINTER(SZ)],
    "ntt_serve_stats_ex": [VP, ctypes.POINTER(SZ), SZ],
}
_lib = None


def lib():
    global _lib
    if _lib is None:
        hip.lib()  # libntt_hip.so first (dependency)
        if not os.path.exists(LIB_PATH):
            raise hip.NttHipError(f"host runtime not built: {LIB_PATH}")
        h = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "ntt_serve_last_error" else INT
        _lib = h
    return _lib


def _check(status, what):
    if status in (hip.OK, hip.NOT_READY):
        return status
    msg = lib().ntt_serve_last_error().decode(errors="replace")
    if status == hip.EINVAL:
        raise hip.ContractError(f"{what}: {msg}")
    raise hip.NttHipError(f"{what}: {msg}")


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(U64P)


class Server:
    def __init__(self, device=0, hbm_budget_bytes=32 << 30, streams=4, max_batch=32):
        h = VP()
        _check(lib().ntt_serve_create(ctypes.byref(h), device, hbm_budget_bytes, streams, max_batch), "serve_create")
        self.handle = h.value
        self.params = {}

    def add_params(self, hp, key_b, key_a):
        q, qp = _u64(hp.q)
        qr, qrp = _u64(hp.q_psi)
        p, pp = _u64(hp.p)
        pr, prp = _u64(hp.p_psi)
        kb, kbp = _u64(key_b)
        ka, kap = _u64(key_a)
        pid = ctypes.c_uint32(0)
        _check(lib().ntt_serve_add_params(self.handle, hp.n, qp, qrp, len(hp.q), pp, prp, len(hp.p), hp.alpha, kbp,
                                          kap, ctypes.byref(pid)), "serve_add_params")
        self.params[pid.value] = hp
        return pid.value

    def output_words(self, pid, op, level):
        w = SZ(0)
        _check(lib().ntt_serve_output_words(self.handle, pid, op, level, ctypes.byref(w)), "serve_output_words")
        return w.value

    def submit(self, pid, op, level, values):
        a, ap = _u64(values)
        t = ctypes.c_uint64(0)
        _check(lib().ntt_serve_submit(self.handle, pid, op, level, ap, a.size, ctypes.byref(t)), "serve_submit")
        return t.value

    def submit_device(self, pid, op, level, in_ptr, out_ptr, out2_ptr=None):
        """Device-resident request: the caller's device buffers (ciphertexts
        stay in HBM); collect with wait_device."""
        t = ctypes.c_uint64(0)
        _check(lib().ntt_serve_submit_device(self.handle, pid, op, level, in_ptr, out_ptr, out2_ptr,
                                             ctypes.byref(t)), "serve_submit_device")
        return t.value

    def wait_device(self, ticket):
        lat = ctypes.c_double(0)
        bs = SZ(0)
        _check(lib().ntt_serve_wait(self.handle, ticket, None, 0, None, 0, ctypes.byref(lat), ctypes.byref(bs)),
               "serve_wait")
        return lat.value, bs.value

    def ready(self, ticket):
        return _check(lib().ntt_serve_ready(self.handle, ticket), "serve_ready") == hip.OK

    def wait(self, ticket, pid, op, level):
        w = self.output_words(pid, op, level)
        out = np.empty(w, dtype=np.uint64)
        out2 = np.empty(w if op == KEYSWITCH else 0, dtype=np.uint64)
        lat = ctypes.c_double(0)
        bs = SZ(0)
        _check(lib().ntt_serve_wait(self.handle, ticket, out.ctypes.data_as(U64P), w,
                                    out2.ctypes.data_as(U64P) if op == KEYSWITCH else None, out2.size,
                                    ctypes.byref(lat), ctypes.byref(bs)), "serve_wait")
        return (out, out2 if op == KEYSWITCH else None), lat.value, bs.value

    def stats(self):
        v = (SZ * 9)()
        _check(lib().ntt_serve_stats_ex(self.handle, v, 9), "serve_stats")
        keys = ("submitted", "completed", "batches", "max_in_flight", "device_bytes", "peak_device_bytes",
                "resident_bytes", "plans_resident", "plans_evicted")
        return dict(zip(keys, list(v)))

    def close(self):
        if getattr(self, "handle", None):
            _check(lib().ntt_serve_destroy(self.handle), "serve_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
