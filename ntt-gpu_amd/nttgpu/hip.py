"""ctypes binding of libntt_hip.so (include/ntt_hip.h).

The product path: every call goes to the hand-written gfx950 kernels. There is
no CPU fallback — if the library is missing or no GPU is visible, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NTT_HIP_LIBRARY selects another build of the library (for example a
# different tuning compiled elsewhere); the default is the in-tree build.
LIB_PATH = os.environ.get("NTT_HIP_LIBRARY") or os.path.join(os.path.dirname(_HERE), "libntt_hip.so")

OK, NOT_READY, EINVAL, ERUNTIME = 0, 1, -1, -2
H2D, D2H, D2D = 1, 2, 3

U64 = ctypes.c_uint64
U64P = ctypes.POINTER(ctypes.c_uint64)
SZ = ctypes.c_size_t
SZP = ctypes.POINTER(ctypes.c_size_t)
VP = ctypes.c_void_p
INT = ctypes.c_int

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "ntt_hip_last_error": [],
    "ntt_hip_abi_version": [],
    "ntt_hip_device_count": [ctypes.POINTER(INT)],
    "ntt_hip_device_name": [INT, ctypes.c_char_p, SZ],
    "ntt_hip_device_properties": [INT, ctypes.c_char_p, SZ, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(INT),
                                  ctypes.POINTER(INT), ctypes.POINTER(INT)],
    "ntt_hip_malloc": [INT, ctypes.POINTER(VP), SZ],
    "ntt_hip_free": [INT, VP],
    "ntt_hip_host_alloc": [ctypes.POINTER(VP), SZ],
    "ntt_hip_host_free": [VP],
    "ntt_hip_memcpy": [INT, VP, VP, SZ, INT, VP],
    "ntt_hip_stream_create": [INT, ctypes.POINTER(VP)],
    "ntt_hip_stream_destroy": [INT, VP],
    "ntt_hip_stream_synchronize": [INT, VP],
    "ntt_hip_event_create": [INT, ctypes.POINTER(VP), INT],
    "ntt_hip_event_destroy": [INT, VP],
    "ntt_hip_event_record": [INT, VP, VP],
    "ntt_hip_event_query": [INT, VP],
    "ntt_hip_event_synchronize": [INT, VP],
    "ntt_hip_event_elapsed": [ctypes.POINTER(ctypes.c_float), VP, VP],
    "ntt_hip_copy_words": [INT, VP, VP, SZ, VP],
    "ntt_hip_context_create": [ctypes.POINTER(VP), INT, U64, U64P, U64P, SZ],
    "ntt_hip_context_destroy": [VP],
    "ntt_hip_context_stream": [VP],
    "ntt_hip_context_device": [VP],
    "ntt_hip_context_table_bytes": [VP],
    "ntt_hip_autotune": [VP, SZ, ctypes.c_char_p, SZ],
    "ntt_hip_set_tuning": [VP, ctypes.c_uint32, ctypes.c_uint32],
    "ntt_hip_get_tuning": [VP, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)],
    "ntt_hip_transform": [VP, VP, VP, SZ, SZ, INT, VP],
    "ntt_hip_transform_batch": [VP, VP, VP, SZ, SZ, SZ, SZ, SZ, INT, VP],
    "ntt_hip_pointwise": [VP, VP, VP, SZ, SZ, VP],
    "ntt_hip_switch_basis": [VP, VP, VP, SZP, SZ, SZP, SZ, VP],
    "ntt_hip_correct": [VP, VP, VP, VP, SZ, SZ, SZP, SZ, VP],
    "ntt_hip_hybrid_mac": [VP, VP, VP, VP, VP, VP, SZ, SZ, INT, VP],
    "ntt_hip_hybrid_mac_digits": [VP, VP, VP, ctypes.POINTER(VP), ctypes.POINTER(VP), ctypes.POINTER(VP), SZ, SZ, SZ,
                                  INT, VP],
    "ntt_hip_hybrid_rotate": [VP, VP, VP, VP, VP, VP, SZ, SZ, SZ, SZP, SZ, ctypes.c_uint32, VP],
    "ntt_hip_keyswitch_create": [ctypes.POINTER(VP), VP, SZ, SZ, SZ, SZ],
    "ntt_hip_keyswitch_destroy": [VP],
    "ntt_hip_keyswitch_upload_key": [VP, U64P, U64P, SZ, U64P],
    "ntt_hip_keyswitch_workspace_bytes": [VP, SZ, SZP],
    "ntt_hip_keyswitch_run": [VP, VP, SZ, VP, VP, VP, VP],
    "ntt_hip_keyswitch_run_static": [VP, VP, SZ, VP, VP, VP],
    "ntt_hip_set_ntt_shape": [VP, INT, ctypes.c_uint32, ctypes.c_uint32],
    "ntt_hip_context_device_bytes": [VP, ctypes.POINTER(SZ)],
    "ntt_hip_keyswitch_autotune": [VP, SZ, ctypes.c_char_p, SZ],
    "ntt_hip_keyswitch_set_shape": [VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32],
    "ntt_hip_keyswitch_get_shape": [VP, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(INT)],
    "ntt_hip_keyswitch_device_bytes": [VP, ctypes.POINTER(SZ)],
    "ntt_hip_transform_jobs": [VP, VP, VP, VP, SZ, VP, INT, VP],
    "ntt_hip_transform_mixed_workspace_bytes": [SZ, ctypes.POINTER(SZ)],
    "ntt_hip_keyswitch_gather_workspace_bytes": [VP, SZ, ctypes.POINTER(SZ)],
    "ntt_hip_keyswitch_run_gather": [VP, VP, SZ, VP, VP, VP, VP],
    "ntt_hip_transform_mixed": [VP, SZ, VP, INT, VP],
    "ntt_hip_get_ntt_shape": [VP, ctypes.POINTER(INT), ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)],
    "ntt_hip_keyswitch_capture": [VP, VP, SZ, VP, VP, VP, VP, ctypes.POINTER(VP)],
    "ntt_hip_graph_launch": [VP, VP],
    "ntt_hip_graph_destroy": [VP],
    "ntt_hip_keyswitch_in_place": [VP, VP, VP, VP, SZ, VP, VP],
    "ntt_hip_relinearize": [ctypes.POINTER(VP), SZ, VP, VP, ctypes.POINTER(VP), SZ, VP, VP],
    "ntt_hip_rotate_workspace_bytes": [VP, SZ, SZP],
    "ntt_hip_keyswitch_rotate": [ctypes.POINTER(VP), SZ, ctypes.POINTER(ctypes.c_uint32), VP, VP, SZ,
                                 ctypes.POINTER(VP), ctypes.POINTER(VP), VP, VP],
    "ntt_hip_modup_workspace_bytes": [VP, SZ, SZP],
    "ntt_hip_modup_run": [VP, VP, INT, SZ, VP, VP, VP],
    "ntt_hip_moddown_workspace_bytes": [VP, SZ, SZP],
    "ntt_hip_moddown_run": [VP, VP, SZ, VP, VP, VP],
}
_RESTYPES = {
    "ntt_hip_last_error": ctypes.c_char_p,
    "ntt_hip_context_stream": VP,
    "ntt_hip_context_table_bytes": SZ,
}

_lib = None


class NttHipError(RuntimeError):
    pass


class ContractError(NttHipError, ValueError):
    """Raised for NTT_HIP_EINVAL: a contract violation caught before device work."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NttHipError(f"HIP extension not built: {LIB_PATH} (run __graft_entry__.build())")
        handle = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, INT)
        _lib = handle
    return _lib


def check(status, what=""):
    if status in (OK, NOT_READY):
        return status
    msg = lib().ntt_hip_last_error().decode(errors="replace")
    if status == EINVAL:
        raise ContractError(f"{what}: {msg}")
    raise NttHipError(f"{what}: {msg}")


def device_count():
    n = INT(0)
    check(lib().ntt_hip_device_count(ctypes.byref(n)), "device_count")
    return n.value


def device_name(device=0):
    buf = ctypes.create_string_buffer(256)
    check(lib().ntt_hip_device_name(device, buf, 256), "device_name")
    return buf.value.decode()


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(U64P)


def _sizes(v):
    arr = (ctypes.c_size_t * len(v))(*[int(x) for x in v])
    return arr


class Stream:
    def __init__(self, device=0, handle=None):
        self.device = device
        self.owned = handle is None
        if handle is None:
            h = VP()
            check(lib().ntt_hip_stream_create(device, ctypes.byref(h)), "stream_create")
            handle = h.value
        self.handle = handle

    def synchronize(self):
        check(lib().ntt_hip_stream_synchronize(self.device, self.handle), "stream_synchronize")

    def close(self):
        if self.owned and self.handle:
            check(lib().ntt_hip_stream_destroy(self.device, self.handle), "stream_destroy")
            self.handle = None


class Event:
    def __init__(self, device=0, timing=True):
        self.device = device
        h = VP()
        check(lib().ntt_hip_event_create(device, ctypes.byref(h), int(timing)), "event_create")
        self.handle = h.value

    def record(self, stream):
        check(lib().ntt_hip_event_record(self.device, self.handle, stream), "event_record")

    def synchronize(self):
        check(lib().ntt_hip_event_synchronize(self.device, self.handle), "event_synchronize")

    def query(self):
        return check(lib().ntt_hip_event_query(self.device, self.handle), "event_query") == OK

    def elapsed_ms(self, later):
        ms = ctypes.c_float(0)
        check(lib().ntt_hip_event_elapsed(ctypes.byref(ms), self.handle, later.handle), "event_elapsed")
        return ms.value

    def close(self):
        if self.handle:
            check(lib().ntt_hip_event_destroy(self.device, self.handle), "event_destroy")
            self.handle = None


class DeviceBuffer:
    """Raw device allocation of uint64 words."""

    def __init__(self, words, device=0):
        self.device = device
        self.words = int(words)
        p = VP()
        check(lib().ntt_hip_malloc(device, ctypes.byref(p), max(1, self.words) * 8), "malloc")
        self.ptr = p.value

    def at(self, word_offset=0):
        return self.ptr + 8 * int(word_offset)

    def upload(self, values, stream=None, offset=0):
        a, _ = _u64(values)
        assert offset + a.size <= self.words
        check(lib().ntt_hip_memcpy(self.device, self.at(offset), a.ctypes.data, a.size * 8, H2D, stream), "upload")
        if stream is not None:
            Stream(self.device, stream).synchronize()
        else:
            check(lib().ntt_hip_stream_synchronize(self.device, None), "sync")
        return self

    def download(self, count=None, offset=0, stream=None):
        count = self.words - offset if count is None else int(count)
        out = np.empty(count, dtype=np.uint64)
        check(lib().ntt_hip_memcpy(self.device, out.ctypes.data, self.at(offset), count * 8, D2H, stream), "download")
        if stream is not None:
            Stream(self.device, stream).synchronize()
        else:
            check(lib().ntt_hip_stream_synchronize(self.device, None), "sync")
        return out

    def free(self):
        if self.ptr:
            check(lib().ntt_hip_free(self.device, self.ptr), "free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """A fixed basis of towers (q, psi) at ring dimension n with device tables."""

    def __init__(self, n, q, psi, device=0):
        self.n, self.q, self.psi, self.device = int(n), [int(x) for x in q], [int(x) for x in psi], device
        qa, qp = _u64(self.q)
        pa, pp = _u64(self.psi)
        h = VP()
        check(lib().ntt_hip_context_create(ctypes.byref(h), device, self.n, qp, pp, len(self.q)), "context_create")
        self.handle = h.value
        self.stream = lib().ntt_hip_context_stream(self.handle)

    @property
    def towers(self):
        return len(self.q)

    def table_bytes(self):
        return lib().ntt_hip_context_table_bytes(self.handle)

    def synchronize(self, stream=None):
        check(lib().ntt_hip_stream_synchronize(self.device, stream or self.stream), "sync")

    def transform(self, data_ptr, first, towers, inverse, src_ptr=None, stream=None):
        check(lib().ntt_hip_transform(self.handle, data_ptr, src_ptr, first, towers, int(inverse), stream),
              "transform")

    def transform_batch(self, data_ptr, polys, poly_stride, first, towers, inverse, src_ptr=None,
                        src_poly_stride=0, stream=None):
        check(lib().ntt_hip_transform_batch(self.handle, data_ptr, src_ptr, polys, poly_stride, src_poly_stride,
                                            first, towers, int(inverse), stream), "transform_batch")

    def pointwise(self, out_ptr, right_ptr, first, towers, stream=None):
        check(lib().ntt_hip_pointwise(self.handle, out_ptr, right_ptr, first, towers, stream), "pointwise")

    def switch_basis(self, out_ptr, in_ptr, src, dst, stream=None):
        check(lib().ntt_hip_switch_basis(self.handle, out_ptr, in_ptr, _sizes(src), len(src), _sizes(dst),
                                         len(dst), stream), "switch_basis")

    def correct(self, out_ptr, left_ptr, right_ptr, first, towers, p, stream=None):
        check(lib().ntt_hip_correct(self.handle, out_ptr, left_ptr, right_ptr, first, towers, _sizes(p), len(p),
                                    stream), "correct")

    def hybrid_mac(self, b_ptr, a_ptr, digit_ptr, kb_ptr, ka_ptr, first, towers, first_digit, stream=None):
        check(lib().ntt_hip_hybrid_mac(self.handle, b_ptr, a_ptr, digit_ptr, kb_ptr, ka_ptr, first, towers,
                                       int(first_digit), stream), "hybrid_mac")

    def hybrid_mac_digits(self, b_ptr, a_ptr, digit_ptrs, kb_ptrs, ka_ptrs, first, towers, first_digit,
                          stream=None):
        count = len(digit_ptrs)
        arr = lambda ptrs: (VP * max(1, count))(*ptrs)
        check(lib().ntt_hip_hybrid_mac_digits(self.handle, b_ptr, a_ptr, arr(digit_ptrs), arr(kb_ptrs), arr(ka_ptrs),
                                              count, first, towers, int(first_digit), stream), "hybrid_mac_digits")

    def hybrid_rotate(self, ob_ptr, oa_ptr, eb_ptr, ea_ptr, c0_ptr, first, towers, active_q, p, automorphism,
                      stream=None):
        check(lib().ntt_hip_hybrid_rotate(self.handle, ob_ptr, oa_ptr, eb_ptr, ea_ptr, c0_ptr, first, towers,
                                          active_q, _sizes(p), len(p), automorphism, stream), "hybrid_rotate")

    def close(self):
        if getattr(self, "handle", None):
            check(lib().ntt_hip_context_destroy(self.handle), "context_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """An instantiated HIP graph (captured key-switch batch)."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream):
        check(lib().ntt_hip_graph_launch(self.handle, stream), "graph_launch")

    def close(self):
        if self.handle:
            check(lib().ntt_hip_graph_destroy(self.handle), "graph_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KeySwitch:
    """Fused batched Hybrid key switch on a context whose basis is Q (L) then P (K)."""

    def __init__(self, ctx, L, K, alpha, max_batch=1):
        self.ctx, self.L, self.K, self.alpha, self.max_batch = ctx, L, K, alpha, max_batch
        self.beta = (L + alpha - 1) // alpha
        h = VP()
        check(lib().ntt_hip_keyswitch_create(ctypes.byref(h), ctx.handle, L, K, alpha, max_batch), "keyswitch_create")
        self.handle = h.value

    def upload_key(self, key_b, key_a, M, q_full):
        b, bp = _u64(key_b)
        a, ap = _u64(key_a)
        qf, qp = _u64(q_full)
        check(lib().ntt_hip_keyswitch_upload_key(self.handle, bp, ap, M, qp), "keyswitch_upload_key")

    def workspace_words(self, batch):
        out = SZ(0)
        check(lib().ntt_hip_keyswitch_workspace_bytes(self.handle, batch, ctypes.byref(out)), "workspace_bytes")
        return out.value // 8

    def run(self, in_ptr, batch, out_b_ptr, out_a_ptr, workspace_ptr, stream=None):
        check(lib().ntt_hip_keyswitch_run(self.handle, in_ptr, batch, out_b_ptr, out_a_ptr, workspace_ptr, stream),
              "keyswitch_run")

    def gather_workspace_words(self, batch):
        out = SZ(0)
        check(lib().ntt_hip_keyswitch_gather_workspace_bytes(self.handle, batch, ctypes.byref(out)),
              "gather_workspace_bytes")
        return out.value // 8

    def run_gather(self, in_ptrs, out_b_ptrs, out_a_ptrs, workspace_ptr, stream=None):
        """Key-switch len(in_ptrs) components, each at its own device addresses
        (lists of device pointers as ints); equal to run() on the same
        components gathered into contiguous buffers."""
        b = len(in_ptrs)
        if len(out_b_ptrs) != b or len(out_a_ptrs) != b:
            raise ValueError("run_gather: one output pair per input component")
        arr = lambda ps: (ctypes.c_void_p * max(b, 1))(*[ctypes.c_void_p(int(p)) for p in ps])
        check(lib().ntt_hip_keyswitch_run_gather(self.handle, arr(in_ptrs), b, arr(out_b_ptrs), arr(out_a_ptrs),
                                                 workspace_ptr, stream), "keyswitch_run_gather")

    def autotune(self, batch):
        """Tune the plan and its context on `batch` key switches; returns the report."""
        import json
        buf = ctypes.create_string_buffer(8192)
        check(lib().ntt_hip_keyswitch_autotune(self.handle, batch, buf, 8192), "keyswitch_autotune")
        return json.loads(buf.value.decode())

    def run_static(self, in_ptr, batch, out_b_ptr, out_a_ptr, stream=None):
        """The key switch over the library's static device workspace (no
        allocation; calls overlapping on one GPU race)."""
        check(lib().ntt_hip_keyswitch_run_static(self.handle, in_ptr, batch, out_b_ptr, out_a_ptr, stream),
              "keyswitch_run_static")

    def modup_workspace_words(self, batch):
        out = SZ(0)
        check(lib().ntt_hip_modup_workspace_bytes(self.handle, batch, ctypes.byref(out)), "modup_workspace")
        return out.value // 8

    def moddown_workspace_words(self, batch):
        out = SZ(0)
        check(lib().ntt_hip_moddown_workspace_bytes(self.handle, batch, ctypes.byref(out)), "moddown_workspace")
        return out.value // 8

    def mod_up(self, in_ptr, batch, out_ptr, workspace_ptr, evaluation=True, stream=None):
        check(lib().ntt_hip_modup_run(self.handle, in_ptr, int(evaluation), batch, out_ptr, workspace_ptr, stream),
              "modup_run")

    def mod_down(self, in_ptr, batch, out_ptr, workspace_ptr, stream=None):
        check(lib().ntt_hip_moddown_run(self.handle, in_ptr, batch, out_ptr, workspace_ptr, stream), "moddown_run")

    def in_place(self, c0_ptr, c1_ptr, batch, workspace_ptr, c2_ptr=None, stream=None):
        """KeySwitchInPlace: two components (c2 None): c0 += B, c1 = A of c1;
        three: c0 += B, c1 += A of c2 (reference hybrid_key_switch_in_place)."""
        check(lib().ntt_hip_keyswitch_in_place(self.handle, c0_ptr, c1_ptr, c2_ptr, batch, workspace_ptr, stream),
              "keyswitch_in_place")

    @staticmethod
    def relinearize(keys, c0_ptr, c1_ptr, higher_ptrs, batch, workspace_ptr, stream=None):
        """RelinearizeInPlace: component 2 + j switched with keys[j], each
        contribution ModDown-reduced and added into (c0, c1)."""
        if len(keys) != len(higher_ptrs):
            raise ContractError("one relinearization key per higher component")
        R = len(keys)
        kh = (VP * max(R, 1))(*[k.handle for k in keys])
        hi = (VP * max(R, 1))(*higher_ptrs)
        check(lib().ntt_hip_relinearize(kh, R, c0_ptr, c1_ptr, hi, batch, workspace_ptr, stream), "relinearize")

    def rotate_workspace_words(self, batch):
        out = SZ(0)
        check(lib().ntt_hip_rotate_workspace_bytes(self.handle, batch, ctypes.byref(out)), "rotate_workspace")
        return out.value // 8

    @staticmethod
    def rotate(keys, automorphisms, c0_ptr, c1_ptr, batch, out_b_ptrs, out_a_ptrs, workspace_ptr, stream=None):
        """Hoisted rotations: digits of c1 once, then per rotation r the MAC with
        keys[r], sigma_k(B + P c0), sigma_k(A) and ModDown into out_b_ptrs[r] /
        out_a_ptrs[r]. keys[0] sizes the workspace (rotate_workspace_words)."""
        R = len(keys)
        if not (len(automorphisms) == len(out_b_ptrs) == len(out_a_ptrs) == R):
            raise ContractError("one automorphism and one output pair per rotation key")
        kh = (VP * R)(*[k.handle for k in keys])
        au = (ctypes.c_uint32 * R)(*automorphisms)
        ob = (VP * R)(*out_b_ptrs)
        oa = (VP * R)(*out_a_ptrs)
        check(lib().ntt_hip_keyswitch_rotate(kh, R, au, c0_ptr, c1_ptr, batch, ob, oa, workspace_ptr, stream),
              "keyswitch_rotate")

    def capture(self, in_ptr, batch, out_b_ptr, out_a_ptr, workspace_ptr, stream):
        g = VP()
        check(lib().ntt_hip_keyswitch_capture(self.handle, in_ptr, batch, out_b_ptr, out_a_ptr, workspace_ptr,
                                              stream, ctypes.byref(g)), "keyswitch_capture")
        return Graph(g.value)

    def close(self):
        if getattr(self, "handle", None):
            check(lib().ntt_hip_keyswitch_destroy(self.handle), "keyswitch_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MixedJob(ctypes.Structure):
    """ntt_hip_mixed_job: one tower of a mixed-ring-dimension transform."""
    _fields_ = [("ctx", VP), ("data", VP), ("src", VP), ("tower", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def transform_mixed(jobs, inverse, stream=None, workspace=None):
    """Transforms of towers of different ring dimensions in one launch per
    pass. jobs: (Context, data_ptr, tower) or (Context, data_ptr, tower,
    src_ptr) tuples. Returns the device workspace holding the job table; it
    must outlive the launch (pass it back in as `workspace` to reuse it)."""
    arr = (MixedJob * len(jobs))()
    for i, job in enumerate(jobs):
        ctx, data, tower = job[0], job[1], job[2]
        arr[i] = MixedJob(ctx.handle, data, job[3] if len(job) > 3 else None, tower, 0)
    need = SZ(0)
    check(lib().ntt_hip_transform_mixed_workspace_bytes(len(jobs), ctypes.byref(need)), "transform_mixed_workspace")
    if workspace is None or workspace.words * 8 < need.value:
        workspace = DeviceBuffer((need.value + 7) // 8)
    check(lib().ntt_hip_transform_mixed(ctypes.cast(arr, VP), len(jobs), workspace.ptr, int(inverse), stream),
          "transform_mixed")
    return workspace


def autotune(ctx, towers):
    """Run the device autotuner for `towers` towers; returns its JSON report."""
    import json
    buf = ctypes.create_string_buffer(4096)
    check(lib().ntt_hip_autotune(ctx.handle, towers, buf, 4096), "autotune")
    return json.loads(buf.value.decode())


def set_tuning(ctx, ntt_chunk, conv_targets):
    check(lib().ntt_hip_set_tuning(ctx.handle, ntt_chunk, conv_targets), "set_tuning")


def set_ntt_shape(ctx, fwd_headroom=-1, row_residues=8, wg_per_cu=0):
    check(lib().ntt_hip_set_ntt_shape(ctx.handle, fwd_headroom, row_residues, wg_per_cu), "set_ntt_shape")


def set_keyswitch_shape(ks, lanes=2, mac_group=0, fuse_mac=2):
    """Key-switch plan shape: concurrent lanes, ModUp/MAC interleave group,
    ModUp row pass fused with the key inner product (0 off, 1 128-bit sums,
    2 64-bit sums)."""
    check(lib().ntt_hip_keyswitch_set_shape(ks.handle, lanes, mac_group, fuse_mac), "keyswitch_set_shape")


def get_keyswitch_shape(ks):
    la, g, f, fused = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0), INT(0)
    check(lib().ntt_hip_keyswitch_get_shape(ks.handle, ctypes.byref(la), ctypes.byref(g), ctypes.byref(f),
                                            ctypes.byref(fused)), "keyswitch_get_shape")
    return {"lanes": la.value, "mac_group": g.value, "fuse_mac": f.value, "fused": bool(fused.value)}


def get_ntt_shape(ctx):
    h, r, w = INT(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
    check(lib().ntt_hip_get_ntt_shape(ctx.handle, ctypes.byref(h), ctypes.byref(r), ctypes.byref(w)),
          "get_ntt_shape")
    return {"fwd_headroom": h.value, "row_residues": r.value, "wg_per_cu": w.value}


def get_tuning(ctx):
    a, b = ctypes.c_uint32(0), ctypes.c_uint32(0)
    check(lib().ntt_hip_get_tuning(ctx.handle, ctypes.byref(a), ctypes.byref(b)), "get_tuning")
    return {"ntt_chunk": a.value, "conv_targets": b.value}
