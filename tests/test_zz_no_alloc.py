"""The allocation-free key switch (ntt_hip_keyswitch_run_static) against the
workspace form (ntt_hip_keyswitch_run) through the C ABI: bit-exact on two
calls in a row, the second on a smaller plan and input, so the static scratch
is reused with another shape; a batch beyond the scratch is rejected."""
import ctypes

import numpy as np
import pytest

from common import log_case, uniform
from nttgpu import hip, params

pytestmark = pytest.mark.gpu


def _plan(n, L, K, alpha, batch, seed):
    hp = params.openfhe_like(n, L, K, alpha)
    ctx = hip.Context(n, hp.q + hp.p, hp.q_psi + hp.p_psi)
    ks = hip.KeySwitch(ctx, L, K, alpha, max(batch, 2))
    full = hp.q + hp.p
    kb = np.concatenate([uniform(n, full, seed + d) for d in range(hp.beta())])
    ka = np.concatenate([uniform(n, full, seed + 50 + d) for d in range(hp.beta())])
    ks.upload_key(kb, ka, hp.M, hp.q)
    x = np.concatenate([uniform(n, hp.q, seed + 100 + c) for c in range(batch)])
    return hp, ctx, ks, x


def _compare(n, L, K, alpha, batch, seed):
    hp, ctx, ks, x = _plan(n, L, K, alpha, batch, seed)
    log_case(f"keyswitch_run_static n={n} L={L} K={K} alpha={alpha} batch={batch}")
    d_in = hip.DeviceBuffer(x.size).upload(x)
    sb, sa = hip.DeviceBuffer(x.size), hip.DeviceBuffer(x.size)
    ks.run_static(d_in.ptr, batch, sb.ptr, sa.ptr)
    ctx.synchronize()
    # the workspace form, its workspace filled with junk first
    ws_words = ks.workspace_words(batch)
    junk = np.random.default_rng(seed).integers(0, 2**63, ws_words, dtype=np.uint64)
    ws = hip.DeviceBuffer(ws_words).upload(junk)
    wb, wa = hip.DeviceBuffer(x.size), hip.DeviceBuffer(x.size)
    ks.run(d_in.ptr, batch, wb.ptr, wa.ptr, ws.ptr)
    ctx.synchronize()
    assert np.array_equal(sb.download(), wb.download()), "B differs from the workspace form"
    assert np.array_equal(sa.download(), wa.download()), "A differs from the workspace form"
    return ks


def test_static_key_switch_equals_workspace_form_twice():
    # the headline shape fills most of the scratch; then a smaller plan and input
    _compare(65536, 24, 6, 6, 1, 11)
    _compare(8192, 6, 2, 2, 2, 12)


def test_static_key_switch_rejects_batches_beyond_the_scratch():
    hp, ctx, ks, x = _plan(65536, 24, 6, 6, 2, 13)
    d_in = hip.DeviceBuffer(x.size).upload(x)
    ob, oa = hip.DeviceBuffer(x.size), hip.DeviceBuffer(x.size)
    log_case("keyswitch_run_static oversize batch")
    status = hip.lib().ntt_hip_keyswitch_run_static(ks.handle, d_in.ptr, 2, ob.ptr, oa.ptr, None)
    assert status == hip.EINVAL
    assert "static scratch" in hip.lib().ntt_hip_last_error().decode()
    # and the entry point still works afterwards
    ks.run_static(d_in.ptr, 1, ob.ptr, oa.ptr)
    ctx.synchronize()
