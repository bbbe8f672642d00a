"""The autotuner and every launch shape it can choose: each shape bit-exact
against the oracle (NTT in every arithmetic mode, the fused key switch and
its ModDown epilogue), a tuning run on fewer towers than the conversion
sweep's width, and the tuned choice not slower than the default."""
import statistics

import numpy as np
import pytest

import oracle as o
from common import log_case, uniform
from nttgpu import hip, params

pytestmark = pytest.mark.gpu
N16 = 65536

SHAPES = [(-1, 8, 0), (0, 8, 0), (2, 8, 4), (-1, 16, 0), (-1, 16, 5), (-1, 8, 6)]


def _top(bits, count, n=N16):
    qs = [params.last_prime(bits, 2 * n)]
    while len(qs) < count:
        qs.append(params.previous_prime(qs[-1], 2 * n))
    return qs


@pytest.mark.parametrize("bits", [62, 59, 50, 30])
@pytest.mark.parametrize("shape", SHAPES)
def test_every_shape_is_bit_exact(bits, shape):
    qs = _top(bits, 3)
    rs = [params.min_root(q, N16) for q in qs]
    ctx = hip.Context(N16, qs, rs)
    hip.set_ntt_shape(ctx, *shape)
    assert hip.get_ntt_shape(ctx) == dict(zip(("fwd_headroom", "row_residues", "wg_per_cu"), shape))
    x = uniform(N16, qs, bits * 7 + shape[1])
    x[:N16] = qs[0] - 1
    log_case(f"shape {shape} {bits}-bit forward/inverse")
    buf = hip.DeviceBuffer(x.size).upload(x)
    ctx.transform(buf.ptr, 0, 3, False)
    ctx.synchronize()
    got = buf.download()
    assert np.array_equal(got, o.ntt_towers(x, N16, qs, rs)), f"forward, shape {shape}"
    ctx.transform(buf.ptr, 0, 3, True)
    ctx.synchronize()
    assert np.array_equal(buf.download(), x), f"round trip, shape {shape}"


@pytest.mark.parametrize("shape", [(0, 16, 4), (-1, 8, 5)])
def test_key_switch_under_tuned_shapes(shape):
    """The fused key switch (complement transforms, ModDown epilogue in the
    16-residue row pass) under non-default shapes, vs the oracle."""
    hp = params.openfhe_like(N16, 6, 2, 2)
    ctx = hip.Context(N16, hp.q + hp.p, hp.q_psi + hp.p_psi)
    hip.set_ntt_shape(ctx, *shape)
    ks = hip.KeySwitch(ctx, 6, 2, 2, 1)
    full = hp.q + hp.p
    kb = np.concatenate([uniform(N16, full, 40 + d) for d in range(hp.beta())])
    ka = np.concatenate([uniform(N16, full, 50 + d) for d in range(hp.beta())])
    ks.upload_key(kb, ka, hp.M, hp.q)
    x = uniform(N16, hp.q, 60)
    log_case(f"key switch under shape {shape}")
    d_in = hip.DeviceBuffer(x.size).upload(x)
    ob, oa = hip.DeviceBuffer(x.size), hip.DeviceBuffer(x.size)
    ws = hip.DeviceBuffer(ks.workspace_words(1))
    ks.run(d_in.ptr, 1, ob.ptr, oa.ptr, ws.ptr)
    ctx.synchronize()
    wb, wa = o.key_switch(x, N16, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
    assert np.array_equal(ob.download(), wb) and np.array_equal(oa.download(), wa)


def test_autotune_with_fewer_towers_than_the_conversion_width():
    """A 30-tower context tuned on 8 towers: the conversion sweep (6 -> 24)
    needs more rows than the NTT's; the scratch holds both."""
    hp = params.openfhe_like(N16, 24, 6, 6)
    ctx = hip.Context(N16, hp.q + hp.p, hp.q_psi + hp.p_psi)
    log_case("autotune 8 towers on a 30-tower context")
    rep = hip.autotune(ctx, 8)
    assert set(rep["chosen"]) == {"fwd_headroom", "row_residues", "wg_per_cu", "ntt_chunk", "conv_targets"}
    x = uniform(N16, hp.q[:4], 3)
    buf = hip.DeviceBuffer(x.size).upload(x)
    ctx.transform(buf.ptr, 0, 4, False)
    ctx.synchronize()
    assert np.array_equal(buf.download(), o.ntt_towers(x, N16, hp.q[:4], hp.q_psi[:4]))


def test_tuned_choice_is_exact_and_not_slower():
    q = params.last_prime(59, 2 * N16)
    r = params.min_root(q, N16)
    ctx = hip.Context(N16, [q], [r])
    towers = 1024
    data = hip.DeviceBuffer(towers * N16)
    for t0 in range(0, towers, 256):
        data.upload(uniform(N16, [q] * 256, 7 + t0), offset=t0 * N16)
    log_case("autotune 59-bit, 512 towers")
    rep = hip.autotune(ctx, 512)
    tuned = (hip.get_ntt_shape(ctx), hip.get_tuning(ctx))
    st = hip.Stream(0)
    e0, e1 = hip.Event(), hip.Event()

    def timed(shape, tuning):
        hip.set_ntt_shape(ctx, shape["fwd_headroom"], shape["row_residues"], shape["wg_per_cu"])
        hip.set_tuning(ctx, tuning["ntt_chunk"], tuning["conv_targets"])
        ctx.transform_batch(data.ptr, towers, N16, 0, 1, False, stream=st.handle)
        e0.record(st.handle)
        for _ in range(4):
            ctx.transform_batch(data.ptr, towers, N16, 0, 1, False, stream=st.handle)
            ctx.transform_batch(data.ptr, towers, N16, 0, 1, True, stream=st.handle)
        e1.record(st.handle)
        e1.synchronize()
        return e0.elapsed_ms(e1)

    # alternating pairs, compared by the median of their ratios: single
    # samples of a few milliseconds spread by up to 20% with the clock, and
    # pairing cancels the drift between samples
    default = ({"fwd_headroom": -1, "row_residues": 8, "wg_per_cu": 0}, {"ntt_chunk": 0, "conv_targets": 0})
    td, tt = [], []
    for _ in range(8):
        td.append(timed(*default))
        tt.append(timed(*tuned))
    ratios = [t / d for t, d in zip(tt, td)]
    assert statistics.median(ratios) <= 1.05, (rep, td, tt)
    # the tuned shape is bit-exact
    x = uniform(N16, [q] * 2, 99)
    buf = hip.DeviceBuffer(x.size).upload(x)
    ctx.transform_batch(buf.ptr, 2, N16, 0, 1, False)
    ctx.synchronize()
    assert np.array_equal(buf.download(), o.ntt_towers(x, N16, [q] * 2, [r] * 2))


def test_keyswitch_autotune_is_exact_and_not_slower():
    """The key switch's own tuner (NTT shape, lanes, interleave group,
    conversion width timed on the batched key switch): afterwards the key
    switch is bit-exact against the oracle and not slower than the defaults."""
    hp = params.openfhe_like(N16, 24, 6, 6)
    ctx = hip.Context(N16, hp.q + hp.p, hp.q_psi + hp.p_psi)
    B = 8
    ks = hip.KeySwitch(ctx, 24, 6, 6, B)
    full = hp.q + hp.p
    kb = np.concatenate([uniform(N16, full, 70 + d) for d in range(hp.beta())])
    ka = np.concatenate([uniform(N16, full, 80 + d) for d in range(hp.beta())])
    ks.upload_key(kb, ka, hp.M, hp.q)
    log_case("keyswitch autotune L=24 batch 8")
    rep = ks.autotune(B)
    assert {"lanes", "fuse_mac", "mac_group", "conv_targets"} <= set(rep["chosen"])
    x = np.concatenate([uniform(N16, hp.q, 90 + c) for c in range(B)])
    comp = 24 * N16
    d_in = hip.DeviceBuffer(x.size).upload(x)
    ob, oa = hip.DeviceBuffer(x.size), hip.DeviceBuffer(x.size)
    ws = hip.DeviceBuffer(ks.workspace_words(B))
    ks.run(d_in.ptr, B, ob.ptr, oa.ptr, ws.ptr)
    ctx.synchronize()
    gb, ga = ob.download(), oa.download()
    for c in (0, B - 1):
        wb, wa = o.key_switch(x[c * comp:(c + 1) * comp], N16, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(gb[c * comp:(c + 1) * comp], wb) and np.array_equal(ga[c * comp:(c + 1) * comp], wa)
    # tuned vs a fresh default plan on the same data, alternating
    ctx2 = hip.Context(N16, hp.q + hp.p, hp.q_psi + hp.p_psi)
    ks2 = hip.KeySwitch(ctx2, 24, 6, 6, B)
    ks2.upload_key(kb, ka, hp.M, hp.q)
    st = hip.Stream(0)
    e0, e1 = hip.Event(), hip.Event()

    def timed(plan):
        plan.run(d_in.ptr, B, ob.ptr, oa.ptr, ws.ptr, st.handle)
        e0.record(st.handle)
        for _ in range(3):
            plan.run(d_in.ptr, B, ob.ptr, oa.ptr, ws.ptr, st.handle)
        e1.record(st.handle)
        e1.synchronize()
        return e0.elapsed_ms(e1)

    td, tt = [], []
    for _ in range(8):
        td.append(timed(ks2))
        tt.append(timed(ks))
    ratios = [t / d for t, d in zip(tt, td)]
    assert statistics.median(ratios) <= 1.05, (rep, td, tt)
