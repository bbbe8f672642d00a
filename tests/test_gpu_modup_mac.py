"""ModUp's last NTT pass fused with the key inner product
(ntt_pass_b8_fwd_mac): every fusion mode (off, 128-bit sums, 64-bit sums)
under interleave groups and lane counts gives the same key switch, bit for
bit, equal to the oracle; digit layouts with a shortened last digit, one digit
wider than the level, a reduced level, and the gather form over scattered
components; plans outside the fused kernel's range report that it does not
apply."""
import numpy as np
import pytest

import oracle as o
from common import log_case, uniform
from nttgpu import hip, params

pytestmark = pytest.mark.gpu
N16 = 65536


def _plan(hp, L, B, seed):
    ctx = hip.Context(N16, hp.q[:L] + hp.p, hp.q_psi[:L] + hp.p_psi)
    ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, B)
    full = hp.q + hp.p
    kb = np.concatenate([uniform(N16, full, seed + d) for d in range(hp.beta())])
    ka = np.concatenate([uniform(N16, full, seed + 50 + d) for d in range(hp.beta())])
    ks.upload_key(kb, ka, hp.M, hp.q)
    return ctx, ks, kb, ka


def _run(ctx, ks, x, L, B):
    d_in = hip.DeviceBuffer(x.size).upload(x)
    ob, oa = hip.DeviceBuffer(B * L * N16), hip.DeviceBuffer(B * L * N16)
    ws = hip.DeviceBuffer(ks.workspace_words(B))
    ks.run(d_in.ptr, B, ob.ptr, oa.ptr, ws.ptr)
    ctx.synchronize()
    return ob.download(), oa.download()


def _oracle(hp, L, x, kb, ka):
    return o.key_switch(x, N16, hp.q[:L], hp.q_psi[:L], hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)


def test_fusion_modes_groups_and_lanes_agree_with_the_oracle():
    """L=24 (60-bit q0, 50-bit scaling primes), K=6, dnum=4, batch 9: the
    fused kernel's tower runs split at the headroom classes (60-bit and
    50-bit towers), inside interleave groups of 2 and 7 towers (7 leaves a
    partial last group), on one lane (grouped: 9 >= 8 components) and two
    (ungrouped sub-batches of 4 and 5)."""
    hp = params.openfhe_like(N16, 24, 6, 6)
    L, B = 24, 9
    ctx, ks, kb, ka = _plan(hp, L, B, 9100)
    x = np.concatenate([uniform(N16, hp.q, 9200 + c) for c in range(B)])
    ref = None
    for fuse in (0, 1, 2):
        for group, lanes in ((0, 1), (2, 1), (7, 1), (2, 2)):
            log_case(f"modup+mac n=65536 L=24 fuse_mac={fuse} mac_group={group} lanes={lanes} batch={B}")
            hip.set_keyswitch_shape(ks, lanes, group, fuse)
            shape = hip.get_keyswitch_shape(ks)
            assert shape["fused"] == (fuse != 0), shape
            got = _run(ctx, ks, x, L, B)
            if ref is None:
                ref = got
            else:
                assert np.array_equal(got[0], ref[0]), f"B channel, fuse_mac={fuse} group={group} lanes={lanes}"
                assert np.array_equal(got[1], ref[1]), f"A channel, fuse_mac={fuse} group={group} lanes={lanes}"
    w = L * N16
    for c in (0, 4, B - 1):
        wb, wa = _oracle(hp, L, x[c * w:(c + 1) * w], kb, ka)
        assert np.array_equal(ref[0][c * w:(c + 1) * w], wb), f"B component {c}"
        assert np.array_equal(ref[1][c * w:(c + 1) * w], wa), f"A component {c}"


@pytest.mark.parametrize("M,L,K,alpha", [(24, 20, 6, 6), (20, 20, 6, 6), (12, 12, 4, 16), (8, 8, 3, 1)])
def test_fused_digit_layouts(M, L, K, alpha):
    """Reduced level (L=20 of M=24: the key's rows at the level's offsets),
    a shortened last digit (20 = 6+6+6+2), one digit wider than the level
    (alpha=16 > L=12: no complement Q towers), eight one-tower digits (sums
    of 8 products of 60-bit primes: 64-bit sums at their bound 2^4 q <= 2^64):
    fused modes equal the unfused path and the oracle."""
    hp = params.openfhe_like(N16, M, K, alpha)
    B = 3
    ctx, ks, kb, ka = _plan(hp, L, B, 9300 + M + alpha)
    x = np.concatenate([uniform(N16, hp.q[:L], 9400 + c) for c in range(B)])
    outs = []
    for fuse in (0, 1, 2):
        log_case(f"modup+mac n=65536 M={M} L={L} K={K} alpha={alpha} fuse_mac={fuse} batch={B}")
        hip.set_keyswitch_shape(ks, 1, 0, fuse)
        assert hip.get_keyswitch_shape(ks)["fused"] == (fuse != 0)
        outs.append(_run(ctx, ks, x, L, B))
    for fuse in (1, 2):
        assert np.array_equal(outs[fuse][0], outs[0][0]) and np.array_equal(outs[fuse][1], outs[0][1]), fuse
    w = L * N16
    wb, wa = _oracle(hp, L, x[:w], kb, ka)
    assert np.array_equal(outs[1][0][:w], wb) and np.array_equal(outs[1][1][:w], wa)


def test_fused_primes_below_2_61():
    """Every prime just below 2^61 (headroom class 0: the row pass reduces
    at every stage, canonical outputs straight into the products), L=3, K=2,
    alpha=2: 64-bit sums at their bound (2 digits, 4 q <= 2^64)."""
    qs = [params.last_prime(61, 2 * N16)]
    while len(qs) < 5:
        qs.append(params.previous_prime(qs[-1], 2 * N16))
    hp = params.HybridParams(n=N16, q=qs[:3], p=qs[3:], alpha=2)
    L, B = 3, 2
    ctx, ks, kb, ka = _plan(hp, L, B, 9800)
    x = np.concatenate([uniform(N16, hp.q, 9900 + c) for c in range(B)])
    w = L * N16
    for fuse in (1, 2):
        log_case(f"modup+mac n=65536 primes below 2^61 L=3 K=2 alpha=2 fuse_mac={fuse}")
        hip.set_keyswitch_shape(ks, 1, 0, fuse)
        assert hip.get_keyswitch_shape(ks)["fused"]
        gb, ga = _run(ctx, ks, x, L, B)
        for c in range(B):
            wb, wa = _oracle(hp, L, x[c * w:(c + 1) * w], kb, ka)
            assert np.array_equal(gb[c * w:(c + 1) * w], wb) and np.array_equal(ga[c * w:(c + 1) * w], wa), (fuse, c)


def test_fused_with_a_prime_below_2_24():
    """A Q basis with a 23-bit prime beside 60- and 50-bit ones (the context
    runs headroom class 0), three one-tower digits, fused 64-bit sums: equal
    to the oracle."""
    m = 2 * N16
    q = [params.last_prime(60, m), params.last_prime(50, m), params.last_prime(23, m)]
    hp = params.HybridParams(n=N16, q=q, p=[params.previous_prime(q[0], m)], alpha=1)
    L, B = 3, 2
    ctx, ks, kb, ka = _plan(hp, L, B, 9850)
    assert hip.get_keyswitch_shape(ks)["fused"]
    x = np.concatenate([uniform(N16, hp.q, 9860 + c) for c in range(B)])
    log_case(f"modup+mac n=65536 q={q} alpha=1")
    gb, ga = _run(ctx, ks, x, L, B)
    w = L * N16
    for c in range(B):
        wb, wa = _oracle(hp, L, x[c * w:(c + 1) * w], kb, ka)
        assert np.array_equal(gb[c * w:(c + 1) * w], wb) and np.array_equal(ga[c * w:(c + 1) * w], wa), c


def test_fused_gather_equals_contiguous_unfused():
    """run_gather (components at scattered addresses: the fused kernel reads
    each digit's own towers through the pointer table) with the fused kernel
    equals run on one contiguous buffer without it."""
    hp = params.openfhe_like(N16, 24, 6, 6)
    L, B = 24, 9
    ctx, ks, kb, ka = _plan(hp, L, B, 9500)
    x = np.concatenate([uniform(N16, hp.q, 9600 + c) for c in range(B)])
    log_case("modup+mac gather n=65536 L=24 batch=9, unfused contiguous reference")
    hip.set_keyswitch_shape(ks, 2, 2, 0)
    gb, ga = _run(ctx, ks, x, L, B)
    w = L * N16
    order = np.random.default_rng(9700).permutation(B)
    ins, obs, oas = [None] * B, [None] * B, [None] * B
    for c in order:
        ins[c] = hip.DeviceBuffer(w).upload(x[c * w:(c + 1) * w])
        obs[c], oas[c] = hip.DeviceBuffer(w), hip.DeviceBuffer(w)
    wsg = hip.DeviceBuffer(ks.gather_workspace_words(B))
    for fuse in (1, 2):
        log_case(f"modup+mac gather n=65536 L=24 batch=9 fuse_mac={fuse}")
        hip.set_keyswitch_shape(ks, 2, 2, fuse)
        ks.run_gather([b.ptr for b in ins], [b.ptr for b in obs], [b.ptr for b in oas], wsg.ptr)
        ctx.synchronize()
        for c in range(B):
            assert np.array_equal(obs[c].download(), gb[c * w:(c + 1) * w]), f"B channel {c}, fuse_mac={fuse}"
            assert np.array_equal(oas[c].download(), ga[c * w:(c + 1) * w]), f"A channel {c}, fuse_mac={fuse}"


def test_fusion_range():
    """The fused kernel applies only to N = 2^16 wide-lazy plans with the
    8-residue row pass and sums that fit: 17 one-tower digits of 60-bit
    primes (17 q >= 2^64) and 63-bit primes keep the unfused path; so does
    N = 2^13 and the 16-residue row pass; the setter rejects bad values."""
    hp = params.openfhe_like(N16, 17, 1, 1)
    ctx = hip.Context(N16, hp.q + hp.p, hp.q_psi + hp.p_psi)
    ks = hip.KeySwitch(ctx, 17, 1, 1, 1)
    for fuse in (1, 2):
        hip.set_keyswitch_shape(ks, 1, 0, fuse)
        assert not hip.get_keyswitch_shape(ks)["fused"], fuse
    q63 = params.find_primes(N16, 4, 63)
    hp63 = params.HybridParams(n=N16, q=q63[:3], p=q63[3:], alpha=2)
    ctx63 = hip.Context(N16, hp63.q + hp63.p, hp63.q_psi + hp63.p_psi)
    ks63 = hip.KeySwitch(ctx63, 3, 1, 2, 1)
    assert not hip.get_keyswitch_shape(ks63)["fused"]
    hp13 = params.openfhe_like(8192, 6, 2, 2)
    ctx13 = hip.Context(8192, hp13.q + hp13.p, hp13.q_psi + hp13.p_psi)
    assert not hip.get_keyswitch_shape(hip.KeySwitch(ctx13, 6, 2, 2, 1))["fused"]
    hp24 = params.openfhe_like(N16, 24, 6, 6)
    ctx24 = hip.Context(N16, hp24.q + hp24.p, hp24.q_psi + hp24.p_psi)
    ks24 = hip.KeySwitch(ctx24, 24, 6, 6, 1)
    assert hip.get_keyswitch_shape(ks24) == {"lanes": 2, "mac_group": 0, "fuse_mac": 2, "fused": True}
    hip.set_ntt_shape(ctx24, -1, 16, 0)
    assert not hip.get_keyswitch_shape(ks24)["fused"]
    for bad in ((0, 2, 1), (9, 2, 1), (2, 2, 3)):
        with pytest.raises(hip.ContractError):
            hip.set_keyswitch_shape(ks24, *bad)
