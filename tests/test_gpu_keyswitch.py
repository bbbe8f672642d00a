"""GPU parity of basis conversion, ModDown correction, key inner product,
automorphism and the fused batched Hybrid key switch against the CPU oracle.

Bit-exact on full arrays: the Driver-seam operations one by one
(src/cuda/resident_adapter.inc:204-280), then the whole KeySwitchHybrid core
(include/ntt/openfhe_hybrid_adapter.hpp:163-181) at the headline shape
N=2^16, L=24, K=6, dnum=4 and at reduced levels (key offset i + (M - L)).
"""
import numpy as np
import pytest

import oracle as o
from common import log_case, uniform
from nttgpu import hip, params

pytestmark = pytest.mark.gpu


def _dev(x):
    return hip.DeviceBuffer(x.size).upload(x)


@pytest.fixture(scope="module")
def small():
    hp = params.openfhe_like(1024, 6, 2, 2)
    ctx = hip.Context(hp.n, hp.q + hp.p, hp.q_psi + hp.p_psi)
    return hp, ctx


@pytest.mark.parametrize("src,dst", [([0, 1], [2, 3, 4, 5, 6, 7]), ([6, 7], [0, 1, 2, 3, 4, 5]),
                                     ([2, 3, 4], [0, 1, 6]), ([5], [0])])
def test_switch_basis(small, src, dst):
    hp, ctx = small
    mods = hp.q + hp.p
    x = uniform(hp.n, [mods[i] for i in src], 7)
    log_case(f"switch_basis {src}->{dst}")
    out = hip.DeviceBuffer(hp.n * len(dst))
    dx = _dev(x)
    ctx.switch_basis(out.ptr, dx.ptr, src, dst)
    ctx.synchronize()
    want = o.basis_switch(x, hp.n, [mods[i] for i in src], [mods[i] for i in dst])
    assert np.array_equal(out.download(), want)


def test_switch_basis_per_term_path():
    """Four 63-bit sources (sum b_i >= 2^64): the register kernel's limb
    form is exact whatever the source widths."""
    n = 256
    qs = params.find_primes(n, 6, 63)
    ctx = hip.Context(n, qs, [params.find_psi(q, n) for q in qs])
    x = np.concatenate([np.full(n, q - 1, np.uint64) for q in qs[:4]])
    x[::7] = 0
    log_case("switch_basis per-term 63-bit")
    out = hip.DeviceBuffer(2 * n)
    dx = _dev(x)
    ctx.switch_basis(out.ptr, dx.ptr, [0, 1, 2, 3], [4, 5])
    ctx.synchronize()
    assert np.array_equal(out.download(), o.basis_switch(x, n, qs[:4], qs[4:]))


@pytest.mark.parametrize("s", [1, 3, 7, 11, 16])
@pytest.mark.parametrize("bits", [63, 50, 31])
def test_switch_basis_register_kernel_widths(s, bits):
    """Every source count of the register kernel (1..16, one specialisation
    each) at 63-, 50- and 31-bit primes, extreme residues included."""
    n = 256
    qs = params.find_primes(n, s + 3, bits)
    ctx = hip.Context(n, qs, [params.find_psi(q, n) for q in qs])
    x = uniform(n, qs[:s], 1000 + 31 * s + bits)
    x[::5] = np.repeat(np.array(qs[:s], np.uint64) - 1, n)[::5]
    x[1::9] = 0
    log_case(f"switch_basis register kernel s={s} {bits}-bit")
    out = hip.DeviceBuffer(3 * n)
    dx = _dev(x)
    ctx.switch_basis(out.ptr, dx.ptr, list(range(s)), [s, s + 1, s + 2])
    ctx.synchronize()
    assert np.array_equal(out.download(), o.basis_switch(x, n, qs[:s], qs[s:]))


@pytest.mark.parametrize("bits,s", [(30, 20), (62, 18)])
def test_switch_basis_wide_source_sets(bits, s):
    """More than 16 sources: the LDS kernel, with one reduction per output
    (20 x 30-bit, sum b_i < 2^64) or per-term reductions (18 x 62-bit)."""
    n = 256
    qs = params.find_primes(n, s + 4, bits)
    ctx = hip.Context(n, qs, [params.find_psi(q, n) for q in qs])
    x = uniform(n, qs[:s], 77 + s)
    x[::3] = np.repeat(np.array(qs[:s], np.uint64) - 1, n)[::3]
    log_case(f"switch_basis LDS kernel s={s} {bits}-bit")
    out = hip.DeviceBuffer(4 * n)
    dx = _dev(x)
    ctx.switch_basis(out.ptr, dx.ptr, list(range(s)), list(range(s, s + 4)))
    ctx.synchronize()
    assert np.array_equal(out.download(), o.basis_switch(x, n, qs[:s], qs[s:]))


def test_switch_basis_rejections(small):
    hp, ctx = small
    buf = hip.DeviceBuffer(hp.n * 2)
    for src, dst in (([0, 0], [1]), ([0], [0]), ([], [1]), ([0], []), ([0], [99])):
        with pytest.raises(hip.ContractError):
            ctx.switch_basis(buf.ptr, buf.ptr, src, dst)


def test_correct_mac_rotate_pointwise(small):
    hp, ctx = small
    n, L, K = hp.n, hp.M, hp.K
    mods = hp.q + hp.p
    T = L + K
    left = uniform(n, hp.q, 1)
    right = uniform(n, hp.q, 2)
    log_case("correct")
    out = hip.DeviceBuffer(L * n)
    dl, dr = _dev(left), _dev(right)
    ctx.correct(out.ptr, dl.ptr, dr.ptr, 0, L, [L + j for j in range(K)])
    ctx.synchronize()
    want = np.empty_like(left)
    for l, q in enumerate(hp.q):
        pinv = o.inv_mod(o.special_product(hp.p, q), q)
        sl = slice(l * n, (l + 1) * n)
        want[sl] = [o.mul_mod(o.sub_mod(int(a), int(b), q), pinv, q) for a, b in zip(left[sl], right[sl])]
    assert np.array_equal(out.download(), want)

    log_case("hybrid_mac per digit")
    beta = 3
    digits = [uniform(n, mods, 10 + d) for d in range(beta)]
    kb = [uniform(n, mods, 20 + d) for d in range(beta)]
    ka = [uniform(n, mods, 30 + d) for d in range(beta)]
    b, a = hip.DeviceBuffer(T * n), hip.DeviceBuffer(T * n)
    keep = []
    for d in range(beta):
        bufs = [_dev(digits[d]), _dev(kb[d]), _dev(ka[d])]
        keep += bufs
        ctx.hybrid_mac(b.ptr, a.ptr, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, 0, T, d == 0)
    ctx.synchronize()
    wb, wa = o.hybrid_mac(np.concatenate(digits), beta, np.concatenate(kb), np.concatenate(ka), n, hp.q, hp.p, L)
    assert np.array_equal(b.download(), wb) and np.array_equal(a.download(), wa)

    # the same digits in one pass (the driver's merged run), then 11 digits
    # (two launches of at most 8) accumulating onto existing outputs, against
    # the per-digit entry point
    log_case("hybrid_mac_digits")
    ptrs = [keep[3 * d:3 * d + 3] for d in range(beta)]
    b2, a2 = hip.DeviceBuffer(T * n), hip.DeviceBuffer(T * n)
    ctx.hybrid_mac_digits(b2.ptr, a2.ptr, [p[0].ptr for p in ptrs], [p[1].ptr for p in ptrs],
                          [p[2].ptr for p in ptrs], 0, T, True)
    ctx.synchronize()
    assert np.array_equal(b2.download(), wb) and np.array_equal(a2.download(), wa)
    many = [[_dev(uniform(n, mods, 50 + 3 * d + j)) for j in range(3)] for d in range(11)]
    b3, a3 = _dev(wb), _dev(wa)
    for d in range(11):
        ctx.hybrid_mac(b3.ptr, a3.ptr, many[d][0].ptr, many[d][1].ptr, many[d][2].ptr, 0, T, False)
    ctx.hybrid_mac_digits(b2.ptr, a2.ptr, [m[0].ptr for m in many], [m[1].ptr for m in many],
                          [m[2].ptr for m in many], 0, T, False)
    ctx.synchronize()
    assert np.array_equal(b2.download(), b3.download()) and np.array_equal(a2.download(), a3.download())
    assert not np.array_equal(b3.download(), wb)
    with pytest.raises(hip.NttHipError):
        ctx.hybrid_mac_digits(b2.ptr, a2.ptr, [], [], [], 0, T, True)

    log_case("hybrid_rotate")
    c0 = uniform(n, hp.q, 40)
    dc0 = _dev(c0)
    for k in (5, 25, 2 * n - 1):
        ob, oa = hip.DeviceBuffer(T * n), hip.DeviceBuffer(T * n)
        ctx.hybrid_rotate(ob.ptr, oa.ptr, b.ptr, a.ptr, dc0.ptr, 0, T, L, [L + j for j in range(K)], k)
        ctx.synchronize()
        rb, ra = o.hybrid_rotate(wb, wa, c0, n, hp.q, hp.p, k)
        assert np.array_equal(ob.download(), rb) and np.array_equal(oa.download(), ra)
    with pytest.raises(hip.ContractError):
        ctx.hybrid_rotate(ob.ptr, oa.ptr, b.ptr, a.ptr, b.ptr, 0, T, L, [L], 4)

    log_case("pointwise")
    x, y = uniform(n, mods, 50), uniform(n, mods, 51)
    bx = _dev(x)
    dy = _dev(y)
    ctx.pointwise(bx.ptr, dy.ptr, 0, T)
    ctx.synchronize()
    assert np.array_equal(bx.download(), o.pointwise_towers(x, y, n, mods))


def _key(hp, seed):
    full = hp.q + hp.p
    beta_full = hp.beta()
    kb = np.concatenate([uniform(hp.n, full, seed + d) for d in range(beta_full)])
    ka = np.concatenate([uniform(hp.n, full, seed + 100 + d) for d in range(beta_full)])
    return kb, ka


def _keyswitch(hp, L, batch, seed, max_batch=None):
    """Run the fused key switch on `batch` components at active level L."""
    qa, pa = hp.q[:L], hp.p
    ctx = hip.Context(hp.n, qa + pa, hp.q_psi[:L] + hp.p_psi)
    ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, max_batch or batch)
    kb, ka = _key(hp, seed)
    ks.upload_key(kb, ka, hp.M, hp.q)
    x = np.concatenate([uniform(hp.n, qa, seed + 1000 + c) for c in range(batch)])
    inp = _dev(x)
    ob, oa = hip.DeviceBuffer(batch * L * hp.n), hip.DeviceBuffer(batch * L * hp.n)
    ws = hip.DeviceBuffer(ks.workspace_words(batch))
    ks.run(inp.ptr, batch, ob.ptr, oa.ptr, ws.ptr)
    ctx.synchronize()
    return x, kb, ka, ob.download(), oa.download()


@pytest.mark.parametrize("L", [6, 5, 3, 1])
def test_keyswitch_small_levels(L):
    hp = params.openfhe_like(1024, 6, 2, 2)
    log_case(f"keyswitch n=1024 M=6 L={L}")
    x, kb, ka, gb, ga = _keyswitch(hp, L, 2, 500 + L)
    n = hp.n
    for c in range(2):
        xs = x[c * L * n:(c + 1) * L * n]
        wb, wa = o.key_switch(xs, n, hp.q[:L], hp.q_psi[:L], hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(gb[c * L * n:(c + 1) * L * n], wb), f"B component {c}"
        assert np.array_equal(ga[c * L * n:(c + 1) * L * n], wa), f"A component {c}"


def _check_keyswitch(hp, L, batch, seed):
    x, kb, ka, gb, ga = _keyswitch(hp, L, batch, seed)
    n = hp.n
    for c in range(batch):
        xs = x[c * L * n:(c + 1) * L * n]
        wb, wa = o.key_switch(xs, n, hp.q[:L], hp.q_psi[:L], hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(gb[c * L * n:(c + 1) * L * n], wb), f"B component {c}"
        assert np.array_equal(ga[c * L * n:(c + 1) * L * n], wa), f"A component {c}"


@pytest.mark.parametrize("L", [12, 17])
def test_keyswitch_many_digits(L):
    """alpha = 1: L single-tower digits (s = 1 conversions, long MAC sums); at
    L = 17, 17 * 2^60 >= 2^64 selects the MAC's per-term reduction."""
    hp = params.openfhe_like(1024, L, 1, 1)
    log_case(f"keyswitch n=1024 L={L} K=1 alpha=1 ({L} digits) batch=3")
    _check_keyswitch(hp, L, 3, 640 + L)


def test_keyswitch_63bit_primes_per_term():
    """63-bit primes: canonical NTT, per-term basis conversion and per-term
    MAC (4 digits * 2^62 >= 2^64), batch of 5."""
    n = 1024
    qs = params.find_primes(n, 6, 63)
    hp = params.HybridParams(n=n, q=qs[:4], p=qs[4:], alpha=1)
    log_case("keyswitch n=1024 63-bit primes L=4 K=2 alpha=1 batch=5")
    _check_keyswitch(hp, 4, 5, 650)


def test_keyswitch_wide_lazy_boundary_primes():
    """Every prime just below 2^61 (the edge of the wide-lazy NTT range) at
    N=8192, where the ModDown correction runs in the NTT's last pass on lazy
    values: z + 4q - v must stay below 2^64."""
    n = 8192
    qs = [params.last_prime(61, 2 * n)]
    while len(qs) < 5:
        qs.append(params.previous_prime(qs[-1], 2 * n))
    hp = params.HybridParams(n=n, q=qs[:3], p=qs[3:], alpha=2)
    log_case("keyswitch n=8192 primes below 2^61 L=3 K=2 alpha=2 batch=2")
    _check_keyswitch(hp, 3, 2, 660)


@pytest.mark.parametrize("n,L,batch", [(4096, 6, 2), (4096, 4, 2), (65536, 6, 1)])
def test_keyswitch_narrow_primes(n, L, batch):
    """Every modulus below 2^30 (q0 = LastPrime(30), scaling primes near 2^28,
    P below q0): the NTTs run the 32-bit butterflies, including the ModDown
    correction fused into the last forward pass; bit-exact vs the oracle."""
    hp = params.openfhe_like(n, 6, 2, 2, first_bits=30, scale_bits=28)
    assert max(hp.q + hp.p) < 1 << 30
    log_case(f"keyswitch narrow primes n={n} L={L} batch={batch}")
    _check_keyswitch(hp, L, batch, 7100 + L)


@pytest.mark.parametrize("n,bits,L,K,alpha,batch", [
    (8192, 63, 4, 2, 1, 2),    # canonical arithmetic, 4 digits
    (16384, 60, 5, 2, 2, 2),   # wide-lazy, shortened last digit
    (32768, 50, 12, 1, 1, 2),  # 12 single-tower digits
    (8192, 30, 6, 2, 3, 3),    # 32-bit butterflies
])
def test_keyswitch_ring_sizes_and_widths(n, bits, L, K, alpha, batch):
    """Rings 2^13..2^15 (two-pass transforms with a column pass of R < 256
    rows) in every NTT arithmetic mode and with up to 12 digits, bit-exact vs
    the oracle."""
    qs = params.find_primes(n, L + K, bits)
    hp = params.HybridParams(n=n, q=qs[:L], p=qs[L:], alpha=alpha)
    log_case(f"keyswitch n={n} {bits}-bit L={L} K={K} alpha={alpha} batch={batch}")
    _check_keyswitch(hp, L, batch, 9000 + n % 97 + L)


def _random_shapes(count, seed):
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(count):
        log_n = int(rng.integers(10, 17))
        bits = int(rng.choice([30, 45, 50, 55, 60, 61, 62]))
        L = int(rng.integers(1, 13))
        K = int(rng.integers(1, 5))
        alpha = int(rng.integers(1, L + 2))
        batch = int(rng.integers(1, 10)) if log_n <= 13 else int(rng.integers(1, 3))
        shapes.append((1 << log_n, bits, L, K, alpha, batch))
    return shapes


@pytest.mark.parametrize("n,bits,L,K,alpha,batch", _random_shapes(10, 0xC0FFEE))
def test_keyswitch_random_shapes(n, bits, L, K, alpha, batch):
    """Seeded random shapes: ring 2^10..2^16, prime width 30..62 bits (every
    arithmetic mode), 1..12 Q towers, 1..4 P towers, any digit size (also
    wider than L), batches up to 9 (two lanes from 8)."""
    qs = params.find_primes(n, L + K, bits)
    hp = params.HybridParams(n=n, q=qs[:L], p=qs[L:], alpha=alpha)
    log_case(f"keyswitch random shape n={n} {bits}-bit L={L} K={K} alpha={alpha} batch={batch}")
    _check_keyswitch(hp, L, batch, 12000 + n % 1000 + L * 10 + K)


def test_keyswitch_n65536_l24_k6_dnum4():
    """N=2^16, L=24 (~50-bit, first 60-bit), K=6 (60-bit), dnum=4 (alpha=6), batch of 2."""
    hp = params.openfhe_like(65536, 24, 6, 6)
    log_case("keyswitch n=65536 L=24 K=6 dnum=4 batch=2")
    x, kb, ka, gb, ga = _keyswitch(hp, 24, 2, 7000)
    n, L = hp.n, 24
    for c in range(2):
        xs = x[c * L * n:(c + 1) * L * n]
        wb, wa = o.key_switch(xs, n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(gb[c * L * n:(c + 1) * L * n], wb), f"B component {c}"
        assert np.array_equal(ga[c * L * n:(c + 1) * L * n], wa), f"A component {c}"


def test_keyswitch_reduced_level_key_offset():
    """L=20 of M=24: P key rows come from index i + (M - L)."""
    hp = params.openfhe_like(65536, 24, 6, 6)
    log_case("keyswitch n=65536 M=24 L=20")
    x, kb, ka, gb, ga = _keyswitch(hp, 20, 1, 8100)
    wb, wa = o.key_switch(x, hp.n, hp.q[:20], hp.q_psi[:20], hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
    assert np.array_equal(gb, wb) and np.array_equal(ga, wa)


def test_keyswitch_batch_equals_alone():
    """A component switched inside a batch of 5 equals the same component switched alone."""
    hp = params.openfhe_like(4096, 8, 3, 3)
    L, n = 8, 4096
    log_case("keyswitch batch-vs-alone n=4096")
    x, kb, ka, gb, ga = _keyswitch(hp, L, 5, 9000)
    qa = hp.q
    ctx = hip.Context(n, qa + hp.p, hp.q_psi + hp.p_psi)
    ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, 1)
    ks.upload_key(kb, ka, hp.M, hp.q)
    ws = hip.DeviceBuffer(ks.workspace_words(1))
    for c in range(5):
        xs = x[c * L * n:(c + 1) * L * n]
        ob, oa = hip.DeviceBuffer(L * n), hip.DeviceBuffer(L * n)
        dxs = _dev(xs)
        ks.run(dxs.ptr, 1, ob.ptr, oa.ptr, ws.ptr)
        ctx.synchronize()
        assert np.array_equal(ob.download(), gb[c * L * n:(c + 1) * L * n])
        assert np.array_equal(oa.download(), ga[c * L * n:(c + 1) * L * n])


def test_keyswitch_lanes_and_graph_replays():
    """A batch of 11 runs as two concurrent lanes (6 + 5 components on two
    streams, forked from and joined back to the caller's stream): equal to each
    component switched alone, eagerly and through a captured graph replayed
    three times with eager work and synchronizations between the replays."""
    hp = params.openfhe_like(4096, 8, 3, 3)
    L, n, B = 8, 4096, 11
    log_case("keyswitch lanes + graph n=4096 L=8 batch=11")
    x, kb, ka, gb, ga = _keyswitch(hp, L, B, 9500)
    ctx = hip.Context(n, hp.q + hp.p, hp.q_psi + hp.p_psi)
    one = hip.KeySwitch(ctx, L, hp.K, hp.alpha, 1)
    one.upload_key(kb, ka, hp.M, hp.q)
    ws1 = hip.DeviceBuffer(one.workspace_words(1))
    for c in (0, 5, 6, 10):  # both ends of both lanes
        sl = slice(c * L * n, (c + 1) * L * n)
        di, ob, oa = _dev(x[sl]), hip.DeviceBuffer(L * n), hip.DeviceBuffer(L * n)
        one.run(di.ptr, 1, ob.ptr, oa.ptr, ws1.ptr)
        ctx.synchronize()
        assert np.array_equal(ob.download(), gb[sl]) and np.array_equal(oa.download(), ga[sl]), f"component {c}"
    wb, wa = o.key_switch(x[:L * n], n, hp.q[:L], hp.q_psi[:L], hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
    assert np.array_equal(gb[:L * n], wb) and np.array_equal(ga[:L * n], wa)
    ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, B)
    ks.upload_key(kb, ka, hp.M, hp.q)
    d_in, d_b, d_a = _dev(x), hip.DeviceBuffer(B * L * n), hip.DeviceBuffer(B * L * n)
    ws = hip.DeviceBuffer(ks.workspace_words(B))
    st = hip.Stream(0)
    graph = ks.capture(d_in.ptr, B, d_b.ptr, d_a.ptr, ws.ptr, st.handle)
    zeros = np.zeros(B * L * n, np.uint64)
    e_in, e_b, e_a = _dev(x[:L * n]), hip.DeviceBuffer(L * n), hip.DeviceBuffer(L * n)
    for replay in range(3):
        d_b.upload(zeros)
        d_a.upload(zeros)
        one.run(e_in.ptr, 1, e_b.ptr, e_a.ptr, ws1.ptr)  # eager work between replays
        ctx.synchronize()
        graph.launch(st.handle)
        st.synchronize()
        assert np.array_equal(d_b.download(), gb) and np.array_equal(d_a.download(), ga), f"replay {replay}"


def test_keyswitch_full_batch_64_graph():
    """The measured configuration: N=2^16, L=24, K=6, dnum=4, 64 components
    per captured-graph launch (two lanes of 32, ModUp transforms and MAC
    interleaved by tower groups). Replayed twice; the components at both ends
    of both lanes equal the same component switched alone, and component 0
    equals the oracle's key switch."""
    hp = params.openfhe_like(65536, 24, 6, 6)
    L, n, B = 24, 65536, 64
    log_case("keyswitch n=65536 L=24 K=6 dnum=4 batch=64 graph")
    ctx = hip.Context(n, hp.q + hp.p, hp.q_psi + hp.p_psi)
    kb, ka = _key(hp, 9900)
    ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, B)
    ks.upload_key(kb, ka, hp.M, hp.q)
    x = np.concatenate([uniform(n, hp.q, 9901 + c) for c in range(B)])
    d_in, d_b, d_a = _dev(x), hip.DeviceBuffer(B * L * n), hip.DeviceBuffer(B * L * n)
    ws = hip.DeviceBuffer(ks.workspace_words(B))
    st = hip.Stream(0)
    graph = ks.capture(d_in.ptr, B, d_b.ptr, d_a.ptr, ws.ptr, st.handle)
    one = hip.KeySwitch(ctx, L, hp.K, hp.alpha, 1)
    one.upload_key(kb, ka, hp.M, hp.q)
    ws1 = hip.DeviceBuffer(one.workspace_words(1))
    alone = {}
    for c in (0, 31, 32, 63):
        sl = slice(c * L * n, (c + 1) * L * n)
        di, ob, oa = _dev(x[sl]), hip.DeviceBuffer(L * n), hip.DeviceBuffer(L * n)
        one.run(di.ptr, 1, ob.ptr, oa.ptr, ws1.ptr)
        ctx.synchronize()
        alone[c] = (ob.download(), oa.download())
    wb, wa = o.key_switch(x[:L * n], n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
    assert np.array_equal(alone[0][0], wb) and np.array_equal(alone[0][1], wa)
    zeros = np.zeros(B * L * n, np.uint64)
    for replay in range(2):
        d_b.upload(zeros)
        d_a.upload(zeros)
        graph.launch(st.handle)
        st.synchronize()
        gb, ga = d_b.download(), d_a.download()
        for c, (eb, ea) in alone.items():
            sl = slice(c * L * n, (c + 1) * L * n)
            assert np.array_equal(gb[sl], eb) and np.array_equal(ga[sl], ea), f"replay {replay} component {c}"
    # components across both lanes and the batch, straight against the oracle
    for c in (1, 13, 31, 32, 47, 63):
        sl = slice(c * L * n, (c + 1) * L * n)
        log_case(f"batch-64 graph component {c} vs oracle")
        wb, wa = o.key_switch(x[sl], n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(gb[sl], wb) and np.array_equal(ga[sl], wa), f"component {c} differs from the oracle"
    graph.close()
    for buf in (d_in, d_b, d_a, ws, ws1):
        buf.free()


def _scattered_keyswitch(hp, n, L, B, seed):
    """Key-switch B components once from one contiguous buffer (run) and once
    from separately allocated per-component buffers, allocated in a shuffled
    order (run_gather); returns both results per component."""
    ctx = hip.Context(n, hp.q[:L] + hp.p, hp.q_psi[:L] + hp.p_psi)
    kb, ka = _key(hp, seed)
    ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, B)
    ks.upload_key(kb, ka, hp.M, hp.q)
    x = np.concatenate([uniform(n, hp.q[:L], seed + 1 + c) for c in range(B)])
    w = L * n
    d_in, d_b, d_a = _dev(x), hip.DeviceBuffer(B * w), hip.DeviceBuffer(B * w)
    ws = hip.DeviceBuffer(ks.workspace_words(B))
    ks.run(d_in.ptr, B, d_b.ptr, d_a.ptr, ws.ptr)
    ctx.synchronize()
    gb, ga = d_b.download(), d_a.download()
    for buf in (d_in, d_b, d_a, ws):
        buf.free()
    order = np.random.default_rng(seed).permutation(B)
    ins, obs, oas = [None] * B, [None] * B, [None] * B
    for c in order:
        ins[c] = _dev(x[c * w:(c + 1) * w])
        obs[c], oas[c] = hip.DeviceBuffer(w), hip.DeviceBuffer(w)
    wsg = hip.DeviceBuffer(ks.gather_workspace_words(B))
    ks.run_gather([b.ptr for b in ins], [b.ptr for b in obs], [b.ptr for b in oas], wsg.ptr)
    ctx.synchronize()
    out = [(obs[c].download(), oas[c].download()) for c in range(B)]
    return x, kb, ka, gb, ga, out, ks, ctx


@pytest.mark.parametrize("n,M,K,alpha,B", [(4096, 8, 3, 3, 1), (4096, 8, 3, 3, 3), (4096, 8, 3, 3, 8),
                                           (4096, 8, 3, 3, 17), (65536, 24, 6, 6, 17)])
def test_keyswitch_gather_equals_contiguous(n, M, K, alpha, B):
    """run_gather on components at scattered device addresses equals run on
    the same components in one contiguous buffer, bit for bit, over single
    components, one lane, two lanes and the interleaved ModUp/MAC schedule,
    at N = 2^12 (one-pass transforms) and N = 2^16 (two-pass)."""
    hp = params.openfhe_like(n, M, K, alpha)
    L = M
    log_case(f"keyswitch gather n={n} L={L} K={K} batch={B}")
    x, kb, ka, gb, ga, out, _, _ = _scattered_keyswitch(hp, n, L, B, 7100 + B)
    w = L * n
    for c in range(B):
        assert np.array_equal(out[c][0], gb[c * w:(c + 1) * w]), f"B channel of component {c}"
        assert np.array_equal(out[c][1], ga[c * w:(c + 1) * w]), f"A channel of component {c}"
    if n == 4096:
        wb, wa = o.key_switch(x[:w], n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(out[0][0], wb) and np.array_equal(out[0][1], wa)


def test_keyswitch_gather_rejections():
    hp = params.openfhe_like(1024, 4, 2, 2)
    ctx = hip.Context(hp.n, hp.q + hp.p, hp.q_psi + hp.p_psi)
    ks = hip.KeySwitch(ctx, 4, hp.K, hp.alpha, 2)
    kb, ka = _key(hp, 7)
    ks.upload_key(kb, ka, hp.M, hp.q)
    w = 4 * hp.n
    bufs = [hip.DeviceBuffer(w) for _ in range(6)]
    ws = hip.DeviceBuffer(ks.gather_workspace_words(2))
    p = [b.ptr for b in bufs]
    with pytest.raises(hip.ContractError, match="null component pointer"):
        ks.run_gather([p[0], 0], p[2:4], p[4:6], ws.ptr)
    with pytest.raises(hip.ContractError, match="capacity"):
        ks.run_gather(p[0:3], p[0:3], p[3:6], ws.ptr)
    with pytest.raises(ValueError):
        ks.run_gather(p[0:2], p[2:3], p[4:6], ws.ptr)


@pytest.mark.parametrize("evaluation", [True, False])
def test_batched_modup_moddown_l24_k6_alpha6(evaluation):
    """N=2^16, Q: 24 towers (60 + 50-bit), P: 6 towers (60-bit): batched ModUp
    (evaluation or coefficient input) and ModDown, each bit-exact vs the oracle."""
    hp = params.openfhe_like(65536, 24, 6, 6)
    n, L, K, B = hp.n, 24, 6, 2
    ctx = hip.Context(n, hp.q + hp.p, hp.q_psi + hp.p_psi)
    ks = hip.KeySwitch(ctx, L, K, hp.alpha, B)
    x = np.concatenate([uniform(n, hp.q, 300 + c) for c in range(B)])
    log_case(f"batched modup/moddown L=24 K=6 alpha=6 evaluation={evaluation}")
    din = _dev(x)
    up = hip.DeviceBuffer(B * (L + K) * n)
    ws = hip.DeviceBuffer(max(ks.modup_workspace_words(B), ks.moddown_workspace_words(B)))
    ks.mod_up(din.ptr, B, up.ptr, ws.ptr, evaluation=evaluation)
    down = hip.DeviceBuffer(B * L * n)
    ks.mod_down(up.ptr, B, down.ptr, ws.ptr)
    ctx.synchronize()
    gu, gd = up.download(), down.download()
    for c in range(B):
        xs = x[c * L * n:(c + 1) * L * n]
        wu = o.mod_up(xs, n, hp.q, hp.q_psi, hp.p, hp.p_psi, evaluation=evaluation)
        assert np.array_equal(gu[c * (L + K) * n:(c + 1) * (L + K) * n], wu), f"ModUp component {c}"
        wd = o.mod_down(wu, n, hp.q, hp.q_psi, hp.p, hp.p_psi)
        assert np.array_equal(gd[c * L * n:(c + 1) * L * n], wd), f"ModDown component {c}"


def _rotation_oracle(hp, L, c0, c1, kb, ka, k):
    """Reference hybrid_rotate_many per rotation (openfhe_hybrid_adapter.hpp:343-355):
    digits -> MAC -> sigma_k(B + P c0), sigma_k(A) -> ModDown of both channels."""
    n, q, qr = hp.n, hp.q[:L], hp.q_psi[:L]
    digits = o.hybrid_digits(c1, n, q, qr, hp.p, hp.p_psi, hp.alpha)
    zb, za = o.hybrid_mac(digits, (L + hp.alpha - 1) // hp.alpha, kb, ka, n, q, hp.p, hp.M)
    rb, ra = o.hybrid_rotate(zb, za, c0, n, q, hp.p, k)
    return o.mod_down(rb, n, q, qr, hp.p, hp.p_psi), o.mod_down(ra, n, q, qr, hp.p, hp.p_psi)


def _hoisted(hp, L, autos, batch, seed):
    n = hp.n
    ctx = hip.Context(n, hp.q[:L] + hp.p, hp.q_psi[:L] + hp.p_psi)
    keys = []
    for r in range(len(autos)):
        ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, batch)
        kb, ka = _key(hp, seed + 37 * r)
        ks.upload_key(kb, ka, hp.M, hp.q)
        keys.append((ks, kb, ka))
    c0 = np.concatenate([uniform(n, hp.q[:L], seed + 500 + c) for c in range(batch)])
    c1 = np.concatenate([uniform(n, hp.q[:L], seed + 700 + c) for c in range(batch)])
    d0, d1 = _dev(c0), _dev(c1)
    ws = hip.DeviceBuffer(keys[0][0].rotate_workspace_words(batch))
    outs = [(hip.DeviceBuffer(batch * L * n), hip.DeviceBuffer(batch * L * n)) for _ in autos]
    hip.KeySwitch.rotate([k[0] for k in keys], autos, d0.ptr, d1.ptr, batch, [b.ptr for b, _ in outs],
                         [a.ptr for _, a in outs], ws.ptr)
    ctx.synchronize()
    return ctx, keys, c0, c1, [(b.download(), a.download()) for b, a in outs]


@pytest.mark.parametrize("n,M,K,alpha,L,batch,autos", [
    (1024, 6, 2, 2, 6, 2, [5, 25, 2047]),
    (1024, 6, 2, 2, 4, 3, [3, 5]),
    (65536, 24, 6, 6, 24, 1, [5, 131071]),
])
def test_hoisted_rotations(n, M, K, alpha, L, batch, autos):
    """One ModUp of c1 shared by every rotation; each rotation bit-exact vs the oracle."""
    hp = params.openfhe_like(n, M, K, alpha)
    log_case(f"hoisted rotations n={n} L={L} batch={batch} k={autos}")
    _, keys, c0, c1, got = _hoisted(hp, L, autos, batch, 4100 + L)
    for r, k in enumerate(autos):
        for c in range(batch):
            sl = slice(c * L * n, (c + 1) * L * n)
            wb, wa = _rotation_oracle(hp, L, c0[sl], c1[sl], keys[r][1], keys[r][2], k)
            assert np.array_equal(got[r][0][sl], wb), f"rotation {k}, component {c}, B"
            assert np.array_equal(got[r][1][sl], wa), f"rotation {k}, component {c}, A"


def test_identity_rotation_is_key_switch_plus_c0():
    """k = 1: ModDown(B + P c0) = ModDown(B) + c0 exactly, so the outputs are the
    plain key switch of c1 with c0 added to B (independent of the oracle)."""
    hp = params.openfhe_like(4096, 8, 3, 3)
    L, n, batch = 8, 4096, 2
    log_case("identity rotation vs key switch + c0")
    ctx, keys, c0, c1, got = _hoisted(hp, L, [1], batch, 4300)
    ks = keys[0][0]
    d1 = _dev(c1)
    ob, oa = hip.DeviceBuffer(batch * L * n), hip.DeviceBuffer(batch * L * n)
    ws = hip.DeviceBuffer(ks.workspace_words(batch))
    ks.run(d1.ptr, batch, ob.ptr, oa.ptr, ws.ptr)
    ctx.synchronize()
    kb_, ka_ = ob.download(), oa.download()
    qs = np.repeat(np.array(hp.q[:L], dtype=object), n)
    want_b = np.array([(int(b) + int(c)) % int(q) for b, c, q in
                       zip(kb_, c0, np.tile(qs, batch))], dtype=np.uint64)
    assert np.array_equal(got[0][0], want_b)
    assert np.array_equal(got[0][1], ka_)


def test_rotation_rejections():
    hp = params.openfhe_like(1024, 6, 2, 2)
    ctx = hip.Context(1024, hp.q + hp.p, hp.q_psi + hp.p_psi)
    a = hip.KeySwitch(ctx, 6, 2, 2, 1)
    other = hip.KeySwitch(ctx, 6, 2, 3, 1)  # same context, other digit size
    buf = hip.DeviceBuffer(8 * 1024)
    with pytest.raises(hip.ContractError):  # key not uploaded
        hip.KeySwitch.rotate([a], [5], buf.ptr, buf.ptr, 1, [buf.ptr], [buf.ptr], buf.ptr)
    kb, ka = _key(hp, 1)
    a.upload_key(kb, ka, hp.M, hp.q)
    hp3 = params.openfhe_like(1024, 6, 2, 3)
    assert hp3.q == hp.q and hp3.p == hp.p
    kb3, ka3 = _key(hp3, 2)
    other.upload_key(kb3, ka3, hp3.M, hp3.q)
    words = 6 * 1024
    c0, c1, o1, o2 = (hip.DeviceBuffer(words) for _ in range(4))
    ws = hip.DeviceBuffer(a.rotate_workspace_words(1))
    for keys, autos in (([a], [4]), ([a], [2048]), ([a, other], [5, 5])):  # even, out of range, plan mismatch
        with pytest.raises(hip.ContractError):
            hip.KeySwitch.rotate(keys, autos, c0.ptr, c1.ptr, 1, [o1.ptr] * len(keys), [o2.ptr] * len(keys),
                                 ws.ptr)
    # outputs over an input or over each other: every rotation re-reads c0 and c1
    for ob, oa in ((c0, o2), (o1, c1), (o1, o1), (c1, c0)):
        with pytest.raises(hip.ContractError, match="alias"):
            hip.KeySwitch.rotate([a], [5], c0.ptr, c1.ptr, 1, [ob.ptr], [oa.ptr], ws.ptr)
    with pytest.raises(hip.ContractError, match="alias"):
        hip.KeySwitch.rotate([a], [5], c0.ptr, c0.ptr, 1, [o1.ptr], [o2.ptr], ws.ptr)
    # a valid call after the rejections still runs
    hip.KeySwitch.rotate([a], [5], c0.ptr, c1.ptr, 1, [o1.ptr], [o2.ptr], ws.ptr)
    ctx.synchronize()


def test_hybrid_rotate_rejects_p_towers_inside_q(small):
    """ntt_hip_hybrid_rotate takes P tower indices of the context: an index past
    the basis or inside the rotated Q range is a contract error."""
    hp, ctx = small
    n, L, K = hp.n, hp.M, hp.K
    T = L + K
    bufs = [hip.DeviceBuffer(T * n) for _ in range(5)]
    for p in ([L + K], [0], [L - 1, L]):
        with pytest.raises(hip.ContractError, match="P tower"):
            ctx.hybrid_rotate(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, bufs[4].ptr, 0, T, L, p, 5)


def _add_towers(a, b, q, n):
    """(a + b) mod q per tower on tower-major arrays of whole components."""
    qv = np.tile(np.repeat(np.asarray(q, dtype=np.uint64), n), a.size // (len(q) * n))
    return (a + b) % qv


@pytest.mark.parametrize("n,M,K,alpha,L,batch,components", [
    (1024, 6, 2, 2, 5, 2, 2),
    (1024, 6, 2, 2, 5, 2, 3),
    (65536, 24, 6, 6, 24, 1, 2),
])
def test_key_switch_in_place(n, M, K, alpha, L, batch, components):
    """KeySwitchInPlace (openfhe_hybrid_adapter.hpp:221-248; operation graph
    doc :350-362): two components -> c0 + B, A; three -> c0 + B, c1 + A, with
    (B, A) the oracle's key switch of the last component."""
    hp = params.openfhe_like(n, M, K, alpha)
    log_case(f"key switch in place n={n} L={L} batch={batch} components={components}")
    q = hp.q[:L]
    ctx = hip.Context(n, q + hp.p, hp.q_psi[:L] + hp.p_psi)
    ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, batch)
    kb, ka = _key(hp, 6100 + L)
    ks.upload_key(kb, ka, hp.M, hp.q)
    cs = [np.concatenate([uniform(n, q, 6200 + 10 * j + c) for c in range(batch)]) for j in range(components)]
    dev = [_dev(c) for c in cs]
    ws = hip.DeviceBuffer(ks.workspace_words(batch))
    ks.in_place(dev[0].ptr, dev[1].ptr, batch, ws.ptr, c2_ptr=dev[2].ptr if components == 3 else None)
    ctx.synchronize()
    g0, g1 = dev[0].download(), dev[1].download()
    for c in range(batch):
        sl = slice(c * L * n, (c + 1) * L * n)
        wb, wa = o.key_switch(cs[-1][sl], n, q, hp.q_psi[:L], hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(g0[sl], _add_towers(cs[0][sl], wb, q, n)), f"c0 of component {c}"
        want1 = wa if components == 2 else _add_towers(cs[1][sl], wa, q, n)
        assert np.array_equal(g1[sl], want1), f"c1 of component {c}"
    if components == 3:
        assert np.array_equal(dev[2].download(), cs[2]), "c2 must be left as it was"


@pytest.mark.parametrize("extra", [0, 1, 2])
def test_relinearize(extra):
    """RelinearizeInPlace (openfhe_hybrid_adapter.hpp:250-286): each higher
    component switched with its own key and ModDown-reduced on its own, the
    contributions added into c0 / c1; extra = 0 leaves the ciphertext as is."""
    hp = params.openfhe_like(1024, 6, 2, 2)
    n, L, batch = hp.n, 6, 2
    log_case(f"relinearize n={n} L={L} batch={batch} extra={extra}")
    q = hp.q[:L]
    ctx = hip.Context(n, q + hp.p, hp.q_psi[:L] + hp.p_psi)
    keys = []
    for j in range(max(extra, 1)):
        ks = hip.KeySwitch(ctx, L, hp.K, hp.alpha, batch)
        kb, ka = _key(hp, 6500 + 41 * j)
        ks.upload_key(kb, ka, hp.M, hp.q)
        keys.append((ks, kb, ka))
    cs = [np.concatenate([uniform(n, q, 6600 + 10 * j + c) for c in range(batch)]) for j in range(2 + extra)]
    dev = [_dev(c) for c in cs]
    ws = hip.DeviceBuffer(keys[0][0].workspace_words(batch))
    hip.KeySwitch.relinearize([k[0] for k in keys[:extra]], dev[0].ptr, dev[1].ptr, [d.ptr for d in dev[2:]], batch,
                              ws.ptr)
    ctx.synchronize()
    g0, g1 = dev[0].download(), dev[1].download()
    for c in range(batch):
        sl = slice(c * L * n, (c + 1) * L * n)
        w0, w1 = cs[0][sl].copy(), cs[1][sl].copy()
        for j in range(extra):
            wb, wa = o.key_switch(cs[2 + j][sl], n, q, hp.q_psi[:L], hp.M, hp.p, hp.p_psi, hp.alpha, keys[j][1],
                                  keys[j][2])
            w0, w1 = _add_towers(w0, wb, q, n), _add_towers(w1, wa, q, n)
        assert np.array_equal(g0[sl], w0), f"c0 of component {c}"
        assert np.array_equal(g1[sl], w1), f"c1 of component {c}"


def test_relinearize_rejections():
    hp = params.openfhe_like(1024, 6, 2, 2)
    ctx = hip.Context(hp.n, hp.q + hp.p, hp.q_psi + hp.p_psi)
    ks = hip.KeySwitch(ctx, 6, 2, 2, 1)
    other = hip.KeySwitch(ctx, 6, 2, 3, 1)
    kb, ka = _key(hp, 6900)
    ks.upload_key(kb, ka, hp.M, hp.q)
    d = hip.DeviceBuffer(6 * hp.n)
    ws = hip.DeviceBuffer(ks.workspace_words(1))
    d2 = hip.DeviceBuffer(6 * hp.n)
    with pytest.raises(hip.ContractError, match="not uploaded"):
        other.in_place(d.ptr, d2.ptr, 1, ws.ptr)
    with pytest.raises(hip.ContractError, match="alias"):
        ks.in_place(d.ptr, d.ptr, 1, ws.ptr)
    with pytest.raises(hip.ContractError, match="alias"):
        ks.in_place(d.ptr, d2.ptr, 1, ws.ptr, c2_ptr=d2.ptr)
    with pytest.raises(hip.ContractError, match="alias"):
        hip.KeySwitch.relinearize([ks], d.ptr, d2.ptr, [d2.ptr], 1, ws.ptr)
    d3 = hip.DeviceBuffer(6 * hp.n)
    with pytest.raises(hip.ContractError, match="share the context"):
        hip.KeySwitch.relinearize([ks, other], d.ptr, d2.ptr, [d3.ptr, d3.ptr], 1, ws.ptr)
    with pytest.raises(hip.ContractError, match="capacity"):
        ks.in_place(d.ptr, d2.ptr, 2, ws.ptr)
    with pytest.raises(hip.ContractError, match="one relinearization key"):
        hip.KeySwitch.relinearize([ks], d.ptr, d2.ptr, [], 1, ws.ptr)
    # every launch must fit the 65536 workgroup rows HIP allows: 80 rows per
    # component here (T = 8, 3 digits x 6 complement towers = 18 jobs)
    assert hip.KeySwitch(ctx, 6, 2, 2, 65535 // 80).max_batch == 65535 // 80
    with pytest.raises(hip.ContractError, match="rows of one launch"):
        hip.KeySwitch(ctx, 6, 2, 2, 65535 // 80 + 1)


@pytest.mark.parametrize("bits", [59, 58, 57, 50])
def test_keyswitch_headroom_classes(bits):
    """Key switch at N=2^16 on bases whose moduli all sit at the top of one
    forward headroom class (the complement transforms and the ModDown
    epilogue, which brings the fewer-reduced output below 8q first), vs the oracle."""
    n, L, K = 65536, 4, 2
    qs = [params.last_prime(bits, 2 * n)]
    while len(qs) < L + K:
        qs.append(params.previous_prime(qs[-1], 2 * n))
    hp = params.HybridParams(n=n, q=qs[:L], p=qs[L:], alpha=2)
    log_case(f"keyswitch headroom {bits}-bit n=65536 L=4 K=2 alpha=2")
    _check_keyswitch(hp, L, 1, 15000 + bits)
