"""GPU parity of the gfx950 NTT and element-wise kernels against the CPU oracle.

Bit-exact on full arrays (integer work): the reference's seven patterns on the
pinned OpenFHE QP towers (tests/cuda_openfhe_tests.cpp:67-155), the headline
NTT configurations, every ring dimension 2..2^16, and 62-63-bit primes that
take the canonical (non-lazy) butterfly.
"""
import numpy as np
import pytest

import oracle as o
from common import PINNED, log_case, patterns, uniform, lcg_values, SEED
from nttgpu import hip, params

pytestmark = pytest.mark.gpu
N16 = 65536


def _gpu_round(ctx, values, first, towers, inverse):
    buf = hip.DeviceBuffer(values.size).upload(values)
    ctx.transform(buf.ptr, first, towers, inverse)
    ctx.synchronize()
    return buf.download()


@pytest.fixture(scope="module")
def pinned_ctx():
    return hip.Context(N16, PINNED["q"], PINNED["psi"])


@pytest.mark.parametrize("subset", ["Q5", "Q4P", "Q1", "QP8"])
def test_seven_patterns_pinned_towers(pinned_ctx, subset):
    q, psi = PINNED["q"], PINNED["psi"]
    sel = {"Q5": list(range(5)), "Q4P": [0, 1, 2, 3, 5, 6, 7], "Q1": [0], "QP8": list(range(8))}[subset]
    # a context per subset keeps towers contiguous, as the reference does per basis
    ctx = pinned_ctx if subset == "QP8" else hip.Context(N16, [q[i] for i in sel], [psi[i] for i in sel])
    qs, rs = [q[i] for i in sel], [psi[i] for i in sel]
    T = len(sel)
    for pattern, x in patterns(N16, qs):
        log_case(f"ntt pinned {subset} pattern {pattern}")
        want_f = o.ntt_towers(x, N16, qs, rs)
        got_f = _gpu_round(ctx, x, 0, T, False)
        assert np.array_equal(got_f, want_f), f"forward pattern {pattern}"
        assert np.array_equal(_gpu_round(ctx, got_f, 0, T, True), x), f"round trip pattern {pattern}"
        want_i = o.ntt_towers(x, N16, qs, rs, inverse=True)
        assert np.array_equal(_gpu_round(ctx, x, 0, T, True), want_i), f"independent inverse pattern {pattern}"
        # pointwise square then inverse == negacyclic ring product x*x
        a = hip.DeviceBuffer(x.size).upload(want_f)
        b = hip.DeviceBuffer(x.size).upload(want_f)
        ctx.pointwise(a.ptr, b.ptr, 0, T)
        ctx.transform(a.ptr, 0, T, True)
        ctx.synchronize()
        want_prod = o.ntt_towers(o.pointwise_towers(want_f, want_f, N16, qs), N16, qs, rs, inverse=True)
        assert np.array_equal(a.download(), want_prod), f"ring product pattern {pattern}"


def test_single_tower_n65536_59bit():
    """One NTT-friendly prime q ~ 2^59, uniform seeded input, forward then inverse."""
    q = params.last_prime(59, 2 * N16)
    psi = params.min_root(q, N16)
    ctx = hip.Context(N16, [q], [psi])
    x = lcg_values(N16, [q])
    log_case("ntt single tower n=65536, 59-bit prime")
    f = _gpu_round(ctx, x, 0, 1, False)
    assert np.array_equal(f, o.ntt_forward(x, q, psi))
    assert np.array_equal(_gpu_round(ctx, f, 0, 1, True), x)


@pytest.mark.parametrize("bits", [59, 30])
def test_full_size_2048_towers(bits):
    """The roofline workload at full size: 2048 towers of one prime (1 GiB) in
    one standalone call, which runs them in Infinity-Cache-sized groups (four
    of 448 and one of 256). The towers on both sides of every group boundary
    are bit-exact against the oracle, and forward then inverse returns every
    tower."""
    q = params.last_prime(bits, 2 * N16)
    psi = params.min_root(q, N16)
    ctx = hip.Context(N16, [q], [psi])
    T = 2048
    x = np.random.default_rng(bits).integers(0, q, T * N16, dtype=np.uint64)
    log_case(f"ntt full size: 2048 towers, {bits}-bit prime")
    buf = hip.DeviceBuffer(x.size).upload(x)
    ctx.transform_batch(buf.ptr, T, N16, 0, 1, False)
    ctx.synchronize()
    f = buf.download()
    for t in (0, 447, 448, 895, 896, 1343, 1344, 1791, 1792, 2047):
        sl = slice(t * N16, (t + 1) * N16)
        assert np.array_equal(f[sl], o.ntt_forward(x[sl], q, psi)), f"tower {t}"
    ctx.transform_batch(buf.ptr, T, N16, 0, 1, True)
    ctx.synchronize()
    assert np.array_equal(buf.download(), x)
    buf.free()


def test_batched_rns_24_primes_2_polys():
    """24 distinct 50-60-bit primes, 2 polynomials = 48 tower transforms in one launch."""
    qs = params.mixed_width_primes(N16, 24)
    rs = [params.min_root(q, N16) for q in qs]
    ctx = hip.Context(N16, qs, rs)
    x = uniform(N16, qs + qs, SEED)
    log_case("ntt batched 24 primes x 2 polynomials (48 towers)")
    buf = hip.DeviceBuffer(x.size).upload(x)
    ctx.transform_batch(buf.ptr, 2, 24 * N16, 0, 24, False)
    ctx.synchronize()
    got = buf.download()
    want = np.concatenate([o.ntt_towers(x[:24 * N16], N16, qs, rs), o.ntt_towers(x[24 * N16:], N16, qs, rs)])
    assert np.array_equal(got, want)
    ctx.transform_batch(buf.ptr, 2, 24 * N16, 0, 24, True)
    ctx.synchronize()
    assert np.array_equal(buf.download(), x)


@pytest.mark.parametrize("log_n", list(range(1, 17)))
def test_every_ring_dimension(log_n):
    n = 1 << log_n
    qs = params.find_primes(max(n, 2), 3, 60)
    rs = [params.find_psi(q, n) for q in qs]
    ctx = hip.Context(n, qs, rs)
    x = uniform(n, qs, log_n)
    log_case(f"ntt n=2^{log_n}")
    f = _gpu_round(ctx, x, 0, 3, False)
    assert np.array_equal(f, o.ntt_towers(x, n, qs, rs))
    assert np.array_equal(_gpu_round(ctx, f, 0, 3, True), x)
    assert np.array_equal(_gpu_round(ctx, x, 0, 3, True), o.ntt_towers(x, n, qs, rs, inverse=True))


@pytest.mark.parametrize("n", [4096, 65536])
def test_wide_primes_canonical_path(n):
    """62- and 63-bit primes disable the lazy [0,4q) butterflies."""
    qs = params.find_primes(n, 2, 63) + params.find_primes(n, 1, 62)
    rs = [params.find_psi(q, n) for q in qs]
    ctx = hip.Context(n, qs, rs)
    x = uniform(n, qs, 99)
    log_case(f"ntt wide primes n={n}")
    f = _gpu_round(ctx, x, 0, 3, False)
    assert np.array_equal(f, o.ntt_towers(x, n, qs, rs))
    assert np.array_equal(_gpu_round(ctx, f, 0, 3, True), x)


@pytest.mark.parametrize("n", [4096, 65536])
def test_wide_lazy_boundary_primes(n):
    """Largest NTT primes below 2^61 (8q just under 2^64: the edge of the
    wide-lazy range) with all-(q-1) and seeded inputs, both directions."""
    qs = [params.last_prime(61, 2 * n)]
    qs.append(params.previous_prime(qs[0], 2 * n))
    rs = [params.find_psi(q, n) for q in qs]
    ctx = hip.Context(n, qs, rs)
    log_case(f"ntt wide-lazy boundary n={n} q={qs}")
    for x in (np.concatenate([np.full(n, q - 1, np.uint64) for q in qs]), uniform(n, qs, 61)):
        f = _gpu_round(ctx, x, 0, 2, False)
        assert np.array_equal(f, o.ntt_towers(x, n, qs, rs))
        assert np.array_equal(_gpu_round(ctx, f, 0, 2, True), x)
        assert np.array_equal(_gpu_round(ctx, x, 0, 2, True), o.ntt_towers(x, n, qs, rs, inverse=True))


@pytest.mark.parametrize("n", [2, 256, 4096, 16384, 65536])
def test_narrow_primes(n):
    """Towers below 2^30 run the 32-bit butterflies (every launch shape: whole
    transform in LDS, column pass for R < 256, the R = 256 column pass): the
    largest NTT prime below 2^30 (4q just under 2^32), 30-bit primes from the
    reference generator and the smallest class member, with all-(q-1) and
    seeded inputs, both directions."""
    m = 2 * max(n, 2)
    qs = [params.last_prime(30, m)] + params.find_primes(max(n, 2), 2, 30)
    small = next(q for q in range(m + 1, 1 << 30, m) if params.is_prime(q))
    qs.append(small)
    rs = [params.find_psi(q, n) for q in qs]
    ctx = hip.Context(n, qs, rs)
    log_case(f"ntt narrow primes n={n} q={qs}")
    T = len(qs)
    for x in (np.concatenate([np.full(n, q - 1, np.uint64) for q in qs]), uniform(n, qs, 30)):
        f = _gpu_round(ctx, x, 0, T, False)
        assert np.array_equal(f, o.ntt_towers(x, n, qs, rs))
        assert np.array_equal(_gpu_round(ctx, f, 0, T, True), x)
        assert np.array_equal(_gpu_round(ctx, x, 0, T, True), o.ntt_towers(x, n, qs, rs, inverse=True))


def test_primes_below_2_24_in_a_wide_context():
    """Primes below 2^24 beside 50- and 60-bit primes at N = 2^16: such a
    context runs headroom class 0 (the quotient estimate of the higher
    classes needs moduli of 24 bits or more), on its own towers and on the
    small towers alone; forward and inverse equal the oracle, all-(q-1) and
    seeded inputs."""
    m = 2 * N16
    qs = [params.last_prime(60, m), params.last_prime(50, m), params.last_prime(23, m), params.last_prime(20, m)]
    rs = [params.find_psi(q, N16) for q in qs]
    ctx = hip.Context(N16, qs, rs)
    log_case(f"ntt n=65536 primes below 2^24 beside wide ones q={qs}")
    T = len(qs)
    for x in (np.concatenate([np.full(N16, q - 1, np.uint64) for q in qs]), uniform(N16, qs, 31)):
        f = _gpu_round(ctx, x, 0, T, False)
        assert np.array_equal(f, o.ntt_towers(x, N16, qs, rs))
        assert np.array_equal(_gpu_round(ctx, f, 0, T, True), x)
        assert np.array_equal(_gpu_round(ctx, x, 0, T, True), o.ntt_towers(x, N16, qs, rs, inverse=True))
        tail = x[2 * N16:]
        assert np.array_equal(_gpu_round(ctx, tail, 2, 2, False), o.ntt_towers(tail, N16, qs[2:], rs[2:]))
        assert np.array_equal(_gpu_round(ctx, tail, 2, 2, True), o.ntt_towers(tail, N16, qs[2:], rs[2:], inverse=True))


def test_narrow_packed_handoff_out_of_place_and_strided():
    """32-bit towers at N = 2^16 hand their values between the passes packed in
    the first half of each tower: out of place in both directions (the source
    stays untouched), and a batch of polynomials whose stride leaves a gap
    tower after each one (the gaps keep their contents), all against the
    oracle."""
    n = N16
    qs = [params.last_prime(30, 2 * n)] + params.find_primes(n, 2, 30)
    rs = [params.find_psi(q, n) for q in qs]
    ctx = hip.Context(n, qs, rs)
    T = len(qs)
    x = uniform(n, qs, 31)
    want = o.ntt_towers(x, n, qs, rs)
    log_case("ntt narrow packed hand-off, out of place")
    src, dst, back = hip.DeviceBuffer(x.size).upload(x), hip.DeviceBuffer(x.size), hip.DeviceBuffer(x.size)
    ctx.transform(dst.ptr, 0, T, False, src_ptr=src.ptr)
    ctx.synchronize()
    assert np.array_equal(dst.download(), want)
    assert np.array_equal(src.download(), x)
    ctx.transform(back.ptr, 0, T, True, src_ptr=dst.ptr)
    ctx.synchronize()
    assert np.array_equal(back.download(), x)
    assert np.array_equal(dst.download(), want)
    log_case("ntt narrow packed hand-off, strided batch with gap towers")
    P, stride = 4, (T + 1) * n
    polys = [uniform(n, qs, 40 + p) for p in range(P)]
    gap = np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    buf = np.concatenate([np.concatenate([polys[p], gap]) for p in range(P)])
    d = hip.DeviceBuffer(buf.size).upload(buf)
    ctx.transform_batch(d.ptr, P, stride, 0, T, False)
    ctx.synchronize()
    got = d.download()
    for p in range(P):
        assert np.array_equal(got[p * stride:p * stride + T * n], o.ntt_towers(polys[p], n, qs, rs)), f"poly {p}"
        assert np.array_equal(got[p * stride + T * n:(p + 1) * stride], gap), f"gap after poly {p}"
    ctx.transform_batch(d.ptr, P, stride, 0, T, True)
    ctx.synchronize()
    assert np.array_equal(d.download(), buf)


def test_pointwise_every_width():
    """Barrett variable x variable products (lazy form below 2^62, exact form
    up to 2^63) on primes of 30, 50, 60, 61.9, 62.5 and ~63 bits: all-(q-1),
    ones and seeded operands against the oracle's 128-bit remainder."""
    n = 4096
    qs = [params.last_prime(30, 2 * n), params.last_prime(50, 2 * n), params.last_prime(60, 2 * n),
          params.last_prime(62, 2 * n), params.find_primes(n, 1, 63)[0], params.last_prime(63, 2 * n)]
    rs = [params.find_psi(q, n) for q in qs]
    ctx = hip.Context(n, qs, rs)
    log_case(f"pointwise widths q={qs}")
    T = len(qs)
    cases = [(np.concatenate([np.full(n, q - 1, np.uint64) for q in qs]),) * 2,
             (np.concatenate([np.full(n, q - 1, np.uint64) for q in qs]), np.ones(T * n, np.uint64)),
             (uniform(n, qs, 3), uniform(n, qs, 4))]
    for x, y in cases:
        a, b = hip.DeviceBuffer(x.size).upload(x), hip.DeviceBuffer(y.size).upload(y)
        ctx.pointwise(a.ptr, b.ptr, 0, T)
        ctx.synchronize()
        assert np.array_equal(a.download(), o.pointwise_towers(x, y, n, qs))


def test_out_of_place_and_subrange():
    qs = params.mixed_width_primes(N16, 6)
    rs = [params.min_root(q, N16) for q in qs]
    ctx = hip.Context(N16, qs, rs)
    x = uniform(N16, qs[2:5], 5)
    log_case("ntt out of place, towers 2..5")
    src = hip.DeviceBuffer(x.size).upload(x)
    dst = hip.DeviceBuffer(x.size)
    ctx.transform(dst.ptr, 2, 3, False, src_ptr=src.ptr)
    ctx.synchronize()
    assert np.array_equal(dst.download(), o.ntt_towers(x, N16, qs[2:5], rs[2:5]))
    assert np.array_equal(src.download(), x)  # source untouched
    ctx.transform(src.ptr, 2, 3, True, src_ptr=dst.ptr)
    ctx.synchronize()
    assert np.array_equal(src.download(), x)


def _top_primes(bits, count, n=N16):
    """The `count` largest NTT primes below 2^bits (q = 1 mod 2n): the top of a
    forward headroom class, where the skipped reductions leave the least room."""
    qs = [params.last_prime(bits, 2 * n)]
    while len(qs) < count:
        qs.append(params.previous_prime(qs[-1], 2 * n))
    return qs


def test_forward_class_follows_the_transformed_towers():
    """A context of 60-, 50-, 40- and 59-bit top-of-class primes: a launch
    over a range of towers takes the headroom class of that range's own widest
    prime (ranges [0,4), [1,3), [2,3), [1,2), [3,4)), bit-exact vs the oracle
    with every residue at q - 1 and seeded."""
    qs = [_top_primes(60, 1)[0], _top_primes(50, 1)[0], _top_primes(40, 1)[0], _top_primes(59, 1)[0]]
    rs = [params.min_root(q, N16) for q in qs]
    ctx = hip.Context(N16, qs, rs)
    for first, towers in [(0, 4), (1, 2), (2, 1), (1, 1), (3, 1)]:
        sub_q, sub_r = qs[first:first + towers], rs[first:first + towers]
        for name, x in [("q-1", np.concatenate([np.full(N16, q - 1, np.uint64) for q in sub_q])),
                        ("uniform", uniform(N16, sub_q, 31 * first + towers))]:
            log_case(f"mixed-class forward towers [{first}, {first + towers}) {name}")
            got = _gpu_round(ctx, x, first, towers, False)  # the buffer holds the range's towers
            assert np.array_equal(got, o.ntt_towers(x, N16, sub_q, sub_r)), f"[{first}, {first + towers}) {name}"


@pytest.mark.parametrize("bits", [61, 60, 59, 58, 57, 56, 50])
def test_forward_headroom_classes(bits):
    """The forward transform at N=2^16 skips x reductions while residues stay
    below 2^64 (every q < 2^(61 - H): 7, 3, 1 or 0 reductions over 16 stages).
    Top-of-class primes, patterns that drive every residue to its largest value
    (q - 1 everywhere) and seeded ones: bit-exact vs the oracle, and the round
    trip."""
    qs = _top_primes(bits, 3)
    rs = [params.min_root(q, N16) for q in qs]
    ctx = hip.Context(N16, qs, rs)
    for name, x in [("q-1", np.concatenate([np.full(N16, q - 1, np.uint64) for q in qs])),
                    ("uniform", uniform(N16, qs, bits)),
                    ("x^(N-1)", np.concatenate([(np.arange(N16) == N16 - 1).astype(np.uint64)] * 3))]:
        log_case(f"forward headroom {bits}-bit {name}")
        got = _gpu_round(ctx, x, 0, 3, False)
        assert np.array_equal(got, o.ntt_towers(x, N16, qs, rs)), f"{bits}-bit forward ({name})"
        assert np.array_equal(_gpu_round(ctx, got, 0, 3, True), x), f"{bits}-bit round trip ({name})"


def _mixed_contexts():
    """Contexts of every two-pass ring dimension, wide-lazy (60- and 50-bit
    top-of-range primes) and 32-bit (30-bit primes)."""
    ctxs = []
    for n in (8192, 16384, 32768, 65536):
        qs = [params.last_prime(60, 2 * n), params.last_prime(50, 2 * n)]
        ctxs.append((n, qs, [params.min_root(q, n) for q in qs]))
    for n in (8192, 65536):
        qs = [params.last_prime(30, 2 * n), params.previous_prime(params.last_prime(30, 2 * n), 2 * n)]
        ctxs.append((n, qs, [params.min_root(q, n) for q in qs]))
    return [(n, qs, rs, hip.Context(n, qs, rs)) for n, qs, rs in ctxs]


def test_mixed_ring_dimensions_in_one_launch():
    """Towers of N = 2^13 .. 2^16 from six contexts (wide-lazy and 32-bit) in
    one mixed launch per pass, job order shuffled, some read out of place:
    every tower equals the oracle's transform of its own N, and the inverse
    launch restores the inputs."""
    ctxs = _mixed_contexts()
    rng = np.random.default_rng(2024)
    jobs, cases = [], []
    for ci, (n, qs, rs, ctx) in enumerate(ctxs):
        for t in range(len(qs)):
            for rep in range(2):
                x = uniform(n, [qs[t]], 100 * ci + 10 * t + rep)
                if rep == 1:
                    x[:] = qs[t] - 1  # every residue at its largest value
                buf = hip.DeviceBuffer(n).upload(x)
                src = hip.DeviceBuffer(n).upload(x) if (ci + t + rep) % 3 == 0 else None
                cases.append((n, qs[t], rs[t], x, buf, src))
                jobs.append((ctx, buf.ptr, t, src.ptr if src else None))
    order = rng.permutation(len(jobs))
    jobs = [jobs[i] for i in order]
    cases = [cases[i] for i in order]
    log_case(f"mixed launch: {len(jobs)} towers over N = 2^13..2^16, 60/50/30-bit")
    st = hip.Stream(0)
    ws = hip.transform_mixed(jobs, inverse=False, stream=st.handle)
    st.synchronize()
    for n, q, r, x, buf, src in cases:
        assert np.array_equal(buf.download(), o.ntt_towers(x, n, [q], [r])), f"forward N={n} q={q}"
    hip.transform_mixed([(j[0], j[1], j[2]) for j in jobs], inverse=True, stream=st.handle, workspace=ws)
    st.synchronize()
    for n, q, r, x, buf, src in cases:
        assert np.array_equal(buf.download(), x), f"round trip N={n} q={q}"


def test_mixed_launch_strides_over_more_work_than_its_grid():
    """300 towers (2^16 and 2^14, interleaved): the persistent grid holds far
    fewer workgroups than the 4000+ work items, so every workgroup strides."""
    ctxs = _mixed_contexts()
    (n1, q1, r1, c1), (n2, q2, r2, c2) = ctxs[3], ctxs[1]
    towers = []
    for k in range(300):
        n, qs, rs, ctx = (n1, q1, r1, c1) if k % 3 else (n2, q2, r2, c2)
        t = k % 2
        towers.append((n, qs[t], rs[t], ctx, t, uniform(n, [qs[t]], 7000 + k)))
    big = hip.DeviceBuffer(sum(n for n, *_ in towers))
    offs, off = [], 0
    for n, q, r, ctx, t, x in towers:
        big.upload(x, offset=off)
        offs.append(off)
        off += n
    log_case("mixed launch: 300 towers, persistent striding")
    st = hip.Stream(0)
    ws = hip.transform_mixed([(ctx, big.at(o_), t) for (n, q, r, ctx, t, x), o_ in zip(towers, offs)],
                             inverse=False, stream=st.handle)
    st.synchronize()  # the job table in `ws` stays allocated until the launch has run
    del ws
    got = big.download()
    for (n, q, r, ctx, t, x), o_ in zip(towers, offs):
        assert np.array_equal(got[o_:o_ + n], o.ntt_towers(x, n, [q], [r])), f"tower at {o_} (N={n})"


def test_mixed_launch_rejections():
    small_q = [params.last_prime(50, 2 * 4096)]
    small = hip.Context(4096, small_q, [params.min_root(small_q[0], 4096)])
    wide_q = params.find_primes(8192, 1, 62)
    wide = hip.Context(8192, wide_q, [params.find_psi(wide_q[0], 8192)])
    ok_q = [params.last_prime(50, 2 * 8192)]
    ok = hip.Context(8192, ok_q, [params.min_root(ok_q[0], 8192)])
    buf = hip.DeviceBuffer(8192)
    with pytest.raises(hip.ContractError, match="2\\^13"):
        hip.transform_mixed([(small, buf.ptr, 0)], inverse=False)
    with pytest.raises(hip.ContractError, match="below 2\\^61"):
        hip.transform_mixed([(ok, buf.ptr, 0), (wide, buf.ptr, 0)], inverse=False)
    with pytest.raises(hip.ContractError, match="outside"):
        hip.transform_mixed([(ok, buf.ptr, 1)], inverse=False)
