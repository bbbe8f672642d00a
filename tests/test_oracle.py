"""Pin the CPU oracle before trusting it.

Every oracle function is checked against the reference's own known answers,
closed forms or an independent arbitrary-precision restatement (Python ints),
chosen so that plausible mistakes (dropped term, wrong sign/index, transposed
operand, canonical instead of unreduced CRT) fail.
"""
import json
import math
import os
import random

import numpy as np
import pytest

import oracle as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
N16 = 65536
PINNED = json.load(open(os.path.join(GOLDEN, "pinned_openfhe_qp.json")))


# ---- arithmetic: tests/unit_tests.cpp:29-39, tests/oracle/test_arithmetic.py:14-94 ----

def test_mod_arith_kats():
    q = 1000000007
    assert o.add_mod(q - 1, 2, q) == 1
    assert o.sub_mod(1, 2, q) == q - 1
    assert o.mul_mod(q - 1, q - 1, q) == 1
    assert o.pow_mod(2, q - 1, q) == 1
    for x in range(1, 50):
        assert o.mul_mod(o.inv_mod(x, q), x, q) == 1
    assert o.pow_mod(5, 0, q) == 1
    assert o.pow_mod(5, 0, 1) == 0


def test_arithmetic_against_bigints():
    rng = random.Random(0xD6)
    moduli = [2, 17, 1000000007]
    for bits in (30, 32, 50, 55, 60, 63):
        moduli += [(1 << bits) - 1, (1 << bits) - 25]
    moduli += [rng.randrange(3, 1 << 63) for _ in range(12)]
    corrections, carries = set(), set()
    mask = (1 << 64) - 1
    for q in moduli:
        edges = sorted({v % q for v in [0, 1, 2, q // 2, q - 2, q - 1, (1 << 32) - 1, 1 << 32, (1 << 32) + 1]})
        pairs = [(a, b) for a in edges for b in edges] + [(rng.randrange(q), rng.randrange(q)) for _ in range(64)]
        mu = (1 << 128) // q
        for a, b in pairs:
            assert o.add_mod(a, b, q) == (a + b) % q
            assert o.sub_mod(a, b, q) == (a - b) % q
            assert o.mul_mod(a, b, q) == a * b % q
            if q > 1:
                assert o.shoup_multiply(a, b, o.shoup_precompute(b, q), q) == a * b % q
                assert o.barrett_multiply(a, b, q) == a * b % q
                est = a * b * mu >> 128
                low, high = (a * b) & mask, (a * b) >> 64
                assert o.barrett_quotient(low, high, q) == est
                corrections.add(a * b // q - est)
                carries.add(((low * (mu & mask) >> 64) + ((low * (mu >> 64)) & mask) +
                             ((high * (mu & mask)) & mask)) >> 64)
        for a in edges + [mask]:
            for e in (0, 1, 2, 63, 64, mask):
                assert o.pow_mod(a, e, q) == pow(a, e, q)
        assert o.barrett_precompute(q) == (mu & mask, mu >> 64)
    # Barrett's estimate is exact or one short; the middle carry takes 0..2.
    assert corrections == {0, 1}
    assert carries == {0, 1, 2}
    for q in [17, 1000000007] + [o.find_primes(N16, 1, b)[0] for b in (50, 55, 60)]:
        for x in (1, q - 1, q + 1, mask):
            if x % q:
                assert o.inv_mod(x, q) == pow(x, -1, q)
        for x in (0, q):
            with pytest.raises(ValueError):
                o.inv_mod(x, q)


def test_bit_reverse_kats():
    # tests/unit_tests.cpp:41-48
    assert [o.bit_reverse(i, 3) for i in range(8)] == [0, 4, 2, 6, 1, 5, 3, 7]
    assert o.bit_reverse(0b110, 3) == 0b011
    assert o.log2(1) == 0 and o.log2(1024) == 10


def _py_is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_primes_and_roots():
    # find_primes walk (include/ntt/param_gen.hpp:64-79): ascending, (2n) | (q-1), in range.
    for n, bits in ((1024, 46), (65536, 50), (65536, 60)):
        primes = o.find_primes(n, 5, bits)
        assert primes == sorted(primes) and len(set(primes)) == 5
        for q in primes:
            assert _py_is_prime(q) and (q - 1) % (2 * n) == 0 and (1 << (bits - 1)) <= q < (1 << bits)
        first = ((1 << (bits - 1)) // (2 * n) + 1) * 2 * n + 1
        assert all(not _py_is_prime(c) for c in range(first, primes[0], 2 * n))
        for q in primes:
            psi = o.find_psi(q, n)
            assert pow(psi, n, q) == q - 1  # order exactly 2n
    for c in (561, 1105, 3215031751, 2**61 - 1, 2**62 - 57, 2**64 - 59):
        assert o.is_prime(c) == _py_is_prime(c)


def test_pinned_openfhe_parameters():
    """The eight pinned OpenFHE QP towers (tests/cuda_openfhe_tests.cpp:139-145):
    LastPrime/PreviousPrime chain and the minimum primitive root convention
    (tests/openfhe_representation_tests.cpp:55-62)."""
    q, psi = PINNED["q"], PINNED["psi"]
    m = 2 * N16
    assert o.last_prime(60, m) == q[0]
    chain = [q[0]]
    for _ in range(3):
        chain.append(o.previous_prime(chain[-1], m))
    assert chain[1:] == q[5:8]
    for qi, r in zip(q, psi):
        assert _py_is_prime(qi) and qi % m == 1
        assert o.min_root(qi, N16) == r
        assert pow(r, N16, qi) == qi - 1


# ---- transform: include/ntt/ntt.hpp, docs/modular-arithmetic-contract.md:200-209 ----

def test_four_point_worked_example():
    fwd = o.ntt_forward([1, 15, 3, 15], 17, 2)
    assert list(fwd) == [10, 16, 3, 9]
    assert list(o.ntt_table([1, 15, 3, 15], 17, 2)) == [10, 16, 3, 9]
    assert list(o.ntt_inverse([10, 16, 3, 9], 17, 2)) == [1, 15, 3, 15]
    assert list(o.ntt_table([10, 16, 3, 9], 17, 2, inverse=True)) == [1, 15, 3, 15]


@pytest.mark.parametrize("n", [2, 8, 64, 256])
def test_forward_matches_naive_definition(n):
    rng = random.Random(n)
    for bits in (30, 50, 60):
        q = o.find_primes(n, 1, bits)[0]
        psi = o.find_psi(q, n)
        a = [rng.randrange(q) for _ in range(n)]
        # independent pure-Python bigint evaluation at the odd bit-reversed points
        lg = n.bit_length() - 1
        want = [sum(a[i] * pow(psi, (2 * o.bit_reverse(k, lg) + 1) * i, q) for i in range(n)) % q
                for k in range(n)]
        assert list(o.ntt_forward(a, q, psi)) == want
        assert list(o.ntt_table(a, q, psi)) == want
        assert list(o.ntt_naive(a, q, psi)) == want
        assert list(o.ntt_inverse(want, q, psi)) == a
        assert list(o.ntt_table(want, q, psi, inverse=True)) == a


def _patterns(n, q, seed):
    rng = np.random.default_rng(seed)
    out = [np.zeros(n, np.uint64), np.full(n, q - 1, np.uint64),
           np.array([q - 1 if i % 2 else 0 for i in range(n)], np.uint64)]
    e = np.zeros(n, np.uint64); e[-1] = 1; out.append(e)
    out.append(np.arange(n, dtype=np.uint64) % np.uint64(q))
    out.append(rng.integers(0, q, n, dtype=np.uint64))
    return out


def test_full_size_transform_contract():
    """N=65536 at 50/55/60 bits (tests/arithmetic_contract_tests.cpp:23-75)."""
    for bits in (50, 55, 60):
        q = o.find_primes(N16, 1, bits)[0]
        psi = o.find_psi(q, N16)
        for k, x in enumerate(_patterns(N16, q, bits)):
            f = o.ntt_table(x, q, psi)
            assert (f < q).all()
            if k == 0:
                assert (f == 0).all()
            if k == 5:
                assert np.array_equal(f, o.ntt_forward(x, q, psi))
            assert np.array_equal(o.ntt_table(f, q, psi, inverse=True), x)
            g = o.ntt_table(x, q, psi, inverse=True)
            assert np.array_equal(o.ntt_table(g, q, psi), x)
        # x^(N-1) * x = -1, not +1
        left = np.zeros(N16, np.uint64); left[-1] = 1
        right = np.zeros(N16, np.uint64); right[1] = 1
        fl, fr = o.ntt_table(left, q, psi), o.ntt_table(right, q, psi)
        prod = np.array([o.mul_mod(int(a), int(b), q) for a, b in zip(fl, fr)], np.uint64)
        back = o.ntt_table(prod, q, psi, inverse=True)
        assert back[0] == q - 1 and not back[1:].any()


def test_monomials_and_inverse_basis_closed_forms():
    """tests/openfhe_representation_tests.cpp:64-112 on two pinned towers."""
    lg = 16
    for q, psi in list(zip(PINNED["q"], PINNED["psi"]))[:2]:
        points = [pow(psi, 2 * o.bit_reverse(k, lg) + 1, q) for k in range(N16)]
        assert points[0] == psi and points[1] == q - psi
        for degree in (0, 1, 17, N16 - 1):
            x = np.zeros(N16, np.uint64); x[degree] = 1
            want = np.array([pow(p, degree, q) for p in points], np.uint64)
            assert np.array_equal(o.ntt_table(x, q, psi), want)
        for slot in (0, 1, N16 // 2, N16 - 1):
            x = np.zeros(N16, np.uint64); x[slot] = 1
            ninv, step = pow(N16, -1, q), pow(points[slot], -1, q)
            want = np.array([ninv * pow(step, i, q) % q for i in range(N16)], np.uint64)
            assert np.array_equal(o.ntt_table(x, q, psi, inverse=True), want)


def test_twiddle_table_layout():
    # transform_stage.hpp:19-30: T[bitrev(e)] = root^e, T[n] = root^n
    q, psi, n = 17, 2, 4
    assert list(o.twiddles(n, psi, q)) == [1, 4, 2, 8, 16]


# ---- basis conversion: tests/oracle/test_basis_conversion.py:14-54 ----

def _bigint_lift(values, dim, source, target):
    product = math.prod(source)
    hats = [product // m for m in source]
    invs = [pow(h, -1, m) for h, m in zip(hats, source)]
    lifted = [sum(h * (values[t * dim + i] * inv % m) for t, (h, inv, m) in enumerate(zip(hats, invs, source)))
              for i in range(dim)]
    return [v % m for m in target for v in lifted], lifted, product


def test_basis_switch_unreduced_crt():
    rng = random.Random(0xB09)
    witnesses = 0
    for bits in (30, 50, 55, 60, 63):
        primes = o.find_primes(N16, 43, bits)
        for count in (1, 2, 3, 8, 16, 24, 32, 40):
            src, dst = primes[:count], primes[-3:]
            for reverse in (False, True):
                s, t = (dst, src) if reverse else (src, dst)
                for pattern in range(4):
                    vals = [(0 if pattern == 0 else m - 1 if pattern == 1 else int(i == 7) if pattern == 2
                             else rng.randrange(m)) for m in s for i in range(8)]
                    want, lifted, product = _bigint_lift(vals, 8, s, t)
                    got = o.basis_switch(vals, 8, s, t)
                    assert [int(v) for v in got] == want
                    witnesses += sum(v % t[0] != (v % product) % t[0] for v in lifted)
    assert witnesses > 0  # canonical CRT would differ here
    # worked case: tests/oracle/test_basis_conversion.py:42
    want, _, _ = _bigint_lift([16] * 8 + [96] * 8, 8, [17, 97], [5, 113])
    assert [int(v) for v in o.basis_switch([16] * 8 + [96] * 8, 8, [17, 97], [5, 113])] == want


@pytest.mark.parametrize("src,dst,vals,dim", [
    ([17, 17], [97], [0] * 16, 8), ([17], [17], [0] * 8, 8), ([15], [97], [0] * 8, 8),
    ([17], [97], [17] * 8, 8), ([17], [97], [0] * 3, 3), ([2], [97], [0] * 8, 8),
])
def test_basis_switch_rejections(src, dst, vals, dim):
    with pytest.raises(ValueError):
        o.basis_switch(vals, dim, src, dst)
    primes = o.find_primes(N16, 42, 30)
    with pytest.raises(ValueError):
        o.basis_switch([0] * (41 * 8), 8, primes[:41], [primes[-1]])


# ---- ModDown / MAC / key switch: docs/openfhe-hybrid-operation-graph.md:250-348 ----

def _small_params(n, L, K, bits_q=40, bits_p=45):
    q = o.find_primes(n, L, bits_q)
    p = o.find_primes(n, K, bits_p)
    return q, [o.find_psi(x, n) for x in q], p, [o.find_psi(x, n) for x in p]


def _naive_inverse(vals, q, psi):
    n = len(vals)
    lg = n.bit_length() - 1
    pts = [pow(psi, 2 * o.bit_reverse(k, lg) + 1, q) for k in range(n)]
    ninv = pow(n, -1, q)
    # a_i = n^-1 sum_k A_k pts[k]^-i
    return [ninv * sum(vals[k] * pow(pts[k], -i, q) for k in range(n)) % q for i in range(n)]


def _naive_forward(vals, q, psi):
    n = len(vals)
    lg = n.bit_length() - 1
    return [sum(vals[i] * pow(psi, (2 * o.bit_reverse(k, lg) + 1) * i, q) for i in range(n)) % q
            for k in range(n)]


def test_mod_down_bigint_restatement():
    n, L, K = 16, 3, 2
    q, qr, p, pr = _small_params(n, L, K)
    rng = random.Random(7)
    z = [rng.randrange(m) for m in q + p for _ in range(n)]
    # INTT(P) -> unreduced lift to Q -> NTT -> (Z - corr) * P^-1, with Python ints
    pc = []
    for j, (m, r) in enumerate(zip(p, pr)):
        pc += _naive_inverse(z[(L + j) * n:(L + j + 1) * n], m, r)
    corr, _, _ = _bigint_lift(pc, n, p, q)
    P = math.prod(p)
    want = []
    for l, (m, r) in enumerate(zip(q, qr)):
        ce = _naive_forward(corr[l * n:(l + 1) * n], m, r)
        want += [(z[l * n + i] - ce[i]) * pow(P, -1, m) % m for i in range(n)]
    assert [int(v) for v in o.mod_down(z, n, q, qr, p, pr)] == want


@pytest.mark.parametrize("evaluation", [True, False])
def test_mod_up_bigint_restatement(evaluation):
    """ModUp (src/resident.cpp:431-447) restated with Python ints: the Q towers
    pass through (evaluation input) or are transformed (coefficient input);
    the P towers are NTT(unreduced CRT lift of the Q coefficients)."""
    n, L, K = 16, 4, 3
    q, qr, p, pr = _small_params(n, L, K)
    rng = random.Random(0x5EED + evaluation)
    x = [rng.randrange(m) for m in q for _ in range(n)]
    if evaluation:
        coeff = []
        for l, (m, r) in enumerate(zip(q, qr)):
            coeff += _naive_inverse(x[l * n:(l + 1) * n], m, r)
        q_part = x
    else:
        coeff = x
        q_part = []
        for l, (m, r) in enumerate(zip(q, qr)):
            q_part += _naive_forward(x[l * n:(l + 1) * n], m, r)
    lift, lifted, product = _bigint_lift(coeff, n, q, p)
    p_part = []
    for j, (m, r) in enumerate(zip(p, pr)):
        p_part += _naive_forward(lift[j * n:(j + 1) * n], m, r)
    got = [int(v) for v in o.mod_up(x, n, q, qr, p, pr, evaluation=evaluation)]
    assert got == q_part + p_part
    # the lift is the unreduced sum: a canonical CRT (reduced mod Q first) differs
    canonical = []
    for j, (m, r) in enumerate(zip(p, pr)):
        canonical += _naive_forward([(v % product) % m for v in lifted], m, r)
    assert canonical != p_part


def _sigma_coefficients(coeffs, k, q, signed=True):
    """a(X) -> a(X^k) mod (X^n + 1): X^i goes to X^(ik mod 2n), negated when
    ik mod 2n >= n (the definition the automorphism map must agree with)."""
    n = len(coeffs)
    out = [0] * n
    for i, a in enumerate(coeffs):
        e = i * k % (2 * n)
        if e < n:
            out[e] = (out[e] + a) % q
        else:
            out[e - n] = (out[e - n] + (-a if signed else a)) % q
    return out


@pytest.mark.parametrize("n", [8, 16, 64])
def test_automorphism_matches_definition(n):
    """or_automorphism_source / or_hybrid_rotate (include/ntt/resident_conversion.hpp:15-29,
    src/resident.cpp:571-617) against sigma_k applied in the coefficient domain and
    evaluated by the naive forward transform: for odd k (3, 5, 25 and 2n-1 among them),
    NTT(a(X^k)) must equal the map applied to NTT(a). A transposed map (k^-1) or a map
    without the negacyclic sign must fail this check."""
    L, K = 2, 2
    q, qr, p, pr = _small_params(n, L, K)
    rng = random.Random(0xA0 + n)
    ks = sorted({3, 5, 25 % (2 * n), 2 * n - 1, 2 * n - 3} | {rng.randrange(3, 2 * n, 2) for _ in range(3)})
    mods, roots = q + p, qr + pr
    c_b = [[rng.randrange(m) for _ in range(n)] for m in mods]
    c_a = [[rng.randrange(m) for _ in range(n)] for m in mods]
    c_0 = [[rng.randrange(m) for _ in range(n)] for m in q]
    ext_b = [v for t, (m, r) in enumerate(zip(mods, roots)) for v in _naive_forward(c_b[t], m, r)]
    ext_a = [v for t, (m, r) in enumerate(zip(mods, roots)) for v in _naive_forward(c_a[t], m, r)]
    c0 = [v for t, (m, r) in enumerate(zip(q, qr)) for v in _naive_forward(c_0[t], m, r)]
    P = math.prod(p)
    discriminated = 0
    for k in ks:
        got_b, got_a = o.hybrid_rotate(ext_b, ext_a, c0, n, q, p, k)
        want_b, want_a = [], []
        for t, (m, r) in enumerate(zip(mods, roots)):
            # B + P*c0 on the Q towers, B alone on the P towers; A unchanged
            bc = [(b + (P * c if t < L else 0)) % m for b, c in zip(c_b[t], c_0[t] if t < L else [0] * n)]
            want_b += _naive_forward(_sigma_coefficients(bc, k, m), m, r)
            want_a += _naive_forward(_sigma_coefficients(c_a[t], k, m), m, r)
        assert [int(v) for v in got_b] == want_b, f"k={k}"
        assert [int(v) for v in got_a] == want_a, f"k={k}"
        # the address map alone, tower 0 of A: out[j] = in[source(j)]
        for j in range(n):
            assert want_a[j] == ext_a[o.automorphism_source(j, n, k)]
        # plausible mistakes give different outputs
        kinv = pow(k, -1, 2 * n)
        m0, r0 = mods[0], roots[0]
        transposed = _naive_forward(_sigma_coefficients(c_a[0], kinv, m0), m0, r0)
        unsigned = _naive_forward(_sigma_coefficients(c_a[0], k, m0, signed=False), m0, r0)
        if kinv != k:
            assert transposed != want_a[:n]
            discriminated += 1
        assert unsigned != want_a[:n]
    assert discriminated >= 2
    # k = 1 is the identity
    ib, ia = o.hybrid_rotate(ext_b, ext_a, [0] * (L * n), n, q, p, 1)
    assert [int(v) for v in ib] == ext_b and [int(v) for v in ia] == ext_a


def test_accumulate_before_mod_down_witness():
    """Down(16)=0, Down(1)=0 but Down(17)=1 (docs/openfhe-hybrid-operation-graph.md:343-348):
    reducing each digit separately differs from reducing the sum."""
    n, L, K = 8, 2, 1
    q, qr, p, pr = _small_params(n, L, K)
    rng = random.Random(3)
    a = [rng.randrange(m) for m in q + p for _ in range(n)]
    b = [rng.randrange(m) for m in q + p for _ in range(n)]
    s = [(x + y) % m for m, (x, y) in zip([m for m in q + p for _ in range(n)], zip(a, b))]
    whole = o.mod_down(s, n, q, qr, p, pr)
    da, db = o.mod_down(a, n, q, qr, p, pr), o.mod_down(b, n, q, qr, p, pr)
    split = [(int(x) + int(y)) % m for m, x, y in zip([m for m in q for _ in range(n)], da, db)]
    assert [int(v) for v in whole] != split


def test_hybrid_mac_key_offset():
    """Reduced level L < M reads P key rows at i + (M - L) (backend_contract.hpp:112-119)."""
    n, M, L, K, alpha = 8, 5, 4, 3, 3
    q, qr, p, pr = _small_params(n, M, K)
    beta_full, beta = 2, 2
    full = M + K
    mod_full = q + p
    kb = [o.mix(0xbbb + i + 1000037 * l + 1000039 * d) % mod_full[l]
          for d in range(beta_full) for l in range(full) for i in range(n)]
    ka = [o.mix(0xaaa + i + 1000003 * l + 1000033 * d) % mod_full[l]
          for d in range(beta_full) for l in range(full) for i in range(n)]
    act = q[:L] + p
    digits = [o.mix(d * 7919 + l * 31 + i) % act[l] for d in range(beta) for l in range(L + K) for i in range(n)]
    zb, za = o.hybrid_mac(digits, beta, kb, ka, n, q[:L], p, M)
    for l in range(L + K):
        kl = l if l < L else l + M - L
        for i in range(n):
            wb = sum(digits[(d * (L + K) + l) * n + i] * kb[(d * full + kl) * n + i] for d in range(beta)) % act[l]
            wa = sum(digits[(d * (L + K) + l) * n + i] * ka[(d * full + kl) * n + i] for d in range(beta)) % act[l]
            assert int(zb[l * n + i]) == wb and int(za[l * n + i]) == wa
    # the same flat index would read q4 (dropped) instead of p0
    wrong = sum(digits[(d * (L + K) + L) * n] * kb[(d * full + L) * n] for d in range(beta)) % act[L]
    assert wrong != int(zb[L * n])


def _crt_centered(residues, moduli):
    Q = math.prod(moduli)
    x = 0
    for r, m in zip(residues, moduli):
        h = Q // m
        x += r * h * pow(h, -1, m)
    x %= Q
    return x - Q if x > Q // 2 else x


def _keygen(n, q_full, qr_full, p, pr, alpha, s_old, s_new, rng):
    """Private-to-private Hybrid key (docs/openfhe-hybrid-operation-graph.md:258-266):
    digit j: b = -a*s_new + e + P*s_old on digit j's Q towers, in evaluation form."""
    M, K = len(q_full), len(p)
    mods, roots = q_full + p, qr_full + pr
    P = math.prod(p)
    beta = (M + alpha - 1) // alpha
    kb, ka = [], []
    for j in range(beta):
        e = [rng.randint(-3, 3) for _ in range(n)]
        for t, (m, r) in enumerate(zip(mods, roots)):
            a = [rng.randrange(m) for _ in range(n)]
            se = o.ntt_table(np.array([v % m for v in s_new], np.uint64), m, r)
            so = o.ntt_table(np.array([v % m for v in s_old], np.uint64), m, r)
            ee = o.ntt_table(np.array([v % m for v in e], np.uint64), m, r)
            in_digit = t < M and j * alpha <= t < (j + 1) * alpha
            b = [(-a[i] * int(se[i]) + int(ee[i]) + (P % m * int(so[i]) if in_digit else 0)) % m for i in range(n)]
            kb += b
            ka += a
    return kb, ka


@pytest.mark.parametrize("L", [6, 4])
def test_key_switch_decrypts(L):
    """End-to-end semantic pin: with a genuine key-switching key from s_old to s_new,
    B + A*s_new == c*s_old up to a small error (mod Q_L)."""
    n, M, K, alpha = 64, 6, 2, 2
    q, qr, p, pr = _small_params(n, M, K, bits_q=40, bits_p=50)
    rng = random.Random(1000 + L)
    s_old = [rng.randint(-1, 1) for _ in range(n)]
    s_new = [rng.randint(-1, 1) for _ in range(n)]
    kb, ka = _keygen(n, q, qr, p, pr, alpha, s_old, s_new, rng)
    c = np.concatenate([np.array([rng.randrange(m) for _ in range(n)], np.uint64) for m in q[:L]])
    ob, oa = o.key_switch(c, n, q[:L], qr[:L], M, p, pr, alpha, kb, ka)
    qs = q[:L]
    coeffs = []
    for l, (m, r) in enumerate(zip(qs, qr[:L])):
        se = o.ntt_table(np.array([v % m for v in s_new], np.uint64), m, r)
        so = o.ntt_table(np.array([v % m for v in s_old], np.uint64), m, r)
        sl = slice(l * n, (l + 1) * n)
        diff = [(int(b) + int(a) * int(x) - int(ci) * int(y)) % m
                for b, a, x, ci, y in zip(ob[sl], oa[sl], se, c[sl], so)]
        coeffs.append(o.ntt_table(np.array(diff, np.uint64), m, r, inverse=True))
    err = [abs(_crt_centered([int(coeffs[l][i]) for l in range(L)], qs)) for i in range(n)]
    assert max(err) < 2 ** 30  # Q_L is ~2^160..2^240; a wrong graph gives ~Q-sized values
    # and the error is not trivially zero (rounding of ModDown is present)
    assert max(err) > 0
