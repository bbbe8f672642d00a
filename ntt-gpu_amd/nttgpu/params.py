"""RNS parameter sets for the CKKS hot path (host logic, pure Python ints).

Primes follow OpenFHE's conventions so that parameters are drop-in
compatible with the pinned fixtures of the reference
(tests/cuda_openfhe_tests.cpp:139-145, tests/openfhe_hybrid_tests.cpp:1050-1053):
the first Q modulus is the largest prime below 2^bits congruent to 1 mod 2N,
special primes P walk down from it, scaling primes alternate below/above
2^scale_bits, and every root is the minimum primitive 2N-th root
(tests/openfhe_representation_tests.cpp:55-62). The reference's own walk
(find_primes / find_psi, include/ntt/param_gen.hpp:64-96) is also provided.
"""
from dataclasses import dataclass, field
from functools import lru_cache

_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n):
    """Miller-Rabin with the reference's 12 bases (param_gen.hpp:29-56)."""
    if n < 2:
        return False
    for p in _BASES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _BASES:
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


def find_primes(n, count, bits=46):
    """Ascending walk from the first multiple of 2n above 2^(bits-1), plus one."""
    step = 2 * n
    q = ((1 << (bits - 1)) // step + 1) * step + 1
    out = []
    while len(out) < count:
        if q >= 1 << bits:
            raise RuntimeError("prime search exhausted the bit range")
        if is_prime(q):
            out.append(q)
        q += step
    return out


def find_psi(q, n):
    """Deterministic xorshift candidate walk (param_gen.hpp:82-96)."""
    s = (q ^ n) or 1
    mask = (1 << 64) - 1
    while True:
        s ^= (s << 13) & mask
        s ^= s >> 7
        s ^= (s << 17) & mask
        g = 2 + s % (q - 3)
        zeta = pow(g, (q - 1) // (2 * n), q)
        if pow(zeta, n, q) == q - 1:
            return zeta


@lru_cache(maxsize=None)
def min_root(q, n):
    """Minimum primitive 2n-th root of unity mod q (OpenFHE's root convention)."""
    psi = find_psi(q, n)
    sq = psi * psi % q
    best, power = psi, psi
    for _ in range(n):
        if power < best:
            best = power
        power = power * sq % q
    return best


def last_prime(bits, m):
    q = (1 << bits) - m + 1
    while not is_prime(q):
        q -= m
    return q


def previous_prime(q, m):
    q -= m
    while not is_prime(q):
        q -= m
    return q


def next_prime(q, m):
    q += m
    while not is_prime(q):
        q += m
    return q


@dataclass
class HybridParams:
    """M Q towers (first modulus then scaling primes), K P towers, alpha per digit."""
    n: int
    q: list
    p: list
    alpha: int
    q_psi: list = field(default_factory=list)
    p_psi: list = field(default_factory=list)

    def __post_init__(self):
        if not self.q_psi:
            self.q_psi = [min_root(x, self.n) for x in self.q]
        if not self.p_psi:
            self.p_psi = [min_root(x, self.n) for x in self.p]

    @property
    def M(self):
        return len(self.q)

    @property
    def K(self):
        return len(self.p)

    def beta(self, L=None):
        L = self.M if L is None else L
        return (L + self.alpha - 1) // self.alpha


def openfhe_like(n, M, K, alpha, first_bits=60, scale_bits=50):
    """OpenFHE-shaped CKKS basis: q0 = LastPrime(first_bits), scaling primes in
    (below, above) pairs around 2^scale_bits filled from the top level down,
    P primes walking down from q0. Reproduces the pinned N=2^16, M=5, K=3 set."""
    m = 2 * n
    q0 = last_prime(first_bits, m)
    scaling = M - 1
    pairs = []
    lo, hi = None, None
    while 2 * len(pairs) < scaling:
        lo = last_prime(scale_bits, m) if lo is None else previous_prime(lo, m)
        hi = next_prime((1 << scale_bits) + 1 - m, m) if hi is None else next_prime(hi, m)
        pairs.append((lo, hi))
    slots = [0] * scaling
    top = scaling
    for below, above in pairs:
        if top >= 2:
            slots[top - 2], slots[top - 1] = below, above
            top -= 2
        elif top == 1:
            slots[0] = below
            top -= 1
    q = [q0] + slots
    p, cur = [], q0
    for _ in range(K):
        cur = previous_prime(cur, m)
        while cur in q:
            cur = previous_prime(cur, m)
        p.append(cur)
    return HybridParams(n=n, q=q, p=p, alpha=alpha)


def mixed_width_primes(n, count, low_bits=50, high_bits=60):
    """`count` distinct primes, alternating ~high_bits and ~low_bits wide."""
    m = 2 * n
    hi, lo, out = None, None, []
    while len(out) < count:
        hi = last_prime(high_bits, m) if hi is None else previous_prime(hi, m)
        out.append(hi)
        if len(out) < count:
            lo = last_prime(low_bits, m) if lo is None else previous_prime(lo, m)
            out.append(lo)
    return out
