"""Served mixed request stream (Poisson arrivals; N in {2^13..2^16}, L uniform
in [3, 30], dnum in {2, 3, 4}, K = alpha = ceil(L / dnum); 40% KeySwitch,
30% NTT/INTT, 30% ModUp/ModDown) with seeded uniform residues.

Each (N, L, dnum) is its own key-switching parameter set (M = L Q towers at
the top level, K P towers, digit size alpha = K, so the request uses exactly
dnum digits or fewer). Primes come from one OpenFHE-shaped pool per N
(q0 = LastPrime(60), alternating ~50-bit scaling primes, P walking down from
q0) with minimum roots, so sets that share a tower share its modulus.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import params, serve


@dataclass
class Request:
    arrival_s: float
    n: int
    level: int
    dnum: int
    op: int
    data: np.ndarray


class PrimePool:
    """Per ring dimension: up to 30 Q primes and 15 P primes with roots."""

    def __init__(self, n, max_q=30, max_p=15):
        base = params.openfhe_like(n, max_q, max_p, 1)
        self.n, self.q, self.q_psi, self.p, self.p_psi = n, base.q, base.q_psi, base.p, base.p_psi

    def param_set(self, level, dnum):
        K = math.ceil(level / dnum)
        return params.HybridParams(n=self.n, q=self.q[:level], p=self.p[:K], alpha=K,
                                   q_psi=self.q_psi[:level], p_psi=self.p_psi[:K])


def seeded_residues(n, moduli, rng):
    return np.concatenate([rng.integers(0, q, n, dtype=np.uint64) for q in moduli])


def make_key(hp, seed):
    rng = np.random.default_rng(seed)
    full = hp.q + hp.p
    kb = np.concatenate([seeded_residues(hp.n, full, rng) for _ in range(hp.beta())])
    ka = np.concatenate([seeded_residues(hp.n, full, rng) for _ in range(hp.beta())])
    return kb, ka


def draw_op(rng):
    u = rng.random()
    if u < 0.4:
        return serve.KEYSWITCH
    if u < 0.7:
        return serve.FORWARD if rng.random() < 0.5 else serve.INVERSE
    return serve.MODUP if rng.random() < 0.5 else serve.MODDOWN


def stream(count, rate_per_s, seed, sizes=(8192, 16384, 32768, 65536), levels=(3, 30), dnums=(2, 3, 4)):
    """Requests with exponential inter-arrival times; data generated per request."""
    rng = np.random.default_rng(seed)
    pools = {}
    out, t = [], 0.0
    for _ in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        n = int(rng.choice(sizes))
        level = int(rng.integers(levels[0], levels[1] + 1))
        dnum = int(rng.choice(dnums))
        op = draw_op(rng)
        if n not in pools:
            pools[n] = PrimePool(n)
        pool = pools[n]
        hp = pool.param_set(level, dnum)
        towers = hp.q + (hp.p if op == serve.MODDOWN else [])
        out.append(Request(t, n, level, dnum, op, seeded_residues(n, towers, rng)))
    return out, pools


class MixedServer:
    """Registers every (N, L, dnum) parameter set of a stream on a server."""

    def __init__(self, server, pools, key_seed=0xC5):
        self.server, self.pools, self.key_seed = server, pools, key_seed
        self.ids, self.sets = {}, {}

    def params_for(self, req):
        key = (req.n, req.level, req.dnum)
        if key not in self.ids:
            hp = self.pools[req.n].param_set(req.level, req.dnum)
            kb, ka = make_key(hp, self.key_seed + hash(key) % 100003)
            self.ids[key] = self.server.add_params(hp, kb, ka)
            self.sets[key] = (hp, kb, ka)
        return self.ids[key]

    def submit(self, req):
        return self.server.submit(self.params_for(req), req.op, req.level, req.data)

    def wait(self, ticket, req):
        return self.server.wait(ticket, self.params_for(req), req.op, req.level)
