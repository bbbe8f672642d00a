"""Shared helpers for the parity tests: seeded inputs and case logging."""
import json
import os
import sys

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PINNED = json.load(open(os.path.join(GOLDEN, "pinned_openfhe_qp.json")))
SEED = 0x7251D9  # reference include/ntt/benchmark_inputs.hpp:9


def log_case(msg):
    """Name the case on stdout before the GPU step runs."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def lcg_values(n, moduli, state=SEED):
    """MMIX LCG residues per tower (reference include/ntt/benchmark_inputs.hpp:9-22)."""
    out = np.empty(n * len(moduli), dtype=np.uint64)
    a, c, mask = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    s = state
    for k, q in enumerate(moduli):
        for i in range(n):
            s = (s * a + c) & mask
            out[k * n + i] = s % q
    return out


def uniform(n, moduli, seed):
    """Fast seeded uniform residues per tower (numpy), for large shapes."""
    rng = np.random.default_rng(seed)
    parts = [rng.integers(0, q, n, dtype=np.uint64) for q in moduli]
    return np.concatenate(parts)


def patterns(n, moduli):
    """The seven GPU patterns of reference tests/cuda_openfhe_tests.cpp:67-83."""
    idx = np.arange(n, dtype=np.uint64)
    for pattern in range(7):
        if pattern == 6:
            yield pattern, lcg_values(n, moduli) if n * len(moduli) <= 1 << 18 else uniform(n, moduli, SEED)
            continue
        parts = []
        for q in moduli:
            q = np.uint64(q)
            if pattern == 0:
                v = np.zeros(n, np.uint64)
            elif pattern == 1:
                v = np.ones(n, np.uint64)
            elif pattern == 2:
                v = np.full(n, q - np.uint64(1), np.uint64)
            elif pattern == 3:
                v = np.where(idx % np.uint64(2) == 1, q - np.uint64(1), np.uint64(0)).astype(np.uint64)
            elif pattern == 4:
                v = (idx == np.uint64(17)).astype(np.uint64)
            else:
                v = (idx == np.uint64(n - 1)).astype(np.uint64)
            parts.append(v)
        yield pattern, np.concatenate(parts)


