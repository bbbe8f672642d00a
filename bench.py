"""Headline benchmark: KeySwitch/s at N=2^16, L=24, K=6, dnum=4 (alpha=6) on one
MI355X, batched (64 components per step, captured HIP graph, resident key)
versus plain (one component per submission, per-stage launches, a sync per
request), with p50/p99 latency, the NTT's HBM roofline fraction and the CPU
oracle timed on this host beside it.

A "step" is one batched key switch of `--batch` components already resident
in HBM. Prints one JSON line.
"""
import argparse
import contextlib
import ctypes
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ntt-gpu_amd"))

from nttgpu import hip, params  # noqa: E402

N = 65536
L, K, ALPHA = 24, 6, 6
HBM_PEAK_GBS = 8000.0
BUTTERFLIES_PER_TOWER = (N // 2) * 16  # one direction of a 2^16-point transform
# HBM bytes of the roofline launch measured by rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE (per dispatch, calibrated on a copy kernel of known size)
HBM_COUNTERS = os.path.join(ROOT, "profiles", "ntt_n65536_hbm_counters.json")
HBM_COUNTERS_30 = os.path.join(ROOT, "profiles", "ntt_n65536_hbm_counters_30bit.json")
VALU_PROBE = os.path.join(ROOT, "ntt-gpu_amd", "bin", "valu_probe")
SESSION_BENCH = os.path.join(ROOT, "ntt-gpu_amd", "bin", "session_keyswitch_bench")


T0 = time.perf_counter()


def progress(msg):
    print(f"[bench {time.perf_counter() - T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


_ROCTX = []


@contextlib.contextmanager
def trace_range(name):
    """A roctx range around a timed section, so that a kernel trace taken with
    rocprofv3 --marker-trace can pick out the dispatches of that section from
    the tuner's and the warm-up's (a no-op when no profiler is attached or the
    marker library is absent)."""
    if not _ROCTX:
        try:
            lib = ctypes.CDLL("librocprofiler-sdk-roctx.so.1")
            lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
            _ROCTX.append(lib)
        except OSError:
            _ROCTX.append(None)
    lib = _ROCTX[0]
    if lib is not None:
        lib.roctxRangePushA(name.encode())
    try:
        yield
    finally:
        if lib is not None:
            lib.roctxRangePop()


def pct(values, q):
    v = sorted(values)
    pos = (len(v) - 1) * q
    lo = int(pos)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (pos - lo)


def seeded(n, moduli, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.integers(0, q, n, dtype=np.uint64) for q in moduli])


def tuned(ctx, towers):
    """Run the device autotuner on a context (before any timed work) and keep
    the short form of its report: the choice and the NTT default/best times."""
    rep = hip.autotune(ctx, towers)
    return {"towers": rep["towers"], "chosen": rep["chosen"],
            "ntt_default_ms": round(rep["ntt"]["default_ms"], 4), "ntt_tuned_ms": round(rep["ntt"]["best_ms"], 4)}


def ntt_roofline(towers, reps, stream, bits=59, autotune=True):
    """Forward NTT over `towers` distinct 2^16-point towers (one prime just below
    2^bits): the key switch's dominant kernel pair, measured with HIP events on
    its stream. bits <= 30 runs the 32-bit butterflies of narrow towers."""
    q = params.last_prime(bits, 2 * N)
    ctx = hip.Context(N, [q], [params.min_root(q, N)])
    tuning = tuned(ctx, min(towers, 1024)) if autotune else None
    data = hip.DeviceBuffer(towers * N)
    chunk = 256
    for t0 in range(0, towers, chunk):
        data.upload(seeded(N, [q] * min(chunk, towers - t0), 17 + t0), offset=t0 * N)
    events = [hip.Event() for _ in range(reps + 1)]
    for w in range(6):  # warm: steady-state clocks and caches before the timed launches
        ctx.transform_batch(data.ptr, towers, N, 0, 1, w % 2 == 1, stream=stream)
    hip.Stream(0, stream).synchronize()
    # One more untimed launch keeps the device busy while the timed ones are
    # enqueued back to back, so each event interval is device time of one
    # launch, not the host's submission latency.
    ctx.transform_batch(data.ptr, towers, N, 0, 1, True, stream=stream)
    with trace_range(f"ntt roofline {bits}-bit timed"):
        events[0].record(stream)
        for r in range(reps):
            ctx.transform_batch(data.ptr, towers, N, 0, 1, r % 2 == 1, stream=stream)
            events[r + 1].record(stream)
        events[-1].synchronize()
    times = [events[r].elapsed_ms(events[r + 1]) for r in range(reps)]
    fwd = [t for i, t in enumerate(times) if i % 2 == 0]
    inv = [t for i, t in enumerate(times) if i % 2 == 1]
    alg_bytes = 16 * N * towers  # read + write residues once per direction
    ms = statistics.mean(fwd)
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    data.free()
    ctx.close()
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(towers, bits),
        "kernel": f"forward NTT (ntt_pass_a256 + ntt_pass_b8_fwd), N=65536, {bits}-bit prime",
        "towers_per_launch": towers,
        "launch_ms": round(ms, 4),
        "inverse_launch_ms": round(statistics.mean(inv), 4),
        "alg_bytes_per_launch": alg_bytes,
        "autotune": tuning,
    }


def measured_traffic(towers, bits=59):
    """HBM bytes of one forward roofline launch over `towers` towers, from the
    committed counter summary of the prime's width (FETCH_SIZE and WRITE_SIZE
    per dispatch of both passes, each scaled by the copy-kernel calibration;
    30-bit towers hand off packed 32-bit words), per tower x towers."""
    try:
        with open(HBM_COUNTERS_30 if bits <= 30 else HBM_COUNTERS) as f:
            c = json.load(f)
    except (OSError, ValueError):
        return None
    per_tower = 0.0
    for kern in ("pass_a_forward", "pass_b_forward"):
        k = c["kernels"][kern]
        per_tower += (k["fetch_kib"] * c["fetch_scale"] + k["write_kib"] * c["write_scale"]) * 1024 / c["towers"]
    return int(per_tower * towers)


def valu_roofline(roof, narrow=None):
    """The forward NTT against the VALU roof measured on this device: the rate
    of the same wide-lazy forward butterfly in independent chains
    (ntt-gpu_amd/bin/valu_probe), versus butterflies per second achieved by
    the roofline launch (16 stages x N/2 butterflies per tower); `narrow`
    (the 30-bit roofline launch) against the 32-bit butterfly's rate."""
    if not os.path.exists(VALU_PROBE):
        return None
    try:
        r = subprocess.run([VALU_PROBE, "--json"], capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return None
    if r.returncode != 0:
        return None
    rows = {}
    for line in r.stdout.splitlines():
        try:
            d = json.loads(line)
            rows[d["probe"]] = d
        except ValueError:
            continue
    if "ct8_forward_butterfly" not in rows:
        return None
    peak = rows["ct8_forward_butterfly"]["per_s"]
    achieved = BUTTERFLIES_PER_TOWER * roof["towers_per_launch"] / (roof["launch_ms"] * 1e-3)
    narrow_frac = None
    if narrow and "ct32_forward_butterfly" in rows:
        rate = BUTTERFLIES_PER_TOWER * narrow["towers_per_launch"] / (narrow["launch_ms"] * 1e-3)
        narrow_frac = {"achieved": round(rate / 1e12, 4),
                       "peak": round(rows["ct32_forward_butterfly"]["per_s"] / 1e12, 4),
                       "frac": round(rate / rows["ct32_forward_butterfly"]["per_s"], 4)}
    return {
        "bound": "valu",
        "achieved": round(achieved / 1e12, 4),
        "peak": round(peak / 1e12, 4),
        "unit": "T butterfly/s",
        "frac": round(achieved / peak, 4),
        "peak_source": "valu_probe: forward butterfly chains on this device",
        "probe": {k: round(v["per_s"] / 1e12, 4) for k, v in rows.items()},
        "narrow": narrow_frac,
    }


def hoisted_rotations(ctx, hp, x, stream, R, B, steps=5):
    """R hoisted rotations (one ModUp shared by all, per rotation MAC +
    automorphism + ModDown: ntt_hip_keyswitch_rotate) over B ciphertexts per
    call; rotations/s counts rotated ciphertexts (R * B per call)."""
    comp = L * N
    keys = []
    for r in range(R):
        ks = hip.KeySwitch(ctx, L, K, ALPHA, max_batch=B)
        full = hp.q + hp.p
        kb = np.concatenate([seeded(N, full, 7000 + 31 * r + d) for d in range(hp.beta())])
        ka = np.concatenate([seeded(N, full, 8000 + 31 * r + d) for d in range(hp.beta())])
        ks.upload_key(kb, ka, hp.M, hp.q)
        keys.append(ks)
    autos = [pow(5, r + 1, 2 * N) for r in range(R)]
    c0 = hip.DeviceBuffer(B * comp).upload(x[: B * comp])
    c1 = hip.DeviceBuffer(B * comp).upload(x[: B * comp][::-1].copy())
    outs = [(hip.DeviceBuffer(B * comp), hip.DeviceBuffer(B * comp)) for _ in range(R)]
    ws = hip.DeviceBuffer(keys[0].rotate_workspace_words(B))

    def call():
        hip.KeySwitch.rotate(keys, autos, c0.ptr, c1.ptr, B, [o[0].ptr for o in outs], [o[1].ptr for o in outs],
                             ws.ptr, stream)
    call()
    hip.Stream(0, stream).synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        call()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    for ks in keys:
        ks.close()
    return {"rotations": R, "ciphertexts": B, "ms_per_call": round(ms, 3),
            "rotations_per_s": round(R * B / (ms * 1e-3), 1),
            "note": "one ModUp of c1 shared by the R rotations; per rotation MAC, automorphism, ModDown"}


def in_place_and_relinearize(ks, x, stream, B, steps=5):
    """The caller-facing wrappers on the batched path: KeySwitchInPlace on B
    two-component ciphertexts (c1 switched; c0 += B, c1 = A, additions fused
    into ModDown's last pass) and relinearization of B three-component
    ciphertexts with one key (c2 switched, c0 += B, c1 += A). Each call
    updates its ciphertexts in place, so the timed calls run back to back on
    the same buffers; KeySwitch/s counts switched components."""
    comp = L * N
    c = [hip.DeviceBuffer(B * comp).upload(x[: B * comp]) for _ in range(3)]
    ws = hip.DeviceBuffer(ks.workspace_words(B))
    e0, e1 = hip.Event(), hip.Event()
    out = {}
    for name, call in (("in_place", lambda: ks.in_place(c[0].ptr, c[1].ptr, B, ws.ptr, stream=stream)),
                       ("relinearize", lambda: hip.KeySwitch.relinearize([ks], c[0].ptr, c[1].ptr, [c[2].ptr], B,
                                                                         ws.ptr, stream))):
        call()
        e0.record(stream)
        for _ in range(steps):
            call()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        out[name] = {"batch": B, "ms_per_call": round(ms, 3), "keyswitch_per_s": round(B / (ms * 1e-3), 1)}
    for b in c + [ws]:
        b.free()
    return out


def session_plain(hp, requests):
    """The reference-API plain path: KeySwitchHybrid one request at a time
    through the include/ntt Session (hybrid_digits, hybrid_accumulate, two
    mod_down; resident input and key), run by session_keyswitch_bench in a
    child process on the same primes."""
    if not os.path.exists(SESSION_BENCH):
        return None
    text = " ".join(str(v) for v in [N, L, K, ALPHA] + hp.q + hp.q_psi + hp.p + hp.p_psi)
    try:
        r = subprocess.run([SESSION_BENCH, str(requests), "2"], input=text, capture_output=True, text=True,
                           timeout=300)
    except (OSError, subprocess.TimeoutExpired):
        return None
    if r.returncode != 0:
        progress(f"session bench failed: {r.stderr.strip()[-200:]}")
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["path"] = ("include/ntt Session chain per request (hybrid_digits -> hybrid_accumulate -> "
                   "mod_down x2), driver ops launched one by one, fresh device buffers per output")
    return out


def served_stream(requests, rates, device_rates, seed=0xC5):
    """The mixed served stream: Poisson arrivals of mixed N/L/dnum/op requests at swept rates
    into the continuous-batching server; p50/p99 latency and throughput per
    rate, and every response compared with the same request served alone.
    Two modes: "host" (requests carry host arrays: upload, compute, download
    per request) at `rates`, and "device" (inputs and outputs resident in HBM,
    submit_device) at `device_rates`."""
    from nttgpu import serve, workload
    results = []
    reqs, pools = workload.stream(requests, 1.0, seed)
    srv = serve.Server(streams=4, max_batch=32, hbm_budget_bytes=48 << 30)
    ms = workload.MixedServer(srv, pools)
    for r in reqs:
        ms.params_for(r)
    alone_srv = serve.Server(streams=1, max_batch=1, hbm_budget_bytes=16 << 30)
    alone = workload.MixedServer(alone_srv, pools)
    # warm every (params, level) plan on both servers outside the timed stream
    seen = set()
    for r in reqs:
        key = (r.n, r.level, r.dnum, r.op)
        if key not in seen:
            seen.add(key)
            ms.wait(ms.submit(r), r)
            alone.wait(alone.submit(r), r)
    progress(f"serve: {len(seen)} warm plans")
    reference = [alone.wait(alone.submit(r), r)[0] for r in reqs]
    progress("serve: alone results ready")

    def run(rate, mode, submit, collect):
        """Poisson arrivals at `rate`; collect(i, ticket) -> (outputs or None, latency ms, batch)."""
        gaps = np.random.default_rng(seed + int(rate)).exponential(1.0 / rate, len(reqs))
        tickets, t0 = [], time.perf_counter()
        due = t0
        for i, g in enumerate(gaps):
            due += g
            while time.perf_counter() < due:
                pass
            tickets.append(submit(i))
        waited = [collect(i, t) for i, t in enumerate(tickets)]
        wall = time.perf_counter() - t0
        lat = [w[1] for w in waited]
        exact = None
        if mode == "host":
            exact = all(np.array_equal(w[0][0], ref[0]) and (ref[1] is None or np.array_equal(w[0][1], ref[1]))
                        for w, ref in zip(waited, reference))
        by_op = {}
        for w, r in zip(waited, reqs):
            by_op.setdefault(serve.OP_NAMES[r.op], []).append(w[2])
        results.append({"mode": mode, "offered_rate_per_s": rate, "requests": len(reqs),
                        "throughput_per_s": round(len(reqs) / wall, 1),
                        "p50_ms": round(pct(lat, 0.5), 3), "p99_ms": round(pct(lat, 0.99), 3),
                        "mean_batch": round(statistics.mean(w[2] for w in waited), 2),
                        "mean_batch_by_op": {k: round(statistics.mean(v), 2) for k, v in sorted(by_op.items())},
                        "all_equal_alone": exact})

    def report():
        r = results[-1]
        progress(f"serve {r['mode']} rate {r['offered_rate_per_s']}/s: p50 {r['p50_ms']} ms p99 {r['p99_ms']} ms "
                 f"exact={r['all_equal_alone']}")

    for rate in rates:
        run(rate, "host", lambda i: ms.submit(reqs[i]), lambda i, t: ms.wait(t, reqs[i]))
        report()
    # device-resident: every request's input and outputs live in HBM before the
    # stream starts (uploaded outside the timed region); outputs are read back
    # after the timed region for the comparison with the alone results
    dev = []
    for r in reqs:
        w = srv.output_words(ms.params_for(r), r.op, r.level)
        dev.append((hip.DeviceBuffer(r.data.size).upload(r.data), hip.DeviceBuffer(w),
                    hip.DeviceBuffer(w) if r.op == serve.KEYSWITCH else None))

    def submit_dev(i):
        d = dev[i]
        return srv.submit_device(ms.params_for(reqs[i]), reqs[i].op, reqs[i].level, d[0].ptr, d[1].ptr,
                                 d[2].ptr if d[2] else None)

    def collect_dev(i, t):
        lat_ms, b = srv.wait_device(t)
        return None, lat_ms, b

    for rate in device_rates:
        run(rate, "device", submit_dev, collect_dev)
        results[-1]["all_equal_alone"] = all(
            np.array_equal(d[1].download(), ref[0]) and (d[2] is None or np.array_equal(d[2].download(), ref[1]))
            for d, ref in zip(dev, reference))
        report()
    stats = srv.stats()
    srv.close()
    alone_srv.close()
    # the alone results against the oracle: two requests of every op
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as o
    o.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0)))
    checked, equal, per_op = 0, True, {}
    for r, ref in zip(reqs, reference):
        if per_op.get(r.op, 0) >= 2:
            continue
        per_op[r.op] = per_op.get(r.op, 0) + 1
        hp, kb, ka = ms.sets[(r.n, r.level, r.dnum)]
        if r.op == serve.KEYSWITCH:
            want = o.key_switch(r.data, r.n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        elif r.op in (serve.FORWARD, serve.INVERSE):
            want = (o.ntt_towers(r.data, r.n, hp.q, hp.q_psi, inverse=r.op == serve.INVERSE), None)
        elif r.op == serve.MODUP:
            want = (o.mod_up(r.data, r.n, hp.q, hp.q_psi, hp.p, hp.p_psi, evaluation=True), None)
        else:
            want = (o.mod_down(r.data, r.n, hp.q, hp.q_psi, hp.p, hp.p_psi), None)
        equal = equal and np.array_equal(ref[0], want[0]) and (want[1] is None or np.array_equal(ref[1], want[1]))
        checked += 1
    progress(f"serve: {checked} alone results vs the oracle: equal={equal}")
    return {"rates": results, "peak_device_bytes": stats["peak_device_bytes"],
            "resident_bytes": stats["resident_bytes"], "plans_resident": stats["plans_resident"],
            "plans_evicted": stats["plans_evicted"], "max_in_flight": stats["max_in_flight"],
            "budget_bytes": 48 << 30, "oracle_sample": {"checked": checked, "equal": equal},
            "mix": "40% KeySwitch, 30% NTT/INTT, 30% ModUp/ModDown; N in 2^13..2^16; L in [3,30]; dnum in {2,3,4}"}


def served_keyswitch(hp, kb, ka, d_in, d_b, d_a, B, requests, rates, seed=0x5E):
    """Continuous batching on one parameter set: Poisson arrivals of
    device-resident KeySwitch requests (N=2^16, L=24, K=6, dnum=4; input i is
    component i mod B of the batched run) into the server (max batch B, four
    streams). Requests that arrive while a batch runs are gathered into the
    next one; every sampled response equals the batched run's component."""
    from nttgpu import serve
    comp = L * N
    srv = serve.Server(streams=4, max_batch=B, hbm_budget_bytes=96 << 30)
    pid = srv.add_params(hp, kb, ka)
    outs = [(hip.DeviceBuffer(comp), hip.DeviceBuffer(comp)) for _ in range(requests)]
    for i in range(2):  # plan, key upload and arenas outside the timed streams
        srv.wait_device(srv.submit_device(pid, serve.KEYSWITCH, L, d_in.at(0), outs[i][0].ptr, outs[i][1].ptr))
    results = []
    for rate in rates:
        gaps = np.random.default_rng(seed + int(rate)).exponential(1.0 / rate, requests)
        tickets, t0 = [], time.perf_counter()
        due = t0
        for i, g in enumerate(gaps):
            due += g
            while time.perf_counter() < due:
                pass
            tickets.append(srv.submit_device(pid, serve.KEYSWITCH, L, d_in.at((i % B) * comp), outs[i][0].ptr,
                                             outs[i][1].ptr))
        waited = [srv.wait_device(t) for t in tickets]
        wall = time.perf_counter() - t0
        sample = range(0, requests, max(1, requests // 8))
        exact = all(np.array_equal(outs[i][0].download(), d_b.download(comp, (i % B) * comp)) and
                    np.array_equal(outs[i][1].download(), d_a.download(comp, (i % B) * comp)) for i in sample)
        lat = [w[0] for w in waited]
        results.append({"offered_rate_per_s": rate, "requests": requests,
                        "throughput_per_s": round(requests / wall, 1), "p50_ms": round(pct(lat, 0.5), 3),
                        "p99_ms": round(pct(lat, 0.99), 3), "mean_batch": round(statistics.mean(w[1] for w in waited), 2),
                        "sampled_equal_batched": exact})
        progress(f"served keyswitch {rate}/s: {results[-1]['throughput_per_s']} KS/s, mean batch "
                 f"{results[-1]['mean_batch']}, p50 {results[-1]['p50_ms']} ms, exact={exact}")
    srv.close()
    for b_, a_ in outs:
        b_.free()
        a_.free()
    return {"workload": "KeySwitch requests on one parameter set, N=65536, L=24, K=6, dnum=4, device-resident",
            "max_batch": B, "rates": results}


# The reference's largest published workload (N=65536, 40 towers of 60-bit
# primes from its generator, zero coefficients in the vector file; operands
# from its fixed-seed stream), RTX 4080 SUPER medians in ms
# (docs/batch40-native-analysis.md:27-34, docs/paired-ntt-native-analysis.md:130-134,182-187).
REFERENCE_PUBLISHED = {
    ("events", "gpu_forward"): 0.552, ("events", "gpu_inverse"): 0.598, ("events", "gpu_pointwise"): 0.048,
    ("events", "gpu_fused_multiply"): 1.723,
    ("public", "gpu_forward"): 8.079, ("public", "gpu_inverse"): 8.045, ("public", "gpu_pointwise"): 11.858,
    ("public", "gpu_fused_multiply"): 12.431,
    ("conversion", "resident_modup_moddown"): 3.776, ("conversion", "resident_conversion_host_boundary"): 20.989,
}
CLI = os.path.join(ROOT, "ntt-gpu_amd", "bin", "ntt_hip_cli")
CONVERSION_BENCH = os.path.join(ROOT, "ntt-gpu_amd", "bin", "session_conversion_bench")


def mixed_ring_dimensions(towers_per_n=64, reps=10):
    """Forward transforms of towers of four ring dimensions (N = 2^13 .. 2^16,
    `towers_per_n` towers each, 60/50-bit primes): one mixed launch per pass
    (ntt_hip_transform_mixed) against one launch per ring dimension, on the
    same stream and data; the two results are compared bit for bit."""
    sizes = (8192, 16384, 32768, 65536)
    st = hip.Stream(0)
    runs = []
    for n in sizes:
        qs = [params.last_prime(60, 2 * n), params.last_prime(50, 2 * n)]
        ctx = hip.Context(n, qs, [params.min_root(q, n) for q in qs])
        x = np.concatenate([seeded(n, qs, 40 + n + t) for t in range(towers_per_n // 2)])
        a, b = hip.DeviceBuffer(x.size).upload(x), hip.DeviceBuffer(x.size).upload(x)
        runs.append((n, ctx, a, b))
    jobs = [(ctx, a.at(t * n), t % 2) for n, ctx, a, b in runs for t in range(towers_per_n)]
    ws = hip.transform_mixed(jobs, inverse=False, stream=st.handle)  # warm
    for n, ctx, a, b in runs:
        ctx.transform_batch(b.ptr, towers_per_n // 2, 2 * n, 0, 2, False, stream=st.handle)
    st.synchronize()
    e0, e1 = hip.Event(), hip.Event()

    def timed(fn):
        e0.record(st.handle)
        for _ in range(reps):
            fn()
        e1.record(st.handle)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    mixed_ms = timed(lambda: hip.transform_mixed(jobs, inverse=False, stream=st.handle, workspace=ws))
    split_ms = timed(lambda: [ctx.transform_batch(b.ptr, towers_per_n // 2, 2 * n, 0, 2, False, stream=st.handle)
                              for n, ctx, a, b in runs])
    equal = all(np.array_equal(a.download(), b.download()) for n, ctx, a, b in runs)
    total = sum(n for n in sizes) * towers_per_n
    out = {"towers_per_ring_dimension": towers_per_n, "ring_dimensions": list(sizes),
           "mixed_launch_ms": round(mixed_ms, 4), "launch_per_ring_dimension_ms": round(split_ms, 4),
           "mixed_residues_per_s": round(total / (mixed_ms * 1e-3), 1),
           "outputs_equal": bool(equal),
           "note": "forward transforms; both forms transform the same towers repeatedly in place (results compared "
                   "after the same number of transforms)"}
    for n, ctx, a, b in runs:
        a.free()
        b.free()
        ctx.close()
    return out


def reference_workloads():
    """The reference's own benchmark commands on its largest workload, run
    through this build's drop-in tools: ntt_hip_cli events-jsonl (resident,
    device events) and gpu-bench-jsonl (public host-vector calls), and the
    resident ModUp -> ModDown pair through the include/ntt Session. Each tool
    checks its results against the CPU reference NTT before reporting."""
    import tempfile
    if not (os.path.exists(CLI) and os.path.exists(CONVERSION_BENCH)):
        return None
    n, towers = 65536, 40
    q = params.find_primes(n, towers, 60)
    lines = [f"{n} {towers}"] + [f"{m} {params.find_psi(m, n)}" for m in q] + ["COEFFS"] + ["0"] * n
    rows = []
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, f"benchmark-vector-n{n}-k{towers}-b60.txt")
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
        for scope, cmd in (("events", [CLI, "events-jsonl", path, "5", "20"]),
                           ("public", [CLI, "gpu-bench-jsonl", path, "2", "10"]),
                           ("conversion", [CONVERSION_BENCH, path, "2", "10"])):
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            except (OSError, subprocess.TimeoutExpired):
                return None
            if r.returncode != 0:
                progress(f"reference workload {scope} failed: {r.stderr.strip()[-200:]}")
                return None
            for line in r.stdout.splitlines():
                rec = json.loads(line)
                if rec.get("type") != "summary":
                    continue
                ref = REFERENCE_PUBLISHED.get((scope, rec["operation"]))
                rows.append({"scope": scope, "operation": rec["operation"], "median_ms": round(rec["median_ms"], 4),
                             "p95_ms": round(rec["p95_ms"], 4), "reference_ms": ref,
                             "reference_over_this": round(ref / rec["median_ms"], 2) if ref else None})
    progress("reference workloads: " + ", ".join(f"{r['scope']}/{r['operation']} {r['median_ms']} ms" for r in rows
                                                  if r["reference_ms"]))
    return {"workload": "N=65536, 40 towers of 60-bit primes (reference generator), fixed-seed operands; "
                        "conversion: + 3 P primes, 3 towers per digit",
            "reference_hardware": "RTX 4080 SUPER (sm_89), the reference's published medians",
            "rows": rows}


def host_description():
    """nproc, the CPU cores this process may use, the CPU model and the compiler
    of the oracle (the reference's CPU-baseline protocol records them)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, subprocess.TimeoutExpired, IndexError):
        cc = None
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "cpu_model": model,
            "compiler": cc, "flags": "-O2 -fopenmp (oracle/Makefile)"}


def cpu_baseline(hp, x0, kb, ka, gpu_b, gpu_a, samples):
    """The oracle's key switch (the reference algorithm in C) on component 0 of
    the same batch, timed on this host's cores by the reference's protocol
    (host wall clock, one warm-up, then samples; median, p5, p95): one thread
    with towers in sequence (the reference's OpenFHE CPU protocol), and towers
    in parallel on every core this process may use (OpenMP; the box's share,
    OMP_NUM_THREADS). Also checks the GPU output bit-exactly."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as o
    exact = True

    def leg(threads):
        nonlocal exact
        o.set_threads(threads)
        o.key_switch(x0, N, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)  # warm-up
        times = []
        for _ in range(samples):
            t0 = time.perf_counter()
            wb, wa = o.key_switch(x0, N, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
            times.append(time.perf_counter() - t0)
            exact = exact and np.array_equal(wb, gpu_b) and np.array_equal(wa, gpu_a)
        med = statistics.median(times)
        return {"value": round(1.0 / med, 4), "unit": "KeySwitch/s", "cores": threads,
                "median_s": round(med, 4), "p5_s": round(pct(times, 0.05), 4), "p95_s": round(pct(times, 0.95), 4)}

    host = host_description()
    all_cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host["usable_cores"]
    one = leg(1)
    many = leg(all_cores)
    progress(f"cpu baseline: 1 core {one['value']} KS/s, {all_cores} cores {many['value']} KS/s")
    return {
        "value": one["value"],
        "unit": "KeySwitch/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{samples} key switch(es) of one component (N=2^16, L=24, K=6, dnum=4), C oracle, "
                  f"1 thread, towers in sequence; median {one['median_s']:.2f} s (OpenFHE CPU baseline "
                  f"unavailable on this box)",
        "all_cores": dict(many, sample="the same key switch, towers and conversion targets in parallel (OpenMP)"),
        "host": host,
    }, exact


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--plain-requests", type=int, default=64)
    ap.add_argument("--ntt-towers", type=int, default=2048)
    ap.add_argument("--cpu-samples", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--roofline-only", action="store_true")
    ap.add_argument("--keyswitch-only", action="store_true",
                    help="for counter runs: only the batched key switch (warm-up and timed graph replays)")
    ap.add_argument("--roofline-bits", type=int, default=59)
    ap.add_argument("--serve-rates", type=str, default="200,1000")
    ap.add_argument("--serve-device-rates", type=str, default="1000,5000,20000")
    ap.add_argument("--serve-keyswitch-requests", type=int, default=256)
    ap.add_argument("--no-reference-workloads", action="store_true")
    ap.add_argument("--serve-keyswitch-rates", type=str, default="2000,6000,12000")
    ap.add_argument("--rotations", type=int, default=8)
    ap.add_argument("--session-requests", type=int, default=16)
    ap.add_argument("--rotation-batch", type=int, default=16)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(data-parallel over ciphertexts), so --gpus must be 1"}))
        sys.exit(2)

    progress(f"device {hip.device_name(0)}")
    if args.roofline_only:
        # For counter runs: the NTT roofline launches plus a plain copy kernel
        # of known size (1 GiB read + 1 GiB written) to calibrate the counters.
        s = hip.Stream(0)
        roof = ntt_roofline(args.ntt_towers, 12, s.handle, bits=args.roofline_bits, autotune=not args.no_autotune)
        words = 1 << 27
        a, b = hip.DeviceBuffer(words), hip.DeviceBuffer(words)
        e0, e1 = hip.Event(), hip.Event()
        hip.check(hip.lib().ntt_hip_copy_words(0, b.ptr, a.ptr, words, s.handle), "copy")
        e0.record(s.handle)
        hip.check(hip.lib().ntt_hip_copy_words(0, b.ptr, a.ptr, words, s.handle), "copy")
        e1.record(s.handle)
        e1.synchronize()
        ms = e0.elapsed_ms(e1)
        print(json.dumps({"roofline": roof, "copy_kernel": {"bytes": 16 * words, "ms": round(ms, 4),
                                                             "GBps": round(16 * words / ms / 1e6, 1)}}), flush=True)
        return
    hp = params.openfhe_like(N, L, K, ALPHA)
    B = args.batch
    ctx = hip.Context(N, hp.q + hp.p, hp.q_psi + hp.p_psi)
    ks = hip.KeySwitch(ctx, L, K, ALPHA, max_batch=B)
    full = hp.q + hp.p
    beta = hp.beta()
    kb = np.concatenate([seeded(N, full, 1000 + d) for d in range(beta)])
    ka = np.concatenate([seeded(N, full, 2000 + d) for d in range(beta)])
    progress("key generated")
    ks.upload_key(kb, ka, hp.M, hp.q)
    progress("key resident")
    ks_tuning = None
    if not args.no_autotune:  # on the batched key switch itself, before the graph is captured
        rep = ks.autotune(B)
        ks_tuning = {"batch": rep["batch"], "chosen": rep["chosen"],
                     "default_ms": round(rep["keyswitch"]["default_ms"], 3),
                     "tuned_ms": round(rep["keyswitch"]["tuned_ms"], 3)}
        progress(f"autotune: {ks_tuning}")
    x = np.concatenate([seeded(N, hp.q, 5000 + c) for c in range(B)])
    comp = L * N
    d_in = hip.DeviceBuffer(B * comp).upload(x)
    d_b, d_a = hip.DeviceBuffer(B * comp), hip.DeviceBuffer(B * comp)
    ws = hip.DeviceBuffer(ks.workspace_words(B))
    stream = hip.Stream(0)
    s = stream.handle

    progress("inputs resident")
    # ---- batched: captured graph, one launch per step
    graph = ks.capture(d_in.ptr, B, d_b.ptr, d_a.ptr, ws.ptr, s)
    for _ in range(args.warmup):
        graph.launch(s)
    stream.synchronize()
    events = [hip.Event() for _ in range(args.steps + 1)]
    with trace_range("batched key switch timed"):
        t0 = time.perf_counter()
        events[0].record(s)
        for i in range(args.steps):
            graph.launch(s)
            events[i + 1].record(s)
        stream.synchronize()
        wall = time.perf_counter() - t0
    step_ms = [events[i].elapsed_ms(events[i + 1]) for i in range(args.steps)]
    batched_kss = B * args.steps / wall
    gpu_b = d_b.download(comp, 0)
    gpu_a = d_a.download(comp, 0)

    progress(f"batched done: {batched_kss:.1f} KS/s")
    if args.keyswitch_only:
        print(json.dumps({"batched_keyswitch_per_s": round(batched_kss, 2), "batch": B, "steps": args.steps,
                          "warmup": args.warmup, "autotune": ks_tuning}), flush=True)
        return
    # ---- plain: one component per submission, per-stage launches, sync per request
    ws1 = hip.DeviceBuffer(ks.workspace_words(1))
    for _ in range(2):
        ks.run(d_in.ptr, 1, d_b.ptr, d_a.ptr, ws1.ptr, s)
    stream.synchronize()
    lat = []
    t1 = time.perf_counter()
    for r in range(args.plain_requests):
        c = r % B
        r0 = time.perf_counter()
        ks.run(d_in.at(c * comp), 1, d_b.at(c * comp), d_a.at(c * comp), ws1.ptr, s)
        stream.synchronize()
        lat.append((time.perf_counter() - r0) * 1e3)
    plain_wall = time.perf_counter() - t1
    plain_kss = args.plain_requests / plain_wall

    progress(f"plain done: {plain_kss:.1f} KS/s")
    sess = session_plain(hp, args.session_requests)
    if sess:
        progress(f"session plain: {sess['keyswitch_per_s']} KS/s")
    rot = hoisted_rotations(ctx, hp, x, s, args.rotations, min(B, args.rotation_batch))
    progress(f"hoisted rotations: {rot['rotations_per_s']} rotations/s")
    wrappers = in_place_and_relinearize(ks, x, s, B)
    progress(f"in-place key switch: {wrappers['in_place']['keyswitch_per_s']} KS/s, "
             f"relinearize: {wrappers['relinearize']['keyswitch_per_s']} KS/s")
    roof = ntt_roofline(args.ntt_towers, 12, s, autotune=not args.no_autotune)
    progress(f"ntt roofline done: {roof['achieved']} GB/s")
    narrow = ntt_roofline(args.ntt_towers, 12, s, bits=30, autotune=not args.no_autotune)
    progress(f"narrow (30-bit) ntt roofline: {narrow['achieved']} GB/s")
    valu = valu_roofline(roof, narrow)
    if valu:
        progress(f"valu roofline: {valu['achieved']} of {valu['peak']} T butterfly/s")
    refw = reference_workloads() if not args.no_reference_workloads else None
    mixed = mixed_ring_dimensions()
    progress(f"mixed ring dimensions: {mixed['mixed_launch_ms']} ms mixed, "
             f"{mixed['launch_per_ring_dimension_ms']} ms per-N, equal={mixed['outputs_equal']}")
    served_ks = None
    if args.serve_keyswitch_requests:
        served_ks = served_keyswitch(hp, kb, ka, d_in, d_b, d_a, B, args.serve_keyswitch_requests,
                                     [float(r) for r in args.serve_keyswitch_rates.split(",")])
    served = None
    if args.serve_requests:
        served = served_stream(args.serve_requests, [float(r) for r in args.serve_rates.split(",")],
                               [float(r) for r in args.serve_device_rates.split(",")])
    cpu, exact = (None, None)
    if not args.no_cpu:
        cpu, exact = cpu_baseline(hp, x[:comp], kb, ka, gpu_b, gpu_a, args.cpu_samples)
        progress(f"cpu baseline done: {cpu['value']} KS/s, exact={exact}")

    out = {
        "metric": "KeySwitch/s N=2^16 L=24 dnum=4 batched vs plain; p50/p99 ms; NTT HBM GB/s",
        "value": round(batched_kss, 2),
        "unit": "KeySwitch/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded uniform residues per prime (inputs and evaluation key)",
        "config": {"workload": "batched Hybrid key switch, N=65536, L=24 Q (60+50-bit), K=6 P (60-bit), "
                               "dnum=4 (alpha=6), OpenFHE-shaped primes and minimum roots",
                   "batch": B, "graph": True},
        "autotune": ks_tuning,
        "batched_p50_ms": round(pct(step_ms, 0.5), 3),
        "batched_p99_ms": round(pct(step_ms, 0.99), 3),
        "plain_keyswitch_per_s": round(plain_kss, 2),
        "plain_p50_ms": round(pct(lat, 0.5), 3),
        "plain_p99_ms": round(pct(lat, 0.99), 3),
        "batched_over_plain": round(batched_kss / plain_kss, 3),
        "plain_session": sess,
        "batched_over_plain_session": round(batched_kss / sess["keyswitch_per_s"], 2) if sess else None,
        "hoisted_rotations": rot,
        "keyswitch_wrappers": wrappers,
        "roofline": roof,
        "valu_roofline": valu,
        "narrow_ntt_roofline": narrow,
        "cpu_baseline": cpu,
        "bit_exact_vs_oracle": exact,
        "served_stream": served,
        "served_keyswitch": served_ks,
        "reference_workloads": refw,
        "mixed_ring_dimensions": mixed,
        "device": hip.device_name(0),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
