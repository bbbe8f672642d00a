"""Continuous batching: a mixed stream of KeySwitch, NTT/INTT and ModUp/ModDown
requests of different N, L and dnum submitted while earlier batches run.
Every response equals the same request served alone (batch of one, one
stream), and responses are checked bit-exactly against the oracle."""
import numpy as np
import pytest

import oracle as o
from common import log_case
from nttgpu import hip, serve, workload

pytestmark = pytest.mark.gpu


def _oracle(req, hp, kb, ka):
    n, L = hp.n, req.level
    if req.op == serve.KEYSWITCH:
        return o.key_switch(req.data, n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
    if req.op in (serve.FORWARD, serve.INVERSE):
        return o.ntt_towers(req.data, n, hp.q, hp.q_psi, inverse=req.op == serve.INVERSE), None
    if req.op == serve.MODUP:
        return o.mod_up(req.data, n, hp.q, hp.q_psi, hp.p, hp.p_psi, evaluation=True), None
    return o.mod_down(req.data, n, hp.q, hp.q_psi, hp.p, hp.p_psi), None


def test_mixed_stream_alone_equals_in_mix():
    reqs, pools = workload.stream(48, rate_per_s=4000.0, seed=5, sizes=(8192, 16384), levels=(3, 12))
    log_case(f"serve mixed stream: {len(reqs)} requests")
    mixed_srv = serve.Server(streams=4, max_batch=8, hbm_budget_bytes=8 << 30)
    mixed = workload.MixedServer(mixed_srv, pools)
    for r in reqs:  # register and warm every parameter set outside the stream
        mixed.params_for(r)
    tickets = [mixed.submit(r) for r in reqs]  # back-to-back: later ones arrive while earlier run
    got = [mixed.wait(t, r) for t, r in zip(tickets, reqs)]
    stats = mixed_srv.stats()
    assert stats["completed"] == len(reqs)
    assert stats["batches"] < len(reqs), "requests of one group must share launches"
    assert any(bs > 1 for _, _, bs in got)

    alone_srv = serve.Server(streams=1, max_batch=1, hbm_budget_bytes=8 << 30)
    alone = workload.MixedServer(alone_srv, pools)
    for i, r in enumerate(reqs):
        log_case(f"serve alone {i} op={serve.OP_NAMES[r.op]} n={r.n} L={r.level} dnum={r.dnum}")
        (a1, a2), _, bs = alone.wait(alone.submit(r), r)
        (m1, m2), _, _ = got[i]
        assert bs == 1
        assert np.array_equal(a1, m1), f"request {i}: mixed result differs from alone"
        if r.op == serve.KEYSWITCH:
            assert np.array_equal(a2, m2)
        hp, kb, ka = mixed.sets[(r.n, r.level, r.dnum)]
        w1, w2 = _oracle(r, hp, kb, ka)
        assert np.array_equal(m1, w1), f"request {i} ({serve.OP_NAMES[r.op]}) differs from the oracle"
        if w2 is not None:
            assert np.array_equal(m2, w2)
    mixed_srv.close()
    alone_srv.close()


def test_admission_by_footprint_queues_instead_of_failing():
    """A budget that fits one key-switch batch at a time still serves every request."""
    reqs, pools = workload.stream(12, rate_per_s=1e6, seed=9, sizes=(8192,), levels=(6, 6), dnums=(3,))
    for r in reqs:
        r.op = serve.KEYSWITCH
        r.data = r.data[: r.level * r.n]
    srv = serve.Server(streams=4, max_batch=4, hbm_budget_bytes=64 << 20)
    ms = workload.MixedServer(srv, pools)
    for r in reqs:
        ms.params_for(r)
    log_case("serve admission under a 64 MiB budget")
    tickets = [ms.submit(r) for r in reqs]
    outs = [ms.wait(t, r) for t, r in zip(tickets, reqs)]
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["peak_device_bytes"] <= 64 << 20
    hp, kb, ka = ms.sets[(reqs[0].n, reqs[0].level, reqs[0].dnum)]
    for r, ((b, a), _, _) in zip(reqs[:3], outs[:3]):
        wb, wa = o.key_switch(r.data, hp.n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
        assert np.array_equal(b, wb) and np.array_equal(a, wa)
    srv.close()


def test_device_resident_requests_equal_host_requests():
    """Ciphertexts resident in HBM (submit_device): a mixed stream plus a burst of
    one group (key-switch batches > 1 run on the callers' buffers through
    pointer tables) give the same residues as the host path and the oracle."""
    reqs, pools = workload.stream(24, rate_per_s=4000.0, seed=15, sizes=(8192,), levels=(3, 8))
    burst, _ = workload.stream(6, rate_per_s=4000.0, seed=16, sizes=(8192,), levels=(5, 5), dnums=(2,))
    for r in burst:
        r.op = serve.KEYSWITCH
        r.data = r.data[: r.level * r.n]
    for r in burst:
        pools.setdefault(r.n, workload.PrimePool(r.n))
    srv = serve.Server(streams=4, max_batch=8, hbm_budget_bytes=8 << 30)
    ms = workload.MixedServer(srv, pools)
    everything = reqs + burst
    for r in everything:
        ms.params_for(r)
    log_case(f"device-resident serve: {len(everything)} requests")
    dev = []
    for r in everything:
        pid = ms.params_for(r)
        w = srv.output_words(pid, r.op, r.level)
        d_in = hip.DeviceBuffer(r.data.size).upload(r.data)
        d_out = hip.DeviceBuffer(w)
        d_out2 = hip.DeviceBuffer(w) if r.op == serve.KEYSWITCH else None
        dev.append((d_in, d_out, d_out2))
    tickets = [srv.submit_device(ms.params_for(r), r.op, r.level, d[0].ptr, d[1].ptr, d[2].ptr if d[2] else None)
               for r, d in zip(everything, dev)]
    sizes = [srv.wait_device(t)[1] for t in tickets]
    assert max(sizes) > 1, "the burst must share launches"
    host = serve.Server(streams=1, max_batch=1, hbm_budget_bytes=8 << 30)
    hs = workload.MixedServer(host, pools)
    for i, (r, d) in enumerate(zip(everything, dev)):
        (h1, h2), _, _ = hs.wait(hs.submit(r), r)
        assert np.array_equal(d[1].download(), h1), f"request {i} ({serve.OP_NAMES[r.op]}) differs from host path"
        if r.op == serve.KEYSWITCH:
            assert np.array_equal(d[2].download(), h2)
    hp, kb, ka = ms.sets[(burst[0].n, burst[0].level, burst[0].dnum)]
    wb, wa = o.key_switch(burst[0].data, hp.n, hp.q[:5], hp.q_psi[:5], hp.M, hp.p, hp.p_psi, hp.alpha, kb, ka)
    i0 = len(reqs)
    assert np.array_equal(dev[i0][1].download(), wb) and np.array_equal(dev[i0][2].download(), wa)
    srv.close()
    host.close()


def test_served_stream_at_the_stated_ranges():
    """The served mixed stream at its full ranges: N in {2^13, 2^14, 2^15, 2^16},
    L uniform in [3, 30], dnum in {2, 3, 4}, K = ceil(L / dnum), ops 40% KeySwitch,
    30% NTT/INTT, 30% ModUp/ModDown, submitted at their Poisson arrival times.
    Every response equals the same request served alone (tolerance 0), and at
    least 8 responses of every op are checked against the oracle."""
    import time

    reqs, pools = workload.stream(96, rate_per_s=3000.0, seed=0xC5C5)
    sizes = {r.n for r in reqs}
    levels = [r.level for r in reqs]
    assert sizes == {8192, 16384, 32768, 65536} and min(levels) <= 6 and max(levels) >= 27
    srv = serve.Server(streams=4, max_batch=16, hbm_budget_bytes=48 << 30)
    ms = workload.MixedServer(srv, pools)
    for r in reqs:
        ms.params_for(r)
    log_case(f"served stream, stated ranges: {len(reqs)} requests")
    t0 = time.perf_counter()
    tickets = []
    for r in reqs:  # Poisson arrivals: submit each request at its arrival time
        delay = t0 + r.arrival_s - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(ms.submit(r))
    got = [ms.wait(t, r) for t, r in zip(tickets, reqs)]
    assert srv.stats()["completed"] == len(reqs)
    alone_srv = serve.Server(streams=1, max_batch=1, hbm_budget_bytes=48 << 30)
    alone = workload.MixedServer(alone_srv, pools)
    checked = {op: 0 for op in serve.OP_NAMES}
    for i, r in enumerate(reqs):
        (a1, a2), _, _ = alone.wait(alone.submit(r), r)
        (m1, m2), _, _ = got[i]
        assert np.array_equal(a1, m1), f"request {i}: in-mix result differs from alone"
        if r.op == serve.KEYSWITCH:
            assert np.array_equal(a2, m2), f"request {i}: in-mix A differs from alone"
        if checked[r.op] < 8:
            log_case(f"stream oracle {i} op={serve.OP_NAMES[r.op]} n={r.n} L={r.level} dnum={r.dnum}")
            hp, kb, ka = ms.sets[(r.n, r.level, r.dnum)]
            w1, w2 = _oracle(r, hp, kb, ka)
            assert np.array_equal(m1, w1), f"request {i} ({serve.OP_NAMES[r.op]}) differs from the oracle"
            if w2 is not None:
                assert np.array_equal(m2, w2)
            checked[r.op] += 1
    assert all(v >= 8 for v in checked.values()), checked
    srv.close()
    alone_srv.close()


def test_transforms_of_different_levels_share_launches():
    """NTT and INTT requests of one ring dimension but different levels and
    parameter sets run in shared launches (one job table over the ring's
    prime pool); every response equals the same request alone and the oracle."""
    reqs, pools = workload.stream(40, rate_per_s=1e6, seed=21, sizes=(8192,), levels=(3, 20))
    reqs = [r for r in reqs if r.op in (serve.FORWARD, serve.INVERSE)]
    assert len({r.level for r in reqs}) > 3
    srv = serve.Server(streams=2, max_batch=16, hbm_budget_bytes=8 << 30)
    ms = workload.MixedServer(srv, pools)
    for r in reqs:
        ms.params_for(r)
    log_case(f"pooled transforms: {len(reqs)} requests over {len({r.level for r in reqs})} levels")
    tickets = [ms.submit(r) for r in reqs]
    got = [ms.wait(t, r) for t, r in zip(tickets, reqs)]
    st = srv.stats()
    assert st["batches"] <= len(reqs) // 2, st
    mixed_levels = False
    alone = workload.MixedServer(serve.Server(streams=1, max_batch=1, hbm_budget_bytes=8 << 30), pools)
    for i, r in enumerate(reqs):
        (m, _), _, bs = got[i]
        (a, _), _, _ = alone.wait(alone.submit(r), r)
        assert np.array_equal(a, m), f"request {i}: pooled result differs from alone"
        hp, _, _ = ms.sets[(r.n, r.level, r.dnum)]
        assert np.array_equal(m, o.ntt_towers(r.data, r.n, hp.q, hp.q_psi, inverse=r.op == serve.INVERSE))
        mixed_levels |= bs > 1
    assert mixed_levels
    srv.close()
    alone.server.close()


def test_resident_plans_count_and_evict_under_the_budget():
    """Resident contexts and keys count against the budget: key switches
    cycling over more parameter sets than fit evict idle plans (least recently
    used first) instead of failing, the peak stays within the budget, and the
    results equal the oracle."""
    pools = {8192: workload.PrimePool(8192)}
    levels = (10, 12, 14, 16, 18)
    sets = [pools[8192].param_set(L, 2) for L in levels]
    rng = np.random.default_rng(33)
    srv = serve.Server(streams=2, max_batch=2, hbm_budget_bytes=40 << 20)
    ids, keys = [], []
    for hp in sets:
        kb, ka = workload.make_key(hp, 1000 + len(hp.q))
        ids.append(srv.add_params(hp, kb, ka))
        keys.append((kb, ka))
    log_case("serve eviction under a 40 MiB budget")
    reqs = []
    for k in range(15):
        j = k % len(sets)
        x = workload.seeded_residues(8192, sets[j].q, rng)
        reqs.append((j, x, srv.submit(ids[j], serve.KEYSWITCH, len(sets[j].q), x)))
    for j, x, t in reqs:
        (b, a), _, _ = srv.wait(t, ids[j], serve.KEYSWITCH, len(sets[j].q))
        hp = sets[j]
        wb, wa = o.key_switch(x, hp.n, hp.q, hp.q_psi, hp.M, hp.p, hp.p_psi, hp.alpha, *keys[j])
        assert np.array_equal(b, wb) and np.array_equal(a, wa)
    st = srv.stats()
    assert st["completed"] == 15 and st["plans_evicted"] > 0, st
    assert st["peak_device_bytes"] <= 40 << 20 and st["resident_bytes"] > 0, st
    srv.close()


def test_submit_and_wait_on_different_threads():
    """One thread submits while two others wait on the tickets as they come
    (the C calls release the GIL, so they overlap): every result equals the
    oracle, and a ticket that has been collected is unknown afterwards."""
    import queue
    import threading

    pools = {8192: workload.PrimePool(8192)}
    hp = pools[8192].param_set(6, 2)
    kb, ka = workload.make_key(hp, 77)
    srv = serve.Server(streams=2, max_batch=8, hbm_budget_bytes=4 << 30)
    pid = srv.add_params(hp, kb, ka)
    rng = np.random.default_rng(91)
    inputs = [workload.seeded_residues(8192, hp.q, rng) for _ in range(48)]
    log_case("serve: submit and wait on different threads, 48 transforms")
    work, results, errors = queue.Queue(), {}, []

    def submitter():
        try:
            for i, x in enumerate(inputs):
                work.put((i, srv.submit(pid, serve.FORWARD, len(hp.q), x)))
        except Exception as e:  # noqa: BLE001 - reported by the main thread
            errors.append(e)
        finally:
            for _ in range(2):
                work.put(None)

    def waiter():
        try:
            while True:
                item = work.get()
                if item is None:
                    return
                i, t = item
                (out, _), _, _ = srv.wait(t, pid, serve.FORWARD, len(hp.q))
                results[i] = (t, out)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=submitter)] + [threading.Thread(target=waiter) for _ in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a submitting or waiting thread did not finish"
    assert not errors, errors
    assert sorted(results) == list(range(len(inputs)))
    for i, x in enumerate(inputs):
        assert np.array_equal(results[i][1], o.ntt_towers(x, 8192, hp.q, hp.q_psi)), f"request {i}"
    with pytest.raises(Exception, match="unknown ticket"):
        srv.wait(results[0][0], pid, serve.FORWARD, len(hp.q))
    srv.close()


def test_transforms_of_different_ring_dimensions_share_launches():
    """NTT and INTT requests of N = 2^13 .. 2^16 (three per ring dimension)
    submitted back to back on one slot: a batch of four or more must span ring
    dimensions (one mixed launch per pass); every response equals the oracle
    and the same request served alone."""
    sizes = (8192, 16384, 32768, 65536)
    pools = {n: workload.PrimePool(n) for n in sizes}
    srv = serve.Server(streams=1, max_batch=16, hbm_budget_bytes=8 << 30)
    alone = serve.Server(streams=1, max_batch=1, hbm_budget_bytes=8 << 30)
    rng = np.random.default_rng(64)
    reqs = []
    for k in range(12):
        n = sizes[k % 4]
        hp = pools[n].param_set(3 + k % 3, 2)
        kb, ka = workload.make_key(hp, 500 + k)
        op = serve.FORWARD if k % 2 == 0 else serve.INVERSE
        reqs.append((hp, op, srv.add_params(hp, kb, ka), alone.add_params(hp, kb, ka),
                     workload.seeded_residues(n, hp.q, rng)))
    log_case("serve: transforms over four ring dimensions")
    tickets = [srv.submit(pid, op, len(hp.q), x) for hp, op, pid, _, x in reqs]
    got = [srv.wait(t, pid, op, len(hp.q)) for t, (hp, op, pid, _, x) in zip(tickets, reqs)]
    sizes_seen = [bs for _, _, bs in got]
    assert max(sizes_seen) >= 4, f"no batch spans ring dimensions: batch sizes {sizes_seen}"
    for (hp, op, pid, apid, x), ((out, _), _, _) in zip(reqs, got):
        want = o.ntt_towers(x, hp.n, hp.q, hp.q_psi, inverse=op == serve.INVERSE)
        assert np.array_equal(out, want), f"N={hp.n} L={len(hp.q)} op={op}"
        (a, _), _, _ = alone.wait(alone.submit(apid, op, len(hp.q), x), apid, op, len(hp.q))
        assert np.array_equal(a, out)
    srv.close()
    alone.close()


def test_device_resident_transforms_of_different_ring_dimensions():
    """Device-resident NTT and INTT requests of N = 2^13 .. 2^16 run in mixed
    launches on the callers' buffers (no copies through the slot arena), some
    in place (input buffer == output buffer): every result equals the oracle,
    and some batch spans ring dimensions."""
    sizes = (8192, 16384, 32768, 65536)
    pools = {n: workload.PrimePool(n) for n in sizes}
    srv = serve.Server(streams=1, max_batch=16, hbm_budget_bytes=8 << 30)
    rng = np.random.default_rng(65)
    reqs = []
    for k in range(16):
        n = sizes[k % 4]
        hp = pools[n].param_set(2 + k % 4, 2)
        kb, ka = workload.make_key(hp, 700 + k)
        op = serve.FORWARD if k % 3 else serve.INVERSE
        x = workload.seeded_residues(n, hp.q, rng)
        d_in = hip.DeviceBuffer(x.size).upload(x)
        d_out = d_in if k % 5 == 0 else hip.DeviceBuffer(x.size)
        reqs.append((hp, op, srv.add_params(hp, kb, ka), x, d_in, d_out))
    log_case("serve: device-resident transforms over four ring dimensions")
    tickets = [srv.submit_device(pid, op, len(hp.q), d_in.ptr, d_out.ptr) for hp, op, pid, x, d_in, d_out in reqs]
    seen = [srv.wait_device(t)[1] for t in tickets]
    assert max(seen) >= 4, f"no batch spans ring dimensions: batch sizes {seen}"
    for hp, op, pid, x, d_in, d_out in reqs:
        want = o.ntt_towers(x, hp.n, hp.q, hp.q_psi, inverse=op == serve.INVERSE)
        assert np.array_equal(d_out.download(), want), f"N={hp.n} L={len(hp.q)} op={op}"
    srv.close()
