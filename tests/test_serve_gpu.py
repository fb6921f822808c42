"""The expansion server (pcgsq_server_*): continuous batching with admission
by device-memory budget.  Needs an MI355X.  Every served result is compared
with the oracle (small N) or with the same request run alone (large N), bit
for bit, and the OLE relation x0 * x1 = z0 + z1 is checked."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = O.P


@pytest.fixture(scope="module")
def pq():
    import pcgsq
    if not pcgsq.device_available():
        pytest.fail("no gfx950 device visible (the GPU tests must run on an MI355X)")
    return pcgsq


def _digest(outs):
    return [O.fnv1a64(v) for v in outs]


def test_served_small_mix_matches_oracle(pq):
    reqs = [(1 << 12, 2, 16, 11), (1 << 13, 2, 16, 12), (1 << 12, 3, 16, 13),
            (1 << 14, 2, 16, 14), (1 << 12, 2, 16, 15), (1 << 13, 3, 16, 16)]
    with pq.Server(0, 0, 3) as srv:
        ids = [srv.submit(N, c, t, seed) for (N, c, t, seed) in reqs]
        for rid, (N, c, t, seed) in zip(ids, reqs):
            srv.wait(rid, 60.0)
            got = srv.outputs(rid, N)
            exp = O.pcg_ole_expand(N, c, t, 1, O.PRG_CHACHA8, seed=seed, pub_seed=0, threads=8)
            for g, e, name in zip(got, exp, ("x0", "z0", "x1", "z1")):
                assert np.array_equal(g, e), f"request N={N} c={c} seed={seed}: {name}"
            srv.release(rid)
        assert srv.stats()["completed"] == len(reqs)


def test_served_result_same_alone_and_in_mix(pq):
    """Requests of 2^16..2^22 submitted together overlap on a 4-slot server;
    each result equals the same request run alone (1 slot, nothing else
    running), and the OLE relation holds for each."""
    mix = [(1 << 20, 101), (1 << 16, 102), (1 << 18, 103), (1 << 22, 104), (1 << 16, 105),
           (1 << 20, 106), (1 << 18, 107), (1 << 16, 108)]
    mixed = {}
    with pq.Server(0, 0, 4) as srv:
        ids = [srv.submit(N, 4, 16, seed) for (N, seed) in mix]
        overlapped = False
        for rid, (N, seed) in zip(ids, mix):
            srv.wait(rid, 120.0)
            mixed[seed] = _digest(srv.outputs(rid, N))
        times = [srv.timing(rid) for rid in ids]
        # some request started before another finished: they overlapped
        for i, (_, si, di) in enumerate(times):
            for j, (_, sj, dj) in enumerate(times):
                if i != j and si < dj and sj < di:
                    overlapped = True
        assert overlapped, "no two requests ran at the same time"
        for rid in ids:
            srv.release(rid)
    with pq.Server(0, 0, 1) as alone, pq.Context(0) as ctx:
        for N, seed in mix:
            rid = alone.submit(N, 4, 16, seed)
            alone.wait(rid, 120.0)
            outs = alone.outputs(rid, N)
            assert _digest(outs) == mixed[seed], f"N={N} seed={seed}: differs alone vs in the mix"
            x0, z0, x1, z1 = outs
            if N <= (1 << 20):
                prod = ctx.poly_mul(x0[None, :], x1[None, :])[0]
                s = z0 + z1
                s = np.where(s >= np.uint64(P), s - np.uint64(P), s)
                assert np.array_equal(prod, s), f"N={N} seed={seed}: x0*x1 != z0+z1"
            alone.release(rid)


def test_served_admission_respects_memory_budget(pq):
    """A budget that holds one 2^18 request at a time: four slots, but the
    requests run one after another and the peak stays within the budget."""
    budget = 8 << 30
    for _ in range(8):                       # the per-plan budget follows the total
        with pq.Server(0, budget, 4) as probe:
            charge = probe.charge(1 << 18, 4, 16)
        if charge <= budget < 2 * charge:
            break
        budget = charge + charge // 2
    assert charge <= budget < 2 * charge
    with pq.Server(0, budget, 4) as srv:
        ids = [srv.submit(1 << 18, 4, 16, 200 + k) for k in range(4)]
        for rid in ids:
            srv.wait(rid, 120.0)
        st = srv.stats()
        assert st["peak"] <= budget
        times = sorted(srv.timing(rid)[1:] for rid in ids)
        for (s0, d0), (s1, d1) in zip(times, times[1:]):
            assert s1 >= d0 - 1e-3, "two requests ran together despite the budget"
        for rid in ids:
            srv.release(rid)


def test_served_rejects_request_beyond_budget(pq):
    with pq.Server(0, 1 << 30, 2) as srv:
        with pytest.raises(pq.PcgsqError):
            srv.submit(1 << 22, 4, 16, 1)


def test_served_poisson_run(pq):
    """A short open-loop Poisson run over the four ring dimensions: every
    request completes and each output of the smallest size equals the oracle."""
    from pcgsq import serving
    checked = []

    def keep(rid, N, seed):
        if N == (1 << 16) and len(checked) < 2:
            checked.append((seed, srv.outputs(rid, N)))

    with pq.Server(0, 0, 4) as srv:
        recs, wall = serving.run_poisson(srv, rate=20.0, duration_s=1.0,
                                         sizes=[1 << 16, 1 << 18, 1 << 20, 1 << 22], c=4,
                                         t=16, seed=3, keep=keep)
    assert len(recs) > 5
    assert all(r["latency_ms"] > 0 for r in recs)
    for seed, outs in checked:
        exp = O.pcg_ole_expand(1 << 16, 4, 16, 1, O.PRG_CHACHA8, seed=seed, pub_seed=0,
                               threads=16)
        for g, e in zip(outs, exp):
            assert np.array_equal(g, e)


def test_served_busy_shape_does_not_block_other_shapes(pq):
    """With one 2^22 plan, a second 2^22 request waits for it, and a 2^16
    request queued behind that second one starts before it (skip-ahead
    admission), while the two 2^22 requests keep their order."""
    with pq.Server(0, 0, 3) as srv:
        srv.reserve(1 << 22, 4, 16, 1)
        srv.reserve(1 << 16, 4, 16, 1)
        big1 = srv.submit(1 << 22, 4, 16, 301)
        big2 = srv.submit(1 << 22, 4, 16, 302)
        small = srv.submit(1 << 16, 4, 16, 303)
        for rid in (big1, big2, small):
            srv.wait(rid, 120.0)
        _, s1, d1 = srv.timing(big1)
        _, s2, _ = srv.timing(big2)
        _, ss, ds = srv.timing(small)
        assert s2 >= d1 - 1e-3, "the second 2^22 request started before the first ended"
        assert ss < s2 and ds <= d1 + 0.05, "the 2^16 request waited behind the busy shape"
        assert srv.plans()["built"] == 2
        for rid in (big1, big2, small):
            srv.release(rid)


def test_served_steady_state_allocates_nothing(pq):
    """After one request per shape, a second round of the same mix builds and
    evicts no plan and charges no new device memory (outputs are pooled), and
    its results equal the first round's for the same seeds."""
    mix = [(1 << 16, 401), (1 << 18, 402), (1 << 20, 403), (1 << 16, 404)]
    with pq.Server(0, 0, 4) as srv:
        digests = []
        for rnd in range(2):
            ids = [srv.submit(N, 4, 16, seed) for N, seed in mix]
            for rid in ids:
                srv.wait(rid, 120.0)
            digests.append([_digest(srv.outputs(rid, N)) for rid, (N, _) in zip(ids, mix)])
            for rid in ids:
                srv.release(rid)
            if rnd == 0:
                used0, plans0 = srv.stats()["used"], srv.plans()
        assert srv.stats()["used"] == used0
        assert srv.plans() == plans0 and plans0["evicted"] == 0
        assert digests[0] == digests[1]


def test_served_window_keeps_large_requests_in_order(pq):
    """Window 2^22 + 2^19: two 2^22 requests and a 2^20 one start in arrival
    order and never overlap one another, while 2^16 requests submitted after
    them run alongside; every result equals the same request run alone."""
    win = (1 << 22) + (1 << 19)
    with pq.Server(0, 0, 4, window=win) as srv:
        for N in (1 << 16, 1 << 20, 1 << 22):
            srv.reserve(N, 4, 16, 2)
        heavy = [srv.submit(1 << 22, 4, 16, 501), srv.submit(1 << 20, 4, 16, 502),
                 srv.submit(1 << 22, 4, 16, 503)]
        small = [srv.submit(1 << 16, 4, 16, 510 + k) for k in range(3)]
        for rid in heavy + small:
            srv.wait(rid, 120.0)
        th = [srv.timing(rid) for rid in heavy]
        for (_, s0, d0), (_, s1, _) in zip(th, th[1:]):
            assert s1 >= d0 - 1e-3, "two large requests overlapped despite the window"
        ts = [srv.timing(rid) for rid in small]
        assert all(d <= th[0][2] + 1e-3 for _, _, d in ts), \
            "small requests waited behind the large ones"
        outs = {rid: _digest(srv.outputs(rid, N)) for rid, N in
                zip(heavy + small, [1 << 22, 1 << 20, 1 << 22] + [1 << 16] * 3)}
        for rid in heavy + small:
            srv.release(rid)
    with pq.Server(0, 0, 1) as alone:
        for rid, (N, seed) in zip(heavy + small, [(1 << 22, 501), (1 << 20, 502), (1 << 22, 503)]
                                  + [(1 << 16, 510 + k) for k in range(3)]):
            r2 = alone.submit(N, 4, 16, seed)
            alone.wait(r2, 120.0)
            assert _digest(alone.outputs(r2, N)) == outs[rid], f"N={N} seed={seed}"
            alone.release(r2)


def test_served_outputs_match_precomputed_cpu_hashes(pq):
    """The served-path gate: every request of tests/golden/served_hashes.json
    (2^16..2^22, hashes written by the CPU oracle) submitted together to a
    4-slot server returns x0, z0, x1, z1 with exactly those FNV-1a-64 hashes."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "served_hashes.json")
    reqs = json.load(open(path))["requests"]
    with pq.Server(0, 0, 4, window=(1 << 22) + (1 << 19)) as srv:
        ids = [srv.submit(r["N"], r["c"], r["t"], r["seed"], r["pub_seed"]) for r in reqs]
        for rid, r in zip(ids, reqs):
            srv.wait(rid, 120.0)
            got = [f"{h:016x}" for h in _digest(srv.outputs(rid, r["N"]))]
            assert got == r["fnv1a64"], f"N={r['N']} seed={r['seed']}: hashes differ from the CPU oracle"
            srv.release(rid)


def test_served_exclusive_class_runs_large_requests_alone(pq):
    """exclusive_N = 2^20: 2^20 and 2^22 requests never overlap one another
    and start in arrival order even with two plans per shape, while 2^16
    requests submitted after them run alongside; every result equals the same
    request run alone."""
    with pq.Server(0, 0, 4, exclusive_N=1 << 20) as srv:
        for N in (1 << 16, 1 << 20, 1 << 22):
            srv.reserve(N, 4, 16, 2)
        heavy = [(1 << 20, 601), (1 << 22, 602), (1 << 20, 603), (1 << 22, 604)]
        small = [(1 << 16, 610 + k) for k in range(4)]
        hid = [srv.submit(N, 4, 16, s) for N, s in heavy]
        sid = [srv.submit(N, 4, 16, s) for N, s in small]
        for rid in hid + sid:
            srv.wait(rid, 120.0)
        th = [srv.timing(rid) for rid in hid]
        for (_, s0, d0), (_, s1, _) in zip(th, th[1:]):
            assert s1 >= d0 - 1e-3, "two exclusive-class requests overlapped"
        ts = [srv.timing(rid) for rid in sid]
        assert all(s < th[0][2] for _, s, _ in ts), "small requests did not run alongside"
        outs = {rid: _digest(srv.outputs(rid, N)) for rid, (N, _) in zip(hid + sid, heavy + small)}
        for rid in hid + sid:
            srv.release(rid)
    with pq.Server(0, 0, 1) as alone:
        for rid, (N, seed) in zip(hid + sid, heavy + small):
            r2 = alone.submit(N, 4, 16, seed)
            alone.wait(r2, 120.0)
            assert _digest(alone.outputs(r2, N)) == outs[rid], f"N={N} seed={seed}"
            alone.release(r2)


def test_served_age_bound_stops_skip_ahead(pq):
    """A 2^22 request B that waits for the 2^22 request A before it (one plan
    of that shape) is passed by 2^16 requests submitted after it only until B
    is older than the age bound: every later request that starts before B
    starts within the bound (plus the worker's polling), and none of those
    submitted after B aged starts before B."""
    bound = 0.010
    with pq.Server(0, 0, 4, age_bound_s=bound) as srv:
        srv.reserve(1 << 22, 4, 16, 1)
        srv.reserve(1 << 16, 4, 16, 2)
        a = srv.submit(1 << 22, 4, 16, 701)
        b = srv.submit(1 << 22, 4, 16, 702)
        smalls = []
        t_end = pq.Server.clock() + 0.15
        k = 0
        while pq.Server.clock() < t_end:   # a stream of small requests behind B
            smalls.append(srv.submit(1 << 16, 4, 16, 710 + k))
            k += 1
            import time
            time.sleep(0.002)
        for rid in [a, b] + smalls:
            srv.wait(rid, 120.0)
        tb_sub, tb_start, _ = srv.timing(b)
        _, _, ta_done = srv.timing(a)
        assert tb_start >= ta_done - 1e-3
        early = 0
        for rid in smalls:
            sub, start, _ = srv.timing(rid)
            if start < tb_start:
                early += 1
                assert start <= tb_sub + bound + 0.005, "a later request passed B after it aged"
            if sub > tb_sub + bound:
                assert start >= tb_start - 1e-3, "a request submitted after B aged started first"
        assert early >= 1, "no small request ran while B waited within its bound"
        for rid in [a, b] + smalls:
            srv.release(rid)
