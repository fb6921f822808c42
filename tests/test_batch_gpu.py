"""Batched NTT of uneven lengths in one persistent launch or grouped by
length (pcgsq_ntt_batch_*), through the C ABI.  Needs an MI355X.  Every result is compared bit for bit
with the per-length launches (the single-launch four-step / one-pass path)
and, for the shorter lengths, with the CPU oracle."""
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


@pytest.fixture(scope="module")
def ctx(pq):
    c = pq.Context(0)
    yield c
    c.close()


def _check_batch(ctx, b, a, oracle_max_log=14):
    f, r = b.read()
    assert np.array_equal(r, a), "inverse(forward(a)) != a"
    fs, as_ = b.split(f), b.split(a)
    for i, L in enumerate(b.logN):
        exp = ctx.ntt_forward(as_[i][None, :])[0]
        assert np.array_equal(fs[i], exp), f"transform {i} (2^{L}) differs from the per-length launch"
        if L <= oracle_max_log and i % 7 == 0:
            assert np.array_equal(fs[i], O.ntt_forward(as_[i])), f"transform {i} vs oracle"


def test_ntt_batch_uneven_lengths(ctx, pq):
    rng = np.random.default_rng(3)
    logN = [int(x) for x in rng.choice([10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22], 48)]
    with pq.NttBatch(ctx, logN) as b:
        a = rng.integers(0, P, b.total, dtype=np.uint64)
        b.load(a)
        b.run()
        b.finish()
        _check_batch(ctx, b, a)


def test_ntt_batch_forward_only_and_inverse_only(ctx, pq):
    rng = np.random.default_rng(4)
    logN = [12, 20, 14, 10, 16, 13]
    with pq.NttBatch(ctx, logN, ops=pq.NTT_BATCH_FORWARD) as b:
        a = rng.integers(0, P, b.total, dtype=np.uint64)
        b.load(a)
        b.run()
        f, _ = b.read()
        for i, x in enumerate(b.split(a)):
            assert np.array_equal(b.split(f)[i], ctx.ntt_forward(x[None, :])[0])
    with pq.NttBatch(ctx, logN, ops=pq.NTT_BATCH_INVERSE) as b:
        b.load(a)
        b.run()
        _, r = b.read()
        for i, x in enumerate(b.split(a)):
            assert np.array_equal(b.split(r)[i], ctx.ntt_inverse(x[None, :])[0])


def test_ntt_batch_every_split(ctx, pq):
    """Every split the batch tuner may pick gives the same transforms."""
    rng = np.random.default_rng(5)
    logN = [13, 14, 17, 20, 12]
    with pq.NttBatch(ctx, logN) as b:
        a = rng.integers(0, P, b.total, dtype=np.uint64)
        b.load(a)
        for L in sorted(set(logN)):
            splits = [s for s in range(7, 12) if 6 <= L - s <= 11] + ([L] if L <= 12 else [])
            for s in splits:
                b.tune(L, s)
                b.run()
                b.finish()
                _check_batch(ctx, b, a, oracle_max_log=0)
            b.tune(L, 0)
        for wgs in (1, 2):
            b.tune(0, 0, wgs)
            b.run()
            b.finish()
            _check_batch(ctx, b, a, oracle_max_log=0)


def test_ntt_batch_graph_replays(ctx, pq):
    """The captured launch replays correctly, with eager work and new inputs
    between replays (every replay checked)."""
    rng = np.random.default_rng(6)
    logN = [12, 18, 14, 20, 16, 12, 15]
    with pq.NttBatch(ctx, logN) as b:
        for rep in range(3):
            a = rng.integers(0, P, b.total, dtype=np.uint64)
            b.load(a)
            b.run(graph=True)
            b.finish()
            _check_batch(ctx, b, a, oracle_max_log=0)
            ctx.ntt_forward(a[:1 << 12][None, :])      # eager work between replays
        b.run(graph=False)
        b.finish()
        _check_batch(ctx, b, a, oracle_max_log=0)


def test_ntt_batch_config3_workload(ctx, pq):
    """256 transforms, lengths drawn from mt19937_64(seed) over {2^12 .. 2^20}
    (steps of 4x), coefficients mt19937_64(seed) % P, forward then inverse:
    bit-exact against the single-launch path per transform."""
    m = O.MT64(2024)
    logN = [12 + 2 * int(x) for x in m.take(256, 5)]
    with pq.NttBatch(ctx, logN) as b:
        a = m.take(b.total, P)
        b.load(a)
        b.run(graph=True)
        b.finish()
        _check_batch(ctx, b, a)


def test_ntt_batch_grouped_mode(ctx, pq):
    """Grouped by length (per-length pass launches over offset tables) gives
    the same transforms as the persistent launch, for both directions, one
    direction only, graph replays and switches between the modes."""
    rng = np.random.default_rng(7)
    logN = [int(x) for x in rng.choice([10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22], 40)]
    with pq.NttBatch(ctx, logN) as b:
        a = rng.integers(0, P, b.total, dtype=np.uint64)
        b.load(a)
        b.set_mode(pq.NTT_BATCH_GROUPED)
        assert b.info()["grid"] == 0
        for graph in (False, True, True):
            b.run(graph=graph)
            b.finish()
            _check_batch(ctx, b, a)
        b.set_mode(pq.NTT_BATCH_PERSISTENT)
        b.run(graph=True)
        b.finish()
        _check_batch(ctx, b, a, oracle_max_log=0)
    for ops, which in ((pq.NTT_BATCH_FORWARD, 0), (pq.NTT_BATCH_INVERSE, 1)):
        with pq.NttBatch(ctx, logN[:12], ops=ops) as b:
            a = rng.integers(0, P, b.total, dtype=np.uint64)
            b.load(a)
            b.set_mode(pq.NTT_BATCH_GROUPED)
            b.run()
            got = b.split(b.read()[which])
            for i, x in enumerate(b.split(a)):
                exp = (ctx.ntt_forward if which == 0 else ctx.ntt_inverse)(x[None, :])[0]
                assert np.array_equal(got[i], exp), f"transform {i} (2^{b.logN[i]})"


def test_ntt_batch_grouped_more_rows_than_one_launch(ctx, pq):
    """65536 transforms of one length (one more than a launch has rows) plus
    one other: the grouped mode splits the length over two launches and
    equals the persistent launch."""
    rng = np.random.default_rng(9)
    logN = [10] * 65536
    logN[12345] = 12
    with pq.NttBatch(ctx, logN) as b:
        a = rng.integers(0, P, b.total, dtype=np.uint64)
        b.load(a)
        b.set_mode(pq.NTT_BATCH_GROUPED)
        b.run(graph=True)
        b.finish()
        f, r = b.read()
        assert np.array_equal(r, a)
        b.set_mode(pq.NTT_BATCH_PERSISTENT)
        b.run()
        b.finish()
        f2, _ = b.read()
        assert np.array_equal(f, f2)
        fs, as_ = b.split(f), b.split(a)
        for i in (0, 12345, 65535):
            assert np.array_equal(fs[i], O.ntt_forward(as_[i])), f"transform {i} vs oracle"


def test_ntt_batch_grouped_follows_context_shapes(pq):
    """The grouped mode runs each length in the context's tuned shape (split,
    grouping, tile widths): unusual shapes give the same transforms."""
    rng = np.random.default_rng(8)
    with pq.Context(0, apply_tuning=False) as c:
        c.ntt_tune(1 << 16, 7, 1, 2, 8)
        c.ntt_tune(1 << 13, 13, 0, 4, 0)
        c.ntt_tune(1 << 20, 11, 2, 0, 4)
        logN = [16, 13, 20, 16, 13, 10]
        with pq.NttBatch(c, logN) as b:
            a = rng.integers(0, P, b.total, dtype=np.uint64)
            b.load(a)
            b.set_mode(pq.NTT_BATCH_GROUPED)
            b.run(graph=True)
            b.finish()
            f, r = b.read()
            assert np.array_equal(r, a)
            for i, x in enumerate(b.split(a)):
                assert np.array_equal(b.split(f)[i], O.ntt_forward(x)), f"transform {i} vs oracle"


def test_ntt_batch_spin_limit_reports_error(ctx, pq):
    """A dependency wait that runs out of spins is reported, not hidden (the
    limit is set from the host; every workgroup still exits normally)."""
    with pq.NttBatch(ctx, [20]) as b:
        b.load(np.zeros(b.total, dtype=np.uint64))
        b.tune(0, 0, 0, 1)
        b.run()
        with pytest.raises(pq.PcgsqError):
            b.finish()
        b.tune(0, 0, 0, 0)                  # default limit again: runs clean
        b.run()
        b.finish()


# ---- batched DPF keys (dealer Gen + sync-free eval + leaf convert) -------------

def _rks(prg):
    return O.aes_tree_round_keys(1) if prg == O.PRG_AES else None


def _check_point_function(b, y0, y1, alpha, beta):
    s = y0 + y1                                   # both < P < 2^62: no overflow
    s = np.where(s >= np.uint64(P), s - np.uint64(P), s)
    nz = np.flatnonzero(s)
    exp = (b.offsets[:-1] + alpha.astype(np.int64))
    assert np.array_equal(nz, exp), "y0 + y1 is not a point function at the keys' alphas"
    assert np.array_equal(s[nz], beta)


@pytest.mark.parametrize("prg", [O.PRG_CHACHA8, O.PRG_AES])
def test_dpf_batch_matches_oracle(ctx, pq, prg):
    rng = np.random.default_rng(10 + prg)
    logn = [int(x) for x in rng.choice([10, 11, 12, 13, 14, 15, 16], 20)]
    K = len(logn)
    alpha = np.array([rng.integers(0, 1 << n) for n in logn], dtype=np.uint64)
    beta = rng.integers(1, P, K, dtype=np.uint64)
    delta = rng.integers(0, 2**63, (K, 2), dtype=np.uint64)
    delta[:, 0] |= np.uint64(1)
    d0 = rng.integers(0, 2**63, (K, 2), dtype=np.uint64)
    with pq.DpfBatch(ctx, logn, prg) as b:
        b.gen(alpha, beta, delta, d0)
        roots, cw, cw_out = b.keys()
        b.run()
        y0, y1 = b.read()
        _check_point_function(b, y0, y1, alpha, beta)
        rks = _rks(prg)
        for k, n in enumerate(logn):
            r0, r1, ecw, eco = O.dpf_key_gen(n, prg, int(alpha[k]), int(beta[k]), delta[k], d0[k],
                                             rks=rks)
            assert tuple(int(v) for v in roots[0, k]) == r0 and tuple(int(v) for v in roots[1, k]) == r1
            assert np.array_equal(cw[b.cw_offsets[k]:b.cw_offsets[k + 1]], ecw)
            assert int(cw_out[k]) == eco
            if k % 3 == 0:
                assert np.array_equal(b.key_slice(y0, k), O.dpf_key_eval_y(n, prg, r0, ecw, eco, 0, rks=rks))
                assert np.array_equal(b.key_slice(y1, k), O.dpf_key_eval_y(n, prg, r1, ecw, eco, 1, rks=rks))


def test_dpf_batch_dealer_spec(ctx, pq):
    """deal(seed) draws alpha, beta, Delta, Delta_0 by the documented dealer
    spec (domains 10-13) and generates the same keys as gen() on them."""
    logn = [10, 12, 11]
    seed = 77
    alpha = np.array([pq.dealer_u64(seed, 10, k, 0) & ((1 << n) - 1) for k, n in enumerate(logn)],
                     dtype=np.uint64)
    beta = np.array([pq.dealer_u64(seed, 11, k, 0) % (P - 1) + 1 for k in range(3)], np.uint64)
    delta = np.array([[pq.dealer_u64(seed, 12, k, 0) | 1, pq.dealer_u64(seed, 12, k, 1)]
                      for k in range(3)], np.uint64)
    d0 = np.array([[pq.dealer_u64(seed, 13, k, 0) | 1, pq.dealer_u64(seed, 13, k, 1)]
                   for k in range(3)], np.uint64)
    with pq.DpfBatch(ctx, logn) as b:
        b.deal(seed)
        a2, b2 = b.points()
        assert np.array_equal(a2, alpha) and np.array_equal(b2, beta)
        k1 = b.keys()
        b.gen(alpha, beta, delta, d0)
        k2 = b.keys()
        for u, v in zip(k1, k2):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("prg", [O.PRG_CHACHA8, O.PRG_AES])
def test_dpf_batch_config4_workload(ctx, pq, prg):
    """4096 key pairs, domains 2^10..2^18 (n uniform from mt19937_64), dealt
    points: both parties' converted domains form beta * e_alpha for every key
    (graph replay), and a sample of keys equals the oracle."""
    m = O.MT64(4096 + prg)
    logn = [10 + int(x) for x in m.take(4096, 9)]
    with pq.DpfBatch(ctx, logn, prg) as b:
        b.deal(5)
        b.run(graph=True)
        y0, y1 = b.read()
        alpha, beta = b.points()
        _check_point_function(b, y0, y1, alpha, beta)
        roots, cw, cw_out = b.keys()
        rks = _rks(prg)
        order = np.argsort(logn)
        for k in list(order[:3]) + [int(order[-1])]:
            n = logn[k]
            ecw = cw[b.cw_offsets[k]:b.cw_offsets[k + 1]]
            assert np.array_equal(b.key_slice(y0, k),
                                  O.dpf_key_eval_y(n, prg, roots[0, k], ecw, cw_out[k], 0, rks=rks))
            assert np.array_equal(b.key_slice(y1, k),
                                  O.dpf_key_eval_y(n, prg, roots[1, k], ecw, cw_out[k], 1, rks=rks))


def test_dpf_batch_graph_replays_new_keys(ctx, pq):
    """The captured eval replays with keys regenerated in between (every
    replay checked as a point function)."""
    rng = np.random.default_rng(12)
    logn = [int(x) for x in rng.integers(10, 15, 40)]
    with pq.DpfBatch(ctx, logn) as b:
        for rep in range(3):
            b.deal(100 + rep)
            b.run(graph=True)
            y0, y1 = b.read()
            alpha, beta = b.points()
            _check_point_function(b, y0, y1, alpha, beta)
