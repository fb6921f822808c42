"""Batched NTT of uneven lengths in one persistent launch (pcgsq_ntt_batch_*),
through the C ABI.  Needs an MI355X.  Every result is compared bit for bit
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
