"""pcgsq_dpf_out_sums_noalloc, the entry point for callers that cannot
allocate device memory: its device sums equal pcgsq_dpf_out_sums's host sums
on the same device leaves, over consecutive calls that shrink (the fixed
scratch array keeps the previous call's partials), and sizes beyond the
scratch array are refused before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pq():
    import pcgsq
    if not pcgsq.device_available():
        pytest.skip("no gfx950 device")
    return pcgsq


def _leaves(ctx, pq, rng, B, n):
    roots = rng.integers(0, 2**63, (B, 2), dtype=np.uint64)
    dlt = rng.integers(0, 2**63, (B * n, 2), dtype=np.uint64)
    return ctx.dpf_batch_full_eval_device(roots, dlt, n, 1, lambda level, mine: np.zeros_like(mine))


def test_noalloc_sums_match_allocating_entry_point(pq):
    rng = np.random.default_rng(2026)
    with pq.Context(0) as ctx:
        outC = pq.DeviceBuffer(ctx, 8 * 4096)
        outT = pq.DeviceBuffer(ctx, 8 * 4096)
        # large, then smaller shapes: 2^17 leaves per tree (64 partials per
        # tree), 2^11 (one), 2^6 (several trees per workgroup)
        for B, n in ((256, 17), (40, 11), (7, 6), (33, 17)):
            d = _leaves(ctx, pq, rng, B, n)
            wantC, wantT = ctx.dpf_out_sums(d, B, 1 << n)
            ctx.dpf_out_sums_noalloc(d, B, 1 << n, outC.ptr, outT.ptr)
            ctx.synchronize()
            gotC = outC.to_host()[:B]
            gotT = outT.to_host()[:B]
            assert np.array_equal(gotC, wantC), (B, n)
            assert np.array_equal(gotT, wantT), (B, n)
        # beyond the scratch array: 2^17-leaf trees take 1 KiB of partials
        # each, so 4097 of them do not fit in 4 MiB; refused before any launch
        with pytest.raises(pq.PcgsqError, match="scratch"):
            ctx.dpf_out_sums_noalloc(d, 4097, 1 << 17, outC.ptr, outT.ptr)
        ctx.synchronize()
        outC.close()
        outT.close()
