"""Autotuner for the four-step NTT pass shape.

For each transform length N it times forward + inverse of a device-resident
batch for every candidate shape -- column/row split (n1 = 2^n1_log), stages
per register group (radix 2^r) and sub-transforms per workgroup -- with the
library's own per-launch event timing, and keeps the fastest.  Shapes never
change results (every candidate is checked bit-exact against the default
shape before it is timed).

    python pcg-square_amd/pcgsq/autotune.py --out tuning.json

The cache maps device name -> {N: [n1_log, radix, cols_max, us_per_transform]}
and is applied by pcgsq.Context(apply_tuning=True).
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(HERE, "ntt_tuning.json")


def candidates(N: int):
    L = N.bit_length() - 1
    base = (L + 1) // 2
    splits = sorted({s for s in (base - 1, base, base + 1) if 1 <= s < L and max(s, L - s) <= 12})
    for s, r, c in itertools.product(splits, (3, 4, 5), (4, 8, 16)):
        yield s, r, c


def time_shape(ctx, buf, batch, N, iters=3) -> float:
    """Average device ms of forward + inverse on `buf` (batch x N)."""
    ctx.ntt_forward_dev(buf.ptr, batch, N)
    ctx.ntt_inverse_dev(buf.ptr, batch, N)
    ctx.synchronize()
    ctx.profile(True)
    for _ in range(iters):
        ctx.ntt_forward_dev(buf.ptr, batch, N)
        ctx.ntt_inverse_dev(buf.ptr, batch, N)
    ctx.synchronize()
    z = ctx.profile_read()
    ctx.profile(False)
    return sum(v[1] for k, v in z.items() if k.startswith("ntt.")) / iters


def tune(ctx, Ns, elems=1 << 24, log=print) -> dict:
    import pcgsq
    res = {}
    rng = np.random.default_rng(1)
    for N in Ns:
        batch = max(1, elems // N)
        a = rng.integers(0, pcgsq.P, (batch, N), dtype=np.uint64)
        ctx.ntt_tune(N)                      # default shape as the reference result
        ref = ctx.ntt_forward(a[:1])
        buf = ctx.to_device(a)
        best = None
        for s, r, c in candidates(N):
            ctx.ntt_tune(N, s, r, c)
            if not np.array_equal(ctx.ntt_forward(a[:1]), ref):
                raise RuntimeError(f"shape {(s, r, c)} changed the result at N={N}")
            ms = time_shape(ctx, buf, batch, N)
            us = ms * 1e3 / batch
            log(f"N=2^{N.bit_length() - 1} n1_log={s} radix={r} cols={c}: {us:.2f} us/transform "
                f"(fwd+inv)")
            if best is None or us < best[3]:
                best = [s, r, c, us]
        buf.close()
        ctx.ntt_tune(N, *best[:3])
        res[str(N)] = best
        log(f"best N=2^{N.bit_length() - 1}: {best}")
    return res


def apply(ctx, path: str = DEFAULT_CACHE) -> int:
    """Apply the cached shapes for this device; returns how many were set."""
    try:
        with open(path) as f:
            cache = json.load(f)
    except (OSError, ValueError):
        return 0
    dev = ctx.device_name()
    entry = cache.get(dev) or cache.get(dev.split(" (")[-1].split(",")[0])
    if not entry:
        return 0
    for N, (s, r, c, _us) in entry.items():
        ctx.ntt_tune(int(N), int(s), int(r), int(c))
    return len(entry)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--logN", default="12,14,16,18,20,22")
    a = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import pcgsq
    Ns = [1 << int(x) for x in a.logN.split(",")]
    with pcgsq.Context(0, apply_tuning=False) as ctx:
        res = tune(ctx, Ns)
        key = ctx.device_name().split(" (")[-1].split(",")[0]   # gfx arch string
    cache = {}
    if os.path.exists(a.out):
        with open(a.out) as f:
            cache = json.load(f)
    cache[key] = res
    with open(a.out, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    print(json.dumps({key: res}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
