"""Autotuner for the NTT pass shapes.

For each transform length N it times candidate shapes -- one pass or a
column/row split (n1 = 2^n1_log), the register grouping of each pass (variant
bits) and the sub-transforms per workgroup tile of each pass (cols1, cols2) --
on the transforms of one PCG-OLE expansion at c = 4: 2c^2 + 2c = 40 forward
transforms of a device-resident batch and 4 inverse ones (tile widths that
fill the GPU at 40 transforms are not those that fill it at 16).  It uses the
library's own per-launch event timing and keeps the fastest.  Shapes never change results:
every candidate is checked bit-exact against the default shape before it is
timed.  Splits and groupings are searched exhaustively, tile widths by
coordinate descent from the best of those.

    python pcg-square_amd/pcgsq/autotune.py --out tuning.json

The cache maps the device's gfx name -> {N: [n1_log, variant, cols1, cols2,
us per forward transform of that mix]} and is applied by
pcgsq.Context(apply_tuning=True).
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


SINGLE_MAX = 13          # longest one-pass transform
PASS_MIN = 6             # shortest four-step pass
ALT_GROUPING = (9, 10)   # pass lengths with a second register grouping


def splits(N: int):
    L = N.bit_length() - 1
    out = [L] if L <= SINGLE_MAX else []
    base = (L + 1) // 2
    for s in (base - 1, base, base + 1):
        if s < L and PASS_MIN <= s <= SINGLE_MAX and PASS_MIN <= L - s <= SINGLE_MAX:
            out.append(s)
    return sorted(set(out))


def variants(L: int, s: int):
    bits = [0]
    if s in ALT_GROUPING:
        bits.append(1)
    if s < L and (L - s) in ALT_GROUPING:
        bits += [b | 2 for b in bits]
    return bits


def time_shape(ctx, buf, batch, inv_batch, N, iters=10) -> float:
    """Average device ms of `batch` forward transforms on `buf` (batch x N)
    and `inv_batch` inverse ones, over enough replays that widths a few
    percent apart order the same way from run to run."""
    ctx.ntt_forward_dev(buf.ptr, batch, N)
    ctx.ntt_inverse_dev(buf.ptr, inv_batch, N)
    ctx.synchronize()
    ctx.profile(True)
    for _ in range(iters):
        ctx.ntt_forward_dev(buf.ptr, batch, N)
        ctx.ntt_inverse_dev(buf.ptr, inv_batch, N)
    ctx.synchronize()
    z = ctx.profile_read()
    ctx.profile(False)
    return sum(v[1] for k, v in z.items() if k.startswith("ntt.")) / iters


def tune(ctx, Ns, batch=40, inv_batch=4, log=print) -> dict:
    import pcgsq
    res = {}
    rng = np.random.default_rng(1)
    for N in Ns:
        L = N.bit_length() - 1
        a = rng.integers(0, pcgsq.P, (batch, N), dtype=np.uint64)
        ctx.ntt_tune(N)                      # default shape gives the reference result
        ref = ctx.ntt_forward(a[:1])
        buf = ctx.to_device(a)
        seen = {}

        def run(shape):
            if shape in seen:
                return seen[shape]
            ctx.ntt_tune(N, *shape)
            if not np.array_equal(ctx.ntt_forward(a[:1]), ref):
                raise RuntimeError(f"shape {shape} changed the result at N={N}")
            us = time_shape(ctx, buf, batch, inv_batch, N) * 1e3 / batch
            log(f"N=2^{L} n1_log={shape[0]} variant={shape[1]} cols={shape[2]},{shape[3]}: "
                f"{us:.2f} us per forward transform ({inv_batch} inverse per {batch})")
            seen[shape] = us
            return us

        best = min(((s, v, 0, 0) for s in splits(N) for v in variants(L, s)), key=run)
        for slot in (2, 3):
            if slot == 3 and best[0] == L:
                break                        # one pass: no second tile width
            for c in (1, 2, 4, 8, 16):
                cand = list(best)
                cand[slot] = c
                if run(tuple(cand)) < seen[best]:
                    best = tuple(cand)
        buf.close()
        ctx.ntt_tune(N, *best)
        res[str(N)] = list(best) + [round(seen[best], 3)]
        log(f"best N=2^{L}: {res[str(N)]}")
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
    n = 0
    for N, shape in entry.items():
        if len(shape) != 5:
            continue                          # an entry of an older layout
        s, v, c1, c2, _us = shape
        ctx.ntt_tune(int(N), int(s), int(v), int(c1), int(c2))
        n += 1
    return n


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--logN", default="10,12,13,14,16,18,20,22")
    ap.add_argument("--batch", type=int, default=40, help="forward transforms timed per shape")
    ap.add_argument("--inv-batch", type=int, default=4, help="inverse transforms timed per shape")
    a = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import pcgsq
    Ns = [1 << int(x) for x in a.logN.split(",")]
    with pcgsq.Context(0, apply_tuning=False) as ctx:
        res = tune(ctx, Ns, batch=a.batch, inv_batch=min(a.inv_batch, a.batch))
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
