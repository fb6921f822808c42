"""Open-loop load generator for the expansion server (pcgsq.Server).

Requests arrive as a Poisson process of a fixed rate; each draws its ring
dimension uniformly from `sizes` and carries a fresh seed.  The generator submits
every request at its arrival time (so the server's submit timestamp is the
arrival time), releases outputs as requests finish, and returns one record
per request with the server-side times.  Latency = completion - arrival,
where completion is the moment the server saw the request's work finish on
the device (host-visible completion).
"""
from __future__ import annotations

import time

import numpy as np


def percentile(xs, q: float) -> float:
    s = sorted(xs)
    if not s:
        return float("nan")
    k = max(0, min(len(s) - 1, int(round(q / 100.0 * (len(s) - 1)))))
    return s[k]


def run_poisson(server, rate: float, duration_s: float, sizes, c: int, t: int, seed: int,
                keep=None, pub_seed: int = 0):
    """Drive `server` for `duration_s` seconds of arrivals at `rate` per
    second.  `keep(rid, N)` may fetch outputs before they are released.
    Returns (records, wall_seconds): records are dicts with N, seed,
    arrival, start, done, latency_ms."""
    import pcgsq

    rng = np.random.default_rng(seed)
    gaps = []
    tot = 0.0
    while True:
        g = rng.exponential(1.0 / rate)
        if tot + g > duration_s:
            break
        tot += g
        gaps.append(tot)
    Ns = [int(sizes[i]) for i in rng.integers(0, len(sizes), len(gaps))]
    seeds = [int(x) for x in rng.integers(1, 2**62, len(gaps))]
    pending = {}
    records = []
    t0 = pcgsq.Server.clock()

    def reap(block: bool):
        for rid in list(pending):
            st = server.poll(rid)
            if st == pcgsq.REQ_QUEUED or st == pcgsq.REQ_RUNNING:
                if not block:
                    continue
                server.wait(rid)
                st = server.poll(rid)
            N, sd = pending.pop(rid)
            if st != pcgsq.REQ_DONE:
                raise RuntimeError(f"served request {rid} (N={N}) failed")
            if keep is not None:
                keep(rid, N, sd)
            a, s, d = server.timing(rid)
            records.append({"N": N, "seed": sd, "arrival": a - t0, "start": s - t0,
                            "done": d - t0, "latency_ms": (d - a) * 1e3})
            server.release(rid)

    for at, N, sd in zip(gaps, Ns, seeds):
        while True:
            now = pcgsq.Server.clock() - t0
            if now >= at:
                break
            reap(False)
            time.sleep(min(0.0005, at - now))
        rid = server.submit(N, c, t, sd, pub_seed)
        pending[rid] = (N, sd)
    reap(True)
    wall = max((r["done"] for r in records), default=0.0)
    records.sort(key=lambda r: r["arrival"])
    return records, wall


def summarize(records, wall: float) -> dict:
    lat = [r["latency_ms"] for r in records]
    oles = sum(r["N"] for r in records)
    return {"requests": len(records), "wall_s": wall,
            "OLEs_per_s": oles / wall if wall > 0 else float("nan"),
            "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
            "max_ms": max(lat) if lat else float("nan"),
            "by_N": {f"2^{int(np.log2(n))}": sum(1 for r in records if r["N"] == n)
                     for n in sorted({r["N"] for r in records})}}


def by_size(records) -> dict:
    """p50 and p99 latency (ms) per ring dimension."""
    out = {}
    for n in sorted({r["N"] for r in records}):
        lat = [r["latency_ms"] for r in records if r["N"] == n]
        out[f"2^{int(n).bit_length() - 1}"] = [percentile(lat, 50), percentile(lat, 99)]
    return out
