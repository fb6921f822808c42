"""PCG-OLE expansion throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one full PCG-OLE expansion for both parties at N=2^20, c=4, t=16,
w=1, ChaCha8 (the shape of the reference's 2PC benchmark,
CPU_baseline/test/bench_pcg_ole_2pc.cpp, pcg_ole_impl.h:46-490): 4096
half-tree DPFs of domain 2^17 per party with the fused correction-word
protocol, the Beaver-corrected output layer and step 4.  The per-step inputs
come from a seeded dealer and are resident in HBM before the timed region.

Prints one JSON line: OLEs/s (N OLEs per expansion), p50/p99 latency of one
request per launch, the roofline of the dominant kernel (the DPF leaf level
fused with the ChaCha8 out-hash and the output layer's gather, VALU-bound),
the NTT passes' HBM roofline, and the CPU oracle timed on this host; then
the other workloads (--parts): batched NTTs, batched DPF keys, DPF arity,
the expansion as two parties (threads, and OS processes over gloo), and the
continuous-batching server.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pcg-square_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "OLEs/sec at N=2^20,c=4,t=16 (bit-exact) + p50/p99 request latency"
CHACHA8_OPS = 384                    # int ops per ChaCha8 block (8 rounds x 4 QR x 12)
VALU_NOMINAL = 256 * 64 * 2.4e9      # 256 CUs x 64 lanes x 2.4 GHz, 1 op/lane/clock
HBM_PEAK = 8.0e12                    # MI355X HBM3E, bytes/s


def percentile(xs, q):
    s = sorted(xs)
    k = max(0, min(len(s) - 1, int(round(q / 100.0 * (len(s) - 1)))))
    return s[k]


def pmc_traffic(path: str, key: str = "hbm_bytes_per_launch"):
    """HBM bytes per launch of one kernel from a committed rocprofv3 --pmc
    summary (FETCH_SIZE / WRITE_SIZE, the line `key = bytes`), if present."""
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"^" + re.escape(key) + r"\s*=\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


def gmw_and_gates(N: int, t: int, w: int = 1) -> int:
    """AND gates of one instance's GMW preprocessing: the Kogge-Stone adder
    over nb = log2(N/t) bits (nb + sum_d 2(nb - d), bool_circuit/adder.h:3-5)
    and the DltSft circuit, levels * (2^w - 1) * (w - 1 + 128)."""
    nb = (N // t).bit_length() - 1
    n = nb + 1
    s, d = nb, 1
    while d < nb:
        s += 2 * (nb - d)
        d <<= 1
    return s + (n // w) * ((1 << w) - 1) * (w - 1 + 128)


def host_cpus() -> dict:
    """CPUs of this host as the OS reports them: `nproc` (the CPUs this
    process may run on), os.cpu_count() (all online CPUs) and the
    OMP_NUM_THREADS share the job was given."""
    import subprocess
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = len(os.sched_getaffinity(0))
    return {"nproc": nproc, "cpu_count": os.cpu_count(),
            "omp_num_threads": int(os.environ.get("OMP_NUM_THREADS", "0")) or None}


def cpu_baseline_full(N: int, c: int, t: int, threads: int) -> dict:
    """The CPU oracle (a C restatement of CPU_baseline's local compute: per-node
    ChaCha8, u128 %P, radix-2 NTT with a cached plan) running one whole
    expansion for both parties, dealt inputs included, on `threads` OpenMP
    threads: the same work as one GPU step, timed in full."""
    import oracle as O
    t0 = time.perf_counter()
    O.pcg_ole_expand(N, c, t, 1, O.PRG_CHACHA8, seed=1, pub_seed=0, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": N / dt, "unit": "OLEs/s", "cores": threads, "kind": "port",
            "host": host_cpus(),
            "sample": (f"one full expansion (both parties: {c * c} DPF blocks + output layer + "
                       f"{c + 2 * c * c} poly-muls of step 4) by the oracle (C, u128 %P, per-node "
                       f"ChaCha8) on {threads} OpenMP threads: {dt:.1f} s")}


def cpu_baseline(N: int, c: int, t: int, threads: int, polymul_sample: int = 0) -> dict:
    """The CPU oracle on a bounded sample: one DPF block of the c^2 (both
    parties, DPF + output layer) and `polymul_sample` of the c + 2c^2
    poly-muls of step 4 (all of them if 0), each with `threads` OpenMP
    threads, scaled to the whole expansion (c^2 blocks + step 4)."""
    import numpy as np

    import oracle as O
    t0 = time.perf_counter()
    O.pcg_ole_blocks(N, c, t, 1, O.PRG_CHACHA8, seed=1, blk_begin=0, blk_end=1, threads=threads)
    t_block = time.perf_counter() - t0
    npm = c + 2 * c * c
    ns = polymul_sample or npm
    rng = np.random.default_rng(0)
    a = rng.integers(0, O.P, (ns, N), dtype=np.uint64)
    b = rng.integers(0, O.P, (ns, N), dtype=np.uint64)
    t0 = time.perf_counter()
    O.poly_mul_batch(a, b, threads=threads)
    t_step4 = (time.perf_counter() - t0) * npm / ns
    full = c * c * t_block + t_step4
    pm = f"{npm} poly-muls" if ns == npm else f"{npm} poly-muls (timed {ns}, scaled)"
    return {"value": N / full, "unit": "OLEs/s", "cores": threads, "kind": "port",
            "host": host_cpus(),
            "sample": (f"oracle (C, u128 %P, per-node ChaCha8) on {threads} host threads: 1 of "
                       f"{c * c} DPF blocks both parties ({t_block:.2f} s) + {pm} of "
                       f"step 4 ({t_step4:.2f} s); expansion = {c * c} x block + step 4 = "
                       f"{full:.1f} s (extrapolated)")}


NTT_BATCH_PROFILE = os.path.join(ROOT, "profiles", "ntt_batch_256_uneven_2e12_2e20.md")


def bench_ntt_batch(ctx, pcgsq, reps: int = 5, mode: str | None = None,
                    splits: dict | None = None, wgs_per_cu: int = 0) -> dict:
    """Batched NTT of uneven lengths: 256 transforms, lengths drawn uniformly
    from {2^12, 2^14, 2^16, 2^18, 2^20}, uniform coefficients in [0, P),
    forward then inverse of every transform in one replay of the batch's
    captured launches.  The autotuner picks the splits of the persistent
    launch, its workgroups per CU and then the mode (persistent, or grouped
    by length over offset tables); against one launch per transform (the
    plain form) and against the same transforms regrouped by length into
    contiguous batches.  `mode` ("persistent" / "grouped") skips the tuner and
    the comparisons: only that mode runs (the profiled command), with the
    persistent launch's splits {L: n1_log} and workgroups per CU if given
    (the shape the tuner picked)."""
    import numpy as np
    rng = np.random.default_rng(2024)
    logN = [int(12 + 2 * x) for x in rng.integers(0, 5, 256)]
    with pcgsq.NttBatch(ctx, logN) as b:
        b.load(rng.integers(0, pcgsq.P, b.total, dtype=np.uint64))
        if mode is None:
            shape = b.autotune()
        else:
            b.set_mode(pcgsq.NTT_BATCH_GROUPED if mode == "grouped" else pcgsq.NTT_BATCH_PERSISTENT)
            for L, s1 in (splits or {}).items():
                b.tune(L, s1, wgs_per_cu)
            b.tune(0, 0, wgs_per_cu)
            shape = {"mode": mode, "wgs_per_cu": wgs_per_cu, "splits": dict(splits or {})}
        b.run(graph=True)
        b.finish()
        ctx.profile(True)
        for _ in range(reps):
            b.run(graph=True)
        b.finish()
        z = ctx.profile_read()
        ctx.profile(False)
        ms = z["ntt.batch"][1] / z["ntt.batch"][0]
        info = b.info()
        pa, _, _, _ = b.dev()
        if mode is not None:
            return {"mode": mode, "shape": shape, "ms": ms, "alg_bytes": info["alg_bytes"],
                    "GBps": info["alg_bytes"] / (ms * 1e-3) / 1e9, "replays": reps + 1}

        def plain():
            for i, L in enumerate(logN):
                ptr = pa + 8 * int(b.offsets[i])
                ctx.ntt_forward_dev(ptr, 1, 1 << L)
                ctx.ntt_inverse_dev(ptr, 1, 1 << L)

        plain()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plain()
        ctx.synchronize()
        ms_plain = (time.perf_counter() - t0) * 1e3 / reps

    # The same transforms regrouped: one contiguous batch per length, each
    # run by the per-length pass launches (summed device time of those
    # launches; the regrouping copy is not counted).
    counts = {L: logN.count(L) for L in sorted(set(logN))}
    groups = {L: ctx.to_device(np.zeros(k << L, dtype=np.uint64)) for L, k in counts.items()}
    try:
        def grouped():
            for L, k in counts.items():
                ctx.ntt_forward_dev(groups[L].ptr, k, 1 << L)
                ctx.ntt_inverse_dev(groups[L].ptr, k, 1 << L)

        grouped()
        ctx.synchronize()
        ctx.profile(True)
        for _ in range(reps):
            grouped()
        ctx.synchronize()
        z = ctx.profile_read()
        ctx.profile(False)
        ms_grouped = sum(v[1] for k, v in z.items() if k.startswith("ntt.")) / reps
    finally:
        for g in groups.values():
            g.close()
    gbps = info["alg_bytes"] / (ms * 1e-3) / 1e9
    grouped_mode = shape["mode"] == "grouped"
    return {"workload": "256 forward+inverse NTTs, lengths uniform over 2^12..2^20 (x4 steps), "
                        "one replay of the batch's captured launches, autotuned mode",
            "lengths": {f"2^{L}": logN.count(L) for L in sorted(set(logN))},
            "ms": ms, "alg_bytes": info["alg_bytes"], "GBps": gbps,
            "hbm_frac": gbps * 1e9 / HBM_PEAK, "work_items": info["items"],
            "grid": info["grid"], "tuned": {"mode": shape["mode"],
                                            "persistent_ms": shape.get("persistent_ms"),
                                            "grouped_ms": shape.get("grouped_ms"),
                                            "wgs_per_cu": shape["wgs_per_cu"],
                                            "splits": {f"2^{k}": v for k, v in
                                                       shape["splits"].items()}},
            "one_launch_per_transform_ms": ms_plain,
            "speedup_vs_one_launch_per_transform": ms_plain / ms,
            "contiguous_by_length_ms": ms_grouped,
            "roofline": {"bound": "hbm",
                         "kernel": ("ntt_pass_kernel, per-length launches over offset tables"
                                    if grouped_mode else "ntt_batch_kernel (persistent)"),
                         "achieved": gbps, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": gbps * 1e9 / HBM_PEAK,
                         "traffic": pmc_traffic(NTT_BATCH_PROFILE,
                                                "grouped_hbm_bytes_per_run" if grouped_mode
                                                else "persistent_hbm_bytes_per_run"),
                         "avg_launch_ms": ms}}


def bench_dpf_batch(ctx, pcgsq, ceiling_blocks: float, reps: int = 5) -> dict:
    """Batched DPF keys: 4096 key pairs, domains 2^10..2^18 (log-uniform),
    dealer Gen with random special points and payloads, both parties' full
    domains with the fused leaf convert in one persistent launch (graph
    replay); ChaCha8 and AES tree PRGs (ChaCha8 out-hash in both)."""
    import numpy as np
    rng = np.random.default_rng(4096)
    logn = [int(x) for x in rng.integers(10, 19, 4096)]
    out = {"workload": "4096 DPF key pairs, domains 2^10..2^18, dealer Gen + sync-free "
                       "full-domain eval of both parties + leaf convert, one persistent launch",
           "leaves_both_parties": 2 * sum(1 << n for n in logn)}
    for prg, name in ((pcgsq.PRG_CHACHA8, "chacha8"), (pcgsq.PRG_AES, "aes")):
        with pcgsq.DpfBatch(ctx, logn, prg) as b:
            b.deal(5)
            b.run(graph=True)
            ctx.synchronize()
            ctx.profile(True)
            for _ in range(reps):
                b.run(graph=True)
            ctx.synchronize()
            z = ctx.profile_read()
            ctx.profile(False)
            ms = z["dpf.batch"][1] / z["dpf.batch"][0]
            info = b.info()
        r = {"ms": ms, "tree_blocks": info["tree_blocks"], "out_hash_blocks": info["out_blocks"],
             "leaves_per_s": info["out_blocks"] / (ms * 1e-3)}
        if prg == pcgsq.PRG_CHACHA8:
            blocks = info["tree_blocks"] + info["out_blocks"]
            ops = blocks * CHACHA8_OPS / (ms * 1e-3)
            r.update({"chacha8_blocks_per_s": blocks / (ms * 1e-3), "Tops_per_s": ops / 1e12,
                      "frac_of_chacha8_ceiling": blocks / (ms * 1e-3) / ceiling_blocks})
        else:
            r.update({"aes_blocks_per_s": info["tree_blocks"] / (ms * 1e-3),
                      "note": "AES tree blocks (T-table AES-128 in LDS) + ChaCha8 out-hash blocks; "
                              "no algorithmic op model for AES"})
        out[name] = r
    return out


# The reference's published poly-mul recipe (GPU_baseline/ntt_batch_bench.cu:
# 47-48,69): N = 2^20, batch 16, a[j] = (0x9e3779b97f4a7c15 (j+1)) mod 2^64 mod P,
# b[j] = (0xc2b2ae3d27d4eb4f (j+7)) mod 2^64 mod P; its XOR of all outputs
# (GPU_baseline/results_l40s/ntt_4step.txt:7) and its ms per product
# (device time between the input and output copies, divided by the batch).
POLYMUL_XOR_2_20 = 0x3969340B039F5D5D
POLYMUL_PUBLISHED_MS = {
    "l40s_four_step_standalone_brev": 0.6889,   # GPU_baseline/final_l40s/b3_ntt.csv:95
    "l40s_four_step_folded_brev": 0.5401,       # b3_ntt.csv:96
    "l40s_naive": 1.6170,                       # b3_ntt.csv:97
    "b200_four_step": 0.2479,                   # pcg_baseline_out_235933_prev/ntt_stages.csv:64
    "b200_naive": 0.2671,                       # ntt_stages.csv:65
}


def bench_polymul_recipe(ctx, pcgsq, reps: int = 10) -> dict:
    """The reference's batch-16 poly-mul at N = 2^20 on both backends, inputs
    and output resident on the device as in its device-time figure: the
    products' XOR checked against the reference's published checksum, and
    ms per product beside the published L40S and B200 figures."""
    import numpy as np
    N, batch = 1 << 20, 16
    j = np.arange(batch * N, dtype=np.uint64)
    with np.errstate(over="ignore"):
        a = (np.uint64(0x9E3779B97F4A7C15) * (j + np.uint64(1))) % np.uint64(pcgsq.P)
        b = (np.uint64(0xC2B2AE3D27D4EB4F) * (j + np.uint64(7))) % np.uint64(pcgsq.P)
    da, db = ctx.to_device(a), ctx.to_device(b)
    dout = pcgsq.DeviceBuffer(ctx, a.nbytes)
    res = {"workload": "reference poly-mul recipe (ntt_batch_bench.cu:47-48,69): N=2^20, "
                       "batch 16, inputs and output on the device",
           "published_ms_per_mul": POLYMUL_PUBLISHED_MS}
    try:
        for name, backend in (("four_step", pcgsq.NTT_FOURSTEP), ("naive", pcgsq.NTT_NAIVE)):
            ctx.poly_mul_dev(dout.ptr, da.ptr, db.ptr, batch, N, backend)
            ctx.synchronize()
            xor = int(np.bitwise_xor.reduce(dout.to_host()))
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.poly_mul_dev(dout.ptr, da.ptr, db.ptr, batch, N, backend)
            ctx.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / reps
            res[name] = {"ms_per_batch": ms, "ms_per_mul": ms / batch,
                         "xor": "%016x" % xor, "xor_matches_reference": xor == POLYMUL_XOR_2_20,
                         "vs_l40s": POLYMUL_PUBLISHED_MS["l40s_four_step_folded_brev"
                                                         if name == "four_step" else
                                                         "l40s_naive"] / (ms / batch),
                         "vs_b200": POLYMUL_PUBLISHED_MS["b200_four_step" if name == "four_step"
                                                         else "b200_naive"] / (ms / batch)}
    finally:
        for x in (da, db, dout):
            x.close()
    return res


# The reference's one-block DPF + convert recipe (GPU_baseline/dpf_real_bench.cu:
# 48-110: ChaCha8, w = 1, one party, closed-form roots and DltSft, the peer's
# correction-word share stubbed to zero, host outputs) and its published L40S
# figures (GPU_baseline/final_l40s/b1_shape.csv:45-46: expand, sums, scatter,
# total ms) and checksum (GPU_baseline/results_l40s/plateau.txt:7, n=18 B=256).
DPF_BLOCK_PUBLISHED = {
    17: {"expand_ms": 4.850, "sums_ms": 0.844, "scatter_ms": 1.346, "total_ms": 7.040},
    18: {"expand_ms": 9.571, "sums_ms": 1.562, "scatter_ms": 2.559, "total_ms": 13.692,
         "checksum": "27cb4298e84744ed"},
}


def bench_dpf_block(ctx, pcgsq, reps: int = 5) -> dict:
    """One DPF block (B = 256 instances, t = 16) of the reference's recipe
    through the drop-in entry points (pcg_cuda::dpf_gpu_batch_full_eval_device,
    dpf_out_sums, dpf_out_scatter_g): one party, a host exchange per level
    returning a zero peer share (in Python here, a C++ lambda there), outputs
    copied to the host as the reference's are.  Wall time per phase, beside the
    published L40S figures; at n = 18 the reference's checksum."""
    import numpy as np
    B, t = 256, 16
    res = {"workload": "reference DPF+convert recipe (dpf_real_bench.cu:48-110): one party, "
                       "B=256, t=16, w=1, ChaCha8, host exchange per level (zero peer share), "
                       "host outputs"}
    for n in (17, 18):
        roots = np.zeros((B, 2), dtype=np.uint64)
        with np.errstate(over="ignore"):
            roots[:, 0] = np.uint64(0x9E3779B97F4A7C15) * (np.arange(B, dtype=np.uint64) +
                                                           np.uint64(1))
        roots[:, 1] = np.arange(B, dtype=np.uint64) + 1
        i = np.arange(B * n, dtype=np.uint64)
        dlt = np.stack([np.uint64(0x1234567) * (i + np.uint64(1)), i], axis=1)
        CW = (np.uint64(0xABCDEF) * (np.arange(B, dtype=np.uint64) + np.uint64(1))) % \
            np.uint64(pcgsq.P)
        D, gN = 1 << n, t * (1 << (n - 1))
        zero = lambda level, mine: np.zeros_like(mine)
        g = np.zeros(gN, dtype=np.uint64)   # one host g for every run, as the reference's
        tt = {"expand_ms": 0.0, "sums_ms": 0.0, "scatter_ms": 0.0}
        for k in range(reps + 1):
            t0 = time.perf_counter()
            d = ctx.dpf_batch_full_eval_device(roots, dlt, n, 1, zero, party=0,
                                               prg=pcgsq.PRG_CHACHA8)
            ctx.synchronize()
            t1 = time.perf_counter()
            sc, st = ctx.dpf_out_sums(d, B, D)
            t2 = time.perf_counter()
            ctx.dpf_out_scatter_g(d, B, D, t, 0, CW, gN, out=g)
            t3 = time.perf_counter()
            if k:
                tt["expand_ms"] += (t1 - t0) * 1e3 / reps
                tt["sums_ms"] += (t2 - t1) * 1e3 / reps
                tt["scatter_ms"] += (t3 - t2) * 1e3 / reps
        cks = int(np.bitwise_xor.reduce(sc ^ st)) ^ int(np.bitwise_xor.reduce(g[::97]))
        tot = sum(tt.values())
        pub = DPF_BLOCK_PUBLISHED[n]
        r = dict(tt, total_ms=tot, leaves=B * D, ns_per_leaf=tot * 1e6 / (B * D),
                 checksum="%016x" % cks, published_l40s=pub,
                 vs_l40s=pub["total_ms"] / tot)
        if "checksum" in pub:
            r["checksum_matches_reference"] = r["checksum"] == pub["checksum"]
        res[f"n{n}"] = r
    return res


# The reference's one-party end-to-end arms (GPU_baseline/e2e_full_bench.cu,
# published in GPU_baseline/final_l40s/e2e_arms.csv:32-36: c=4, t=16, N=2^20, one
# party, peer stubbed, PCG_RTT_US injected per exchange; wall ms).
E2E_PUBLISHED_MS = {
    "serial": {"5": 192.1137, "25.2": 198.5512, "2000": 839.7925},
    "batched": {"5": 189.6091, "25.2": 190.3887},
}


def bench_one_party_stub(pcgsq, N: int, c: int, t: int, reps: int = 3) -> dict:
    """One party's whole expansion on its own, as the reference's end-to-end
    bench times it: a one-party plan dealt by the trusted dealer, every
    exchange answered by a stub peer with zeros after the injected round trip
    (a busy wait), the serial and the batched arm.  Timing numbers only: with a
    zero peer the outputs are not the 2PC outputs (the reference's rows are
    the same)."""
    import numpy as np
    res = {"workload": f"N=2^{N.bit_length() - 1}, c={c}, t={t}, ChaCha8: party 0 alone, "
                       "stub peer (zeros), injected round trip per exchange",
           "published_l40s_ms": E2E_PUBLISHED_MS}
    nx = [0]

    def stub(rtt_us):
        def exchange(phase, level, batch, send, nrecv):
            nx[0] += 1
            _spin_us(rtt_us)
            return np.zeros(nrecv, dtype=np.uint8) if nrecv else None
        return exchange

    with pcgsq.Context(0) as cx, pcgsq.OleParty(cx, N, c, t, 0) as op:
        for arm, mode in (("serial", pcgsq.EXCHANGE_PER_BLOCK),
                          ("batched", pcgsq.EXCHANGE_BATCHED)):
            op.set_exchange_mode(mode)
            for rtt in (5.0, 25.2, 2000.0):
                if rtt == 2000.0 and arm == "batched":
                    continue
                n_rep = 1 if rtt >= 1000 else reps
                times = []
                for k in range(n_rep + 1):
                    op.deal(1)
                    nx[0] = 0
                    t0 = time.perf_counter()
                    op.run(stub(rtt))
                    if k:
                        times.append(time.perf_counter() - t0)
                ms = 1e3 * sum(times) / len(times)
                key = ("%g" % rtt)
                pub = E2E_PUBLISHED_MS[arm].get(key)
                res.setdefault(arm, {})[f"rtt_{key}us"] = {
                    "ms": ms, "exchanges": nx[0], "published_l40s_ms": pub,
                    "vs_l40s": pub / ms if pub else None}
    return res


def bench_arity(ctx, pcgsq, reps: int = 5) -> dict:
    """Higher DPF arity (up to 15 children per ChaCha8 block) at N = 2^20,
    c = 4: t = 32 (DPF domain 2^16) with w = 1, 2, 4 and t = 8 (domain 2^18)
    with w = 1, 2, 3, the arities that divide the domain's bits (the
    headline's 2^17 admits only w = 1)."""
    N, c = 1 << 20, 4
    out = {"workload": "pcg-ole expansion N=2^20, c=4, ChaCha8, both parties; "
                       "t=32 (w1, w2, w4) and t=8 (t8_w1, t8_w2, t8_w3)"}
    for t, ws, tag in ((32, (1, 2, 4), ""), (8, (1, 2, 3), "t8_")):
        for w in ws:
            with pcgsq.OlePlan(ctx, N, c, t, w) as pl:
                pl.deal(1)
                pl.run()
                ctx.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    pl.run()
                ctx.synchronize()
                ms = (time.perf_counter() - t0) * 1e3 / reps
            out[f"{tag}w{w}"] = {"ms": ms, "OLEs_per_s": N / (ms * 1e-3)}
    return out


def _spin_us(us: float) -> None:
    """Busy-wait `us` microseconds (an injected network round trip)."""
    if us <= 0:
        return
    end = time.perf_counter_ns() + int(us * 1e3)
    while time.perf_counter_ns() < end:
        pass


def _two_party_threads(pcgsq, N, c, t, mode: str, rtt_us: float, reps: int) -> dict:
    """The expansion as two party threads, each with its own context and
    one-party plan, exchanging through an in-memory queue pair; party 0
    waits rtt_us more per exchange (a network round trip: the reference's
    PCG_RTT_US, GPU_baseline/dpf_real_bench.cu:64-81).  mode: "dealer" (the
    trusted-dealer simulation: the run starts at the DPF; per-block exchanges,
    the reference's serial arm), "dealer_batched" (the same in its batched
    arm: one exchange per level for all blocks), "own" / "own_batched" (each
    party from its own seed, the 2PC preprocessing first, serial / batched)."""
    import queue
    import threading
    inbox = (queue.Queue(), queue.Queue())
    stats = [{"exchanges": 0, "bytes": 0}, {"exchanges": 0, "bytes": 0}]
    times, errs = [], []
    go = [threading.Barrier(2) for _ in range(reps + 1)]

    def party(p):
        try:
            def exchange(phase, level, batch, send, nrecv):
                stats[p]["exchanges"] += 1
                stats[p]["bytes"] += int(send.nbytes)
                if send.size:
                    inbox[1 - p].put(send)
                got = inbox[p].get(timeout=120) if nrecv else None
                if p == 0:
                    _spin_us(rtt_us)
                return got
            with pcgsq.Context(0) as cx, pcgsq.OleParty(cx, N, c, t, p) as op:
                if mode.endswith("batched"):
                    op.set_exchange_mode(pcgsq.EXCHANGE_BATCHED)
                for k in range(reps + 1):   # the first round warms up (and runs the base OTs)
                    if mode.startswith("own"):
                        op.deal_own(1)
                    else:
                        op.deal(1)
                    go[k].wait()
                    if k == 1:
                        stats[p]["exchanges"] = stats[p]["bytes"] = 0
                    t0 = time.perf_counter()
                    op.run(exchange)
                    if p == 0 and k > 0:
                        times.append(time.perf_counter() - t0)
                op.outputs()
        except Exception as e:  # reported in the result
            errs.append(repr(e))
            for b in go:
                b.abort()

    th = [threading.Thread(target=party, args=(p,)) for p in (0, 1)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    if errs or not times:
        return {"error": errs[:1]}
    ms = 1e3 * sum(times) / len(times)
    return {"ms": ms, "OLEs_per_s": N / (ms * 1e-3),
            "exchanges_per_party_per_run": stats[0]["exchanges"] // reps,
            "bytes_sent_per_party_per_run": [stats[p]["bytes"] // reps for p in (0, 1)]}


def bench_two_parties(pcgsq, N: int, c: int, t: int, reps: int = 3) -> dict:
    """The 2PC expansion as two party processes run it (pcgsq_ole_party_*):
    two threads with a context and one-party plan each, both on this one GPU
    (with a party per GPU each would have a GPU of its own): the
    trusted-dealer simulation, then each party from its own seed with the 2PC
    preprocessing (base OTs, IKNP OT extension, GMW adder and DltSft, Gilboa)
    in the serial and the batched arm, at injected round trips of 0, 5 (the
    reference's tier), 50 and 500 us; and the own-seed serial arm as two OS
    processes over gloo (bench_two_processes)."""
    res = {"workload": f"N=2^{N.bit_length() - 1}, c={c}, t={t}, ChaCha8: party 0 and party 1 "
                       "as two threads with a context and one-party plan each, exchanges "
                       "through a queue pair; both on this one GPU"}
    for rtt in (0, 5, 50, 500):
        for mode, key in (("dealer", "dealer_serial"), ("dealer_batched", "dealer_batched"),
                          ("own", "own_serial"), ("own_batched", "own_batched")):
            r = _two_party_threads(pcgsq, N, c, t, mode, rtt, 2 if rtt >= 50 else reps)
            res.setdefault(key, {})[f"rtt_{rtt}us"] = r
    res["as_os_processes_over_gloo"] = {
        mode: bench_two_processes(N, c, t, mode, reps)
        for mode in ("dealer_batched", "own_batched", "own")}
    return res


_PARTY_PROCESS = """
import json, os, sys, time
sys.path.insert(0, os.environ["PCGSQ_PKG"])
import torch.distributed as dist
import pcgsq
rank = int(os.environ["RANK"])
N, c, t, reps = (int(v) for v in os.environ["PCGSQ_CASE"].split(","))
dist.init_process_group("gloo", init_method="tcp://127.0.0.1:" + os.environ["MASTER_PORT"],
                        rank=rank, world_size=2)
ex = pcgsq.dist_exchange(1 - rank)
times = []
mode = os.environ["PCGSQ_MODE"]
with pcgsq.Context(0) as cx, pcgsq.OleParty(cx, N, c, t, rank) as op:
    if mode.endswith("batched"):
        op.set_exchange_mode(pcgsq.EXCHANGE_BATCHED)
    for k in range(reps + 1):   # the first round warms up (and runs the base OTs)
        if mode.startswith("own"):
            op.deal_own(1)
        else:
            op.deal(1)
        dist.barrier()
        t0 = time.perf_counter()
        op.run(ex)
        if k:
            times.append(time.perf_counter() - t0)
    op.outputs()
dist.barrier()
if rank == 0:
    print("RESULT " + json.dumps(times), flush=True)
dist.destroy_process_group()
"""


def bench_two_processes(N: int, c: int, t: int, mode: str, reps: int = 3) -> dict:
    """The 2PC expansion as the reference deploys it: one OS process per
    party, every exchange over a network transport (torch.distributed gloo
    on 127.0.0.1, pcgsq.dist_exchange, in place of emp::NetIO), both
    processes on this GPU.  mode as in _two_party_threads."""
    import socket
    import subprocess
    with socket.socket() as so:
        so.bind(("127.0.0.1", 0))
        port = so.getsockname()[1]
    here = os.path.dirname(os.path.abspath(__file__))
    procs = []
    for rank in (0, 1):
        env = dict(os.environ, RANK=str(rank), MASTER_PORT=str(port),
                   PCGSQ_PKG=os.path.join(here, "pcg-square_amd"),
                   PCGSQ_CASE=f"{N},{c},{t},{reps}", PCGSQ_MODE=mode)
        procs.append(subprocess.Popen([sys.executable, "-c", _PARTY_PROCESS], env=env,
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    logs = []
    for p in procs:
        try:
            logs.append(p.communicate(timeout=240)[0])
        except subprocess.TimeoutExpired:
            for q in procs:
                q.kill()
            return {"error": "timed out"}
    res = [ln for ln in logs[0].splitlines() if ln.startswith("RESULT ")]
    if any(p.returncode for p in procs) or not res:
        return {"error": (logs[0] + logs[1])[-400:]}
    times = json.loads(res[0][len("RESULT "):])
    ms = 1e3 * sum(times) / len(times)
    how = ("from a dealt session (trusted-dealer simulation)" if mode.startswith("dealer")
           else "each from its own seed with the 2PC preprocessing")
    arm = "batched" if mode.endswith("batched") else "serial"
    return {"workload": f"N=2^{N.bit_length() - 1}, c={c}, t={t}, ChaCha8: party 0 and party 1 "
                        f"as two OS processes, {how}, {arm} arm, exchanges over "
                        "torch.distributed gloo (TCP, 127.0.0.1); both on this one GPU",
            "ms": ms, "OLEs_per_s": N / (ms * 1e-3)}


def bench_served(pcgsq, duration_s: float = 8.0) -> dict:
    """Served PCG-OLE: Poisson arrivals (rate = 75% of the measured
    one-at-a-time capacity), N uniform over {2^16, 2^18, 2^20, 2^22}, c=4,
    t=16, a fresh seed per request; a 4-slot continuous-batching server
    (GPU-filling requests of 2^20 and 2^22 one at a time in arrival order,
    smaller ones alongside, skipping ahead bounded at 100 ms) against one
    request per launch (1 slot) on the same arrival trace."""
    from pcgsq import serving
    sizes = [1 << 16, 1 << 18, 1 << 20, 1 << 22]
    service = {}
    with pcgsq.Server(0, 0, 1) as one:
        for N in sizes:
            one.reserve(N, 4, 16, 1)
            lat = []
            for k in range(4):
                rid = one.submit(N, 4, 16, 1000 + k)
                one.wait(rid)
                a, s, d = one.timing(rid)
                lat.append(d - s)
                one.release(rid)
            service[N] = min(lat[1:])              # the first run builds the plan
    mean_service = sum(service.values()) / len(sizes)
    rate = 0.75 / mean_service
    res = {"workload": "Poisson arrivals, N uniform over 2^16..2^22 (x4 steps), c=4, t=16, "
                       "fresh seed per request", "rate_per_s": rate, "duration_s": duration_s,
           "service_ms_alone": {f"2^{N.bit_length() - 1}": v * 1e3 for N, v in service.items()}}
    # plans up front: for the 4-slot server two per small shape (requests of
    # one shape overlap) and one per GPU-filling shape (2^20, 2^22: the
    # exclusive class runs them one at a time, in arrival order), with an age
    # bound of 100 ms on skipping ahead; one plan per shape for one request
    # per launch
    excl, age = 1 << 20, 0.100
    res["policy"] = {"exclusive_N": excl, "age_bound_s": age}
    for slots, key in ((4, "continuous_batching_4_slots"), (1, "one_request_per_launch")):
        with pcgsq.Server(0, 0, slots, exclusive_N=excl if slots > 1 else 0,
                          age_bound_s=age if slots > 1 else 0.0) as srv:
            for N in sizes:
                srv.reserve(N, 4, 16, 2 if slots > 1 and N < excl else 1)
            recs, wall = serving.run_poisson(srv, rate, duration_s, sizes, 4, 16, seed=7)
            st, pl = srv.stats(), srv.plans()
        summ = serving.summarize(recs, wall)
        summ["p50_p99_ms_by_N"] = serving.by_size(recs)
        summ["peak_device_bytes"] = st["peak"]
        summ["plans"] = pl
        res[key] = summ
    # saturation: the same 64-request mix submitted at once (capacity, OLEs/s)
    import numpy as np
    rng = np.random.default_rng(11)
    mix = [int(sizes[i]) for i in rng.integers(0, len(sizes), 64)]
    for slots, key in ((4, "saturated_4_slots"), (1, "saturated_one_request_per_launch")):
        with pcgsq.Server(0, 0, slots, exclusive_N=excl if slots > 1 else 0,
                          age_bound_s=age if slots > 1 else 0.0) as srv:
            for N in sizes:
                srv.reserve(N, 4, 16, 2 if slots > 1 and N < excl else 1)
            t0 = pcgsq.Server.clock()
            ids = [srv.submit(N, 4, 16, 5000 + k) for k, N in enumerate(mix)]
            for rid in ids:
                srv.wait(rid)
            wall = max(srv.timing(rid)[2] for rid in ids) - t0
            for rid in ids:
                srv.release(rid)
        res[key] = {"requests": len(mix), "wall_s": wall, "OLEs_per_s": sum(mix) / wall}
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=1 << 20)
    ap.add_argument("--c", type=int, default=4)
    ap.add_argument("--t", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--parts",
                    default="expand,ntt_batch,polymul,dpf_block,dpf_batch,arity,one_party,parties,"
                            "served",
                    help="comma list of expand, ntt_batch, polymul, dpf_block, dpf_batch, arity, "
                         "one_party, parties, served; "
                         "without expand only the named parts run and print (for profiling "
                         "one workload)")
    ap.add_argument("--ntt-batch-mode", choices=("persistent", "grouped"), default=None,
                    help="batched NTT: run only this mode, untuned, with no comparisons "
                         "(for profiling it)")
    ap.add_argument("--ntt-batch-splits", default="",
                    help="with --ntt-batch-mode persistent: splits L:n1_log,... and "
                         "workgroups per CU as wgs:K (e.g. 20:10,18:10,wgs:2)")
    args = ap.parse_args()
    parts = set(args.parts.split(","))
    unknown = parts - {"expand", "ntt_batch", "polymul", "dpf_block", "dpf_batch", "arity",
                       "one_party", "parties", "served"}
    if unknown:
        ap.error(f"unknown parts: {sorted(unknown)}")
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); the expansion "
              "shards by DPF block across GPUs but no multi-GPU run is provided", file=sys.stderr)
        return 2

    import pcgsq

    N, c, t = args.N, args.c, args.t
    if "expand" not in parts:
        res = {}
        with pcgsq.Context(0) as ctx:
            res["hbm_copy_GBps"] = ctx.probe_copy(1 << 30) / 1e9   # PMC calibration kernel
            if "ntt_batch" in parts:
                sp = dict(tuple(int(v) if v != "wgs" else v for v in kv.split(":"))
                          for kv in args.ntt_batch_splits.split(",") if kv)
                wgs = sp.pop("wgs", 0)
                res["batched_ntt"] = bench_ntt_batch(ctx, pcgsq, mode=args.ntt_batch_mode,
                                                     splits=sp, wgs_per_cu=wgs)
            if "polymul" in parts:
                res["polymul_recipe"] = bench_polymul_recipe(ctx, pcgsq)
            if "dpf_block" in parts:
                res["dpf_block_recipe"] = bench_dpf_block(ctx, pcgsq)
            if "dpf_batch" in parts:
                res["batched_dpf"] = bench_dpf_batch(ctx, pcgsq, ctx.probe_chacha8())
            if "arity" in parts:
                res["arity"] = bench_arity(ctx, pcgsq)
        if "one_party" in parts:
            res["one_party_stub_peer"] = bench_one_party_stub(pcgsq, N, c, t)
        if "parties" in parts:
            res["two_party_processes"] = bench_two_parties(pcgsq, N, c, t)
        if "served" in parts:
            res["served"] = bench_served(pcgsq)
        print(json.dumps(res), flush=True)
        return 0
    with pcgsq.Context(0) as ctx, pcgsq.OlePlan(ctx, N, c, t, 1, pcgsq.PRG_CHACHA8) as pl:
        pl.deal(1)
        for _ in range(args.warmup):
            pl.run()
        # ---- throughput: K back-to-back expansions ----
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            pl.run()
        ctx.synchronize()
        dt = time.perf_counter() - t0
        # ---- latency: one request per launch, host-visible completion ----
        lats = []
        for _ in range(max(5, min(args.steps, 20))):
            s = time.perf_counter()
            pl.run()
            ctx.synchronize()
            lats.append((time.perf_counter() - s) * 1e3)
        # ---- the 2PC preprocessing under GMW on the device (both parties) ----
        for gmw in (False, True):
            pl.deal(2, gmw=gmw)
        ctx.synchronize()
        deal_ms = {}
        for gmw in (False, True):
            t0 = time.perf_counter()
            for _ in range(5):
                pl.deal(2, gmw=gmw)
            ctx.synchronize()
            deal_ms[gmw] = (time.perf_counter() - t0) * 1e3 / 5
        pl.deal(1)
        # ---- host-buffer form: each expansion also copies x0, z0, x1, z1 into
        # fresh host arrays (after one untimed round, which sets up HIP's
        # staging for pageable copies) ----
        pl.run()
        pl.outputs()
        t0 = time.perf_counter()
        for _ in range(5):
            pl.run()
            pl.outputs()
        dt_pcie = (time.perf_counter() - t0) / 5
        # ... and into page-locked host arrays that the caller keeps
        pinned = tuple(pcgsq.pinned_array(N) for _ in range(4))
        pl.run()
        pl.outputs(out=pinned)
        t0 = time.perf_counter()
        for _ in range(5):
            pl.run()
            pl.outputs(out=pinned)
        dt_pinned = (time.perf_counter() - t0) / 5
        del pinned
        # ---- live per-launch timing of every kernel (HIP events on the stream) ----
        ctx.profile(True)
        nprof = 2
        for _ in range(nprof):
            pl.run()
        ctx.synchronize()
        zones = ctx.profile_read()
        ctx.profile(False)
        ceiling = ctx.probe_chacha8()
        hbm = ctx.probe_hbm(1 << 30)
        copy_bps = hbm["copy"]
        ntt_batch = bench_ntt_batch(ctx, pcgsq) if "ntt_batch" in parts else None
        polymul = bench_polymul_recipe(ctx, pcgsq) if "polymul" in parts else None
        dpf_block = bench_dpf_block(ctx, pcgsq) if "dpf_block" in parts else None
        dpf_batch = bench_dpf_batch(ctx, pcgsq, ceiling) if "dpf_batch" in parts else None
        arity = bench_arity(ctx, pcgsq) if "arity" in parts else None
    one_party = bench_one_party_stub(pcgsq, N, c, t) if "one_party" in parts else None
    two_parties = bench_two_parties(pcgsq, N, c, t) if "parties" in parts else None
    served = bench_served(pcgsq) if "served" in parts else None

    leaves = 2 * c * c * t * t * (2 * N // t)        # both parties, all instances
    trees = 2 * c * c * t * t
    tree_blocks = trees * ((2 * N // t) - 1)         # internal nodes, one ChaCha8 block each
    if "dpf.leaf_deep" in zones:
        # the leaf level with the last three tree levels fused in: per tree
        # 2^(n-3) + 2^(n-2) + 2^(n-1) tree blocks and 2^n out-hash blocks
        leaf_zone = "dpf.leaf_deep"
        leaf_blocks = leaves + trees * 7 * (2 * N // t) // 8
        kernel = ("deep_leaf_kernel (last three tree levels + leaf level + ChaCha8 out-hash + "
                  "output-layer gather, in registers)")
        algorithmic = (f"{leaves} out-hash + {leaf_blocks - leaves} tree ChaCha8 blocks x "
                       f"{CHACHA8_OPS} ops per launch")
    else:
        leaf_zone = "dpf.leaf_gather" if "dpf.leaf_gather" in zones else "dpf.leaf_convert"
        leaf_blocks = leaves
        kernel = ("leaf_gather_kernel<w=1> (leaf level + ChaCha8 out-hash + output-layer "
                  "gather)" if leaf_zone == "dpf.leaf_gather" else
                  "dpf_level_kernel<w=1,ChaCha8,convert> (leaf level + out-hash)")
        algorithmic = f"{leaves} ChaCha8 out-hash blocks x {CHACHA8_OPS} ops per launch"
    n_leaf, ms_leaf = zones[leaf_zone]
    avg_leaf = ms_leaf / n_leaf
    ach = leaf_blocks * CHACHA8_OPS / (avg_leaf * 1e-3)
    peak_meas = ceiling * CHACHA8_OPS
    peak = max(VALU_NOMINAL, peak_meas)
    traffic_file = os.path.join(ROOT, "profiles", "pcg_ole_expand_N20_c4_t16_hbm_traffic.md")
    traffic = pmc_traffic(traffic_file)
    roofline = {"bound": "valu", "kernel": kernel,
                "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "Tops/s",
                "frac": ach / peak, "traffic": traffic,
                "peak_nominal": VALU_NOMINAL / 1e12, "peak_measured": peak_meas / 1e12,
                "algorithmic": algorithmic, "avg_launch_ms": avg_leaf}
    # every DPF launch of the expansion: all its ChaCha8 blocks over their time
    dpf_ms = sum(v[1] for k, v in zones.items() if k.startswith("dpf.")) / nprof
    dpf_rate = (tree_blocks + leaves) / (dpf_ms * 1e-3)
    # NTT: the forward batch of step 4 (2c^2 + 2c transforms of length N)
    nf = 2 * c * c + 2 * c
    f1 = zones.get("ntt.fwd_pass1", (1, 0.0))
    f2 = zones.get("ntt.fwd_pass2", (1, 0.0))
    ntt_bytes_per_pass = nf * N * 16                 # read + write 8 B per coefficient
    ntt_avg = (f1[1] / f1[0] + f2[1] / f2[0]) / 2
    roof_ntt = {"bound": "hbm", "kernel": "ntt_pass_kernel (four-step, forward, batch %d, N=%d)"
                % (nf, N), "achieved": ntt_bytes_per_pass / (ntt_avg * 1e-3) / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ntt_bytes_per_pass / (ntt_avg * 1e-3) / HBM_PEAK,
                "traffic": pmc_traffic(traffic_file, "ntt_fwd_pass_hbm_bytes_per_launch"),
                "avg_launch_ms": ntt_avg}
    # output layer after Beaver (gather_fix_kernel): per coefficient of each
    # party's c^2 blocks, g read and written and one tau word per 64 landing
    # slots read (the +-CW of the set bits added in)
    roof_conv = None
    if "out.gather_fix" in zones:
        n_gf, ms_gf = zones["out.gather_fix"]
        gf_avg = ms_gf / n_gf
        conv_bytes = 2 * c * c * N * (16 + 8 * ((4 * t + 63) // 64))
        roof_conv = {"bound": "hbm",
                     "kernel": "gather_fix_kernel (output layer: the +-CW terms of the set tau "
                               "bits into g, both parties, %d blocks, N=%d)" % (c * c, N),
                     "achieved": conv_bytes / (gf_avg * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": conv_bytes / (gf_avg * 1e-3) / HBM_PEAK,
                     "traffic": pmc_traffic(traffic_file, "gather_fix_hbm_bytes_per_launch"),
                     "algorithmic_bytes": conv_bytes, "avg_launch_ms": gf_avg}
    out = {
        "metric": METRIC,
        "value": N * args.steps / dt,
        "unit": "OLEs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded dealer inputs (noise, Delta/DltSft, beta, Beaver shares)",
        "config": {"workload": "pcg-ole expansion, both parties on one GPU", "N": N, "c": c,
                   "t": t, "w": 1, "prg": "chacha8", "instances_per_party": c * c * t * t,
                   "dpf_domain": 2 * N // t},
        "latency_ms": {"p50": percentile(lats, 50), "p99": percentile(lats, 99),
                       "form": "one request per launch, host-visible completion",
                       "requests": len(lats)},
        "per_party_OLEs_per_s": N * args.steps / dt / 2,
        "pcie_inclusive_OLEs_per_s": N / dt_pcie,
        "pcie_inclusive_pinned_OLEs_per_s": N / dt_pinned,
        "preprocessing_gmw": {
            "what": "both parties' 2PC preprocessing in one plan (pcgsq_ole_deal_gmw): IKNP random "
                    "OTs both ways (base OTs once per plan), position adder (Kogge-Stone) and "
                    "DltSft circuit under GMW with AND triples from the OTs, beta and Beaver rs "
                    "by Gilboa OT products, every message a device copy; replaces the dealer's "
                    "DltSft, beta and rs (same outputs)",
            "instances": c * c * t * t, "and_gates": c * c * t * t * gmw_and_gates(N, t),
            "deal_ms": deal_ms[True], "dealer_deal_ms": deal_ms[False],
            "and_gates_per_s": c * c * t * t * gmw_and_gates(N, t) / (deal_ms[True] * 1e-3)},
        "roofline": roofline,
        "roofline_ntt": roof_ntt,
        "roofline_convert": roof_conv,
        "phases_ms": {k: v[1] / nprof for k, v in sorted(zones.items())},
        "dpf_chacha8_rate_Gblocks_per_s": dpf_rate / 1e9,
        "dpf_ms_per_expansion": dpf_ms,
        "chacha8_ceiling_Gblocks_per_s": ceiling / 1e9,
        "hbm_copy_GBps": copy_bps / 1e9,
        "hbm_probe_GBps": {"copy": hbm["copy"] / 1e9, "read": hbm["read"] / 1e9,
                           "write": hbm["write"] / 1e9, "copy_shape": hbm["copy_shape"]},
        "batched_ntt": ntt_batch,
        "polymul_recipe": polymul,
        "dpf_block_recipe": dpf_block,
        "batched_dpf": dpf_batch,
        "arity": arity,
        "one_party_stub_peer": one_party,
        "two_party_processes": two_parties,
        "served": served,
    }
    if not args.no_cpu:
        # the host share this job was given (OMP_NUM_THREADS), else every CPU
        # the process may run on; `host` records nproc and the online count
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        out["cpu_baseline"] = cpu_baseline_full(N, c, t, threads)
        # the reference-faithful mode: one thread per party, both parties
        out["cpu_baseline_one_thread_per_party"] = cpu_baseline(N, c, t, 2, polymul_sample=4)
        out["cpu_published_context"] = ("CPU_baseline 2PC at N=2^20, c=4, t=16 on a Xeon Gold "
                                        "6442Y: 90.8 s wall, 11.5 k OLEs/s (reference results)")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
