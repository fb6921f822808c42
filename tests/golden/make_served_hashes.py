"""Writes tests/golden/served_hashes.json: FNV-1a-64 hashes of x0, z0, x1, z1
for a fixed set of expansion requests (ring dimension N in 2^16..2^22, c=4,
t=16, w=1, ChaCha8, pub_seed 0), computed with the CPU oracle on the dealt
inputs of each seed.  The served-path GPU test submits the same requests to
the server and compares the hashes of what it returns.

Run from the repository root:  python tests/golden/make_served_hashes.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

REQUESTS = [(1 << 16, 7001), (1 << 16, 7002), (1 << 16, 7003), (1 << 18, 7004),
            (1 << 18, 7005), (1 << 20, 7006), (1 << 20, 7007), (1 << 22, 7008)]


def main() -> None:
    threads = min(8, os.cpu_count() or 1)
    out = []
    for N, seed in REQUESTS:
        t0 = time.time()
        x0, z0, x1, z1 = O.pcg_ole_expand(N, 4, 16, 1, O.PRG_CHACHA8, seed=seed, pub_seed=0,
                                          threads=threads)
        out.append({"N": N, "c": 4, "t": 16, "w": 1, "prg": "chacha8", "seed": seed,
                    "pub_seed": 0,
                    "fnv1a64": [f"{O.fnv1a64(v):016x}" for v in (x0, z0, x1, z1)]})
        print(f"N=2^{N.bit_length() - 1} seed={seed}: {time.time() - t0:.1f} s", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "served_hashes.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_served_hashes.py (CPU oracle)",
                   "requests": out}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
