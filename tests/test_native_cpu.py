"""Native pieces that can be checked without a GPU: the device field
arithmetic compiled for the host, the C ABI's exported symbols, and that the
C++ drop-in header builds as plain C++ against the library."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pcg-square_amd", "pcgsq", "libpcgsq.so")


def _hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def test_field_arithmetic_host(tmp_path):
    """fmul / shoup / freduce64 / mulhi64 / mulP of device_common.hpp against
    exact u128 arithmetic (edge values + 2M random pairs)."""
    exe = tmp_path / "field_host_test"
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "pcg-square_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "field_host_test.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe), "2000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pcgsq.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pcgsq_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libpcgsq.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(pcgsq_\w+)", out))
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if s not in exported]
    assert not missing, missing


def test_library_loads_via_ctypes():
    import pcgsq
    L = pcgsq.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    # no device here: pure-host entry points still work
    assert pcgsq.dealer_u64(1, 2, 3, 4) == pcgsq.dealer_u64(1, 2, 3, 4)
    rk = pcgsq.aes128_expand_key(bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c"))
    assert rk[160:176].hex() == "d014f9a8c9ee2589e13f0cc8b6630ca6"
    ok, why = pcgsq.poly_mul_supported(1 << 25)
    assert not ok and "2^24" in why


def test_dealer_matches_oracle():
    import oracle as O
    import pcgsq
    for args in [(1, 1, 0, 0), (7, 4, 123, 456), (2**63 + 5, 9, 3, 2**40)]:
        assert pcgsq.dealer_u64(*args) == O.dealer_u64(*args)


def test_dropin_header_builds_with_gxx(tmp_path):
    exe = tmp_path / "dropin_test"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "native", "dropin_test.cpp"),
                        "-L" + os.path.dirname(LIB), "-lpcgsq", "-pthread", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


_EXCHANGE_PROCESS = """
import os, sys
import numpy as np
sys.path.insert(0, os.environ["PCGSQ_PKG"])
import torch.distributed as dist
import pcgsq as pq
rank = int(os.environ["RANK"])
dist.init_process_group("gloo", init_method="tcp://127.0.0.1:" + os.environ["MASTER_PORT"],
                        rank=rank, world_size=2)
ex = pq.dist_exchange(1 - rank)
# equal sizes, then uneven ones (the IKNP matrices), then one-way messages
sizes = [(16, 16), (4096, 4096), (1 << 20, 1 << 20), (100, 3000), (0, 64), (64, 0)]
for k, (n0, n1) in enumerate(sizes):
    ns, nr = (n0, n1) if rank == 0 else (n1, n0)
    mine = np.full(ns, 17 * rank + k, dtype=np.uint8)
    got = ex("dpf_gen", k, 1, mine, nr)
    assert got.dtype == np.uint8 and got.shape == (nr,), (got.dtype, got.shape)
    assert np.all(got == 17 * (1 - rank) + k), (rank, k)
# base OTs between the two processes: the chosen seeds are the other party's
b = pq.base_ots_run(rank, 1000 + rank, ex)
def blk(x):
    return int(x.lo) | (int(x.hi) << 64)
mine = [blk(v) for v in b.k0] + [blk(v) for v in b.k1] + [blk(b.s)] + [blk(v) for v in b.ks]
buf = np.frombuffer(b"".join(v.to_bytes(16, "little") for v in mine), dtype=np.uint8)
got = ex("check", 0, 1, buf, buf.size)
peer = [int.from_bytes(got[16 * i:16 * i + 16].tobytes(), "little") for i in range(len(mine))]
pk0, pk1, ps, pks = peer[:128], peer[128:256], peer[256], peer[257:]
mk0, mk1, ms, mks = mine[:128], mine[128:256], mine[256], mine[257:]
for i in range(128):
    assert mks[i] == (pk1[i] if (ms >> i) & 1 else pk0[i]), ("mine", i)
    assert pks[i] == (mk1[i] if (ps >> i) & 1 else mk0[i]), ("peer's", i)
print("ok", rank, flush=True)
dist.destroy_process_group()
"""


def test_dist_exchange_swaps_between_two_processes():
    """pcgsq.dist_exchange, the OleParty exchange over torch.distributed
    (gloo, CPU tensors): two processes on 127.0.0.1 each receive exactly the
    other's bytes, in order, for messages of 16 B to 1 MiB, of uneven sizes
    and one-way; and the base OTs run over it between the processes give
    each party, in each direction, the column seed its choice bit selects
    from the other party's pair."""
    import socket
    import subprocess
    import sys

    with socket.socket() as so:
        so.bind(("127.0.0.1", 0))
        port = so.getsockname()[1]
    procs = []
    for rank in (0, 1):
        env = dict(os.environ, RANK=str(rank), MASTER_PORT=str(port),
                   PCGSQ_PKG=os.path.join(ROOT, "pcg-square_amd"))
        procs.append(subprocess.Popen([sys.executable, "-c", _EXCHANGE_PROCESS], env=env,
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    logs = []
    for p in procs:
        try:
            logs.append(p.communicate(timeout=120)[0])
        except subprocess.TimeoutExpired:
            for q in procs:
                q.kill()
            raise
    for rank, (p, log) in enumerate(zip(procs, logs)):
        assert p.returncode == 0, log
        assert f"ok {rank}" in log.splitlines(), log


def test_plan_estimate_rejects_instance_counts_past_int_range():
    """c^2 t^2 DPF instances index device arrays as int: parameter sets whose
    count would wrap (c = 9, t = 2^13 gives 5.4e9) are rejected, so the plan
    estimate is 0 and no plan or served request is sized from a wrapped
    count; a valid set right below the bound still gets an estimate."""
    import pcgsq
    assert pcgsq.ole_plan_estimate(1 << 24, 9, 1 << 13) == 0
    assert pcgsq.ole_plan_estimate(1 << 20, 16, 1 << 12) == 0
    assert pcgsq.ole_plan_estimate(1 << 24, 1, 1 << 15) == 0      # t^2 = 2^30
    assert pcgsq.ole_plan_estimate(1 << 20, 4, 16) > 0
    assert pcgsq.ole_plan_estimate(1 << 20, 2, 1 << 13) > 0         # 2^28 instances


def test_oracle_aes128_matches_libcrypto():
    """The oracle's AES-128 against OpenSSL's libcrypto (an independent
    implementation, loaded through ctypes: EVP aes-128-ecb without padding)
    on random keys and blocks; and the AES tree PRG's sub-key derivation
    sub_k = AES_S(makeBlock(0, k)) xor makeBlock(0, k) and its CCR hash
    H_k(x) = AES_{sub_k}(x) xor x (half_tree_dpf/tree_prg.h:42-68,
    mkey_prp.h:24-79) recomputed from libcrypto's cipher."""
    import ctypes
    import ctypes.util

    import numpy as np

    import oracle as O
    name = ctypes.util.find_library("crypto")
    if not name:
        pytest.skip("libcrypto not present")
    L = ctypes.CDLL(name)
    L.EVP_CIPHER_CTX_new.restype = ctypes.c_void_p
    L.EVP_aes_128_ecb.restype = ctypes.c_void_p
    L.EVP_EncryptInit_ex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_char_p, ctypes.c_char_p]
    L.EVP_CIPHER_CTX_set_padding.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.EVP_EncryptUpdate.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                    ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    L.EVP_CIPHER_CTX_free.argtypes = [ctypes.c_void_p]

    def ssl_aes(key: bytes, pt: bytes) -> bytes:
        cx = L.EVP_CIPHER_CTX_new()
        try:
            assert L.EVP_EncryptInit_ex(cx, L.EVP_aes_128_ecb(), None, key, None) == 1
            L.EVP_CIPHER_CTX_set_padding(cx, 0)
            out = ctypes.create_string_buffer(len(pt) + 16)
            n = ctypes.c_int()
            assert L.EVP_EncryptUpdate(cx, out, ctypes.byref(n), pt, len(pt)) == 1
            assert n.value == len(pt)
            return out.raw[:len(pt)]
        finally:
            L.EVP_CIPHER_CTX_free(cx)

    rng = np.random.default_rng(197)
    for _ in range(64):
        key = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
        pt = rng.integers(0, 256, 16 * 8, dtype=np.uint8).tobytes()
        rk = O.aes128_expand_key(key)
        got = b"".join(O.aes128_encrypt(rk, pt[i:i + 16]) for i in range(0, len(pt), 16))
        assert got == ssl_aes(key, pt)
    master = O.FIPS197_MASTER
    for nk in (1, 3, 15):
        rks = O.aes_tree_round_keys(nk, master)
        parents = rng.integers(0, 2**63, (5, 2), dtype=np.uint64)
        out = np.empty((5 * nk, 2), dtype=np.uint64)
        O.lib().orc_tree_hash_aes(rks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                  out.ctypes.data, parents.ctypes.data, 5, nk)
        for k in range(nk):
            mk = k.to_bytes(8, "little") + bytes(8)
            sub = bytes(a ^ b for a, b in zip(ssl_aes(master, mk), mk))
            for j in range(5):
                x = parents[j].tobytes()
                want = bytes(a ^ b for a, b in zip(ssl_aes(sub, x), x))
                assert out[j * nk + k].tobytes() == want, (nk, k, j)


def test_base_ots_two_threads_choose_the_right_seeds():
    """Chou-Orlandi base OTs of the two parties (host code, libcrypto P-256)
    in two threads over a queue pair: for every column i of both directions
    the receiver's seed is the sender's k_{s_i}, the other seed differs from
    it, seeds are distinct across columns, the choice bits are not constant,
    and the same seeds give the same results (seeded draws)."""
    import queue
    import threading

    import numpy as np

    import pcgsq

    def run(seeds):
        inbox = (queue.Queue(), queue.Queue())
        out, errs = [None, None], []

        def party(p):
            try:
                def ex(phase, level, batch, send, nrecv):
                    inbox[1 - p].put(bytes(send))
                    got = inbox[p].get(timeout=30)
                    assert len(got) == nrecv
                    return np.frombuffer(got, dtype=np.uint8)
                out[p] = pcgsq.base_ots_run(p, seeds[p], ex)
            except Exception as e:
                errs.append(e)
        th = [threading.Thread(target=party, args=(p,)) for p in (0, 1)]
        for x in th:
            x.start()
        for x in th:
            x.join(60)
        assert not errs, errs
        return out

    def blk(x):
        return int(x.lo) | (int(x.hi) << 64)

    a = run((11, 22))
    for snd, rcv in ((a[1], a[0]), (a[0], a[1])):   # direction 0: party 0 sends
        s = blk(rcv.s)
        assert 0 < s < (1 << 128) - 1
        k0 = [blk(v) for v in snd.k0]
        k1 = [blk(v) for v in snd.k1]
        ks = [blk(v) for v in rcv.ks]
        assert len(set(k0 + k1)) == 256
        for i in range(128):
            want, other = (k1[i], k0[i]) if (s >> i) & 1 else (k0[i], k1[i])
            assert ks[i] == want and ks[i] != other, i
    b = run((11, 22))
    assert [blk(v) for v in b[0].ks] == [blk(v) for v in a[0].ks]
    c = run((12, 22))
    assert [blk(v) for v in c[0].ks] != [blk(v) for v in a[0].ks]
