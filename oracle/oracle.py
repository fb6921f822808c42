"""ctypes view of the CPU oracle (liboracle.so) and of the compiled reference
headers (_ref/libpcgref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product never imports
it.  Function-level citations live in oracle.h and the C sources.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = 4611686018326724609
OUT_DOMAIN = 0x6F75745F68617368
PRG_AES, PRG_CHACHA8 = 0, 1

_lib = None
_ref = None

u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)


class Blk(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


class PartyInputs(C.Structure):
    """orc_party_inputs: one party's dealt inputs (pointers)."""
    _fields_ = [(n, C.c_void_p) for n in ("pos", "pay", "root", "dltsft", "beta", "r", "s", "rs")]


class Params(C.Structure):
    _fields_ = [("N", C.c_int), ("c", C.c_int), ("t", C.c_int), ("w", C.c_int),
                ("prg", C.c_int)]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    so = os.path.join(HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/CPU_baseline"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        so = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _lib = C.CDLL(so)
        L = _lib
        L.orc_fmul.restype = C.c_uint64
        L.orc_fmul.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_gilboa.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32,
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_fpow.restype = C.c_uint64
        L.orc_fpow.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_finv.restype = C.c_uint64
        L.orc_finv.argtypes = [C.c_uint64]
        L.orc_primitive_root.restype = C.c_uint64
        L.orc_ntt_new.restype = C.c_void_p
        L.orc_ntt_new.argtypes = [C.c_int]
        L.orc_ntt_free.argtypes = [C.c_void_p]
        L.orc_ntt_psi.restype = C.c_uint64
        L.orc_ntt_psi.argtypes = [C.c_void_p]
        L.orc_ntt_forward.argtypes = [C.c_void_p, u64p]
        L.orc_ntt_inverse.argtypes = [C.c_void_p, u64p]
        L.orc_poly_mul_schoolbook.argtypes = [u64p, u64p, u64p, C.c_int]
        L.orc_poly_mul.argtypes = [u64p, u64p, u64p, C.c_int]
        L.orc_poly_mul_batch.argtypes = [u64p, u64p, u64p, C.c_int, C.c_int, C.c_int]
        L.orc_chacha8_block.argtypes = [u8p, C.c_uint64, C.c_uint64, u8p]
        L.orc_tree_hash_chacha8.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.orc_out_hash.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.orc_aes128_expand_key.argtypes = [u8p, u8p]
        L.orc_aes128_encrypt.argtypes = [u8p, u8p, u8p]
        L.orc_aes_tree_round_keys.argtypes = [u8p, C.c_int, u8p]
        L.orc_tree_hash_aes.argtypes = [u8p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.orc_dpf_batch_full_eval_both.argtypes = [
            C.c_int, C.c_int, C.c_int, C.c_int, u8p, C.c_void_p, C.c_void_p,
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_dpf_batch_full_eval_both.restype = C.c_int
        L.orc_dpf_eval.restype = Blk
        L.orc_dpf_eval.argtypes = [C.c_int, C.c_int, C.c_int, u8p, Blk, C.c_void_p, C.c_uint64]
        L.orc_public_W.restype = Blk
        L.orc_dpf_key_gen.argtypes = [C.c_int, C.c_int, u8p, C.c_uint64, C.c_uint64, Blk, Blk,
                                      Blk, C.POINTER(Blk), C.POINTER(Blk), C.c_void_p,
                                      C.POINTER(C.c_uint64)]
        L.orc_dpf_key_eval_y.argtypes = [C.c_int, C.c_int, u8p, Blk, C.c_void_p, C.c_uint64,
                                         C.c_int, C.c_void_p]
        L.orc_dpf_path_cw.argtypes = [C.c_int, C.c_int, C.c_int, u8p, Blk, Blk,
                                      C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.orc_dpf_digit.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64]
        L.orc_out_sums.argtypes = [C.c_void_p, C.c_int, C.c_size_t, u64p, u64p]
        L.orc_out_scatter_g.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_int, C.c_int,
                                        u64p, u64p, C.c_int]
        L.orc_output_layer_both.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t,
                                            C.c_int] + [u64p] * 8 + [u64p, u64p, C.c_int, u64p]
        L.orc_dealer_u64.restype = C.c_uint64
        L.orc_dealer_u64.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64]
        L.orc_public_a.argtypes = [C.c_uint64, C.c_int, C.c_int, u64p]
        L.orc_pcg_ole_expand.argtypes = [C.POINTER(Params), C.c_uint64, C.c_uint64,
                                         u64p, u64p, u64p, u64p, C.c_int]
        L.orc_pcg_ole_expand.restype = C.c_int
        L.orc_generate_triples.argtypes = [C.POINTER(Params), C.c_uint64, C.c_uint64, C.c_uint64,
                                           u64p, u64p, u64p, u64p, u64p, u64p, C.c_int]
        L.orc_generate_triples.restype = C.c_int
        L.orc_pcg_ole_blocks.argtypes = [C.POINTER(Params), C.c_uint64, C.c_int, C.c_int,
                                         u64p, u64p, C.c_int]
        L.orc_pcg_ole_blocks.restype = C.c_int
        L.orc_fnv1a64.restype = C.c_uint64
        L.orc_fnv1a64.argtypes = [u64p, C.c_size_t, C.c_uint64]
        L.orc_mt64_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_mt64_next.restype = C.c_uint64
        L.orc_mt64_next.argtypes = [C.c_void_p]
        L.orc_mt64_fill.argtypes = [C.c_void_p, u64p, C.c_size_t, C.c_uint64]
        L.orc_dpf_real_bench_checksum.restype = C.c_uint64
        L.orc_dpf_real_bench_checksum.argtypes = [C.c_int] * 5
        L.orc_dealer_sizes.restype = C.c_size_t
        L.orc_dealer_sizes.argtypes = [C.POINTER(Params), C.POINTER(C.c_size_t),
                                       C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.orc_dealer.argtypes = [C.POINTER(Params), C.c_uint64, C.POINTER(PartyInputs),
                                 C.POINTER(PartyInputs), u64p]
        L.orc_dealer.restype = C.c_int
        L.orc_gmw_preprocess.argtypes = [C.POINTER(Params), C.c_uint64, C.POINTER(PartyInputs),
                                         C.POINTER(PartyInputs), u64p, u64p]
        L.orc_gmw_preprocess.restype = C.c_int
        L.orc_gmw_adder_ops.argtypes = [C.c_int]
        L.orc_gmw_and_gates.argtypes = [C.c_int, C.c_int, C.c_int]
    return _lib


def ref():
    """Compiled reference headers, or None where they are not built (no reference sources)."""
    global _ref
    if _ref is None:
        so = os.path.join(HERE, "_ref", "libpcgref.so")
        if not os.path.exists(so):
            return None
        _ref = C.CDLL(so)
        _ref.ref_ntt_forward.argtypes = [u64p, C.c_int]
        _ref.ref_ntt_inverse.argtypes = [u64p, C.c_int]
        _ref.ref_poly_mul.argtypes = [u64p, u64p, u64p, C.c_int]
        _ref.ref_negacyclic_psi.restype = C.c_uint64
        _ref.ref_negacyclic_psi.argtypes = [C.c_int]
        _ref.ref_chacha8_blocks.argtypes = [u8p, C.c_uint64, C.c_uint64, u8p, C.c_size_t]
    return _ref


def _p(a: np.ndarray, t=u64p):
    return a.ctypes.data_as(t)


def u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


# ---- field / NTT -------------------------------------------------------------

def fmul(a: int, b: int) -> int:
    return lib().orc_fmul(a, b)


def ntt_forward(a) -> np.ndarray:
    a = u64(a).copy()
    t = lib().orc_ntt_new(len(a))
    lib().orc_ntt_forward(t, _p(a))
    lib().orc_ntt_free(t)
    return a


def ntt_inverse(a) -> np.ndarray:
    a = u64(a).copy()
    t = lib().orc_ntt_new(len(a))
    lib().orc_ntt_inverse(t, _p(a))
    lib().orc_ntt_free(t)
    return a


def ntt_psi(N: int) -> int:
    t = lib().orc_ntt_new(N)
    v = lib().orc_ntt_psi(t)
    lib().orc_ntt_free(t)
    return v


def poly_mul(a, b) -> np.ndarray:
    a, b = u64(a), u64(b)
    c = np.empty_like(a)
    lib().orc_poly_mul(_p(c), _p(a), _p(b), len(a))
    return c


def poly_mul_schoolbook(a, b) -> np.ndarray:
    a, b = u64(a), u64(b)
    c = np.empty_like(a)
    lib().orc_poly_mul_schoolbook(_p(c), _p(a), _p(b), len(a))
    return c


def poly_mul_batch(a, b, threads: int = 8) -> np.ndarray:
    a, b = u64(a), u64(b)
    batch, N = a.shape
    c = np.empty_like(a)
    lib().orc_poly_mul_batch(_p(c), _p(a), _p(b), batch, N, threads)
    return c


# ---- ChaCha8 / AES ---------------------------------------------------------------

def chacha8_block(key: bytes, nonce: int = 0, ctr: int = 0) -> bytes:
    k = (C.c_uint8 * 32).from_buffer_copy(key.ljust(32, b"\0"))
    out = (C.c_uint8 * 64)()
    lib().orc_chacha8_block(k, nonce, ctr, out)
    return bytes(out)


def blocks_from(a: np.ndarray) -> np.ndarray:
    """(n,2) uint64 [lo, hi] array, contiguous."""
    return np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 2)


def tree_hash_chacha8(parents: np.ndarray, nk: int) -> np.ndarray:
    parents = blocks_from(parents)
    out = np.empty((len(parents) * nk, 2), dtype=np.uint64)
    lib().orc_tree_hash_chacha8(out.ctypes.data, parents.ctypes.data, len(parents), nk)
    return out


def out_hash(leaves: np.ndarray) -> np.ndarray:
    leaves = blocks_from(leaves)
    out = np.empty_like(leaves)
    lib().orc_out_hash(out.ctypes.data, leaves.ctypes.data, len(leaves))
    return out


def aes128_expand_key(key: bytes) -> bytes:
    k = (C.c_uint8 * 16).from_buffer_copy(key)
    rk = (C.c_uint8 * 176)()
    lib().orc_aes128_expand_key(k, rk)
    return bytes(rk)


def aes128_encrypt(rk: bytes, pt: bytes) -> bytes:
    r = (C.c_uint8 * 176).from_buffer_copy(rk)
    i = (C.c_uint8 * 16).from_buffer_copy(pt)
    o = (C.c_uint8 * 16)()
    lib().orc_aes128_encrypt(r, i, o)
    return bytes(o)


FIPS197_MASTER = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")


def aes_tree_round_keys(nk: int, master: bytes = FIPS197_MASTER) -> np.ndarray:
    m = (C.c_uint8 * 16).from_buffer_copy(master)
    rks = np.zeros(nk * 176, dtype=np.uint8)
    lib().orc_aes_tree_round_keys(m, nk, _p(rks, u8p))
    return rks


# ---- DPF ---------------------------------------------------------------------------

def dpf_full_eval_both(n, w, prg, roots0, roots1, dlt0, dlt1, rks=None, want_cw=False):
    roots0, roots1 = blocks_from(roots0), blocks_from(roots1)
    dlt0, dlt1 = blocks_from(dlt0), blocks_from(dlt1)
    B = len(roots0)
    D = 1 << n
    l0 = np.empty((B * D, 2), dtype=np.uint64)
    l1 = np.empty((B * D, 2), dtype=np.uint64)
    m = 1 << w
    cw = np.empty((B * (n // w) * m, 2), dtype=np.uint64) if want_cw else None
    r = None if rks is None else _p(np.ascontiguousarray(rks, dtype=np.uint8), u8p)
    rc = lib().orc_dpf_batch_full_eval_both(
        B, n, w, prg, r, roots0.ctypes.data, roots1.ctypes.data, dlt0.ctypes.data,
        dlt1.ctypes.data, l0.ctypes.data, l1.ctypes.data,
        cw.ctypes.data if want_cw else None)
    if rc:
        raise ValueError("orc_dpf_batch_full_eval_both failed")
    return (l0, l1, cw) if want_cw else (l0, l1)


def dpf_eval(n, w, prg, root, cw, j, rks=None) -> tuple[int, int]:
    cw = blocks_from(cw)
    r = None if rks is None else _p(np.ascontiguousarray(rks, dtype=np.uint8), u8p)
    b = lib().orc_dpf_eval(n, w, prg, r, Blk(int(root[0]), int(root[1])), cw.ctypes.data, j)
    return (b.lo, b.hi)


def dpf_path_cw(n, w, prg, root0, root1, dlt0, dlt1, alpha, rks=None) -> np.ndarray:
    dlt0, dlt1 = blocks_from(dlt0), blocks_from(dlt1)
    m = 1 << w
    out = np.empty(((n // w) * m, 2), dtype=np.uint64)
    r = None if rks is None else _p(np.ascontiguousarray(rks, dtype=np.uint8), u8p)
    lib().orc_dpf_path_cw(n, w, prg, r, Blk(int(root0[0]), int(root0[1])),
                          Blk(int(root1[0]), int(root1[1])), dlt0.ctypes.data,
                          dlt1.ctypes.data, alpha, out.ctypes.data)
    return out


def public_W() -> tuple[int, int]:
    """The public W block of the dealt expansion (the F_RAND_SEED constant)."""
    b = lib().orc_public_W()
    return (b.lo, b.hi)


def dpf_key_gen(n, prg, alpha, beta, delta, d0, W=None, rks=None):
    """Dealer Gen of a w = 1 key pair: (root0, root1, cw[n][2], cw_out)."""
    W = public_W() if W is None else W
    r0, r1, co = Blk(), Blk(), C.c_uint64()
    cw = np.empty((n, 2), dtype=np.uint64)
    r = None if rks is None else _p(np.ascontiguousarray(rks, dtype=np.uint8), u8p)
    lib().orc_dpf_key_gen(n, prg, r, int(alpha), int(beta), Blk(int(delta[0]), int(delta[1])),
                          Blk(int(d0[0]), int(d0[1])), Blk(int(W[0]), int(W[1])), C.byref(r0),
                          C.byref(r1), cw.ctypes.data, C.byref(co))
    return (r0.lo, r0.hi), (r1.lo, r1.hi), cw, co.value


def dpf_key_eval_y(n, prg, root, cw, cw_out, party, rks=None) -> np.ndarray:
    """Party `party`'s converted full domain y[2^n] from its key."""
    cw = np.ascontiguousarray(cw, dtype=np.uint64).reshape(-1, 2)
    y = np.empty(1 << n, dtype=np.uint64)
    r = None if rks is None else _p(np.ascontiguousarray(rks, dtype=np.uint8), u8p)
    lib().orc_dpf_key_eval_y(n, prg, r, Blk(int(root[0]), int(root[1])), cw.ctypes.data,
                             int(cw_out), party, y.ctypes.data)
    return y


# ---- output layer ----------------------------------------------------------------

def out_sums(leaves, B, D):
    leaves = blocks_from(leaves)
    sc = np.empty(B, dtype=np.uint64)
    st = np.empty(B, dtype=np.uint64)
    lib().orc_out_sums(leaves.ctypes.data, B, D, _p(sc), _p(st))
    return sc, st


def out_scatter_g(leaves, B, D, t, party, cw, N):
    leaves = blocks_from(leaves)
    cw = u64(cw)
    g = np.empty(N, dtype=np.uint64)
    lib().orc_out_scatter_g(leaves.ctypes.data, B, D, t, party, _p(cw), _p(g), N)
    return g


# ---- dealer / expansion ---------------------------------------------------------

def dealer_u64(seed, dom, a, b) -> int:
    return lib().orc_dealer_u64(seed, dom, a, b)


def public_a(pub_seed: int, c: int, N: int) -> np.ndarray:
    out = np.zeros((max(c - 1, 1), N), dtype=np.uint64)
    if c > 1:
        lib().orc_public_a(pub_seed, c, N, _p(out))
    return out[: c - 1]


def dealt_inputs(N, c, t, w=1, seed=1, gmw=False):
    """Both parties' dealt inputs: dict of pos, pay [2][c*t], root [2][ninst][2],
    dltsft [2][ninst][levels][nk][2], beta, r, s, rs [2][ninst], and alpha
    ([ninst], dealer) or alpha_sh ([2][ninst], the GMW adder's XOR shares).
    gmw=True: DltSft from the GMW preprocessing circuits, beta and rs from
    Gilboa OT products, instead of the dealer."""
    p = Params(N, c, t, w, PRG_CHACHA8)
    npos, ninst, ndlt = C.c_size_t(), C.c_size_t(), C.c_size_t()
    lib().orc_dealer_sizes(C.byref(p), C.byref(npos), C.byref(ninst), C.byref(ndlt))
    npos, ninst, ndlt = npos.value, ninst.value, ndlt.value
    out = {"pos": np.empty((2, npos), np.uint64), "pay": np.empty((2, npos), np.uint64),
           "root": np.empty((2, ninst, 2), np.uint64), "dltsft": np.empty((2, ndlt, 2), np.uint64)}
    for k in ("beta", "r", "s", "rs"):
        out[k] = np.empty((2, ninst), np.uint64)
    pi = [PartyInputs(*[out[k][s].ctypes.data for k in ("pos", "pay", "root", "dltsft", "beta",
                                                         "r", "s", "rs")]) for s in range(2)]
    if gmw:
        out["alpha_sh"] = np.empty((2, ninst), np.uint64)
        rc = lib().orc_gmw_preprocess(C.byref(p), seed, C.byref(pi[0]), C.byref(pi[1]),
                                      _p(out["alpha_sh"][0]), _p(out["alpha_sh"][1]))
    else:
        out["alpha"] = np.empty(ninst, np.uint64)
        rc = lib().orc_dealer(C.byref(p), seed, C.byref(pi[0]), C.byref(pi[1]), _p(out["alpha"]))
    if rc:
        raise ValueError("dealer: invalid parameters")
    n = int(2 * N // t).bit_length() - 1
    levels, nk = n // w, (1 << w) - 1
    out["dltsft"] = out["dltsft"].reshape(2, ninst, levels, nk, 2)
    return out


def gilboa(x: int, y: int, seed: int, g: int, q: int) -> tuple:
    """Gilboa's OT product (bool_circuit/circuit.cpp:91-153), product q of
    instance g: both parties' additive shares of x * y mod P."""
    s0, s1 = C.c_uint64(), C.c_uint64()
    lib().orc_gilboa(x, y, seed, g, q, C.byref(s0), C.byref(s1))
    return s0.value, s1.value


def gmw_and_gates(nb: int, w: int, n: int) -> int:
    return lib().orc_gmw_and_gates(nb, w, n)


def gmw_adder_ops(nb: int) -> int:
    return lib().orc_gmw_adder_ops(nb)


def pcg_ole_expand(N, c, t, w=1, prg=PRG_CHACHA8, seed=1, pub_seed=0, threads=8):
    p = Params(N, c, t, w, prg)
    x0, z0, x1, z1 = (np.empty(N, dtype=np.uint64) for _ in range(4))
    rc = lib().orc_pcg_ole_expand(C.byref(p), seed, pub_seed, _p(x0), _p(z0), _p(x1),
                                  _p(z1), threads)
    if rc:
        raise ValueError("orc_pcg_ole_expand: invalid parameters")
    return x0, z0, x1, z1


def generate_triples(N, c, t, seed1, seed2, w=1, prg=PRG_CHACHA8, pub_seed=0, threads=8):
    """(a0, b0, c0, a1, b1, c1), each [N]: the triple shares of triple.h."""
    p = Params(N, c, t, w, prg)
    out = [np.empty(N, dtype=np.uint64) for _ in range(6)]
    rc = lib().orc_generate_triples(C.byref(p), seed1, seed2, pub_seed,
                                    *[_p(o) for o in out], threads)
    if rc:
        raise ValueError("orc_generate_triples: invalid parameters")
    return tuple(out)


def pcg_ole_blocks(N, c, t, w=1, prg=PRG_CHACHA8, seed=1, blk_begin=0, blk_end=1,
                   threads=8, want_g=True):
    p = Params(N, c, t, w, prg)
    nb = blk_end - blk_begin
    g0 = np.empty((nb, N), dtype=np.uint64) if want_g else None
    g1 = np.empty((nb, N), dtype=np.uint64) if want_g else None
    rc = lib().orc_pcg_ole_blocks(C.byref(p), seed, blk_begin, blk_end,
                                  _p(g0) if want_g else None, _p(g1) if want_g else None,
                                  threads)
    if rc:
        raise ValueError("orc_pcg_ole_blocks: invalid parameters")
    return g0, g1


def fnv1a64(v, h=1469598103934665603) -> int:
    v = u64(v).ravel()
    return lib().orc_fnv1a64(_p(v), len(v), h)


class MT64:
    def __init__(self, seed: int):
        self._st = (C.c_uint8 * (312 * 8 + 16))()
        lib().orc_mt64_seed(C.addressof(self._st), seed)

    def next(self) -> int:
        return lib().orc_mt64_next(C.addressof(self._st))

    def take(self, n: int, mod: int = 0) -> np.ndarray:
        out = np.empty(n, dtype=np.uint64)
        lib().orc_mt64_fill(C.addressof(self._st), _p(out), n, mod)
        return out


def dpf_real_bench_checksum(n, B, t=0, w=1, threads=8) -> int:
    return lib().orc_dpf_real_bench_checksum(n, B, t, w, threads)
