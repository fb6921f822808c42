"""Python host interface of libpcgsq (ctypes over include/pcgsq.h).

Mirrors the reference's device plugin seam (GPU_baseline/common/*.h,
pcg_cuda::*) and the local-compute entries of CPU_baseline (NTT<P>,
poly_mul_batch, PCG_OLE<P>::gen_and_expand) with the same argument meaning.
Errors raise PcgsqError (status code + message); an unsupported shape raises
Unsupported, the reference's "exit 77" case.  There is no CPU fallback: if
libpcgsq.so is missing or no gfx950 device is present, calls fail.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpcgsq.so")
P = 4611686018326724609
PRG_AES, PRG_CHACHA8 = 0, 1
NTT_FOURSTEP, NTT_NAIVE = 0, 1

OK, EINVAL, EUNSUPPORTED, EDEVICE, ENOMEM, ECALLBACK, EINTERNAL = range(7)


class PcgsqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pcgsq error {code}: {msg}")
        self.code = code


class Unsupported(PcgsqError):
    pass


class Blk(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


class Params(C.Structure):
    _fields_ = [("N", C.c_int), ("c", C.c_int), ("t", C.c_int), ("w", C.c_int),
                ("prg", C.c_int)]


EXCHANGE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t)
OLE_EXCHANGE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_void_p,
                              C.c_size_t, C.c_void_p, C.c_size_t)
EXCHANGE_PER_BLOCK, EXCHANGE_BATCHED = 0, 1


class BaseOts(C.Structure):
    """One party's base-OT results (pcgsq_base_ots)."""
    _fields_ = [("k0", Blk * 128), ("k1", Blk * 128), ("s", Blk), ("ks", Blk * 128)]

_lib = None
u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
vp = C.c_void_p


def lib() -> C.CDLL:
    """Load the in-tree libpcgsq.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PcgsqError(EINTERNAL, f"{LIB_PATH} not built: run pcg-square_amd/build_native.py")
    L = C.CDLL(LIB_PATH)
    L.pcgsq_last_error.restype = C.c_char_p
    L.pcgsq_device_available.restype = C.c_int
    L.pcgsq_device_name.argtypes = [C.c_char_p, C.c_size_t]
    L.pcgsq_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pcgsq_ctx_destroy.argtypes = [vp]
    L.pcgsq_ctx_stream.argtypes = [vp]
    L.pcgsq_ctx_stream.restype = vp
    L.pcgsq_ctx_synchronize.argtypes = [vp]
    L.pcgsq_ctx_profile.argtypes = [vp, C.c_int]
    L.pcgsq_ctx_profile_read.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.pcgsq_probe_chacha8.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.pcgsq_probe_copy.argtypes = [vp, C.c_size_t, C.POINTER(C.c_double)]
    L.pcgsq_probe_hbm.argtypes = [vp, C.c_size_t, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.pcgsq_field_op.argtypes = [vp, C.c_int, C.c_size_t, u64p, u64p, u64p, u64p]
    L.pcgsq_dev_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    L.pcgsq_dev_free.argtypes = [vp, vp]
    L.pcgsq_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
    L.pcgsq_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
    L.pcgsq_dpf_batch_full_eval.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int,
                                            EXCHANGE_FN, vp, vp, C.c_int]
    L.pcgsq_dpf_batch_full_eval_device.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                                   C.c_int, EXCHANGE_FN, vp, C.POINTER(vp),
                                                   C.c_int]
    L.pcgsq_dpf_batch_full_eval_both.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp,
                                                 vp, vp, vp, vp, C.c_int]
    L.pcgsq_dpf_out_sums.argtypes = [vp, vp, C.c_int, C.c_size_t, C.c_uint64, u64p, u64p]
    L.pcgsq_dpf_out_sums_noalloc.argtypes = [vp, vp, C.c_int, C.c_size_t, C.c_uint64, vp, vp]
    L.pcgsq_dpf_out_scatter_g.argtypes = [vp, vp, C.c_int, C.c_size_t, C.c_int, C.c_int,
                                          C.c_uint64, u64p, u64p, C.c_int]
    L.pcgsq_poly_mul_supported.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_char_p, C.c_size_t]
    L.pcgsq_poly_mul_u64.argtypes = [vp, u64p, u64p, u64p, C.c_int, C.c_int, C.c_uint64,
                                     C.c_int, C.POINTER(C.c_double)]
    L.pcgsq_poly_mul_u64_dev.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int]
    for f in ("pcgsq_ntt_forward", "pcgsq_ntt_inverse"):
        getattr(L, f).argtypes = [vp, u64p, C.c_int, C.c_int, C.c_int]
    for f in ("pcgsq_ntt_forward_dev", "pcgsq_ntt_inverse_dev"):
        getattr(L, f).argtypes = [vp, vp, C.c_int, C.c_int, C.c_int]
    L.pcgsq_ntt_tune.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.pcgsq_aes128_expand_key.argtypes = [u8p, u8p]
    L.pcgsq_aes128_ecb_encrypt.argtypes = [vp, u8p, u8p, u8p, C.c_int]
    L.pcgsq_aes_tree_round_keys.argtypes = [u8p, C.c_int, u8p]
    L.pcgsq_ole_plan_create.argtypes = [vp, C.POINTER(Params), C.c_uint64, C.c_size_t,
                                        C.POINTER(vp)]
    L.pcgsq_ole_plan_destroy.argtypes = [vp]
    L.pcgsq_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.pcgsq_host_free.argtypes = [vp]
    L.pcgsq_ole_party_create.argtypes = [vp, C.POINTER(Params), C.c_int, C.c_uint64, C.c_size_t,
                                         C.POINTER(vp)]
    L.pcgsq_ole_party_run.argtypes = [vp, OLE_EXCHANGE_FN, vp]
    L.pcgsq_ole_party_outputs.argtypes = [vp, u64p, u64p]
    L.pcgsq_ole_party_deal_own.argtypes = [vp, C.c_uint64]
    L.pcgsq_ole_party_set_exchange_mode.argtypes = [vp, C.c_int]
    L.pcgsq_base_ots_run.argtypes = [C.c_int, C.c_uint64, OLE_EXCHANGE_FN, vp, C.POINTER(BaseOts)]
    L.pcgsq_ole_preprocessing_read.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t,
                                               C.POINTER(C.c_size_t)]
    L.pcgsq_ole_plan_bytes.argtypes = [vp]
    L.pcgsq_ole_plan_bytes.restype = C.c_size_t
    L.pcgsq_ole_deal.argtypes = [vp, C.c_uint64]
    L.pcgsq_ole_deal_gmw.argtypes = [vp, C.c_uint64]
    L.pcgsq_ole_dealt_dltsft.argtypes = [vp, C.c_void_p]
    L.pcgsq_ole_dealt_products.argtypes = [vp, C.c_void_p]
    L.pcgsq_ole_run.argtypes = [vp]
    L.pcgsq_ole_run_blocks.argtypes = [vp]
    L.pcgsq_ole_outputs.argtypes = [vp, u64p, u64p, u64p, u64p]
    L.pcgsq_ole_outputs_dev.argtypes = [vp]
    L.pcgsq_ole_outputs_dev.restype = vp
    L.pcgsq_ole_g.argtypes = [vp, u64p]
    L.pcgsq_ole_expand.argtypes = [vp, C.POINTER(Params), C.c_uint64, C.c_uint64, u64p, u64p,
                                   u64p, u64p]
    L.pcgsq_ntt_batch_create.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(vp)]
    L.pcgsq_ntt_batch_destroy.argtypes = [vp]
    L.pcgsq_ntt_batch_tune.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_uint32]
    L.pcgsq_ntt_batch_set_mode.argtypes = [vp, C.c_int]
    L.pcgsq_ntt_batch_dev.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                      C.POINTER(C.c_size_t)]
    L.pcgsq_ntt_batch_load.argtypes = [vp, u64p]
    L.pcgsq_ntt_batch_run.argtypes = [vp, C.c_int]
    L.pcgsq_ntt_batch_finish.argtypes = [vp]
    L.pcgsq_ntt_batch_read.argtypes = [vp, u64p, u64p]
    L.pcgsq_ntt_batch_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
    L.pcgsq_dpf_batch_create.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(vp)]
    L.pcgsq_dpf_batch_destroy.argtypes = [vp]
    L.pcgsq_dpf_batch_deal.argtypes = [vp, C.c_uint64]
    L.pcgsq_dpf_batch_gen.argtypes = [vp, u64p, u64p, vp, vp]
    L.pcgsq_dpf_batch_points.argtypes = [vp, u64p, u64p]
    L.pcgsq_dpf_batch_keys.argtypes = [vp, vp, vp, u64p]
    L.pcgsq_dpf_batch_run.argtypes = [vp, C.c_int]
    L.pcgsq_dpf_batch_read.argtypes = [vp, u64p, u64p]
    L.pcgsq_dpf_batch_dev.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_size_t)]
    L.pcgsq_dpf_batch_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pcgsq_triples_generate.argtypes = [vp, C.c_uint64, C.c_uint64, u64p, u64p, u64p, u64p,
                                         u64p, u64p]
    L.pcgsq_ole_plan_estimate.argtypes = [C.POINTER(Params), C.c_size_t]
    L.pcgsq_ole_plan_estimate.restype = C.c_size_t
    L.pcgsq_server_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(vp)]
    L.pcgsq_server_destroy.argtypes = [vp]
    L.pcgsq_server_submit.argtypes = [vp, C.POINTER(Params), C.c_uint64, C.c_uint64,
                                      C.POINTER(C.c_uint64)]
    L.pcgsq_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(C.c_int)]
    L.pcgsq_server_wait.argtypes = [vp, C.c_uint64, C.c_double]
    L.pcgsq_server_outputs.argtypes = [vp, C.c_uint64, u64p, u64p, u64p, u64p]
    L.pcgsq_server_outputs_dev.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.pcgsq_server_release.argtypes = [vp, C.c_uint64]
    L.pcgsq_server_timing.argtypes = [vp, C.c_uint64, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pcgsq_server_stats.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_size_t), C.POINTER(C.c_long),
                                     C.POINTER(C.c_int)]
    L.pcgsq_server_clock.restype = C.c_double
    L.pcgsq_server_charge.argtypes = [vp, C.POINTER(Params), C.POINTER(C.c_size_t)]
    L.pcgsq_server_plans.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_long),
                                     C.POINTER(C.c_long)]
    L.pcgsq_server_set_window.argtypes = [vp, C.c_uint64]
    L.pcgsq_server_set_policy.argtypes = [vp, C.c_uint64, C.c_double]
    L.pcgsq_server_reserve.argtypes = [vp, C.POINTER(Params), C.c_uint64, C.c_int]
    L.pcgsq_dealer_u64.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64]
    L.pcgsq_dealer_u64.restype = C.c_uint64
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != OK:
        msg = lib().pcgsq_last_error().decode(errors="replace")
        raise (Unsupported if rc == EUNSUPPORTED else PcgsqError)(rc, msg)


def _p(a: np.ndarray, t=u64p):
    return a.ctypes.data_as(t)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def _blocks(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 2)


def _out_array(a, n: int, name: str) -> np.ndarray:
    """An output array the C side writes n uint64 words into: refuse anything
    else (a short, strided or non-uint64 array would be written past)."""
    if not isinstance(a, np.ndarray) or a.dtype != np.uint64 or a.size != n \
            or not a.flags.c_contiguous or not a.flags.writeable:
        raise ValueError(f"{name}: expected a writable C-contiguous uint64 array of {n} "
                         f"elements")
    return a


def _dpf_inputs(roots, dltsft, n: int, w: int):
    """roots [B] and dltsft [B][n/w][2^w - 1] blocks, checked against each
    other (the C side reads B*(n/w)*(2^w-1) DltSft blocks)."""
    roots, dltsft = _blocks(roots), _blocks(dltsft)
    if w < 1 or w > 4 or n % w:
        raise ValueError(f"dpf: w must be in [1, 4] and divide n (n={n}, w={w})")
    want = len(roots) * (n // w) * ((1 << w) - 1)
    if len(dltsft) != want:
        raise ValueError(f"dpf: dltsft has {len(dltsft)} blocks, expected B*(n/w)*(2^w-1) = "
                         f"{want}")
    return roots, dltsft


def _round_keys(round_keys, w: int, prg: int):
    """AES tree round keys, nk*176 bytes (None for ChaCha8)."""
    if round_keys is None:
        return None
    rk = np.ascontiguousarray(round_keys, dtype=np.uint8).reshape(-1)
    if prg == PRG_AES and rk.size != ((1 << w) - 1) * 176:
        raise ValueError(f"dpf: AES round keys must be (2^w - 1) * 176 = {((1 << w) - 1) * 176} "
                         f"bytes, got {rk.size}")
    return rk


def device_available() -> bool:
    """pcg_cuda::is_cuda_available (GPU_baseline/common/poly_mul_cuda.h:9)."""
    return bool(lib().pcgsq_device_available())


def dealer_u64(seed: int, domain: int, a: int, b: int) -> int:
    return lib().pcgsq_dealer_u64(seed, domain, a, b)


class Context:
    """One device + stream + owned scratch (replaces the reference's
    process-global grow-only pools, dpf_gpu.cu:42-75)."""

    def __init__(self, device: int = 0, apply_tuning: bool = True):
        h = vp()
        _check(lib().pcgsq_ctx_create(device, C.byref(h)))
        self.h = h
        if apply_tuning:
            from . import autotune
            autotune.apply(self)

    def close(self) -> None:
        if self.h:
            lib().pcgsq_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().pcgsq_ctx_stream(self.h)

    def synchronize(self) -> None:
        _check(lib().pcgsq_ctx_synchronize(self.h))

    def profile(self, enable: bool = True) -> None:
        """Start (and reset) or stop per-launch device timing."""
        _check(lib().pcgsq_ctx_profile(self.h, 1 if enable else 0))

    def profile_read(self) -> dict:
        """{zone: (launches, total_ms)} of the launches since profile(True)."""
        import json
        buf = C.create_string_buffer(1 << 16)
        _check(lib().pcgsq_ctx_profile_read(self.h, buf, 1 << 16))
        return {k: (v[0], v[1]) for k, v in json.loads(buf.value.decode()).items()}

    def probe_chacha8(self, iters: int = 2048) -> float:
        """Register-only ChaCha8 blocks/s of this device."""
        v = C.c_double()
        _check(lib().pcgsq_probe_chacha8(self.h, iters, C.byref(v)))
        return v.value

    def probe_copy(self, nbytes: int = 1 << 30) -> float:
        """Achievable HBM bytes/s of a plain copy (read + write)."""
        v = C.c_double()
        _check(lib().pcgsq_probe_copy(self.h, nbytes, C.byref(v)))
        return v.value

    def probe_hbm(self, nbytes: int = 1 << 30) -> dict:
        """Achievable HBM rates: copy (read + write counted, best probe shape),
        read-only and write-only, in bytes/s, and the copy's best shape."""
        cp, rd, wr = C.c_double(), C.c_double(), C.c_double()
        shape = (C.c_int * 3)()
        _check(lib().pcgsq_probe_hbm(self.h, nbytes, C.byref(cp), C.byref(rd), C.byref(wr),
                                     shape))
        return {"copy": cp.value, "read": rd.value, "write": wr.value,
                "copy_shape": {"loads_per_lane": shape[0], "nontemporal": bool(shape[1]),
                               "wg_per_cu": shape[2]}}

    def field_op(self, op: int, a, b, c=None) -> np.ndarray:
        """Device F_p arithmetic elementwise (op 0 mulhi64, 1 the butterflies'
        Shoup multiply, 2 the general Shoup multiply, 3 a*b mod P)."""
        a, b = _u64(a), _u64(b)
        cc = None if c is None else _u64(c)
        out = np.empty_like(a)
        _check(lib().pcgsq_field_op(self.h, op, len(a), _p(a), _p(b),
                                    None if cc is None else _p(cc), _p(out)))
        return out

    def to_device(self, arr: np.ndarray) -> "DeviceBuffer":
        """Copy a host array into a new device buffer owned by this context."""
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        _check(lib().pcgsq_memcpy_h2d(self.h, buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        _check(lib().pcgsq_device_name(buf, 256))
        return buf.value.decode()

    # ---- DPF (GPU_baseline/common/dpf_gpu.h:55-75) -------------------------
    def dpf_batch_full_eval(self, roots, dltsft, n, w, exchange, party=0, prg=PRG_CHACHA8,
                            round_keys=None):
        """One party's batched full-domain half-tree DPF with a host exchange
        callable(level, mine: ndarray[(count,2)]) -> peer ndarray."""
        roots, dltsft = _dpf_inputs(roots, dltsft, n, w)
        B = len(roots)
        out = np.empty((B << n, 2), dtype=np.uint64)
        cb, err = self._exchange(exchange)
        rk = _round_keys(round_keys, w, prg)
        rc = lib().pcgsq_dpf_batch_full_eval(
            self.h, B, n, w, roots.ctypes.data, dltsft.ctypes.data,
            None if rk is None else rk.ctypes.data, party, cb, None, out.ctypes.data, prg)
        if err:
            raise err[0]
        _check(rc)
        return out

    def dpf_batch_full_eval_device(self, roots, dltsft, n, w, exchange, party=0,
                                   prg=PRG_CHACHA8, round_keys=None) -> int:
        roots, dltsft = _dpf_inputs(roots, dltsft, n, w)
        cb, err = self._exchange(exchange)
        rk = _round_keys(round_keys, w, prg)
        d = vp()
        rc = lib().pcgsq_dpf_batch_full_eval_device(
            self.h, len(roots), n, w, roots.ctypes.data, dltsft.ctypes.data,
            None if rk is None else rk.ctypes.data, party, cb, None, C.byref(d), prg)
        if err:
            raise err[0]
        _check(rc)
        return d.value

    def dpf_full_eval_both(self, roots0, roots1, dlt0, dlt1, n, w, prg=PRG_CHACHA8,
                           round_keys=None, want_cw=False):
        roots0, dlt0 = _dpf_inputs(roots0, dlt0, n, w)
        roots1, dlt1 = _dpf_inputs(roots1, dlt1, n, w)
        if len(roots1) != len(roots0):
            raise ValueError("dpf: the two parties' batches differ in size")
        B = len(roots0)
        l0 = np.empty((B << n, 2), dtype=np.uint64)
        l1 = np.empty((B << n, 2), dtype=np.uint64)
        nk = (1 << w) - 1
        cw = np.empty(((n // w) * B * nk, 2), dtype=np.uint64) if want_cw else None
        rk = _round_keys(round_keys, w, prg)
        _check(lib().pcgsq_dpf_batch_full_eval_both(
            self.h, B, n, w, roots0.ctypes.data, roots1.ctypes.data, dlt0.ctypes.data,
            dlt1.ctypes.data, None if rk is None else rk.ctypes.data, l0.ctypes.data,
            l1.ctypes.data, cw.ctypes.data if want_cw else None, prg))
        return (l0, l1, cw) if want_cw else (l0, l1)

    def _exchange(self, fn):
        """A C trampoline for the exchange callable, and the list that
        receives the exception the callable raised (the caller re-raises it
        after the C call, so the callable's own error reaches the user)."""
        err = []

        def tramp(user, level, mine, peer, count):
            try:
                m = np.ctypeslib.as_array(C.cast(mine, u64p), shape=(count * 2,)).reshape(-1, 2)
                res = _blocks(fn(level, m.copy()))
                if res.shape != (count, 2):
                    raise ValueError(f"exchange returned {res.shape[0]} blocks, expected {count}")
                C.memmove(peer, res.ctypes.data, count * 16)
                return 0
            except BaseException as e:
                err.append(e)
                return 1
        self._cb = EXCHANGE_FN(tramp)   # keep alive for the call
        return self._cb, err

    # ---- output layer (GPU_baseline/common/leaf_convert_cuda.h:43-52) -------
    def dpf_out_sums(self, d_leaves: int, B: int, D: int):
        sc = np.empty(B, dtype=np.uint64)
        st = np.empty(B, dtype=np.uint64)
        _check(lib().pcgsq_dpf_out_sums(self.h, d_leaves, B, D, P, _p(sc), _p(st)))
        return sc, st

    def dpf_out_sums_noalloc(self, d_leaves: int, B: int, D: int, d_sumC: int,
                             d_sumT: int) -> None:
        """dpf_out_sums into device arrays d_sumC[B], d_sumT[B], enqueued on
        the context stream with no device allocation (partials in the
        library's fixed per-GPU scratch array)."""
        _check(lib().pcgsq_dpf_out_sums_noalloc(self.h, d_leaves, B, D, P, d_sumC, d_sumT))

    def dpf_out_scatter_g(self, d_leaves: int, B: int, D: int, t: int, party: int, CW, N: int,
                          out=None):
        """g [N] on the host; `out` (optional) is a caller-owned uint64 array of N
        words that receives it (reused across calls, it saves the page faults of
        a fresh array)."""
        cw = _u64(CW).reshape(-1)
        if cw.size != B:
            raise ValueError(f"dpf_out_scatter_g: CW has {cw.size} words, expected B = {B}")
        g = np.empty(N, dtype=np.uint64) if out is None else _out_array(out, N,
                                                                         "dpf_out_scatter_g")
        _check(lib().pcgsq_dpf_out_scatter_g(self.h, d_leaves, B, D, t, party, P, _p(cw), _p(g),
                                             N))
        return g

    # ---- poly-mul / NTT (poly_mul_cuda.h, CPU_baseline/ntt.h) --------------
    def poly_mul(self, a, b, backend=NTT_FOURSTEP):
        a, b = _u64(a), _u64(b)
        if a.shape != b.shape or a.ndim == 0:
            raise ValueError(f"poly_mul: operand shapes differ ({a.shape} vs {b.shape})")
        shape = a.shape
        a2, b2 = a.reshape(-1, shape[-1]), b.reshape(-1, shape[-1])
        out = np.empty_like(a2)
        ms = C.c_double()
        _check(lib().pcgsq_poly_mul_u64(self.h, _p(out), _p(a2), _p(b2), a2.shape[0],
                                        a2.shape[1], P, backend, C.byref(ms)))
        self.last_device_ms = ms.value
        return out.reshape(shape)

    def ntt_forward(self, a, backend=NTT_FOURSTEP):
        x = _u64(a).copy()
        x2 = x.reshape(-1, x.shape[-1])
        _check(lib().pcgsq_ntt_forward(self.h, _p(x2), x2.shape[0], x2.shape[1], backend))
        return x

    def ntt_inverse(self, a, backend=NTT_FOURSTEP):
        x = _u64(a).copy()
        x2 = x.reshape(-1, x.shape[-1])
        _check(lib().pcgsq_ntt_inverse(self.h, _p(x2), x2.shape[0], x2.shape[1], backend))
        return x

    def ntt_tune(self, N: int, n1_log: int = 0, variant: int = 0, cols1: int = 0,
                 cols2: int = 0) -> None:
        """Shape of the NTT passes for length N (results do not change): split
        n1 = 2^n1_log (log2 N = one pass), register-grouping variant bits and
        columns per tile of the n1 / n2 passes (0 = automatic)."""
        _check(lib().pcgsq_ntt_tune(self.h, N, n1_log, variant, cols1, cols2))

    def poly_mul_dev(self, out_ptr: int, a_ptr: int, b_ptr: int, batch: int, N: int,
                     backend=NTT_FOURSTEP):
        _check(lib().pcgsq_poly_mul_u64_dev(self.h, out_ptr, a_ptr, b_ptr, batch, N, backend))

    def ntt_forward_dev(self, ptr: int, batch: int, N: int, backend=NTT_FOURSTEP):
        _check(lib().pcgsq_ntt_forward_dev(self.h, ptr, batch, N, backend))

    def ntt_inverse_dev(self, ptr: int, batch: int, N: int, backend=NTT_FOURSTEP):
        _check(lib().pcgsq_ntt_inverse_dev(self.h, ptr, batch, N, backend))

    # ---- AES (aes_gpu.h) ------------------------------------------------------
    def aes128_ecb_encrypt(self, keys: bytes, pt: bytes) -> bytes:
        n = len(pt) // 16
        k = (C.c_uint8 * len(keys)).from_buffer_copy(keys)
        i = (C.c_uint8 * len(pt)).from_buffer_copy(pt)
        o = (C.c_uint8 * len(pt))()
        _check(lib().pcgsq_aes128_ecb_encrypt(self.h, k, i, o, n))
        return bytes(o)


class DeviceBuffer:
    """Raw device memory on a context's GPU (freed on close/GC)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        h = vp()
        _check(lib().pcgsq_dev_alloc(ctx.h, nbytes, C.byref(h)))
        self.ptr = h.value

    def to_host(self, dtype=np.uint64) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _check(lib().pcgsq_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def close(self):
        if self.ptr and self.ctx.h:
            lib().pcgsq_dev_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pinned_array(n: int, dtype=np.uint64) -> np.ndarray:
    """A numpy array of n elements in page-locked host memory (freed when the
    array is collected): device-to-host copies into it run by DMA."""
    import weakref
    nbytes = max(1, int(n) * np.dtype(dtype).itemsize)
    h = vp()
    _check(lib().pcgsq_host_alloc(nbytes, C.byref(h)))
    buf = (C.c_uint8 * nbytes).from_address(h.value)
    arr = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(dtype)[:n]
    weakref.finalize(buf, lib().pcgsq_host_free, h.value)
    return arr


def poly_mul_supported(N: int, backend=NTT_FOURSTEP, prime: int = P) -> tuple[bool, str]:
    buf = C.create_string_buffer(256)
    ok = lib().pcgsq_poly_mul_supported(prime, N, backend, buf, 256)
    return bool(ok), buf.value.decode()


def aes128_expand_key(key: bytes) -> bytes:
    k = (C.c_uint8 * 16).from_buffer_copy(key)
    rk = (C.c_uint8 * 176)()
    _check(lib().pcgsq_aes128_expand_key(k, rk))
    return bytes(rk)


def aes_tree_round_keys(nk: int, master: bytes = bytes.fromhex(
        "2b7e151628aed2a6abf7158809cf4f3c")) -> np.ndarray:
    m = (C.c_uint8 * 16).from_buffer_copy(master)
    rks = np.zeros(nk * 176, dtype=np.uint8)
    _check(lib().pcgsq_aes_tree_round_keys(m, nk, _p(rks, u8p)))
    return rks


class OlePlan:
    """PCG_OLE<P>::gen_and_expand for both parties on one GPU
    (CPU_baseline/pcg_ole_2pc/pcg_ole_impl.h:46-490)."""

    def __init__(self, ctx: Context, N: int, c: int, t: int, w: int = 1, prg: int = PRG_CHACHA8,
                 pub_seed: int = 0, hbm_budget: int = 0):
        self.ctx = ctx
        self.N, self.c, self.t, self.w, self.prg = N, c, t, w, prg
        p = Params(N, c, t, w, prg)
        h = vp()
        _check(lib().pcgsq_ole_plan_create(ctx.h, C.byref(p), pub_seed, hbm_budget, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().pcgsq_ole_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return lib().pcgsq_ole_plan_bytes(self.h)

    def deal(self, seed: int, gmw: bool = False) -> None:
        """Deal the inputs of `seed`; gmw=True runs the two parties' 2PC
        preprocessing in-process (base OTs, IKNP random OTs, the position
        adder and DltSft circuits under GMW, Gilboa products for beta and the
        Beaver triples, every message a device copy) instead of taking
        DltSft, beta and rs from the dealer (same outputs)."""
        _check((lib().pcgsq_ole_deal_gmw if gmw else lib().pcgsq_ole_deal)(self.h, seed))

    def dealt_dltsft(self) -> np.ndarray:
        """Both parties' dealt DltSft shares, [2][c*c*t*t][levels][2^w - 1][2]."""
        n = (2 * self.N // self.t).bit_length() - 1
        levels, nk = n // self.w, (1 << self.w) - 1
        out = np.empty((2, self.c * self.c * self.t * self.t, levels, nk, 2), dtype=np.uint64)
        _check(lib().pcgsq_ole_dealt_dltsft(self.h, out.ctypes.data))
        return out

    def dealt_products(self) -> dict:
        """Both parties' dealt F_p shares: beta, r, s, rs, each [2][c*c*t*t]."""
        out = np.empty((4, 2, self.c * self.c * self.t * self.t), dtype=np.uint64)
        _check(lib().pcgsq_ole_dealt_products(self.h, out.ctypes.data))
        return dict(zip(("beta", "r", "s", "rs"), out))

    PRE_ITEMS = {"choice": (0, np.uint64), "received": (1, np.uint64), "sent0": (2, np.uint64),
                 "sent1": (3, np.uint64), "adder_triples": (4, np.uint64),
                 "delta_triples": (5, np.uint64), "tree_triples": (6, np.uint32)}

    def preprocessing(self, party: int, item: str) -> np.ndarray:
        """What `party`'s last 2PC preprocessing left on the device (for
        checking the protocol): "choice" words of the OTs it received,
        "received" / "sent0" / "sent1" messages ([m][2] words), its
        "adder_triples" [ninst][ops][3], "delta_triples" [ninst][L*nk][5],
        "tree_triples" [ninst][L*nk] (uint32)."""
        what, dt = self.PRE_ITEMS[item]
        n = C.c_size_t()
        _check(lib().pcgsq_ole_preprocessing_read(self.h, party, what, None, 0, C.byref(n)))
        out = np.empty(n.value // np.dtype(dt).itemsize, dtype=dt)
        _check(lib().pcgsq_ole_preprocessing_read(self.h, party, what, out.ctypes.data,
                                                  out.nbytes, C.byref(n)))
        return out

    def run(self) -> None:
        _check(lib().pcgsq_ole_run(self.h))

    def run_blocks(self) -> None:
        _check(lib().pcgsq_ole_run_blocks(self.h))

    def outputs(self, out=None):
        """(x0, z0, x1, z1) in new arrays, or copied into `out` (four uint64
        arrays of N, e.g. from pinned_array) and returned."""
        if out is None:
            out = tuple(np.empty(self.N, dtype=np.uint64) for _ in range(4))
        x0, z0, x1, z1 = (_out_array(a, self.N, "OlePlan.outputs") for a in out)
        _check(lib().pcgsq_ole_outputs(self.h, _p(x0), _p(z0), _p(x1), _p(z1)))
        return x0, z0, x1, z1

    def g(self) -> np.ndarray:
        out = np.empty((2, self.c * self.c, self.N), dtype=np.uint64)
        _check(lib().pcgsq_ole_g(self.h, _p(out)))
        return out

    def outputs_dev(self) -> int:
        return lib().pcgsq_ole_outputs_dev(self.h)

    def triples(self, seed1: int, seed2: int):
        """generate_triples (triple.h): (a0, b0, c0, a1, b1, c1), N Z_p triple
        shares per party from the expansions of seed1 and seed2."""
        out = [np.empty(self.N, dtype=np.uint64) for _ in range(6)]
        a0, b0, c0, a1, b1, c1 = out
        _check(lib().pcgsq_triples_generate(self.h, seed1, seed2, _p(a0), _p(b0), _p(c0),
                                            _p(a1), _p(b1), _p(c1)))
        return a0, b0, c0, a1, b1, c1


def _ole_trampoline(exchange, err):
    """A C exchange callback calling exchange(phase, level, batch, send,
    nrecv) -> the nrecv bytes received; the callable's exception is kept in
    `err` for the caller to re-raise after the C call."""
    def tramp(user, phase, level, batch, send, send_bytes, recv, recv_bytes):
        try:
            m = (np.ctypeslib.as_array((C.c_uint8 * send_bytes).from_address(send)).copy()
                 if send_bytes else np.empty(0, dtype=np.uint8))
            got = exchange(phase.decode(), level, batch, m, recv_bytes)
            got = (np.empty(0, dtype=np.uint8) if got is None else
                   np.ascontiguousarray(got, dtype=np.uint8).reshape(-1))
            if got.size != recv_bytes:
                raise ValueError(f"exchange returned {got.size} bytes, expected {recv_bytes}")
            if recv_bytes:
                C.memmove(recv, got.ctypes.data, recv_bytes)
            return 0
        except BaseException as e:  # re-raised after the run
            err.append(e)
            return 1
    return OLE_EXCHANGE_FN(tramp)


class OleParty(OlePlan):
    """One party of the 2PC expansion, as one party's process runs it
    (pcgsq_ole_party_*; pcg_ole_impl.h:46-490 with emp::NetIO replaced by a
    callable).  run(exchange) calls exchange(phase, level, batch, send, nrecv)
    for every message of the protocol, in order ("base_ot", "ot_ext",
    "gilboa", "circuit_and" when dealt with deal_own; "dpf_gen" per DPF
    level and block, "beaver_de" and "beaver_cw" per block): send is a uint8
    array for the other party, and the callable returns the nrecv bytes the
    other party sent (party 0 sends first).

    deal_own(own_seed) draws only this party's inputs, and the run computes
    the shared ones with the other party (OT extension, GMW, Gilboa); deal(seed)
    is the trusted-dealer simulation (both parties' shares from one seed)."""

    def __init__(self, ctx: Context, N: int, c: int, t: int, party: int, w: int = 1,
                 prg: int = PRG_CHACHA8, pub_seed: int = 0, hbm_budget: int = 0):
        self.ctx = ctx
        self.N, self.c, self.t, self.w, self.prg, self.party = N, c, t, w, prg, party
        p = Params(N, c, t, w, prg)
        h = vp()
        _check(lib().pcgsq_ole_party_create(ctx.h, C.byref(p), party, pub_seed, hbm_budget,
                                            C.byref(h)))
        self.h = h

    def deal_own(self, own_seed: int) -> None:
        """This party's own draws from its own seed; the next run starts with
        the 2PC preprocessing (pcg_ole_impl.h:62-199) through the exchange."""
        _check(lib().pcgsq_ole_party_deal_own(self.h, own_seed))

    def set_exchange_mode(self, mode: int) -> None:
        """EXCHANGE_PER_BLOCK (the reference's serial arm) or EXCHANGE_BATCHED
        (one exchange per DPF level, Beaver open and circuit depth for a whole
        group of blocks, its batched arm)."""
        _check(lib().pcgsq_ole_party_set_exchange_mode(self.h, mode))

    def run(self, exchange) -> None:
        err = []
        cb = _ole_trampoline(exchange, err)
        rc = lib().pcgsq_ole_party_run(self.h, cb, None)
        if err:
            raise err[0]
        _check(rc)

    def outputs(self, out=None):
        """(x, z) of this party."""
        if out is None:
            out = tuple(np.empty(self.N, dtype=np.uint64) for _ in range(2))
        x, z = (_out_array(a, self.N, "OleParty.outputs") for a in out)
        _check(lib().pcgsq_ole_party_outputs(self.h, _p(x), _p(z)))
        return x, z


def base_ots_run(party: int, seed: int, exchange) -> BaseOts:
    """One party's base OTs through exchange(phase, level, batch, send,
    nrecv) (no device needed)."""
    err = []
    cb = _ole_trampoline(exchange, err)
    out = BaseOts()
    rc = lib().pcgsq_base_ots_run(party, seed, cb, None, C.byref(out))
    if err:
        raise err[0]
    _check(rc)
    return out


def dist_exchange(peer: int, group=None):
    """An OleParty exchange over torch.distributed point-to-point messages,
    for one party per process as the reference's 2PC runs: the pair of
    emp::NetIO calls of pcg_ole_impl.h:247-260 / dpf.cpp:539-551, the lower
    rank sending first.  The messages are CPU uint8 tensors, so the process
    group is gloo (TCP between the parties' hosts, or between two processes
    of one host)."""
    import torch
    import torch.distributed as dist

    me = dist.get_rank(group)

    def exchange(phase, level, batch, send, nrecv):
        out = torch.from_numpy(np.ascontiguousarray(send, dtype=np.uint8))
        got = torch.empty(nrecv, dtype=torch.uint8)
        if me < peer:
            if out.numel():
                dist.send(out, peer, group=group)
            if nrecv:
                dist.recv(got, peer, group=group)
        else:
            if nrecv:
                dist.recv(got, peer, group=group)
            if out.numel():
                dist.send(out, peer, group=group)
        return got.numpy()

    return exchange


class CommTrace:
    """Rows of the reference's comm_trace.h CSV (party, context, phase, level,
    batch, send_bytes, recv_bytes, t_begin_ns, t_mid_ns, t_end_ns,
    io_counter_delta) for the exchanges of an OleParty run: wrap(exchange)
    returns a traced exchange."""

    COLUMNS = ("party,context,phase,level,batch,send_bytes,recv_bytes,"
               "t_begin_ns,t_mid_ns,t_end_ns,io_counter_delta")

    def __init__(self, party: int, context: str = ""):
        self.party, self.context, self.rows = party, context, []

    def wrap(self, exchange):
        import time as _t

        def traced(phase, level, batch, send, nrecv):
            t0 = _t.perf_counter_ns()
            peer = exchange(phase, level, batch, send, nrecv)
            t1 = _t.perf_counter_ns()
            n = int(np.asarray(send).nbytes)
            self.rows.append((self.party, self.context, phase, level, batch, n, int(nrecv),
                              t0, t0, t1, n + int(nrecv)))
            return peer
        return traced

    def to_csv(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.COLUMNS + "\n")
            for r in self.rows:
                f.write(",".join(str(v) for v in r) + "\n")


NTT_BATCH_FORWARD, NTT_BATCH_INVERSE = 1, 2
NTT_BATCH_PERSISTENT, NTT_BATCH_GROUPED = 0, 1


class NttBatch:
    """Transforms of uneven lengths (2^10..2^22) in one persistent launch, or
    grouped by length into per-length launches (pcgsq_ntt_batch_*).
    Transform i occupies words offsets[i] .. offsets[i] + 2^logN[i] of the
    concatenated buffers."""

    def __init__(self, ctx: Context, logN, ops: int = NTT_BATCH_FORWARD | NTT_BATCH_INVERSE):
        self.ctx = ctx
        self.logN = [int(x) for x in logN]
        self.ops = ops
        arr = (C.c_int * len(self.logN))(*self.logN)
        h = vp()
        _check(lib().pcgsq_ntt_batch_create(ctx.h, arr, len(self.logN), ops, C.byref(h)))
        self.h = h
        self.offsets = np.concatenate([[0], np.cumsum([1 << x for x in self.logN])]).astype(np.int64)
        self.total = int(self.offsets[-1])

    def close(self):
        if self.h:
            lib().pcgsq_ntt_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, logN: int = 0, n1_log: int = 0, wgs_per_cu: int = 0, spin_limit: int = 0):
        _check(lib().pcgsq_ntt_batch_tune(self.h, logN, n1_log, wgs_per_cu, spin_limit))

    def set_mode(self, mode: int) -> None:
        """NTT_BATCH_PERSISTENT (one persistent launch) or NTT_BATCH_GROUPED
        (per-length pass launches over offset tables); results are equal."""
        _check(lib().pcgsq_ntt_batch_set_mode(self.h, mode))

    def load(self, concat: np.ndarray) -> None:
        a = _u64(concat)
        assert a.size == self.total
        _check(lib().pcgsq_ntt_batch_load(self.h, _p(a)))

    def run(self, graph: bool = False) -> None:
        _check(lib().pcgsq_ntt_batch_run(self.h, int(graph)))

    def finish(self) -> None:
        _check(lib().pcgsq_ntt_batch_finish(self.h))

    def read(self):
        f = np.empty(self.total, dtype=np.uint64)
        r = np.empty(self.total, dtype=np.uint64)
        _check(lib().pcgsq_ntt_batch_read(self.h, _p(f), _p(r)))
        return f, r

    def dev(self):
        a, f, r, n = vp(), vp(), vp(), C.c_size_t()
        _check(lib().pcgsq_ntt_batch_dev(self.h, C.byref(a), C.byref(f), C.byref(r), C.byref(n)))
        return a.value, f.value, r.value, n.value

    def info(self) -> dict:
        n, g, lds, b = C.c_int(), C.c_int(), C.c_size_t(), C.c_double()
        _check(lib().pcgsq_ntt_batch_info(self.h, C.byref(n), C.byref(g), C.byref(lds),
                                          C.byref(b)))
        return {"items": n.value, "grid": g.value, "lds_bytes": lds.value,
                "alg_bytes": b.value}

    def autotune(self, reps: int = 3, log=None) -> dict:
        """Coordinate search over the split of every length in the batch and
        the resident workgroups per CU of the persistent launch, then the
        grouped-by-length mode, timing whole replays of the captured launch
        on the loaded inputs; keeps the fastest.  Shapes and modes never
        change results."""
        def timed():
            self.run(graph=True)
            self.finish()
            self.ctx.profile(True)
            for _ in range(reps):
                self.run(graph=True)
            self.finish()
            z = self.ctx.profile_read()
            self.ctx.profile(False)
            return z["ntt.batch"][1] / z["ntt.batch"][0]

        self.set_mode(NTT_BATCH_PERSISTENT)
        best = {"wgs_per_cu": 0, "splits": {}, "ms": timed()}
        for L in sorted(set(self.logN), reverse=True):
            cands = [s for s in range(7, 12) if 6 <= L - s <= 11] + ([L] if L <= 12 else [])
            keep = 0
            for sp in cands:
                self.tune(L, sp, best["wgs_per_cu"])
                ms = timed()
                if log:
                    log(f"batch 2^{L} split {sp}: {ms:.3f} ms")
                if ms < best["ms"]:
                    best["ms"], keep = ms, sp
            self.tune(L, keep, best["wgs_per_cu"])
            if keep:
                best["splits"][L] = keep
        for w in (1, 2, 3, 4):
            self.tune(0, 0, w)
            ms = timed()
            if log:
                log(f"batch wgs_per_cu {w}: {ms:.3f} ms")
            if ms < best["ms"]:
                best["ms"], best["wgs_per_cu"] = ms, w
        self.tune(0, 0, best["wgs_per_cu"])
        best["mode"] = "persistent"
        best["persistent_ms"] = best["ms"]
        self.set_mode(NTT_BATCH_GROUPED)
        ms = timed()
        best["grouped_ms"] = ms
        if log:
            log(f"batch grouped by length: {ms:.3f} ms")
        if ms < best["ms"]:
            best["ms"], best["mode"] = ms, "grouped"
        else:
            self.set_mode(NTT_BATCH_PERSISTENT)
        return best

    def split(self, arr: np.ndarray):
        """Per-transform views of a concatenated array."""
        return [arr[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.logN))]


class DpfBatch:
    """DPF key pairs of uneven domains (2^10..2^30, w = 1): dealer Gen and
    both parties' converted full domains in one persistent launch
    (pcgsq_dpf_batch_*)."""

    def __init__(self, ctx: Context, logn, prg: int = PRG_CHACHA8):
        self.ctx = ctx
        self.logn = [int(x) for x in logn]
        self.prg = prg
        arr = (C.c_int * len(self.logn))(*self.logn)
        h = vp()
        _check(lib().pcgsq_dpf_batch_create(ctx.h, arr, len(self.logn), prg, C.byref(h)))
        self.h = h
        self.offsets = np.concatenate([[0], np.cumsum([1 << x for x in self.logn])]).astype(np.int64)
        self.total = int(self.offsets[-1])
        self.cw_offsets = np.concatenate([[0], np.cumsum(self.logn)]).astype(np.int64)

    def close(self):
        if self.h:
            lib().pcgsq_dpf_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def deal(self, seed: int) -> None:
        _check(lib().pcgsq_dpf_batch_deal(self.h, seed))

    def gen(self, alpha, beta, delta, delta0) -> None:
        a, b = _u64(alpha), _u64(beta)
        d, d0 = _blocks(delta), _blocks(delta0)
        _check(lib().pcgsq_dpf_batch_gen(self.h, _p(a), _p(b), d.ctypes.data, d0.ctypes.data))

    def points(self):
        a = np.empty(len(self.logn), dtype=np.uint64)
        b = np.empty(len(self.logn), dtype=np.uint64)
        _check(lib().pcgsq_dpf_batch_points(self.h, _p(a), _p(b)))
        return a, b

    def keys(self):
        K = len(self.logn)
        roots = np.empty((2, K, 2), dtype=np.uint64)
        cw = np.empty((int(self.cw_offsets[-1]), 2), dtype=np.uint64)
        co = np.empty(K, dtype=np.uint64)
        _check(lib().pcgsq_dpf_batch_keys(self.h, roots.ctypes.data, cw.ctypes.data, _p(co)))
        return roots, cw, co

    def run(self, graph: bool = False) -> None:
        _check(lib().pcgsq_dpf_batch_run(self.h, int(graph)))

    def read(self):
        y0 = np.empty(self.total, dtype=np.uint64)
        y1 = np.empty(self.total, dtype=np.uint64)
        _check(lib().pcgsq_dpf_batch_read(self.h, _p(y0), _p(y1)))
        return y0, y1

    def dev(self):
        y, n = vp(), C.c_size_t()
        _check(lib().pcgsq_dpf_batch_dev(self.h, C.byref(y), C.byref(n)))
        return y.value, n.value

    def info(self) -> dict:
        n, g, tb, ob = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        _check(lib().pcgsq_dpf_batch_info(self.h, C.byref(n), C.byref(g), C.byref(tb),
                                          C.byref(ob)))
        return {"items": n.value, "grid": g.value, "tree_blocks": tb.value,
                "out_blocks": ob.value}

    def key_slice(self, arr, i):
        return arr[self.offsets[i]:self.offsets[i + 1]]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = range(4)


class Server:
    """Continuous-batching server of PCG-OLE expansion requests on one GPU
    (pcgsq_server_*): `slots` concurrent execution slots, admission by
    device-memory budget, outputs kept on the device until read/released."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, slots: int = 4, window: int = 0,
                 exclusive_N: int = 0, age_bound_s: float = 0.0):
        h = vp()
        _check(lib().pcgsq_server_create(device, hbm_budget, slots, C.byref(h)))
        self.h = h
        if window:
            self.set_window(window)
        if exclusive_N or age_bound_s:
            self.set_policy(exclusive_N, age_bound_s)

    def set_policy(self, exclusive_N: int = 0, age_bound_s: float = 0.0) -> None:
        """Requests with N >= exclusive_N run one at a time, in arrival order
        among themselves; a request waiting longer than age_bound_s is passed
        by no later one.  0 switches either off."""
        _check(lib().pcgsq_server_set_policy(self.h, exclusive_N, age_bound_s))

    def set_window(self, max_running_N: int) -> None:
        """Most ring dimension (sum of N) running at once; 0 = no limit.  A
        request that would exceed it waits (unless nothing runs), and so does
        every later request at least as large."""
        _check(lib().pcgsq_server_set_window(self.h, max_running_N))

    def close(self):
        if self.h:
            _check(lib().pcgsq_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, N: int, c: int, t: int, seed: int, pub_seed: int = 0, w: int = 1,
               prg: int = PRG_CHACHA8) -> int:
        p = Params(N, c, t, w, prg)
        rid = C.c_uint64()
        _check(lib().pcgsq_server_submit(self.h, C.byref(p), seed, pub_seed, C.byref(rid)))
        return rid.value

    def poll(self, rid: int) -> int:
        st = C.c_int()
        _check(lib().pcgsq_server_poll(self.h, rid, C.byref(st)))
        return st.value

    def wait(self, rid: int, timeout: float = 0.0) -> None:
        _check(lib().pcgsq_server_wait(self.h, rid, timeout))

    def outputs(self, rid: int, N: int):
        x0, z0, x1, z1 = (np.empty(N, dtype=np.uint64) for _ in range(4))
        _check(lib().pcgsq_server_outputs(self.h, rid, _p(x0), _p(z0), _p(x1), _p(z1)))
        return x0, z0, x1, z1

    def outputs_dev(self, rid: int) -> int:
        p = vp()
        _check(lib().pcgsq_server_outputs_dev(self.h, rid, C.byref(p)))
        return p.value

    def release(self, rid: int) -> None:
        _check(lib().pcgsq_server_release(self.h, rid))

    def timing(self, rid: int):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().pcgsq_server_timing(self.h, rid, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def stats(self) -> dict:
        b, u, pk, done, q = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_long(), C.c_int()
        _check(lib().pcgsq_server_stats(self.h, C.byref(b), C.byref(u), C.byref(pk),
                                        C.byref(done), C.byref(q)))
        return {"budget": b.value, "used": u.value, "peak": pk.value, "completed": done.value,
                "queued": q.value}

    def plans(self) -> dict:
        """Plans in the pool, plans built and plans evicted so far."""
        n, built, ev = C.c_int(), C.c_long(), C.c_long()
        _check(lib().pcgsq_server_plans(self.h, C.byref(n), C.byref(built), C.byref(ev)))
        return {"plans": n.value, "built": built.value, "evicted": ev.value}

    def reserve(self, N: int, c: int, t: int, count: int = 1, pub_seed: int = 0, w: int = 1,
                prg: int = PRG_CHACHA8) -> None:
        """Create plans of this shape ahead of requests (until `count` exist);
        requests of one shape overlap only on plans reserved this way."""
        _check(lib().pcgsq_server_reserve(self.h, C.byref(Params(N, c, t, w, prg)), pub_seed,
                                          count))

    def charge(self, N: int, c: int, t: int, w: int = 1, prg: int = PRG_CHACHA8) -> int:
        """Device bytes this server charges such a request."""
        b = C.c_size_t()
        _check(lib().pcgsq_server_charge(self.h, C.byref(Params(N, c, t, w, prg)), C.byref(b)))
        return b.value

    @staticmethod
    def clock() -> float:
        return lib().pcgsq_server_clock()


def ole_plan_estimate(N: int, c: int, t: int, w: int = 1, prg: int = PRG_CHACHA8,
                      hbm_budget: int = 0) -> int:
    return lib().pcgsq_ole_plan_estimate(C.byref(Params(N, c, t, w, prg)), hbm_budget)


def ole_expand(ctx: Context, N, c, t, w=1, prg=PRG_CHACHA8, seed=1, pub_seed=0):
    with OlePlan(ctx, N, c, t, w, prg, pub_seed) as pl:
        pl.deal(seed)
        pl.run()
        return pl.outputs()
