"""The CPU oracle against the reference's golden vectors, published known
answers, the compiled reference headers (oracle/_ref) and the invariants the
mathematics fixes.  CPU only."""
import numpy as np
import pytest

import oracle as O

P = O.P


def test_field_constants(golden):
    f = golden["field"]
    assert P == int(f["P"])
    assert P == 2**62 - 3 * 2**25 + 1
    assert (P - 1) == 2**25 * 47189 * 2912521
    assert O.lib().orc_primitive_root() == f["primitive_root"]
    psi = O.ntt_psi(1 << 20)
    assert psi == int(f["psi_2_20"])
    assert pow(psi, 1 << 20, P) == P - 1          # psi^N = -1
    assert O.lib().orc_finv(12345) * 12345 % P == 1


def test_mt19937_64_standard_value():
    # The C++ standard fixes the 10000th output of a default-seeded engine.
    assert int(O.MT64(5489).take(10000)[-1]) == 9981545732273789042


def test_chacha8_kat(golden):
    ks = O.chacha8_block(bytes(32))
    assert ks.hex() == golden["chacha8_zero_key"]["keystream_hex_prefix"]


def test_aes_fips197(golden):
    for case in golden["aes128_fips197"]["cases"]:
        rk = O.aes128_expand_key(bytes.fromhex(case["key"]))
        assert O.aes128_encrypt(rk, bytes.fromhex(case["pt"])).hex() == case["ct"]


def test_aes_key_schedule_fips197_a1():
    # FIPS-197 A.1: last round key of 2b7e1516... is d014f9a8c9ee2589e13f0cc8b6630ca6
    rk = O.aes128_expand_key(O.FIPS197_MASTER)
    assert rk[160:176].hex() == "d014f9a8c9ee2589e13f0cc8b6630ca6"


@pytest.mark.parametrize("logN", [1, 2, 3, 5, 8])
def test_ntt_matches_schoolbook(logN):
    rng = np.random.default_rng(logN)
    N = 1 << logN
    a = rng.integers(0, P, N, dtype=np.uint64)
    b = rng.integers(0, P, N, dtype=np.uint64)
    assert np.array_equal(O.poly_mul(a, b), O.poly_mul_schoolbook(a, b))
    # forward is evaluation at psi^(2k+1) in natural order
    psi = O.ntt_psi(N)
    fa = O.ntt_forward(a)
    for k in (0, 1, N - 1):
        x = pow(psi, 2 * k + 1, P)
        assert int(fa[k]) == sum(int(a[i]) * pow(x, i, P) for i in range(N)) % P
    assert np.array_equal(O.ntt_inverse(fa), a)


def test_poly_mul_patterns():
    # zero / one / impulse / max patterns (GPU_baseline/test/poly_mul_cuda.cpp:56-71)
    N = 64
    rng = np.random.default_rng(3)
    a = rng.integers(0, P, N, dtype=np.uint64)
    one = np.zeros(N, np.uint64); one[0] = 1
    zero = np.zeros(N, np.uint64)
    assert np.array_equal(O.poly_mul(a, one), a)
    assert np.array_equal(O.poly_mul(a, zero), zero)
    imp = np.zeros(N, np.uint64); imp[N - 1] = 1       # X^(N-1): negacyclic shift
    got = O.poly_mul(one, imp)
    assert int(got[N - 1]) == 1
    x = np.zeros(N, np.uint64); x[1] = 1
    assert int(O.poly_mul(imp, x)[0]) == P - 1          # X^N = -1
    mx = np.full(N, P - 1, np.uint64)
    assert np.array_equal(O.poly_mul(mx, mx), O.poly_mul_schoolbook(mx, mx))


def test_ntt_naive_recipe_checksum(golden):
    N, batch = 1 << 20, golden["ntt_naive_fnv"]["batch"]
    m = O.MT64(11)
    a = m.take(batch * N, P).reshape(batch, N)
    b = m.take(batch * N, P).reshape(batch, N)
    c = O.poly_mul_batch(a, b, threads=8)
    assert "%016x" % O.fnv1a64(c) == golden["ntt_naive_fnv"]["by_logN"]["20"]


def test_ntt_4step_recipe_checksum(golden):
    N, batch = 1 << 20, golden["ntt_4step_xor"]["batch"]
    j = np.arange(batch * N, dtype=np.uint64)
    with np.errstate(over="ignore"):
        a = (np.uint64(0x9E3779B97F4A7C15) * (j + np.uint64(1))) % np.uint64(P)
        b = (np.uint64(0xC2B2AE3D27D4EB4F) * (j + np.uint64(7))) % np.uint64(P)
    c = O.poly_mul_batch(a.reshape(batch, N), b.reshape(batch, N), threads=8)
    assert "%016x" % int(np.bitwise_xor.reduce(c.ravel())) == \
        golden["ntt_4step_xor"]["by_logN"]["20"]


@pytest.mark.parametrize("case", [0, 1, 2])
def test_dpf_plateau_checksum(golden, case):
    c = golden["dpf_plateau"]["cases"][case]
    got = O.dpf_real_bench_checksum(c["n"], c["B"], c["t"], 1, 8)
    assert "%016x" % got == c["checksum"]


@pytest.mark.slow
@pytest.mark.parametrize("case", [3, 4])
def test_dpf_plateau_checksum_large(golden, case):
    c = golden["dpf_plateau"]["cases"][case]
    got = O.dpf_real_bench_checksum(c["n"], c["B"], c["t"], 1, 8)
    assert "%016x" % got == c["checksum"]


def _dealt_instance(n, w, seed=5, B=3):
    rng = np.random.default_rng(seed)
    m, levels = 1 << w, n // w
    nk = m - 1
    alphas = rng.integers(0, 1 << n, B)
    d0 = rng.integers(0, 2**63, (B, 2), dtype=np.uint64)
    d1 = rng.integers(0, 2**63, (B, 2), dtype=np.uint64)
    d0[:, 0] |= np.uint64(1)
    d1[:, 0] &= ~np.uint64(1)
    W = rng.integers(0, 2**63, 2, dtype=np.uint64)
    delta = d0 ^ d1
    dl0 = rng.integers(0, 2**63, (B, levels, nk, 2), dtype=np.uint64)
    dl1 = dl0.copy()
    for b in range(B):
        for i in range(levels):
            k = O.lib().orc_dpf_digit(n, w, i, int(alphas[b]))
            if k < nk:
                dl1[b, i, k] ^= delta[b]
    return alphas, d0 ^ W, d1 ^ W, dl0, dl1, delta


@pytest.mark.parametrize("w,prg", [(1, O.PRG_CHACHA8), (2, O.PRG_CHACHA8), (3, O.PRG_CHACHA8),
                                   (1, O.PRG_AES), (2, O.PRG_AES)])
def test_dpf_point_function_and_eval(w, prg):
    n = 6 if w != 3 else 6
    rks = O.aes_tree_round_keys((1 << w) - 1) if prg == O.PRG_AES else None
    alphas, r0, r1, dl0, dl1, delta = _dealt_instance(n, w)
    B = len(alphas)
    l0, l1, cw = O.dpf_full_eval_both(n, w, prg, r0, r1, dl0, dl1, rks=rks, want_cw=True)
    D = 1 << n
    m, levels = 1 << w, n // w
    x = (l0 ^ l1).reshape(B, D, 2)
    for b in range(B):
        for j in range(D):
            exp = delta[b] if j == alphas[b] else np.zeros(2, np.uint64)
            assert np.array_equal(x[b, j], exp)      # leaf0 ^ leaf1 = Delta*[j=alpha]
        cwb = cw.reshape(B, levels * m, 2)[b]
        # brute force: eval() of every point walks to the same leaves
        for j in range(D):
            assert O.dpf_eval(n, w, prg, r0[b], cwb, j, rks=rks) == tuple(int(v) for v in l0[b * D + j])
            assert O.dpf_eval(n, w, prg, r1[b], cwb, j, rks=rks) == tuple(int(v) for v in l1[b * D + j])
        # the fused protocol's CWs equal the path-only CWs (off-path nodes agree)
        pcw = O.dpf_path_cw(n, w, prg, r0[b], r1[b], dl0[b], dl1[b], int(alphas[b]), rks=rks)
        assert np.array_equal(pcw, cwb)


@pytest.mark.parametrize("prg", [O.PRG_CHACHA8, O.PRG_AES])
@pytest.mark.parametrize("n", [1, 4, 7])
def test_dpf_dealer_gen_and_key_eval(prg, n):
    """Dealer Gen (path-only CW) + sync-free full eval of the batched DPF keys:
    the CWs equal the fused protocol's on the same roots and DltSft shares,
    the key eval reproduces the fused leaves (through the point eval of every
    j), and the converted shares sum to beta at alpha and 0 elsewhere."""
    rks = O.aes_tree_round_keys(1) if prg == O.PRG_AES else None
    rng = np.random.default_rng(40 + n)
    for _ in range(4):
        alpha = int(rng.integers(0, 1 << n))
        beta = int(rng.integers(1, O.P))
        delta = [int(rng.integers(0, 2**63)) | 1, int(rng.integers(0, 2**63))]
        d0 = [int(rng.integers(0, 2**63)) | 1, int(rng.integers(0, 2**63))]
        r0, r1, cw, cw_out = O.dpf_key_gen(n, prg, alpha, beta, delta, d0, rks=rks)
        # the fused protocol from the same roots with DltSft shares
        dl = np.zeros((1, n, 1, 2), dtype=np.uint64)
        for i in range(n):
            if ((alpha >> (n - 1 - i)) & 1) == 0:
                dl[0, i, 0] = delta
        sh = rng.integers(0, 2**63, (1, n, 1, 2), dtype=np.uint64)
        l0, l1, fcw = O.dpf_full_eval_both(n, 1, prg, np.array([r0], np.uint64),
                                           np.array([r1], np.uint64), sh, sh ^ dl, rks=rks,
                                           want_cw=True)
        fcw = fcw.reshape(n, 2, 2)
        assert np.array_equal(fcw[:, 0], cw)              # CW_{i,0}
        assert np.array_equal(fcw[:, 1], cw)              # fold word = CW_{i,0} at w = 1
        y0 = O.dpf_key_eval_y(n, prg, r0, cw, cw_out, 0, rks=rks)
        y1 = O.dpf_key_eval_y(n, prg, r1, cw, cw_out, 1, rks=rks)
        full_cw = np.repeat(cw, 2, axis=0)                 # [i][k], k = 0, 1
        for j in range(1 << n):
            leaf0 = O.dpf_eval(n, 1, prg, r0, full_cw, j, rks=rks)
            assert leaf0 == tuple(int(v) for v in l0[j])
            C = int(O.out_hash(np.array([leaf0], np.uint64))[0][0]) % O.P
            e0 = (C + (cw_out if leaf0[0] & 1 else 0)) % O.P
            assert int(y0[j]) == e0
            total = (int(y0[j]) + int(y1[j])) % O.P
            assert total == (beta if j == alpha else 0)


def test_tree_prg_chacha8_is_keystream_chunks():
    rng = np.random.default_rng(9)
    par = rng.integers(0, 2**63, (4, 2), dtype=np.uint64)
    h = O.tree_hash_chacha8(par, 7)
    for j in range(4):
        key = int(par[j, 0]).to_bytes(8, "little") + int(par[j, 1]).to_bytes(8, "little")
        ks = O.chacha8_block(key, 0, 0) + O.chacha8_block(key, 0, 1)
        for k in range(7):
            chunk = ks[16 * k:16 * k + 16]
            assert int(h[j * 7 + k, 0]) == int.from_bytes(chunk[:8], "little")
            assert int(h[j * 7 + k, 1]) == int.from_bytes(chunk[8:], "little")
    oh = O.out_hash(par)
    key = int(par[0, 0]).to_bytes(8, "little") + int(par[0, 1]).to_bytes(8, "little")
    assert int(oh[0, 0]) == int.from_bytes(O.chacha8_block(key, O.OUT_DOMAIN, 0)[:8], "little")


def test_aes_tree_prg_is_ccr():
    rks = O.aes_tree_round_keys(3)
    mrk = O.aes128_expand_key(O.FIPS197_MASTER)
    for k in range(3):
        tweak = k.to_bytes(8, "little") + bytes(8)
        sub = bytes(a ^ b for a, b in zip(O.aes128_encrypt(mrk, tweak), tweak))
        assert bytes(rks[k * 176:(k + 1) * 176]) == O.aes128_expand_key(sub)


@pytest.mark.parametrize("N,c,t,prg", [(1 << 12, 2, 16, O.PRG_CHACHA8), (1 << 10, 2, 4, O.PRG_AES),
                                       (1 << 11, 3, 8, O.PRG_CHACHA8)])
def test_ole_relation(N, c, t, prg):
    x0, z0, x1, z1 = O.pcg_ole_expand(N, c, t, 1, prg, seed=7, threads=8)
    prod = O.poly_mul(x0, x1)
    s = (z0.astype(object) + z1.astype(object)) % P
    assert all(int(p) == int(q) for p, q in zip(prod, s))
    assert int((x0 != 0).sum()) > N // 2            # x is dense (a_i e_i)


@pytest.mark.parametrize("N,c", [(1 << 12, 2), (1 << 13, 3)])
def test_generate_triples_are_triples(N, c):
    """triple.h:45-110 restated: every slot is a Z_p multiplication triple
    (c0 + c1 = (a0 + a1)(b0 + b1)), the a/b shares are the forward NTTs of the
    right OLE outputs, and a planted error in one c share is caught."""
    a0, b0, c0, a1, b1, c1 = O.generate_triples(N, c, 16, 3, 4, threads=8)
    a = (a0.astype(object) + a1.astype(object)) % O.P
    b = (b0.astype(object) + b1.astype(object)) % O.P
    cc = (c0.astype(object) + c1.astype(object)) % O.P
    assert ((a * b) % O.P == cc).all()
    x01, _, x11, _ = O.pcg_ole_expand(N, c, 16, seed=3, threads=8)
    x02, _, x12, _ = O.pcg_ole_expand(N, c, 16, seed=4, threads=8)
    assert np.array_equal(a0, O.ntt_forward(x01)) and np.array_equal(b0, O.ntt_forward(x02))
    assert np.array_equal(a1, O.ntt_forward(x12)) and np.array_equal(b1, O.ntt_forward(x11))
    bad = c0.copy()
    bad[N // 3] = (int(bad[N // 3]) + 1) % O.P
    cc = (bad.astype(object) + c1.astype(object)) % O.P
    assert not ((a * b) % O.P == cc).all()


def test_ole_relation_fails_on_planted_error():
    # a planted mistake in the oracle's output must be caught by the relation
    N = 1 << 10
    x0, z0, x1, z1 = O.pcg_ole_expand(N, 2, 4, 1, O.PRG_CHACHA8, seed=3, threads=4)
    z0 = z0.copy()
    z0[5] = (int(z0[5]) + 1) % P
    prod = O.poly_mul(x0, x1)
    s = (z0.astype(object) + z1.astype(object)) % P
    assert any(int(p) != int(q) for p, q in zip(prod, s))


def test_ref_headers_agree_with_oracle():
    ref = O.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(1)
    for logN in (4, 9, 12):
        N = 1 << logN
        a = rng.integers(0, P, N, dtype=np.uint64)
        b = rng.integers(0, P, N, dtype=np.uint64)
        ra = a.copy()
        ref.ref_ntt_forward(O._p(ra), N)
        assert np.array_equal(ra, O.ntt_forward(a))
        ri = ra.copy()
        ref.ref_ntt_inverse(O._p(ri), N)
        assert np.array_equal(ri, a)
        rc = np.empty_like(a)
        ref.ref_poly_mul(O._p(rc), O._p(a), O._p(b), N)
        assert np.array_equal(rc, O.poly_mul(a, b))
        assert ref.ref_negacyclic_psi(N) == O.ntt_psi(N)
    import ctypes as C
    key = bytes(rng.integers(0, 256, 32, dtype=np.uint8))
    out = (C.c_uint8 * 128)()
    ref.ref_chacha8_blocks((C.c_uint8 * 32).from_buffer_copy(key), O.OUT_DOMAIN, 5, out, 2)
    assert bytes(out)[:64] == O.chacha8_block(key, O.OUT_DOMAIN, 5)
    assert bytes(out)[64:] == O.chacha8_block(key, O.OUT_DOMAIN, 6)


def test_served_hash_fixture_first_entry():
    """tests/golden/served_hashes.json was written by the oracle: its first
    (2^16) entry recomputes to the same four hashes, and every entry's OLE
    parameters are the served workload's (c=4, t=16, ChaCha8)."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "served_hashes.json")
    reqs = json.load(open(path))["requests"]
    assert {r["N"] for r in reqs} == {1 << 16, 1 << 18, 1 << 20, 1 << 22}
    assert all(r["c"] == 4 and r["t"] == 16 and r["prg"] == "chacha8" for r in reqs)
    r = reqs[0]
    outs = O.pcg_ole_expand(r["N"], r["c"], r["t"], r["w"], O.PRG_CHACHA8, seed=r["seed"],
                            pub_seed=r["pub_seed"], threads=8)
    assert [f"{O.fnv1a64(v):016x}" for v in outs] == r["fnv1a64"]


@pytest.mark.parametrize("N,c,t,w", [(1 << 11, 2, 16, 1), (1 << 11, 2, 16, 2), (1 << 11, 2, 16, 4),
                                     (1 << 12, 2, 16, 3), (1 << 13, 3, 8, 1)])
def test_gmw_preprocessing_reconstructs(N, c, t, w):
    """The GMW preprocessing (pcg_ole_impl.h:114-199) evaluated gate by gate:
    for every instance the adder's XOR shares reconstruct alpha = off0 + off1
    (a closed form, here the dealer's alpha) and the DltSft circuit's shares
    reconstruct [digit_i(alpha) = k] * Delta (the dealer's XOR); the shares
    themselves are fresh, and every other dealt value is the dealer's."""
    d = O.dealt_inputs(N, c, t, w, seed=17)
    g = O.dealt_inputs(N, c, t, w, seed=17, gmw=True)
    assert np.array_equal(g["alpha_sh"][0] ^ g["alpha_sh"][1], d["alpha"])
    assert np.array_equal(g["dltsft"][0] ^ g["dltsft"][1], d["dltsft"][0] ^ d["dltsft"][1])
    assert not np.array_equal(g["dltsft"][0], d["dltsft"][0])
    assert not np.array_equal(g["alpha_sh"][0], d["alpha"])
    # beta and rs by Gilboa products: fresh shares of the same field values
    # (beta = pay0 * pay1, rs = (r0 + r1)(s0 + s1), the dealer's closed forms)
    for k in ("beta", "rs"):
        assert np.all(g[k] < P), k
        assert [(int(a) + int(b)) % P for a, b in zip(*g[k])] == \
            [(int(a) + int(b)) % P for a, b in zip(*d[k])], k
        assert not np.array_equal(g[k][0], d[k][0]), k
    for k in ("pos", "pay", "root", "r", "s"):
        assert np.array_equal(g[k], d[k]), k


def test_gilboa_product_shares():
    """Gilboa's OT product (bool_circuit/circuit.cpp:91-153): the two shares
    are in [0, P) and add up to x * y mod P, at the edges (zero, one, P - 1,
    y with bit 61 set, the top bit the 62-bit decomposition must reach) and
    on random operands; different products of one instance use different
    sender randomness."""
    rng = np.random.default_rng(5)
    cases = [(0, 0), (1, 1), (P - 1, P - 1), (7, P - 1), (P - 1, 1 << 61), (1 << 61, 3)]
    cases += [(int(a), int(b)) for a, b in rng.integers(0, P, (64, 2), dtype=np.uint64)]
    for q, (x, y) in enumerate(cases):
        s0, s1 = O.gilboa(x, y, 11, 5, q % 3)
        assert s0 < P and s1 < P
        assert (s0 + s1) % P == x * y % P, (x, y)
    assert O.gilboa(3, 4, 11, 5, 0)[0] != O.gilboa(3, 4, 11, 5, 1)[0]


def test_gmw_and_gate_counts():
    """AND gates per instance: the Kogge-Stone adder's nb + sum_d 2(nb - d)
    (bool_circuit/adder.h:3-5) plus levels * (2^w - 1) * (w - 1 + 128) for
    the DltSft circuit (dpf.cpp:64-70)."""
    def ks(nb):
        s, d = nb, 1
        while d < nb:
            s += 2 * (nb - d)
            d <<= 1
        return s
    for nb, w, n in ((8, 1, 9), (16, 1, 17), (7, 2, 8), (11, 4, 12), (29, 1, 30)):
        assert O.gmw_and_gates(nb, w, n) == ks(nb) + (n // w) * ((1 << w) - 1) * (w - 1 + 128)
        assert O.gmw_adder_ops(nb) == 1 + 2 * max(0, (nb - 1).bit_length())
