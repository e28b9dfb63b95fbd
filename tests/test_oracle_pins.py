"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against something that does not restate its
own formula: the reference's own host C++ (built from /root/reference into
oracle/_ref/ when available), the reference's known answers, a textbook
routine (numpy FFT / linear algebra, numpy's IEEE fp16), brute force, or an
invariant the mathematics fixes.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from oracle_lib import OracleGeom, OrCeParams, OrTbPlan, cref, gold, or_tb_plan, ptr
from puschrx import profiles as P
from puschrx import abi


def cfg_of(profile, slot=0):
    return profile.config(slot)


# ------------------------------------------------------------------ fp16
def test_fp16_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-5, 5000), rng.normal(0, 3e4, 5000),
                           [0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e9, 6.1e-5, 5.96e-8, 2.98e-8]]).astype(np.float32)
    got = np.array([oracle.or_f2h(float(v)) for v in vals], np.uint16)
    want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    back = np.array([oracle.or_h2f(int(h)) for h in want[:2000]], np.float32)
    assert np.array_equal(back, want[:2000].view(np.float16).astype(np.float32))


# ---------------------------------------------------------- runtime config
def _rt_bytes(rt):
    return bytes(memoryview(rt))


def test_runtime_config_known_bytes(oracle):
    rt = P.REFERENCE_FD16X64.runtime_config()
    b = _rt_bytes(rt)
    assert len(b) == 192 and struct.unpack_from("<2H", b) == (2, 192)
    # common/tests/test_mimo_config_compiler.py:35-36
    assert struct.unpack_from("<10H", b, 152) == (0, 321, 321, 12345, 0, 14, 0, 0, 0, 0)
    why = C.create_string_buffer(128)
    assert oracle.or_validate_runtime_config(cref(rt), why, 128) == 0, why.value


MUTATIONS = [
    ("num_layers", 5), ("num_rx_antennas", 65), ("num_tx_antennas", 17), ("max_layers", 16),
    ("data_re_per_layer", None), ("rnti", 0), ("qm", 3), ("fft_size", 8192), ("reserved16", 1),
    ("dmrs_type", 3), ("n_scid", 2), ("num_dmrs_symbols", 3), ("channel_gain", 0.0),
]


@pytest.mark.parametrize("field,value", MUTATIONS)
def test_runtime_config_rejects_like_reference(oracle, reflib, field, value):
    rt = P.REFERENCE_FD16X64.runtime_config()
    if value is None:
        setattr(rt, field, getattr(rt, field) + 1)
    else:
        setattr(rt, field, value)
    w1, w2 = C.create_string_buffer(128), C.create_string_buffer(128)
    got = oracle.or_validate_runtime_config(cref(rt), w1, 128)
    want = reflib.ref_validate_runtime_config(cref(rt), w2, 128)
    assert got == want != 0
    assert w1.value == w2.value


def test_runtime_config_hash_reserved_adapter(oracle, reflib):
    # mutations of common/tests/mimo_runtime_config_contract.cpp:38-68
    for mut in ("hash", "reserved", "adapter", "alloc"):
        rt = P.REFERENCE_FD16X64.runtime_config()
        if mut == "hash":
            for i in range(32):
                rt.profile_sha256[i] = 0
        elif mut == "reserved":
            rt.reserved[0] = 1
        elif mut == "adapter":
            rt.flags &= ~2
        else:
            rt.start_symbol, rt.num_allocated_symbols = 13, 2
        w = C.create_string_buffer(128)
        got = oracle.or_validate_runtime_config(cref(rt), w, 128)
        assert got == reflib.ref_validate_runtime_config(cref(rt), C.create_string_buffer(128), 128) != 0


def test_derive_config_matches_reference(oracle, reflib):
    for prof in (P.REFERENCE_FD16X64, P.FULL_CARRIER_4RX, P.MID_4RX_RANK2):
        rt = prof.runtime_config()
        for slot in (0, 7):
            a, b = abi.PuschMimoConfig(), abi.PuschMimoConfig()
            assert oracle.or_derive_config(cref(rt), prof.rx, slot, cref(a), None, 0) == 0
            assert reflib.ref_derive_config(cref(rt), prof.rx, slot, cref(b)) == 0
            assert bytes(memoryview(a)) == bytes(memoryview(b))


# -------------------------------------------------------------------- Gold
def test_gold_matches_reference_basis(oracle, reflib):
    """dmrs_ref.py:87-98 equivalence: c = x1 xor (xor of the x2 unit responses
    selected by the set bits of c_init), using the reference's BuildGoldBasis."""
    gmat = np.zeros(31 * 2 * 896, np.uint16)
    g1 = np.zeros(2 * 896, np.uint16)
    reflib.ref_dmrs_gold_basis(ptr(gmat), ptr(g1))
    one = 0x3C00

    def plane_bits(p):
        out = np.zeros(2 * 798, np.uint8)
        out[0::2] = p[:798] == one
        out[1::2] = p[896:896 + 798] == one
        return out

    x1 = plane_bits(g1)
    basis = [plane_bits(gmat[b * 1792:(b + 1) * 1792]) for b in range(31)]
    rng = np.random.default_rng(7)
    for cinit in list(rng.integers(0, 2**31, 20)) + [0, 1, 2**31 - 1]:
        want = x1.copy()
        for b in range(31):
            if (int(cinit) >> b) & 1:
                want ^= basis[b]
        assert np.array_equal(gold(oracle, int(cinit), 0, 1596), want)


def test_gold_offset_consistency(oracle):
    full = gold(oracle, 0x1234567, 0, 5000)
    assert np.array_equal(gold(oracle, 0x1234567, 1234, 1000), full[1234:2234])


# ---------------------------------------------------------------- RE demap
def test_re_demap_matches_reference(oracle, reflib):
    cfg = cfg_of(P.REFERENCE_FD16X64)
    g = OracleGeom(cfg)
    idx_o = np.zeros(g.pad, np.uint32)
    idx_r = np.zeros(g.pad, np.uint32)
    oracle.or_re_demap_gather_index(cref(cfg), ptr(idx_o))
    assert reflib.ref_re_demap_gather_index(ptr(idx_r), g.pad) == 0
    assert np.array_equal(idx_o, idx_r)
    rng = np.random.default_rng(3)
    fre = rng.integers(0, 65535, g.nr * 14 * g.fft, dtype=np.uint16)
    fim = rng.integers(0, 65535, g.nr * 14 * g.fft, dtype=np.uint16)
    o = [np.zeros(g.nr * 14 * g.pad, np.uint16) for _ in range(4)]
    oracle.or_re_demap(cref(cfg), ptr(fre), ptr(fim), ptr(o[0]), ptr(o[1]))
    assert reflib.ref_re_demap_reference(ptr(fre), ptr(fim), g.nr, ptr(idx_r), ptr(o[2]), ptr(o[3])) == 0
    assert np.array_equal(o[0], o[2]) and np.array_equal(o[1], o[3])


def test_re_demap_index_is_fftshift_of_stage4(oracle):
    """Closed form: used index u is FFT bin (u - U/2) mod N, stored at
    stage-4 element (bin % P) * 64 + bin // P."""
    cfg = cfg_of(P.FULL_CARRIER_4RX)
    g = OracleGeom(cfg)
    idx = np.zeros(g.pad, np.uint32)
    oracle.or_re_demap_gather_index(cref(cfg), ptr(idx))
    P_ = g.fft // 64
    x = np.arange(g.fft)
    stage4 = np.empty(g.fft, np.int64)
    stage4[(x % P_) * 64 + x // P_] = x  # element -> bin
    bins = stage4[idx[:g.used] // 2]
    assert np.array_equal(bins, (np.arange(g.used) - g.used // 2) % g.fft)


# ------------------------------------------------------------- layer demap
@pytest.mark.parametrize("rank", [1, 2, 3, 4])
def test_layer_demap_matches_reference(oracle, reflib, rank):
    prof = P.Profile(name="r", rx=64, rank=rank, num_rb=133, fft=2048, qm=8, tx=16, padded_subcarriers=1664)
    cfg = cfg_of(prof)
    g = OracleGeom(cfg)
    rng = np.random.default_rng(rank)
    llr = rng.integers(-2560, 2561, g.nl * g.qm * g.qam_plane, dtype=np.int16)
    a = np.zeros(g.qm * g.cw_stride, np.int16)
    b = np.full(g.qm * g.cw_stride, 77, np.int16)
    oracle.or_layer_demap(cref(cfg), ptr(llr), ptr(a))
    assert reflib.ref_layer_demap_reference(ptr(llr), cref(cfg), ptr(b)) == 0
    assert np.array_equal(a, b)


# -------------------------------------------------------------------- DMRS
@pytest.mark.parametrize("rank", [1, 2, 3, 4])
def test_dmrs_gen_matches_reference_profile(oracle, reflib, rank):
    prof = P.Profile(name="r", rx=64, rank=rank, num_rb=133, fft=2048, qm=8, tx=16, padded_subcarriers=1664,
                     slot_number=5)
    cfg = cfg_of(prof, 5)
    g = OracleGeom(cfg)
    cinit = np.zeros(16, np.int32)
    lay = abi.PuschMimoLayout()
    assert reflib.ref_dmrs_gen_profile(cref(cfg), cref(lay), ptr(cinit)) == 0
    for d, s in enumerate(g.dmrs_sym):
        assert oracle.or_dmrs_cinit(cref(cfg), s) == cinit[d]
    re = np.zeros(g.nl * g.ndmrs * g.ref_pad, np.uint16)
    im = np.zeros_like(re)
    oracle.or_dmrs_gen(cref(cfg), ptr(re), ptr(im))
    # TS 38.211 6.4.1.1: r(m) = ((1-2c(2m)) + j(1-2c(2m+1)))/sqrt(2), w_f = (-1)^m on odd ports.
    h = np.float16(0.70710678)
    for l, port in enumerate(prof.port_list()):
        for d in range(g.ndmrs):
            c = gold(oracle, int(cinit[d]), 0, 2 * g.n_dmrs_re)
            wf = np.where((np.arange(g.n_dmrs_re) % 2 == 1) & ((port - 1000) % 2 == 1), -1, 1)
            base = (l * g.ndmrs + d) * g.ref_pad
            er = (h * (1 - 2 * c[0::2].astype(np.float16)) * wf).astype(np.float16)
            ei = (h * (1 - 2 * c[1::2].astype(np.float16)) * wf).astype(np.float16)
            assert np.array_equal(re[base:base + g.n_dmrs_re].view(np.float16), er)
            assert np.array_equal(im[base:base + g.n_dmrs_re].view(np.float16), ei)
            assert not re[base + g.n_dmrs_re:base + g.ref_pad].any()


# ---------------------------------------------------------------- DMRS LS
@pytest.mark.parametrize("rank", [1, 2, 3, 4])
def test_dmrs_ls_profile_matches_reference(oracle, reflib, rank):
    prof = P.Profile(name="r", rx=64, rank=rank, num_rb=133, fft=2048, qm=8, tx=16, padded_subcarriers=1664)
    cfg = cfg_of(prof)
    g = OracleGeom(cfg)
    cnt_r = np.zeros(16, np.uint16)
    sc_r = np.zeros(4 * 2 * 832, np.uint16)
    lay = abi.PuschMimoLayout()
    assert reflib.ref_dmrs_ls_profile(cref(cfg), cref(lay), ptr(cnt_r), ptr(sc_r)) == 0
    rng = np.random.default_rng(rank)
    n = g.nr * 14 * g.pad
    rxr = rng.normal(0, 0.3, n).astype(np.float16).view(np.uint16)
    rxi = rng.normal(0, 0.3, n).astype(np.float16).view(np.uint16)
    ref_re = np.zeros(g.nl * g.ndmrs * g.ref_pad, np.uint16)
    ref_im = np.zeros_like(ref_re)
    oracle.or_dmrs_gen(cref(cfg), ptr(ref_re), ptr(ref_im))
    hr = np.zeros(g.nr * g.nl * g.ndmrs * g.pilot_pad, np.uint16)
    hi = np.zeros_like(hr)
    sc = np.zeros(g.nl * g.ndmrs * g.pilot_pad, np.uint16)
    cnt = np.zeros(16, np.uint16)
    noise = np.zeros(g.nr)
    oracle.or_dmrs_ls(cref(cfg), ptr(rxr), ptr(rxi), ptr(ref_re), ptr(ref_im), ptr(hr), ptr(hi), ptr(sc),
                      ptr(cnt), ptr(noise))
    assert np.array_equal(cnt, cnt_r)
    assert np.array_equal(sc, sc_r[:sc.size])
    # fp16 operation order of mimo_dmrs_ls/scripts/gen_data.py:94-107 with numpy's float16.
    y = (rxr.view(np.float16).reshape(g.nr, 14, g.pad), rxi.view(np.float16).reshape(g.nr, 14, g.pad))
    x = (ref_re.view(np.float16).reshape(g.nl, g.ndmrs, g.ref_pad),
         ref_im.view(np.float16).reshape(g.nl, g.ndmrs, g.ref_pad))
    H = (hr.view(np.float16).reshape(g.nr, g.nl, g.ndmrs, g.pilot_pad),
         hi.view(np.float16).reshape(g.nr, g.nl, g.ndmrs, g.pilot_pad))
    for l, port in enumerate(prof.port_list()):
        comb = (port - 1000) // 2
        for d, s in enumerate(g.dmrs_sym):
            yr = y[0][:, s, comb:comb + 2 * g.n_dmrs_re:2]
            yi = y[1][:, s, comb:comb + 2 * g.n_dmrs_re:2]
            xr, xi = x[0][l, d, :g.n_dmrs_re], x[1][l, d, :g.n_dmrs_re]
            er = (yr * xr) + (yi * xi)
            ei = (yi * xr) - (yr * xi)
            if cnt[l * g.ndmrs + d] == g.n_occ:
                er = (er[:, 0::2] + er[:, 1::2]) * np.float16(0.5)
                ei = (ei[:, 0::2] + ei[:, 1::2]) * np.float16(0.5)
            k = er.shape[1]
            assert np.array_equal(H[0][:, l, d, :k], er)
            assert np.array_equal(H[1][:, l, d, :k], ei)


def test_dmrs_ls_noise_estimate_statistics(oracle):
    """A flat unit channel plus AWGN of variance s2 on every RE: the
    adjacent-difference estimate returns s2 (both observation models)."""
    for rank in (1, 4):
        prof = P.Profile(name="n", rx=4, rank=rank, num_rb=273, fft=4096, qm=8, padded_subcarriers=3328)
        cfg = cfg_of(prof)
        g = OracleGeom(cfg)
        ref_re = np.zeros(g.nl * g.ndmrs * g.ref_pad, np.uint16)
        ref_im = np.zeros_like(ref_re)
        oracle.or_dmrs_gen(cref(cfg), ptr(ref_re), ptr(ref_im))
        s2 = 0.01
        rng = np.random.default_rng(rank)
        y = np.zeros((g.nr, 14, g.pad), np.complex128)
        xr = ref_re.view(np.float16).astype(np.float64).reshape(g.nl, g.ndmrs, g.ref_pad)
        xi = ref_im.view(np.float16).astype(np.float64).reshape(g.nl, g.ndmrs, g.ref_pad)
        for l, port in enumerate(prof.port_list()):
            comb = (port - 1000) // 2
            for d, s in enumerate(g.dmrs_sym):
                y[:, s, comb:comb + 2 * g.n_dmrs_re:2] += xr[l, d, :g.n_dmrs_re] + 1j * xi[l, d, :g.n_dmrs_re]
        y += np.sqrt(s2 / 2) * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))
        yr = y.real.astype(np.float16).view(np.uint16).ravel()
        yi = y.imag.astype(np.float16).view(np.uint16).ravel()
        hr = np.zeros(g.nr * g.nl * g.ndmrs * g.pilot_pad, np.uint16)
        hi = np.zeros_like(hr)
        sc = np.zeros(g.nl * g.ndmrs * g.pilot_pad, np.uint16)
        cnt = np.zeros(16, np.uint16)
        noise = np.zeros(g.nr)
        oracle.or_dmrs_ls(cref(cfg), ptr(yr), ptr(yi), ptr(ref_re), ptr(ref_im), ptr(hr), ptr(hi), ptr(sc),
                          ptr(cnt), ptr(noise))
        assert np.allclose(noise, s2, rtol=0.15)


# -------------------------------------------------------------------- QAM
@pytest.mark.parametrize("qm", [2, 4, 6, 8])
def test_qam_axis_is_maxlog(oracle, qm):
    """The 256QAM integer closed form (gen_data.py:20-40) equals max-log by
    exhaustion over the 16 levels of the reference mapper, for the reference
    test noises; Qm 2/4/6 are defined by that exhaustion."""
    rng = np.random.default_rng(0x256BA7C)
    m = qm // 2
    out16 = (C.c_int16 * 4)()
    out32 = (C.c_int32 * 4)()
    for noise in (0.0125, 0.02, 0.05, 0.2, 0.7):
        for y in rng.normal(0.0, 0.7, 400):
            oracle.or_qam_axis_llr(qm, float(y), noise, out16)
            oracle.or_qam_axis_llr_bruteforce(qm, float(y), noise, out32)
            for i in range(m):
                assert out16[i] == max(-2560, min(2560, out32[i]))


def test_qam_levels_are_gray_and_polarity(oracle):
    """Noiseless points of the reference mapper (qam256_mod_ref.py:41-45)
    decode to their own bits: LLR > 0 exactly where the bit is 0."""
    out = (C.c_int16 * 4)()
    D = 1 / np.sqrt(170)
    for code in range(16):
        c = [(code >> i) & 1 for i in range(4)]
        d = [2 * b - 1 for b in c]
        level = d[0] * (8 - d[1] * (4 - d[2] * (2 - d[3])))
        oracle.or_qam_axis_llr(8, float(level * D), 0.01, out)
        for i in range(4):
            assert (out[i] > 0) == (c[i] == 0)


# ------------------------------------------------------------- descramble
def test_descramble_continues_gold_across_slots(oracle):
    cfg = cfg_of(P.REFERENCE_FD16X64)
    g = OracleGeom(cfg)
    slots = 2
    rng = np.random.default_rng(5)
    x = rng.integers(-2560, 2561, slots * g.qm * g.cw_stride, dtype=np.int16)
    out = np.zeros_like(x)
    oracle.or_descramble(cref(cfg), slots, 0, ptr(x), ptr(out))
    cinit = (cfg.rnti << 15) | cfg.data_scrambling_id
    c = gold(oracle, cinit, 0, slots * g.cw_symbols * g.qm).reshape(slots, g.cw_symbols, g.qm)
    stream_of_bit = [0, 4, 1, 5, 2, 6, 3, 7]  # descramble_mimo.h:30
    xv = x.reshape(slots, g.qm, g.cw_stride)
    ov = out.reshape(slots, g.qm, g.cw_stride)
    for b in range(8):
        want = xv[:, stream_of_bit[b], :g.cw_symbols] * (1 - 2 * c[:, :, b].astype(np.int16))
        assert np.array_equal(ov[:, b, :g.cw_symbols], want)
    assert not ov[:, :, g.cw_symbols:].any()


# --------------------------------------------------------- TBS / rate dematch
def test_tbs_known_answers(oracle):
    # 273 PRB, 2 DMRS symbols (N_RE' = 144), 256QAM, 4 layers, R = 873/1024.
    assert oracle.or_tbs(273, 144, 8, 4, 873 / 1024) == 1081512
    p = or_tb_plan(oracle, 1081512, 39312 * 4 * 8, 8, 4)
    assert (p.C, p.Z, p.K, p.Kprime, p.filler) == (129, 384, 8448, 8408, 40)
    assert (p.E_low, p.E_high, p.n_low) == (9728, 9760, 33)
    # 106 PRB, rank 2, 64QAM, R = 0.6: TBS 110632, C = 14.
    t = oracle.or_tbs(106, 144, 6, 2, 0.6)
    assert t == 110632
    assert or_tb_plan(oracle, t, 106 * 144 * 6 * 2, 6, 2).C == 14


def _reference_plan(layers, slots):
    p = OrTbPlan()
    p.C, p.Z, p.K, p.Kprime, p.filler, p.N, p.Ncb, p.k0 = 143, 384, 8448, 8448, 0, 25344, 25344, 0
    p.tbs = 143 * 8448
    units = slots * 19152
    p.n_low = 143 - units % 143
    p.E_low, p.E_high = 8 * layers * (units // 143), 8 * layers * (-(-units // 143))
    return p


def test_rate_dematch_reference_distribution(oracle):
    """rate_dematch_ref.py:71-76: SISO 23-slot E distribution, full coverage,
    and invariance to poison in the codeword tail."""
    p = _reference_plan(1, 23)
    assert (p.n_low, p.E_low, p.E_high) == (87, 24640, 24648)
    cfg = cfg_of(P.Profile(name="r", rx=64, rank=4, num_rb=133, fft=2048, qm=8, tx=16, padded_subcarriers=1664))
    g = OracleGeom(cfg)
    slots = 2
    p = _reference_plan(4, slots)
    rng = np.random.default_rng(0x52444D31)
    base = rng.integers(-900, 901, (slots, 8, g.cw_stride), dtype=np.int16)
    other = base.copy()
    base[:, :, g.cw_symbols:] = 0x1111
    other[:, :, g.cw_symbols:] = 0x6A6A
    o1 = np.zeros(143 * 68 * 384, np.int16)
    o2 = np.zeros_like(o1)
    oracle.or_rate_dematch(C.byref(p), 4, 8, g.cw_symbols, g.cw_stride, slots, ptr(base), ptr(o1))
    oracle.or_rate_dematch(C.byref(p), 4, 8, g.cw_symbols, g.cw_stride, slots, ptr(other), ptr(o2))
    assert np.array_equal(o1, o2)
    assert not o1.reshape(143, -1)[:, :768].any()


def test_rate_dematch_inverts_rate_match_bruteforce(oracle):
    """Brute force: place distinct values through the TS 38.212 bit-selection +
    interleaver forward map and check the dematcher returns each value x8 at
    its circular-buffer position (filler at +5120, punctured at 0)."""
    p = or_tb_plan(oracle, 3752, 3 * 12 * 14 * 8, 8, 1)  # small TB: 1 CB, Z = 192? no: check below
    Z, Kp, F, Ncb = p.Z, p.Kprime, p.filler, p.Ncb
    E = p.E_low
    assert p.C == 1
    cw_symbols = E // 8
    inp = np.zeros((1, 8, cw_symbols), np.int16)
    expect = np.zeros(68 * Z, np.int32)
    fstart = Kp - 2 * Z
    nonfill = [k for k in range(Ncb) if not (fstart <= k < fstart + F)]
    rng = np.random.default_rng(9)
    vals = rng.integers(-300, 301, E)
    for k in range(E):  # e_k -> circular position of the k-th non-filler bit
        pos = nonfill[k % len(nonfill)]
        j, b = k % cw_symbols, k // cw_symbols
        inp[0, b, j] = vals[k]
        expect[2 * Z + pos] += int(np.clip(vals[k] * 8, -5120, 5120))
    expect[2 * Z + fstart:2 * Z + fstart + F] = 5120
    expect = np.clip(expect, -5120, 5120)
    out = np.zeros(68 * Z, np.int16)
    oracle.or_rate_dematch(C.byref(p), 1, 8, cw_symbols, cw_symbols, 1, ptr(inp), ptr(out))
    assert np.array_equal(out, expect.astype(np.int16))


# -------------------------------------------------------------------- LDPC
def test_bg1_table_structure():
    import re
    from pathlib import Path
    text = (Path(__file__).resolve().parents[1] / "include" / "airan_ldpc_bg1.h").read_text()

    def arr(name):
        body = re.search(r"AIRAN_BG1_TABLE \w+ " + name + r"\[[^\]]*\] = \{(.*?)\};", text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body)
        return [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]

    start, col, sh = arr("AIRAN_BG1_ROW_START"), arr("AIRAN_BG1_COL"), arr("AIRAN_BG1_SHIFT")
    assert len(start) == 47 and start[-1] == 316 == len(col) == len(sh)
    deg = np.diff(start)
    assert deg.max() == 19 and set(deg.tolist()) == {3, 4, 5, 6, 7, 8, 9, 10, 19}
    rows = [col[start[r]:start[r + 1]] for r in range(46)]
    assert all(r == sorted(set(r)) and max(r) < 68 for r in rows)
    assert all(0 <= s < 384 for s in sh)
    # Extension row r >= 4 owns parity column 22 + r with shift 0; core parity double diagonal.
    for r in range(4, 46):
        assert rows[r][-1] == 22 + r and sh[start[r + 1] - 1] == 0
    # ldpc_decode_kernel.cpp:38-42: greedy left-to-right fusion of adjacent
    # equal-degree rows with disjoint columns gives exactly this set.
    fused, r = [], 0
    while r < 45:
        if deg[r] == deg[r + 1] and not set(rows[r]) & set(rows[r + 1]):
            fused.append(r)
            r += 2
        else:
            r += 1
    assert fused == [16, 20, 22, 25, 28, 30, 32, 34, 38, 43]


def _encode(oracle, Z, info):
    cw = np.zeros(68 * Z, np.uint8)
    oracle.or_ldpc_encode(Z, ptr(info), ptr(cw))
    return cw


@pytest.mark.parametrize("Z", [384, 96, 24])
def test_ldpc_encode_syndrome(oracle, Z):
    rng = np.random.default_rng(Z)
    for _ in range(3):
        info = rng.integers(0, 2, 22 * Z, dtype=np.uint8)
        cw = _encode(oracle, Z, info)
        assert oracle.or_ldpc_syndrome_ok(Z, ptr(cw)) == 1
        assert np.array_equal(cw[:22 * Z], info)
        bad = cw.copy()
        bad[rng.integers(0, 68 * Z)] ^= 1
        assert oracle.or_ldpc_syndrome_ok(Z, ptr(bad)) == 0


def _llr_from_cw(cw, Z, snr_db, rng, punct_2z=True):
    x = 1.0 - 2.0 * cw  # bit 0 -> +1
    s2 = 10 ** (-snr_db / 10)
    y = x + rng.normal(0, np.sqrt(s2), x.size)
    llr = np.clip(np.rint(2 * y / s2 * 256), -5120, 5120).astype(np.int16)
    if punct_2z:
        llr[:2 * Z] = 0
    return llr


@pytest.mark.parametrize("mode", [0, 1])
def test_ldpc_decode_corrects_noise(oracle, mode):
    Z, ncb = 96, 4
    rng = np.random.default_rng(11 + mode)
    infos, lams = [], []
    for _ in range(ncb):
        info = rng.integers(0, 2, 22 * Z, dtype=np.uint8)
        infos.append(info)
        lams.append(_llr_from_cw(_encode(oracle, Z, info), Z, 3.0, rng))
    lam = np.concatenate(lams)
    bits = np.zeros(ncb * 22 * Z, np.int8)
    oracle.or_ldpc_decode(Z, ncb, ptr(lam), mode, 8, 46, 22 * Z, 0, ptr(bits), None)
    assert np.array_equal(bits.reshape(ncb, -1), np.stack(infos).astype(np.int8))


def test_ldpc_wide_skipping_untransmitted_rows_is_exact(oracle):
    """Rows whose extension parity column is punctured carry zero messages in
    the integer decoder: decoding with active_rows equals decoding all 46."""
    Z = 384
    p = or_tb_plan(oracle, 1081512, 39312 * 32, 8, 4)
    rows = oracle.or_ldpc_active_rows(Z, p.Kprime, p.E_high, 0, p.Ncb)
    assert rows == 6
    rng = np.random.default_rng(4)
    info = rng.integers(0, 2, 22 * Z, dtype=np.uint8)
    info[p.Kprime:] = 0
    lam = _llr_from_cw(_encode(oracle, Z, info), Z, 4.0, rng)
    lam[2 * Z + p.E_high + p.filler:] = 0  # punctured tail
    a, b = lam.copy(), lam.copy()
    ba = np.zeros(22 * Z, np.int8)
    bb = np.zeros(22 * Z, np.int8)
    oracle.or_ldpc_decode(Z, 1, ptr(a), 1, 8, rows, 22 * Z, 0, ptr(ba), None)
    oracle.or_ldpc_decode(Z, 1, ptr(b), 1, 8, 46, 22 * Z, 0, ptr(bb), None)
    assert np.array_equal(ba, bb)
    assert np.array_equal(a[:22 * Z], b[:22 * Z])


# --------------------------------------------------------------------- CRC
def test_crc_parity_property(oracle):
    rng = np.random.default_rng(2)
    for kind, L in ((0, 24), (1, 24), (2, 16)):
        for n in (1, 7, 100, 8424):
            data = rng.integers(0, 2, n, dtype=np.uint8)
            r = oracle.or_crc(kind, ptr(data), n)
            parity = np.array([(r >> (L - 1 - i)) & 1 for i in range(L)], np.uint8)
            msg = np.concatenate([data, parity])
            assert oracle.or_crc(kind, ptr(msg), msg.size) == 0
            msg[rng.integers(0, msg.size)] ^= 1
            assert oracle.or_crc(kind, ptr(msg), msg.size) != 0


def test_crc24a_polynomial_division_matches_numpy_gf2(oracle):
    """Remainder equals GF(2) long division of M(x) x^24 by g(x) done with
    numpy polynomial arithmetic mod 2."""
    g = np.array([1] + [(0x864CFB >> (23 - i)) & 1 for i in range(24)], np.int64)
    rng = np.random.default_rng(3)
    data = rng.integers(0, 2, 200, dtype=np.uint8)
    num = np.concatenate([data.astype(np.int64), np.zeros(24, np.int64)])
    rem = num.copy()
    for i in range(len(data)):
        if rem[i]:
            rem[i:i + 25] ^= g
    want = int("".join(str(int(v)) for v in rem[-24:]), 2)
    assert oracle.or_crc(0, ptr(data), data.size) == want


# -------------------------------------------------------------------- OFDM
def test_ofdm_demod_is_unitary_dft(oracle):
    prof = P.Profile(name="o", rx=2, rank=1, num_rb=52, fft=1024, qm=8, padded_subcarriers=640)
    cfg = cfg_of(prof)
    g = OracleGeom(cfg)
    rng = np.random.default_rng(1)
    iq = rng.normal(size=(g.nr, g.samples, 2)).astype(np.float32)
    grid = np.zeros((g.nr, 14, g.fft, 2))
    oracle.or_ofdm_demod(cref(cfg), ptr(iq), ptr(grid))
    x = iq[..., 0] + 1j * iq[..., 1]
    pos = 0
    for s in range(14):
        pos += g.cp_long if s == 0 else g.cp_short
        want = np.fft.fft(x[:, pos:pos + g.fft], axis=-1) / np.sqrt(g.fft)
        got = grid[:, s, :, 0] + 1j * grid[:, s, :, 1]
        assert np.allclose(got, want, atol=1e-9)
        pos += g.fft
    assert pos == g.samples == 15 * g.fft


# --------------------------------------------------------------- detection
def test_detect_lmmse_closed_form(oracle):
    prof = P.Profile(name="d", rx=4, rank=3, num_rb=2, fft=128, qm=8, padded_subcarriers=24)
    cfg = cfg_of(prof)
    g = OracleGeom(cfg)
    rng = np.random.default_rng(8)
    H = rng.normal(size=(g.nr, g.nl, 14, g.used)) + 1j * rng.normal(size=(g.nr, g.nl, 14, g.used))
    y = rng.normal(size=(g.nr, 14, g.used)) + 1j * rng.normal(size=(g.nr, 14, g.used))
    s2 = np.array([0.3])
    hh = np.stack([H.real, H.imag], -1).copy()
    yy = np.stack([y.real, y.imag], -1).copy()
    xh = np.zeros((g.nl, 14, g.used, 2))
    ne = np.zeros((g.nl, 14, g.used))
    oracle.or_detect_lmmse(cref(cfg), ptr(hh), ptr(yy), ptr(s2), ptr(xh), ptr(ne))
    for s in range(14):
        for k in range(g.used):
            Hm = H[:, :, s, k]
            A = Hm.conj().T @ Hm + s2[0] * np.eye(g.nl)
            Ai = np.linalg.inv(A)
            x = np.linalg.solve(A, Hm.conj().T @ y[:, s, k])
            d = np.maximum(1 - s2[0] * np.real(np.diag(Ai)), 1e-3)
            assert np.allclose(xh[:, s, k, 0] + 1j * xh[:, s, k, 1], x / d, rtol=1e-9, atol=1e-12)
            assert np.allclose(ne[:, s, k], (1 - d) / d, rtol=1e-9, atol=1e-12)


# -------------------------------------------------------------------- CE
def _ce_case(oracle, prof, h_obs_fn, noise_rel, tau=300e-9, fd=100.0):
    cfg = cfg_of(prof)
    g = OracleGeom(cfg)
    hls = np.zeros((g.nr, g.nl, g.ndmrs, g.pilot_pad, 2))
    for rx in range(g.nr):
        for l, port in enumerate(prof.port_list()):
            comb = (port - 1000) // 2
            occ = sum(((q - 1000) // 2) == comb for q in prof.port_list()) == 2
            n = g.n_occ if occ else g.n_dmrs_re
            sc = comb + (1 + 4 * np.arange(n) if occ else 2 * np.arange(n))
            for d in range(g.ndmrs):
                v = h_obs_fn(rx, l, d, sc)
                hls[rx, l, d, :n, 0], hls[rx, l, d, :n, 1] = v.real, v.imag
    out = np.zeros((g.nr, g.nl, 14, g.used, 2))
    prm = OrCeParams(tau, fd, 30e3)
    nr = np.full(g.nl, noise_rel)
    assert oracle.or_ce_lmmse(cref(cfg), C.byref(prm), ptr(nr), ptr(hls), ptr(out)) == 0
    return g, hls, out[..., 0] + 1j * out[..., 1]


def test_ce_lmmse_matches_eigen_low_rank_formula(oracle):
    """The reference's make_factors (channel_est_lmmse_ref.py:91-117) at full
    rank, done with numpy eigh, equals the oracle's Cholesky-based solve
    (frequency part, single DMRS symbol, no Doppler)."""
    prof = P.Profile(name="c", rx=1, rank=1, num_rb=8, fft=128, qm=8, dmrs_symbols=(2,), padded_subcarriers=128)
    rng = np.random.default_rng(4)
    obs = rng.normal(size=48) + 1j * rng.normal(size=48)
    g, hls, H = _ce_case(oracle, prof, lambda rx, l, d, sc: obs, 0.05, fd=0.0)
    U, tau, scs = g.used, 300e-9, 30e3
    dsc = np.arange(U)[:, None] - np.arange(U)[None, :]
    rf = 1.0 / (1.0 + 2j * np.pi * dsc * scs * tau)
    sup = 2 * np.arange(48)
    rhp, rpp = rf[:, sup], rf[np.ix_(sup, sup)]
    lam, ur = np.linalg.eigh(rpp)
    order = np.argsort(lam)[::-1]
    lam, ur = np.maximum(lam[order], 0), ur[:, order]
    diag = 1 / (lam + 0.05)
    fa, fb = rhp @ ur, diag[:, None] * ur.conj().T
    raa = np.ones(U)
    hv = np.real(np.sum((fa * diag[None, :]) * fa.conj(), axis=1))
    err = np.maximum(raa - hv, 0)
    scale = 2 * raa / (hv + raa - err)
    want = (fa * scale[:, None]) @ (fb @ obs)
    # One DMRS symbol, f_d = 0: every symbol equals the frequency estimate times
    # the time weight J0(0)/(J0(0) + err') = 1/(1 + err').
    err2 = np.maximum(scale * (scale - 1) * hv + (1 - scale) * raa + scale * err, 0)
    for s in range(14):
        assert np.allclose(H[0, 0, s], want / (1 + err2), rtol=1e-8, atol=1e-10)


def test_ce_lmmse_noiseless_smooth_channel(oracle):
    """A channel drawn from the prior itself (exponential PDP, tau = 300 ns) is
    reproduced closely from noiseless pilots."""
    prof = P.Profile(name="c", rx=2, rank=2, num_rb=24, fft=512, qm=8, padded_subcarriers=384)
    rng = np.random.default_rng(5)
    taps = np.arange(16) * 100e-9
    pw = np.exp(-taps / 300e-9)
    pw /= pw.sum()
    hcoef = (rng.normal(size=(2, 16)) + 1j * rng.normal(size=(2, 16))) * np.sqrt(pw / 2)

    def freq(l, sc):
        return (hcoef[l][None, :] * np.exp(-2j * np.pi * np.asarray(sc)[:, None] * 30e3 * taps[None, :])).sum(1)

    g, hls, H = _ce_case(oracle, prof, lambda rx, l, d, sc: freq(l, sc), 1e-4, fd=0.0)
    for l in range(2):
        true = freq(l, np.arange(g.used))
        for s in range(14):
            nmse = np.sum(np.abs(H[0, l, s] - true) ** 2) / np.sum(np.abs(true) ** 2)
            assert nmse < 10 ** (-20 / 10)


def _occ_pair_case(oracle, h_fns, sigma2, tau, rng):
    """An FD-OCC pair (ports 1000/1001) observed physically: z_m = h0(k_m) +
    (-1)^m h1(k_m) + noise on k_m = 2m, and h_ls0,1[p] = (z_2p +- z_2p+1)/2
    (the DMRS-LS pair mean)."""
    prof = P.Profile(name="occ", rx=2, rank=2, num_rb=24, fft=512, qm=8, dmrs_symbols=(2,),
                     padded_subcarriers=384, ports=(1000, 1001))
    cfg = cfg_of(prof)
    g = OracleGeom(cfg)
    k = 2 * np.arange(g.n_dmrs_re)
    z = h_fns[0](k) + (-1.0) ** np.arange(g.n_dmrs_re) * h_fns[1](k)
    z = z + np.sqrt(sigma2 / 2) * (rng.normal(size=z.size) + 1j * rng.normal(size=z.size))
    hls = np.zeros((2, 2, 1, g.pilot_pad, 2))
    for l, sgn in ((0, 1.0), (1, -1.0)):
        v = 0.5 * (z[0::2] + sgn * z[1::2])
        hls[:, l, 0, : g.n_occ, 0], hls[:, l, 0, : g.n_occ, 1] = v.real, v.imag
    nr = np.zeros(4)
    assert oracle.or_ce_noise_rel(cref(cfg), ptr(hls), max(sigma2, 1e-6), ptr(nr)) == 2
    out = np.zeros((2, 2, 14, g.used, 2))
    prm = OrCeParams(tau, 0.0, 30e3)
    assert oracle.or_ce_lmmse(cref(cfg), C.byref(prm), ptr(nr), ptr(hls), ptr(out)) == 0
    return g, out[..., 0] + 1j * out[..., 1]


def _tdl(rng, tau, taps=24, spacing=100e-9):
    d = np.arange(taps) * spacing
    pw = np.exp(-d / tau)
    pw /= pw.sum()
    c = (rng.normal(size=taps) + 1j * rng.normal(size=taps)) * np.sqrt(pw / 2)
    return lambda k: (c[None, :] * np.exp(-2j * np.pi * np.asarray(k)[:, None] * 30e3 * d[None, :])).sum(1)


def test_ce_occ_pair_joint_model_reconstructs_selective_channels(oracle):
    """Two frequency-selective layers sharing a comb (600 ns RMS, where the
    flat-pair approximation would leak about -19 dB of the partner) are
    reconstructed from nearly noiseless pilots to better than -30 dB."""
    rng = np.random.default_rng(12)
    h = [_tdl(rng, 600e-9), _tdl(rng, 600e-9)]
    g, H = _occ_pair_case(oracle, h, 1e-7, 600e-9, rng)
    for l in range(2):
        true = h[l](np.arange(g.used))
        nmse = np.sum(np.abs(H[0, l, 0] - true) ** 2) / np.sum(np.abs(true) ** 2)
        assert nmse < 10 ** (-30 / 10)


def test_ce_occ_pair_does_not_leak(oracle):
    """With the +,+ layer silent, its estimate stays 30 dB below the other."""
    rng = np.random.default_rng(13)
    h1 = _tdl(rng, 300e-9)
    g, H = _occ_pair_case(oracle, [lambda k: np.zeros(len(k), complex), h1], 1e-7, 300e-9, rng)
    e0 = np.sum(np.abs(H[0, 0, 0]) ** 2)
    e1 = np.sum(np.abs(H[0, 1, 0]) ** 2)
    assert e0 < 1e-3 * e1


def _small_ref_llrs(Z, ncb, seed):
    """Q8.8 inputs of magnitude <= 3 (a few thousandths of an LLR): after one
    REF-mode row pass many posteriors are fractions such as -1 + 0.75."""
    rng = np.random.default_rng(seed)
    lam = rng.integers(-3, 4, (ncb, 68 * Z)).astype(np.int16)
    lam[:, :2 * Z] = 0
    return lam.reshape(-1)


def test_ldpc_ref_bits_follow_the_int16_posterior(oracle):
    """fec/ldpc_decode/ldpc_decode_kernel.cpp:735-767: the posterior is
    clamped, converted to int16 with CAST_RINT (the exported lam_out),
    converted back, and the bit is (value < 0).  So in REF mode every bit
    equals lam_out < 0, and a posterior in [-0.5, 0) decodes to 0; the
    inputs here put thousands of posteriors there."""
    Z, ncb = 384, 3
    lam = _small_ref_llrs(Z, ncb, 11)
    for iters in (1, 2):
        out = lam.copy()
        bits = np.zeros(ncb * 22 * Z, np.int8)
        post = np.zeros(ncb * 68 * Z, np.float32)
        oracle.or_ldpc_decode_posterior(Z, ncb, ptr(out), 0, iters, 46, 22 * Z, 0, ptr(bits), None, ptr(post))
        info_out = out.reshape(ncb, 68 * Z)[:, :22 * Z].reshape(-1)
        assert np.array_equal(bits, (info_out < 0).astype(np.int8))
        p = post.reshape(ncb, 68 * Z)[:, :22 * Z].reshape(-1)
        near = (p >= -0.5) & (p < 0)
        assert near.sum() > 1000, near.sum()
        assert not bits[near].any()
        assert np.array_equal(bits[p < -0.5], np.ones(int((p < -0.5).sum()), np.int8))


# ------------------------------------------------------------ SISO LS / ZF
def _siso_ls(oracle, cfg, rx_re, rx_im, ref_re, ref_im):
    g = OracleGeom(cfg)
    hr = np.zeros(14 * g.pad, np.uint16)
    hi = np.zeros_like(hr)
    oracle.or_ls_siso(cref(cfg), ptr(rx_re), ptr(rx_im), ptr(ref_re), ptr(ref_im), ptr(hr), ptr(hi))
    return (hr.view(np.float16).astype(np.float64) + 1j * hi.view(np.float16).astype(np.float64)).reshape(14, g.pad)


def _siso_grid(oracle, cfg, h, rng):
    """Received fp16 grid [14, pad] of data symbols and the DMRS through the
    per-RE channel h [14, used], with the DMRS reference of airan_mimo_dmrs_gen."""
    g = OracleGeom(cfg)
    ref_re = np.zeros(g.ndmrs * g.ref_pad, np.uint16)
    ref_im = np.zeros_like(ref_re)
    oracle.or_dmrs_gen(cref(cfg), ptr(ref_re), ptr(ref_im))
    x = (ref_re.view(np.float16).astype(np.float64) + 1j * ref_im.view(np.float16).astype(np.float64))
    x = x.reshape(g.ndmrs, g.ref_pad)
    s = (rng.choice([-1.0, 1.0], (14, g.used)) + 1j * rng.choice([-1.0, 1.0], (14, g.used))) / np.sqrt(2)
    for d in range(g.ndmrs):
        s[g.dmrs_sym[d], 0::2] = x[d, : g.n_dmrs_re]
    y = np.zeros((14, g.pad), np.complex128)
    y[:, : g.used] = h * s
    return (y.real.astype(np.float16).view(np.uint16).ravel(), y.imag.astype(np.float16).view(np.uint16).ravel(),
            ref_re, ref_im)


@pytest.mark.parametrize("prof", [P.SISO_CONFIG1, P.SISO_133PRB], ids=lambda p: p.name)
def test_ls_siso_pilots_are_dmrs_ls(oracle, prof):
    """The SISO estimate at the DMRS symbols' even subcarriers is the DMRS-LS
    observation h_ls of port 1000 (the same y x* in fp16, pinned against the
    reference's DMRS-LS profile), and its odd subcarriers the midpoints."""
    cfg = cfg_of(prof)
    g = OracleGeom(cfg)
    rng = np.random.default_rng(11)
    h = (rng.normal(size=(14, g.used)) + 1j * rng.normal(size=(14, g.used))) * 0.7
    yr, yi, xr, xi = _siso_grid(oracle, cfg, h, rng)
    H = _siso_ls(oracle, cfg, yr, yi, xr, xi)
    hls_r = np.zeros(g.ndmrs * g.pilot_pad, np.uint16)
    hls_i = np.zeros_like(hls_r)
    sc = np.zeros(g.ndmrs * g.pilot_pad, np.uint16)
    cnt = np.zeros(16, np.uint16)
    noise = np.zeros(1)
    oracle.or_dmrs_ls(cref(cfg), ptr(yr), ptr(yi), ptr(xr), ptr(xi), ptr(hls_r), ptr(hls_i), ptr(sc), ptr(cnt),
                      ptr(noise))
    hls = (hls_r.view(np.float16).astype(np.float64) + 1j * hls_i.view(np.float16).astype(np.float64))
    hls = hls.reshape(g.ndmrs, g.pilot_pad)[:, : g.n_dmrs_re]
    for d in range(g.ndmrs):
        sym = g.dmrs_sym[d]
        assert np.array_equal(H[sym, 0:g.used:2], hls[d])
        mid = np.append(hls[d][1:], 0.0)  # the pilot past the last one is zero
        want = ((hls[d] + mid).astype(np.complex64))
        want = want.real.astype(np.float16).astype(np.float64) * 0.5 + 1j * want.imag.astype(np.float16) * 0.5
        assert np.allclose(H[sym, 1:g.used:2], want, rtol=1e-3, atol=1e-3)
    assert not H[:, g.used:].any()


def test_ls_siso_exact_on_linear_channels(oracle):
    """Linear interpolation reproduces a channel linear in frequency and in
    time (DMRS symbols {2, 11}: interpolated between, extrapolated outside)
    to fp16 accuracy; the orientation of the time weights is fixed by it."""
    cfg = cfg_of(P.SISO_133PRB)
    g = OracleGeom(cfg)
    rng = np.random.default_rng(12)
    sc = np.arange(g.used)
    sym = np.arange(14)[:, None]
    h = (0.6 + 0.2j) + (1e-4 - 2e-4j) * sc[None, :] + (0.03 + 0.02j) * (sym - 6)
    yr, yi, xr, xi = _siso_grid(oracle, cfg, h, rng)
    H = _siso_ls(oracle, cfg, yr, yi, xr, xi)
    inner = slice(0, g.used - 2)  # the last odd subcarrier meets the zero-extended pilot
    err = np.abs(H[:, inner] - h[:, inner]) / np.abs(h[:, inner])
    assert err.max() < 6e-3
    # One DMRS symbol: the same estimate on every symbol.
    cfg1 = cfg_of(P.SISO_CONFIG1)
    g1 = OracleGeom(cfg1)
    h1 = np.broadcast_to((0.5 - 0.3j) + (2e-4 + 1e-4j) * np.arange(g1.used)[None, :], (14, g1.used))
    yr, yi, xr, xi = _siso_grid(oracle, cfg1, h1, rng)
    H1 = _siso_ls(oracle, cfg1, yr, yi, xr, xi)
    assert all(np.array_equal(H1[s], H1[2]) for s in range(14))
    assert np.abs(H1[:, : g1.used - 2] - h1[:, : g1.used - 2]).max() < 4e-3


def test_zf_siso_closed_form(oracle):
    """ZF recovers x = y / h and no_eff = n0 / |h|^2 (to fp16), with the
    channel power floored at 1e-4; the fp32 arithmetic equals numpy's IEEE
    fp32 on the same fp16 inputs."""
    rng = np.random.default_rng(13)
    n = 4096
    h = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 0.8
    h[:64] *= 1e-3  # |h|^2 below the floor
    x = (rng.choice([-3, -1, 1, 3], n) + 1j * rng.choice([-3, -1, 1, 3], n)) / np.sqrt(10)
    y = h * x
    f16 = lambda a: a.astype(np.float16)
    hr, hi, yr, yi = f16(h.real), f16(h.imag), f16(y.real), f16(y.imag)
    n0 = f16(rng.uniform(0.001, 0.1, n))
    out = [np.zeros(n, np.uint16) for _ in range(3)]
    oracle.or_zf_siso(n, ptr(yr.view(np.uint16)), ptr(yi.view(np.uint16)), ptr(hr.view(np.uint16)),
                      ptr(hi.view(np.uint16)), ptr(n0.view(np.uint16)), *[ptr(o) for o in out])
    xr, xi, ne = (o.view(np.float16).astype(np.float64) for o in out)
    hq = hr.astype(np.float64) + 1j * hi.astype(np.float64)
    g = np.maximum(np.abs(hq) ** 2, 1e-4)
    ok = np.abs(hq) ** 2 > 1e-3
    assert np.abs((xr + 1j * xi)[ok] - x[ok]).max() < 2e-2
    assert np.allclose(ne, n0.astype(np.float64) / g, rtol=2e-3)
    assert np.allclose(ne[:64], n0[:64].astype(np.float64) / 1e-4, rtol=2e-3)
    # fp32 arithmetic (numpy float32 ops are IEEE single with one rounding each).
    h32r, h32i, y32r, y32i = (a.astype(np.float32) for a in (hr, hi, yr, yi))
    g32 = np.maximum(h32r * h32r + h32i * h32i, np.float32(1e-4))
    want_r = ((h32r * y32r + h32i * y32i) / g32).astype(np.float16)
    want_i = ((h32r * y32i - h32i * y32r) / g32).astype(np.float16)
    assert np.array_equal(out[0], want_r.view(np.uint16))
    assert np.array_equal(out[1], want_i.view(np.uint16))
