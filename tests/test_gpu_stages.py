"""GPU parity of the per-operator entry points against the oracle, through the
C ABI.  Integer/index stages are bit-exact; fp16 stages that the reference
defines operation by operation (DMRS, DMRS-LS h_ls, QAM inputs) are
bit-exact too; the noise reduction is fp32 on the GPU vs double in the
oracle (relative tolerance 1e-3 after fp16 rounding)."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import OracleGeom, cref, gold, or_tb_plan, ptr
from puschrx import abi, ops
from puschrx import profiles as P

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

PROFILES = {
    "reference_fd16x64": P.REFERENCE_FD16X64,
    "full_carrier_4rx": P.FULL_CARRIER_4RX,
    "mid_4rx_rank2": P.MID_4RX_RANK2,
    "siso_52prb": P.SISO_52PRB,
}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("name", list(PROFILES))
def test_re_demap_bit_exact(oracle, gpu_stream, name):
    cfg = PROFILES[name].config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(1)
    fre = rng.integers(0, 65535, g.nr * 14 * g.fft, dtype=np.uint16)
    fim = rng.integers(0, 65535, g.nr * 14 * g.fft, dtype=np.uint16)
    want_r = np.zeros(g.nr * 14 * g.pad, np.uint16)
    want_i = np.zeros_like(want_r)
    oracle.or_re_demap(cref(cfg), ptr(fre), ptr(fim), ptr(want_r), ptr(want_i))
    rr, ri = ops.re_demap(cfg, dev(fre.view(np.int16)), dev(fim.view(np.int16)))
    assert np.array_equal(host(rr).view(np.uint16), want_r)
    assert np.array_equal(host(ri).view(np.uint16), want_i)


@pytest.mark.parametrize("name", list(PROFILES))
@pytest.mark.parametrize("slot", [0, 9])
def test_dmrs_gen_bit_exact(oracle, gpu_stream, name, slot):
    prof = PROFILES[name]
    cfg = prof.config(slot)
    g = OracleGeom(cfg)
    want_r = np.zeros(g.nl * g.ndmrs * g.ref_pad, np.uint16)
    want_i = np.zeros_like(want_r)
    oracle.or_dmrs_gen(cref(cfg), ptr(want_r), ptr(want_i))
    re, im = ops.dmrs_gen(cfg)
    assert np.array_equal(host(re).view(np.uint16), want_r)
    assert np.array_equal(host(im).view(np.uint16), want_i)


@pytest.mark.parametrize("name", list(PROFILES))
def test_dmrs_ls_bit_exact(oracle, gpu_stream, name):
    cfg = PROFILES[name].config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(2)
    n = g.nr * 14 * g.pad
    rxr = rng.normal(0, 0.4, n).astype(np.float16)
    rxi = rng.normal(0, 0.4, n).astype(np.float16)
    ref_r = np.zeros(g.nl * g.ndmrs * g.ref_pad, np.uint16)
    ref_i = np.zeros_like(ref_r)
    oracle.or_dmrs_gen(cref(cfg), ptr(ref_r), ptr(ref_i))
    hr = np.zeros(g.nr * g.nl * g.ndmrs * g.pilot_pad, np.uint16)
    hi = np.zeros_like(hr)
    sc = np.zeros(g.nl * g.ndmrs * g.pilot_pad, np.uint16)
    cnt = np.zeros(16, np.uint16)
    noise = np.zeros(g.nr)
    oracle.or_dmrs_ls(cref(cfg), ptr(rxr.view(np.uint16)), ptr(rxi.view(np.uint16)), ptr(ref_r), ptr(ref_i),
                      ptr(hr), ptr(hi), ptr(sc), ptr(cnt), ptr(noise))
    ghr, ghi, gsc, gcnt, gnoise = ops.dmrs_ls(cfg, dev(rxr), dev(rxi), dev(ref_r.view(np.float16)),
                                              dev(ref_i.view(np.float16)))
    assert np.array_equal(host(ghr).view(np.uint16), hr)
    assert np.array_equal(host(ghi).view(np.uint16), hi)
    assert np.array_equal(host(gsc).view(np.uint16), sc)
    assert np.array_equal(host(gcnt).view(np.uint16), cnt)
    np.testing.assert_allclose(host(gnoise).astype(np.float64), noise, rtol=2e-3)


@pytest.mark.parametrize("qm", [2, 4, 6, 8])
def test_qam_demod_bit_exact(oracle, gpu_stream, qm):
    prof = P.Profile(name="q", rx=4, rank=4, num_rb=133, fft=2048, qm=qm, padded_subcarriers=1664)
    cfg = prof.config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(0x256BA7C)
    lcap = 16
    xr = rng.normal(0.0, 0.7, (lcap, 14, g.pad)).astype(np.float16)
    xi = rng.normal(0.0, 0.7, (lcap, 14, g.pad)).astype(np.float16)
    noises = np.array([0.0125, 0.02, 0.05, 0.2], np.float16)
    ne = np.empty_like(xr)
    for l in range(lcap):
        ne[l] = noises[l % 4]
    ne[0, 0, :8] = 0.0  # clamp path: non-positive noise
    want = np.zeros(g.nl * g.qm * g.qam_plane, np.int16)
    oracle.or_qam_demod(cref(cfg), ptr(xr.view(np.uint16)), ptr(xi.view(np.uint16)), ptr(ne.view(np.uint16)),
                        ptr(want))
    got = ops.qam_demod(cfg, dev(xr), dev(xi), dev(ne))
    assert np.array_equal(host(got), want)


@pytest.mark.parametrize("rank", [1, 2, 3, 4])
def test_layer_demap_bit_exact(oracle, gpu_stream, rank):
    prof = P.Profile(name="l", rx=4, rank=rank, num_rb=273, fft=4096, qm=8, padded_subcarriers=3328)
    cfg = prof.config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(rank)
    llr = rng.integers(-2560, 2561, g.nl * g.qm * g.qam_plane, dtype=np.int16)
    want = np.zeros(g.qm * g.cw_stride, np.int16)
    oracle.or_layer_demap(cref(cfg), ptr(llr), ptr(want))
    out = torch.full((g.qm * g.cw_stride,), 0x5A5A, dtype=torch.int16, device="cuda")
    got = ops.layer_demap(cfg, dev(llr), out)
    assert np.array_equal(host(got), want)


@pytest.mark.parametrize("name,slots,restart", [("reference_fd16x64", 3, False), ("full_carrier_4rx", 2, True),
                                                ("mid_4rx_rank2", 2, False)])
def test_descramble_bit_exact(oracle, gpu_stream, name, slots, restart):
    cfg = PROFILES[name].config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(3)
    x = rng.integers(-2560, 2561, slots * g.qm * g.cw_stride, dtype=np.int16)
    x[:7] = -32768
    want = np.zeros_like(x)
    if restart:
        per = g.qm * g.cw_stride
        for s in range(slots):
            oracle.or_descramble(cref(cfg), 1, 0, ptr(x[s * per:]), ptr(want[s * per:]))
    else:
        oracle.or_descramble(cref(cfg), slots, 0, ptr(x), ptr(want))
    got = ops.descramble(cfg, dev(x), slots, restart_per_slot=restart)
    assert np.array_equal(host(got), want)


def test_descramble_host_sign_table_matches(lib, gpu_stream):
    cfg = P.REFERENCE_FD16X64.config()
    g = OracleGeom(cfg)
    lay = ops.layout(cfg)
    slots = 2
    sign = np.zeros(slots * g.qm * g.cw_stride, np.int16)
    assert lib.airan_descramble_mimo_BuildGoldSign(C.byref(cfg), C.byref(lay), slots,
                                                   sign.ctypes.data_as(C.POINTER(C.c_int16)), sign.size) == 0
    rng = np.random.default_rng(4)
    x = dev(rng.integers(-2560, 2561, sign.size, dtype=np.int16))
    a = ops.descramble(cfg, x, slots)
    b = ops.descramble(cfg, x, slots, sign=dev(sign))
    assert torch.equal(a, b)


def _plan_pair(oracle, tbs, G, qm, layers):
    return ops.tb_plan(tbs, G, qm, layers), or_tb_plan(oracle, tbs, G, qm, layers)


@pytest.mark.parametrize("name", ["full_carrier_4rx", "mid_4rx_rank2", "reference_fd16x64"])
def test_rate_dematch_bit_exact(oracle, gpu_stream, name):
    prof = PROFILES[name]
    cfg = prof.config()
    g = OracleGeom(cfg)
    slots = 1
    G = slots * g.qm * g.cw_symbols
    if prof.code_rate is None:
        plan = ops.reference_tb_plan(cfg, slots)
        from test_oracle_pins import _reference_plan
        oplan = _reference_plan(g.nl, slots)
        oplan.G = G
    else:
        tbs = ops.tbs(prof.num_rb, 12 * (14 - g.ndmrs), g.qm, g.nl, prof.code_rate)
        plan, oplan = _plan_pair(oracle, tbs, G, g.qm, g.nl)
    rng = np.random.default_rng(5)
    x = rng.integers(-2560, 2561, slots * g.qm * g.cw_stride, dtype=np.int16)
    want = np.zeros(plan.C * 68 * plan.Z, np.int16)
    oracle.or_rate_dematch(C.byref(oplan), g.nl, g.qm, g.cw_symbols, g.cw_stride, slots, ptr(x), ptr(want))
    got = ops.rate_dematch(cfg, dev(x), slots, plan)
    assert np.array_equal(host(got), want)


def _encoded_llrs(oracle, Z, ncb, snr_db, seed, kprime=None):
    rng = np.random.default_rng(seed)
    infos, lams = [], []
    for _ in range(ncb):
        info = rng.integers(0, 2, 22 * Z, dtype=np.uint8)
        if kprime is not None:
            info[kprime:] = 0
        cw = np.zeros(68 * Z, np.uint8)
        oracle.or_ldpc_encode(Z, ptr(info), ptr(cw))
        s2 = 10 ** (-snr_db / 10)
        y = (1.0 - 2.0 * cw) + rng.normal(0, np.sqrt(s2), cw.size)
        llr = np.clip(np.rint(2 * y / s2 * 256), -5120, 5120).astype(np.int16)
        llr[:2 * Z] = 0
        infos.append(info)
        lams.append(llr)
    return np.stack(infos), np.concatenate(lams)


@pytest.mark.parametrize("mode", [abi.LDPC_MODE_REF, abi.LDPC_MODE_FP16])
@pytest.mark.parametrize("rows,iters,early", [(46, 3, False), (46, 8, True), (6, 8, False), (8, 5, True)])
def test_ldpc_bit_exact(oracle, gpu_stream, mode, rows, iters, early):
    Z, ncb = 384, 6
    infos, lam = _encoded_llrs(oracle, Z, ncb, 2.5, 100 + rows)
    want_lam = lam.copy()
    want_bits = np.zeros(ncb * 22 * Z, np.int8)
    want_it = np.zeros(ncb, np.uint8)
    oracle.or_ldpc_decode(Z, ncb, ptr(want_lam), mode, iters, rows, 22 * Z, 1 if early else 0, ptr(want_bits),
                          ptr(want_it))
    bits, lam_out, it = ops.ldpc_decode(dev(lam), ncb, Z, 22 * Z, iters, rows, mode, early, want_lam=True)
    assert np.array_equal(host(bits), want_bits)
    assert np.array_equal(host(lam_out), want_lam)
    assert np.array_equal(host(it), want_it)
    if rows == 46 and iters == 8:
        assert np.array_equal(want_bits.reshape(ncb, -1), infos.astype(np.int8))


def test_ldpc_ref_hard_decision_near_zero(oracle, gpu_stream):
    """REF mode decides a bit on rint(clamp(posterior)) < 0, as the reference
    does (ldpc_decode_kernel.cpp:735-767): inputs of a few Q8.8 units drive
    thousands of posteriors into [-0.5, 0), which must decode to 0 on the
    device exactly as in the oracle."""
    Z, ncb = 384, 4
    rng = np.random.default_rng(12)
    lam = rng.integers(-3, 4, (ncb, 68 * Z)).astype(np.int16)
    lam[:, :2 * Z] = 0
    lam = lam.reshape(-1)
    for rows, iters, early in ((46, 1, False), (46, 2, False), (8, 2, False), (46, 3, True)):
        print(f"ref near-zero rows={rows} iters={iters} early={early}", flush=True)
        want_lam = lam.copy()
        want_bits = np.zeros(ncb * 22 * Z, np.int8)
        want_it = np.zeros(ncb, np.uint8)
        post = np.zeros(ncb * 68 * Z, np.float32)
        oracle.or_ldpc_decode_posterior(Z, ncb, ptr(want_lam), abi.LDPC_MODE_REF, iters, rows, 22 * Z,
                                        1 if early else 0, ptr(want_bits), ptr(want_it), ptr(post))
        p = post.reshape(ncb, 68 * Z)[:, :22 * Z].reshape(-1)
        near = (p >= -0.5) & (p < 0)
        assert near.sum() > 1000 and not want_bits[near].any()
        bits, lam_out, it = ops.ldpc_decode(dev(lam), ncb, Z, 22 * Z, iters, rows, abi.LDPC_MODE_REF, early,
                                            want_lam=True)
        assert np.array_equal(host(bits), want_bits)
        assert np.array_equal(host(lam_out), want_lam)
        assert np.array_equal(host(it), want_it)


@pytest.mark.parametrize("rows,early", [(46, False), (6, False), (8, True), (4, True)])
def test_ldpc_small_lifting_sizes(oracle, gpu_stream, rows, early):
    for Z in (24, 48, 96, 192):
        ncb = 3
        infos, lam = _encoded_llrs(oracle, Z, ncb, 3.0, Z + rows)
        want = np.zeros(ncb * 22 * Z, np.int8)
        want_lam = lam.copy()
        want_it = np.zeros(ncb, np.uint8)
        oracle.or_ldpc_decode(Z, ncb, ptr(want_lam), 1, 6, rows, 22 * Z, 1 if early else 0, ptr(want), ptr(want_it))
        bits, lam_out, it = ops.ldpc_decode(dev(lam), ncb, Z, 22 * Z, 6, rows, abi.LDPC_MODE_FP16, early,
                                            want_lam=True)
        assert np.array_equal(host(bits), want), Z
        assert np.array_equal(host(lam_out), want_lam), Z
        assert np.array_equal(host(it), want_it), Z


def test_ldpc_fp16_saturated_posteriors(oracle, gpu_stream):
    """Clean high-SNR blocks decoded for many iterations without early stop:
    the layered posteriors grow to their largest values (about 480 at 6 rows
    and 2,250 at 46 rows in the oracle; the parity columns of degree one or
    two cap every row's messages, so they stay far below the fp16 range) and
    the decoder must still follow the oracle's fp16 arithmetic there, where
    the rounding of the large messages matters most."""
    Z, ncb = 384, 4
    _, lam = _encoded_llrs(oracle, Z, ncb, 12.0, 77)
    for rows in (6, 46):
        want_lam = lam.copy()
        want_bits = np.zeros(ncb * 22 * Z, np.int8)
        want_it = np.zeros(ncb, np.uint8)
        oracle.or_ldpc_decode(Z, ncb, ptr(want_lam), abi.LDPC_MODE_FP16, 12, rows, 22 * Z, 0, ptr(want_bits),
                              ptr(want_it))
        bits, lam_out, it = ops.ldpc_decode(dev(lam), ncb, Z, 22 * Z, 12, rows, abi.LDPC_MODE_FP16, False,
                                            want_lam=True)
        assert np.array_equal(host(bits), want_bits), rows
        assert np.array_equal(host(lam_out), want_lam), rows
        assert np.array_equal(host(it), want_it), rows


def test_crc_check_matches_oracle(oracle, gpu_stream):
    tbs = 1081512
    plan = ops.tb_plan(tbs, 39312 * 32, 8, 4)
    rng = np.random.default_rng(6)
    ntb = 3
    bits = np.zeros((ntb, plan.C, plan.K), np.int8)
    payloads = []
    for t in range(ntb):
        pay = rng.integers(0, 2, tbs, dtype=np.uint8)
        r = oracle.or_crc(0, ptr(pay), tbs)
        tb = np.concatenate([pay, np.array([(r >> (23 - i)) & 1 for i in range(24)], np.uint8)])
        L = plan.Kprime - 24
        for c in range(plan.C):
            seg = tb[c * L:(c + 1) * L]
            rb = oracle.or_crc(1, ptr(np.ascontiguousarray(seg)), L)
            bits[t, c, :L] = seg
            bits[t, c, L:L + 24] = [(rb >> (23 - i)) & 1 for i in range(24)]
        payloads.append(pay)
    bits[1, 5, 17] ^= 1     # corrupts CB 5 of TB 1 (and TB 1)
    bits[2, 7, L + 3] ^= 1  # corrupts only the CB CRC of CB 7 of TB 2
    cb_ok, tb_ok, tb_bytes = ops.crc_check(plan, ntb, dev(bits.reshape(-1)), want_bytes=True)
    cb = host(cb_ok).reshape(ntb, plan.C)
    assert host(tb_ok).tolist() == [1, 0, 1]
    assert cb.sum() == ntb * plan.C - 2 and cb[1, 5] == 0 and cb[2, 7] == 0
    got = host(tb_bytes).reshape(ntb, -1)
    assert np.array_equal(np.unpackbits(got[0])[:tbs], payloads[0])
