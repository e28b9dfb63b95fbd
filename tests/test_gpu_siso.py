"""The SISO receiver on the GPU: DMRS least squares with linear
frequency/time interpolation (channel_est/channel_est_ls_siso) and
single-tap zero forcing (equalization/equalize_zf_siso) against the oracle
bit for bit, and the plain SISO workload end to end (1 Rx, 52 PRB in a
1024-point FFT, one DMRS symbol, 256QAM at R 0.75, 8-tap Rayleigh at 100 ns
block-fading per slot, AWGN 23 dB, 8 slots, 8 LDPC iterations) through the
batched chain, equal to the op-by-op chain."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import OracleGeom, cref, ptr
from puschrx import abi, ops, synth
from puschrx import profiles as P
from puschrx.chain import ChainOptions, NativeChain, OpChain

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("prof,batch", [(P.SISO_CONFIG1, 3), (P.SISO_133PRB, 2)], ids=["config1", "133prb"])
def test_ls_siso_bit_exact(oracle, gpu_stream, prof, batch):
    cfg = prof.config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(7)
    plane = 14 * g.pad
    y = (rng.normal(size=(batch, plane)) + 1j * rng.normal(size=(batch, plane))) * 0.6
    y[:, :].reshape(batch, 14, g.pad)[:, :, g.used:] = 0  # re-demap zero tail
    yr = y.real.astype(np.float16).view(np.uint16)
    yi = y.imag.astype(np.float16).view(np.uint16)
    refs = []
    for s in range(batch):
        c = prof.config(slot_number=s)
        rr = np.zeros(g.ndmrs * g.ref_pad, np.uint16)
        ri = np.zeros_like(rr)
        oracle.or_dmrs_gen(cref(c), ptr(rr), ptr(ri))
        refs.append((rr, ri))
    ref_re = np.concatenate([r[0] for r in refs])
    ref_im = np.concatenate([r[1] for r in refs])
    for per_slot in (False, True):
        print(f"{prof.name} ls_siso batch {batch} ref_per_slot {per_slot}", flush=True)
        hr, hi = ops.ls_siso(cfg, dev(yr.view(np.float16)), dev(yi.view(np.float16)),
                             dev(ref_re.view(np.float16)), dev(ref_im.view(np.float16)), batch=batch,
                             ref_per_slot=per_slot)
        got_r, got_i = host(hr).view(np.uint16).reshape(batch, plane), host(hi).view(np.uint16).reshape(batch, plane)
        for s in range(batch):
            rr, ri = refs[s] if per_slot else refs[0]
            wr = np.zeros(plane, np.uint16)
            wi = np.zeros_like(wr)
            oracle.or_ls_siso(cref(cfg), ptr(np.ascontiguousarray(yr[s])), ptr(np.ascontiguousarray(yi[s])),
                              ptr(rr), ptr(ri), ptr(wr), ptr(wi))
            assert np.array_equal(got_r[s], wr), s
            assert np.array_equal(got_i[s], wi), s


def test_ls_siso_rejects_mimo(gpu_stream):
    lib = abi.load_library()
    cfg = P.MID_4RX_RANK2.config()
    a = abi.ChannelEstLsSisoOpArgsV1()
    a.abi_version, a.struct_size = abi.ABI_VERSION, C.sizeof(a)
    a.rx_grid_re, a.rx_grid_im, a.dmrs_ref_re, a.dmrs_ref_im = 0x1000, 0x2000, 0x3000, 0x4000
    a.h_re, a.h_im, a.config, a.batch, a.stream = 0x5000, 0x6000, C.pointer(cfg), 1, 0x1
    assert lib.airan_channel_est_ls_siso_ValidateOpArgs(C.byref(a)) == abi.UNSUPPORTED_PROFILE
    cfg1 = P.SISO_CONFIG1.config()
    a.config = C.pointer(cfg1)
    assert lib.airan_channel_est_ls_siso_ValidateOpArgs(C.byref(a)) == abi.OK
    a.h_re = a.rx_grid_re
    assert lib.airan_channel_est_ls_siso_ValidateOpArgs(C.byref(a)) == abi.INVALID_ARGUMENT


def test_zf_siso_bit_exact(oracle, gpu_stream):
    rng = np.random.default_rng(8)
    n = 3 * 14 * 640 + 7
    h = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 0.7
    h[:100] *= 1e-3
    h[100:110] = 0.0
    y = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 0.7
    arrs = [a.astype(np.float16) for a in (y.real, y.imag, h.real, h.imag, rng.uniform(1e-4, 0.2, n))]
    xr, xi, ne = ops.zf_siso(*[dev(a) for a in arrs])
    want = [np.zeros(n, np.uint16) for _ in range(3)]
    oracle.or_zf_siso(n, *[ptr(a.view(np.uint16)) for a in arrs], *[ptr(w) for w in want])
    for got, w in zip((xr, xi, ne), want):
        assert np.array_equal(host(got).view(np.uint16), w)


def _plan(prof):
    cfg = prof.config()
    g = ops.Geometry(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    return cfg, g, ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl)


def test_siso_config1_chain(gpu_stream):
    """8 slots of the plain SISO workload at 23 dB: the batched chain with
    the LS/ZF receiver equals the op-by-op chain on every slot (H, codeword
    LLRs, bits, CRCs), and the BLER is reported."""
    prof = P.SISO_CONFIG1
    cfg, g, plan = _plan(prof)
    assert (cfg.rnti, cfg.data_scrambling_id) == (0x1234, 0)
    rng = np.random.default_rng(0x5150)
    S = 8
    spec = synth.ChannelSpec(taps=8, delay_spread_s=100e-9, doppler_hz=0.0, rx_corr=0.0, snr_db=23.0)
    sn = np.arange(S, dtype=np.uint16)
    payload = synth.random_payload(rng, S, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload, slot_numbers=sn)
    H = synth.channel(rng, cfg, spec, S)
    iq = synth.receive(rng, cfg, grid, H, spec.snr_db, iq_format=abi.IQ_COMPLEX64)
    opts = ChainOptions(ldpc_iterations=8, early_stop=False, iq_format=abi.IQ_COMPLEX64,
                        receiver=abi.CHAIN_RX_LS_ZF)
    iq_dev = torch.from_numpy(iq).cuda()
    sn_dev = torch.from_numpy(sn.astype(np.int16)).cuda()
    print("config 1: native chain (LS + linear interpolation, ZF), 8 slots", flush=True)
    nat = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True)
    assert not nat.outputs().noise_ce
    nat.enqueue(iq_dev, S, sn_dev)
    plane = 14 * g.pad
    h_nat = nat.fetch("h_re", torch.float16, S * plane)
    cw_nat = nat.fetch("cw_llr", torch.int16, S * g.qm * g.cw_stride)
    bits = nat.fetch("bits", torch.int8, S * plan.C * plan.K)
    tb_ok = nat.fetch("tb_ok", torch.uint8, S)
    cb_ok = nat.fetch("cb_ok", torch.uint8, S * plan.C)
    torch.cuda.synchronize()
    h_nat, cw_nat, bits, tb_ok, cb_ok = (t.cpu().numpy() for t in (h_nat, cw_nat, bits, tb_ok, cb_ok))
    nat.close()
    op = OpChain(cfg, plan, opts)
    per_cb = plan.Kprime - plan.crc_cb
    for s in range(S):
        print(f"config 1: op chain slot {s}", flush=True)
        r = op.run_slot(iq_dev[s].contiguous(), keep=True, slot_number=s)
        torch.cuda.synchronize()
        assert np.array_equal(h_nat.reshape(S, -1)[s].view(np.uint16), r["h_re"].cpu().numpy().view(np.uint16))
        assert np.array_equal(cw_nat.reshape(S, -1)[s], r["cw_nr"].cpu().numpy())
        assert np.array_equal(bits.reshape(S, -1)[s], r["bits"].cpu().numpy())
        assert np.array_equal(cb_ok.reshape(S, -1)[s], r["cb_ok"].cpu().numpy())
        assert int(tb_ok[s]) == int(r["tb_ok"].item())
        if tb_ok[s]:
            dec = bits.reshape(S, plan.C, plan.K)[s, :, :per_cb].reshape(-1)[: plan.tbs]
            assert np.array_equal(dec, payload[s])
    op.close()
    bler = 1.0 - float(tb_ok.mean())
    print(f"config 1: BLER {bler:.3f} over {S} slots at {spec.snr_db} dB (C = {plan.C}, TBS = {plan.tbs})")


def test_siso_config1_snr_sweep(gpu_stream):
    """BLER of the LS/ZF receiver over SNR on the config-1 channel (16 slots
    per point), beside the LMMSE receiver (LMMSE CE and detection) on the same
    slots: LS interpolation and ZF cost SNR, the decoder still works."""
    prof = P.SISO_CONFIG1
    cfg, g, plan = _plan(prof)
    S = 16
    sn = (np.arange(S) % 20).astype(np.uint16)
    sn_dev = torch.from_numpy(sn.astype(np.int16)).cuda()
    chains = {}
    for name, rx in (("ls_zf", abi.CHAIN_RX_LS_ZF), ("lmmse", abi.CHAIN_RX_LMMSE)):
        opts = ChainOptions(ldpc_iterations=8, early_stop=False, iq_format=abi.IQ_COMPLEX64, receiver=rx,
                            tau_s=100e-9, fd_hz=0.0)
        chains[name] = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=False)
    bler = {name: {} for name in chains}
    iq_dev = torch.empty((S, g.nr, g.samples, 2), dtype=torch.float32, device="cuda")
    for snr in (23.0, 27.0, 31.0, 35.0):
        rng = np.random.default_rng(int(snr) * 7919)
        spec = synth.ChannelSpec(taps=8, delay_spread_s=100e-9, doppler_hz=0.0, rx_corr=0.0, snr_db=snr)
        payload = synth.random_payload(rng, S, plan.tbs)
        grid = synth.tx_grid(cfg, plan, payload, slot_numbers=sn)
        H = synth.channel(rng, cfg, spec, S)
        iq_dev.copy_(torch.from_numpy(synth.receive(rng, cfg, grid, H, snr, iq_format=abi.IQ_COMPLEX64)))
        for name, ch in chains.items():
            ch.enqueue(iq_dev, S, sn_dev)
            ok = ch.fetch("tb_ok", torch.uint8, S)
            torch.cuda.synchronize()
            bler[name][snr] = 1.0 - float(ok.cpu().numpy().mean())
        print(f"config 1 sweep: {snr:.0f} dB  BLER LS/ZF {bler['ls_zf'][snr]:.3f}  LMMSE {bler['lmmse'][snr]:.3f}",
              flush=True)
    for ch in chains.values():
        ch.close()
    assert bler["ls_zf"][35.0] <= 0.25
    assert bler["lmmse"][35.0] <= bler["ls_zf"][35.0]
