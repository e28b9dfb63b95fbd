"""CPU checks of the synthetic transmitter (include/airan_pusch_tx.h): its CRC
and LDPC encoder agree with the oracle's, and a transmitted slot read back
through the oracle's receive stages (QAM demap -> layer demap -> descramble
-> rate dematch -> LDPC decode) returns the payload with passing CRCs, so the
transmitter is the inverse of the receive chain's bit-level conventions."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import OracleGeom, cref, or_tb_plan, ptr
from puschrx import abi, ops, synth
from puschrx import profiles as P


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_crc_matches_oracle(lib, oracle, kind):
    rng = np.random.default_rng(kind)
    for n in (1, 23, 24, 1000, 8424):
        bits = rng.integers(0, 2, n, dtype=np.uint8)
        assert lib.airan_crc_Compute(kind, ptr(bits), n) == oracle.or_crc(kind, ptr(bits), n)


@pytest.mark.parametrize("Z", [3, 24, 96, 384])
def test_ldpc_encoder_matches_oracle(lib, oracle, Z):
    rng = np.random.default_rng(Z)
    info = rng.integers(0, 2, 22 * Z, dtype=np.uint8)
    a = np.zeros(68 * Z, np.uint8)
    b = np.zeros(68 * Z, np.uint8)
    assert lib.airan_ldpc_bg1_Encode(Z, ptr(info), ptr(a)) == abi.OK
    oracle.or_ldpc_encode(Z, ptr(info), ptr(b))
    assert np.array_equal(a, b)
    assert oracle.or_ldpc_syndrome_ok(Z, ptr(a)) == 1


def _plan_for(prof):
    cfg = prof.config()
    g = OracleGeom(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    G = g.cw_symbols * g.qm
    return cfg, g, ops.tb_plan(tbs, G, g.qm, g.nl)


def _oracle_receive(oracle, cfg, g, plan, grid_slot, no_eff=0.01):
    """Perfect-detection readback of one slot through the oracle stages."""
    x_re = np.zeros((g.nl, 14, g.pad), np.float16)
    x_im = np.zeros_like(x_re)
    x_re[:, :, : g.used] = grid_slot.real
    x_im[:, :, : g.used] = grid_slot.imag
    ne = np.full_like(x_re, no_eff)
    llr = np.zeros(g.nl * g.qm * g.qam_plane, np.int16)
    oracle.or_qam_demod(cref(cfg), ptr(x_re.view(np.uint16)), ptr(x_im.view(np.uint16)), ptr(ne.view(np.uint16)),
                        ptr(llr))
    cw = np.zeros(g.qm * g.cw_stride, np.int16)
    oracle.or_layer_demap(cref(cfg), ptr(llr), ptr(cw))
    ds = np.zeros_like(cw)
    oracle.or_descramble(cref(cfg), 1, 0, ptr(cw), ptr(ds))
    op = or_tb_plan(oracle, plan.tbs, plan.G, g.qm, g.nl)
    lam = np.zeros(op.C * 68 * op.Z, np.int16)
    oracle.or_rate_dematch(C.byref(op), g.nl, g.qm, g.cw_symbols, g.cw_stride, 1, ptr(ds), ptr(lam))
    bits = np.zeros(op.C * op.K, np.int8)
    oracle.or_ldpc_decode(op.Z, op.C, ptr(lam), 1, 8, 46, op.K, 1, ptr(bits), None)
    return op, bits.reshape(op.C, op.K).astype(np.uint8), ds


@pytest.mark.parametrize("prof", [P.SISO_52PRB, P.MID_4RX_RANK2, P.FULL_CARRIER_4RX], ids=lambda p: p.name)
def test_tx_loopback_through_oracle(lib, oracle, prof):
    cfg, g, plan = _plan_for(prof)
    rng = np.random.default_rng(7)
    payload = synth.random_payload(rng, 1, plan.tbs)
    grid, coded = synth.tx_grid(cfg, plan, payload, want_coded=True)
    assert grid.shape == (1, g.nl, 14, g.used)
    # Data REs carry unit-power QAM; DMRS REs carry |r| = 1 on the layer's comb.
    data = grid[0][:, g.data_sym]
    assert abs(np.mean(np.abs(data) ** 2) - 1.0) < 0.02
    op, bits, ds = _oracle_receive(oracle, cfg, g, plan, grid[0])
    # Descrambled LLR signs are the unscrambled coded bits.
    per_cb = op.Kprime - op.crc_cb
    tb = np.concatenate([bits[r, :per_cb] for r in range(op.C)])
    assert np.array_equal(tb[: plan.tbs], payload[0])
    if op.crc_tb:
        assert oracle.or_crc(0 if op.crc_tb == 24 else 2, ptr(tb[: plan.tbs + op.crc_tb].copy()),
                             plan.tbs + op.crc_tb) == 0
    for r in range(op.C if op.crc_cb else 0):
        assert oracle.or_crc(1, ptr(bits[r, : op.Kprime].copy()), op.Kprime) == 0


def test_tx_dmrs_matches_oracle_generator(lib, oracle):
    """The grid's DMRS equals the oracle's DMRS sequence on each layer's comb."""
    prof = P.FULL_CARRIER_4RX
    cfg, g, plan = _plan_for(prof)
    payload = synth.random_payload(np.random.default_rng(1), 1, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)[0]
    ref_r = np.zeros(g.nl * g.ndmrs * g.ref_pad, np.uint16)
    ref_i = np.zeros_like(ref_r)
    oracle.or_dmrs_gen(cref(cfg), ptr(ref_r), ptr(ref_i))
    ref = (ref_r.view(np.float16).astype(np.float32) + 1j * ref_i.view(np.float16).astype(np.float32))
    ref = ref.reshape(g.nl, g.ndmrs, g.ref_pad)
    for l in range(g.nl):
        comb = (cfg.dmrs_ports[l] - 1000) // 2
        for d, s in enumerate(g.dmrs_sym):
            got = grid[l, s, comb::2][: g.n_dmrs_re]
            assert np.allclose(got, ref[l, d, : g.n_dmrs_re], atol=1e-3)
            assert np.all(grid[l, s, 1 - comb::2] == 0)


@pytest.mark.parametrize("prof", [P.SISO_52PRB, P.MID_4RX_RANK2], ids=lambda p: p.name)
def test_oracle_chain_decodes_a_fading_slot(lib, oracle, prof):
    """The CPU chain (the CPU baseline) decodes a transmitted slot."""
    from oracle_lib import OrCeParams
    cfg, g, plan = _plan_for(prof)
    rng = np.random.default_rng(3)
    payload = synth.random_payload(rng, 1, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    spec = synth.ChannelSpec(taps=4, delay_spread_s=50e-9, doppler_hz=5.0, rx_corr=0.3, snr_db=32.0)
    H = synth.channel(rng, cfg, spec, 1)
    iq = synth.receive(rng, cfg, grid, H, spec.snr_db, iq_format=abi.IQ_COMPLEX64)
    op = or_tb_plan(oracle, plan.tbs, plan.G, g.qm, g.nl)
    bits = np.zeros(op.C * op.K, np.int8)
    ok = np.zeros(1, np.uint8)
    prm = OrCeParams(spec.delay_spread_s, spec.doppler_hz, 30e3)
    assert oracle.or_chain_slot(cref(cfg), C.byref(op), C.byref(prm), 8, ptr(np.ascontiguousarray(iq[0])),
                                ptr(bits), ptr(ok)) == 0
    assert ok[0] == 1
    b = bits.reshape(op.C, op.K).astype(np.uint8)
    tb = np.concatenate([b[r, : op.Kprime - op.crc_cb] for r in range(op.C)])[: plan.tbs]
    assert np.array_equal(tb, payload[0])


def test_tx_dmrs_follows_slot_numbers(lib, oracle):
    """A two-slot transport block sent with slot numbers (5, 6): the data
    resource elements are those of the same TB sent with the config's slot
    number, and each slot's DMRS is the oracle's DMRS of its own slot number
    (c_init depends on it, mimo_dmrs_gen.cpp:38-45)."""
    prof = P.MID_4RX_RANK2
    cfg = prof.config()
    g = OracleGeom(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata * 2, g.qm, g.nl, prof.code_rate)
    plan = ops.tb_plan(tbs, 2 * g.cw_symbols * g.qm, g.qm, g.nl)
    payload = synth.random_payload(np.random.default_rng(3), 1, plan.tbs)
    base = synth.tx_grid(cfg, plan, payload, slots_per_tb=2)
    moved = synth.tx_grid(cfg, plan, payload, slots_per_tb=2, slot_numbers=[5, 6])
    data = [s for s in range(14) if s not in g.dmrs_sym]
    assert np.array_equal(base[:, :, data], moved[:, :, data])
    for k, sn in enumerate((5, 6)):
        c = type(cfg).from_buffer_copy(cfg)
        c.slot_number = sn
        ref_re = np.zeros(g.nl * g.ndmrs * g.ref_pad, np.uint16)
        ref_im = np.zeros_like(ref_re)
        oracle.or_dmrs_gen(cref(c), ptr(ref_re), ptr(ref_im))
        ref = (ref_re.view(np.float16).astype(np.float32) + 1j * ref_im.view(np.float16).astype(np.float32))
        ref = ref.reshape(g.nl, g.ndmrs, g.ref_pad)
        for l in range(g.nl):
            comb = (cfg.dmrs_ports[l] - 1000) // 2
            for d, sym in enumerate(g.dmrs_sym):
                got = moved[k, l, sym, comb::2][: g.n_dmrs_re]
                assert np.allclose(got, ref[l, d, : g.n_dmrs_re], atol=1e-3)
                assert not np.allclose(got, base[k, l, sym, comb::2][: g.n_dmrs_re], atol=1e-3)
