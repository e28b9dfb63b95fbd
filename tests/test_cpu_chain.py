"""The FP32 CPU chain timed as the CPU baseline (oracle/cpu_chain.cpp): it
decodes what the transmitter sent, and its threads do not change its bits."""
import numpy as np
import pytest

from oracle_lib import CpuChain, OrCeParams, host_cpu, load_oracle, or_tb_plan
from puschrx import abi, ops, synth
from puschrx import profiles as P


def _grant(prof):
    cfg = prof.config()
    g = ops.Geometry(cfg)
    plan = ops.tb_plan(ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate), g.cw_symbols * g.qm, g.qm,
                       g.nl)
    return cfg, g, plan


@pytest.mark.skipif(host_cpu()["isa"] is None, reason="host CPU lacks x86-64-v3")
def test_cpu_chain_decodes_mid_rank2():
    cfg, g, plan = _grant(P.MID_4RX_RANK2)
    spec = synth.ChannelSpec(taps=12, delay_spread_s=300e-9, doppler_hz=50.0, rx_corr=0.5)
    rng = np.random.default_rng(11)
    n = 3
    payload = synth.random_payload(rng, n, plan.tbs)
    iq = synth.receive(rng, cfg, synth.tx_grid(cfg, plan, payload), synth.channel(rng, cfg, spec, n), 30.0,
                       iq_format=abi.IQ_COMPLEX64)
    op = or_tb_plan(load_oracle(), plan.tbs, plan.G, g.qm, g.nl)
    cc = CpuChain(cfg, op, OrCeParams(spec.delay_spread_s, spec.doppler_hz, 30e3), 8)
    try:
        bits1, ok1 = cc.run(iq, 1)
        bits2, ok2 = cc.run(iq, 3)
    finally:
        cc.close()
    assert np.array_equal(bits1, bits2) and np.array_equal(ok1, ok2)
    per_cb = plan.Kprime - plan.crc_cb
    for s in range(n):
        assert ok1[s] == 1
        assert np.array_equal(bits1[s, :, :per_cb].reshape(-1)[: plan.tbs].astype(np.uint8), payload[s])


@pytest.mark.skipif(host_cpu()["isa"] != "v4", reason="host CPU lacks x86-64-v4")
def test_cpu_chain_isa_levels_agree():
    """The v3 and v4 builds decode the same slot to the same outcome."""
    cfg, g, plan = _grant(P.MID_4RX_RANK2)
    spec = synth.ChannelSpec(taps=4, delay_spread_s=100e-9, doppler_hz=10.0, rx_corr=0.3)
    rng = np.random.default_rng(12)
    payload = synth.random_payload(rng, 1, plan.tbs)
    iq = synth.receive(rng, cfg, synth.tx_grid(cfg, plan, payload), synth.channel(rng, cfg, spec, 1), 30.0,
                       iq_format=abi.IQ_COMPLEX64)
    op = or_tb_plan(load_oracle(), plan.tbs, plan.G, g.qm, g.nl)
    prm = OrCeParams(spec.delay_spread_s, spec.doppler_hz, 30e3)
    oks = []
    for isa in ("v3", "v4"):
        cc = CpuChain(cfg, op, prm, 8, isa=isa)
        try:
            oks.append(cc.run(iq, 1)[1])
        finally:
            cc.close()
    assert oks[0][0] == 1 and oks[1][0] == 1
