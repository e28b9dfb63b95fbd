"""The transmitter and channel on the device (airan_pusch_txgpu_*): the
encoder equals the host transmitter bit for bit (coded bits and resource
grid), the channel and OFDM modulation equal the numpy channel model on the
same gains, the noise has the stated variance, and slots generated on the
device decode through the chain."""
import numpy as np
import pytest

from puschrx import abi, ops, synth
from puschrx import profiles as P
from puschrx.txgpu import GpuTransmitter, device_channel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def plan_for(prof, slots_per_tb=1):
    cfg = prof.config()
    g = ops.Geometry(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    return cfg, g, ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl)


ENC_CASES = [P.FULL_CARRIER_4RX, P.MID_4RX_RANK2, P.SISO_52PRB, P.SISO_CONFIG1]


@pytest.mark.parametrize("prof", ENC_CASES, ids=lambda p: p.name)
def test_gpu_encoder_matches_host(gpu_stream, prof):
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(31)
    n = 3
    payload = synth.random_payload(rng, n, plan.tbs)
    sn = np.array([0, 7, 19], np.uint16)
    want_grid, want_coded = synth.tx_grid(cfg, plan, payload, want_coded=True, slot_numbers=sn)
    print(f"{prof.name}: {n} TBs of {plan.tbs} bits, C = {plan.C}, Z = {plan.Z}", flush=True)
    tx = GpuTransmitter(cfg, plan, max_tbs=n)
    grid, coded = tx.encode(torch.from_numpy(payload).cuda(), torch.from_numpy(sn.astype(np.int16)).cuda(),
                            want_coded=True)
    torch.cuda.synchronize()
    tx.close()
    assert np.array_equal(coded.cpu().numpy(), want_coded)
    got = grid.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want_grid.view(np.uint32))


def test_gpu_encoder_multi_slot_tb(gpu_stream):
    """A TB over several slots (the reference profile's layout, 2 slots here):
    scrambling runs on across the slots and code blocks span them."""
    prof = P.MID_4RX_RANK2
    cfg = prof.config()
    g = ops.Geometry(cfg)
    T = 2
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata * T, g.qm, g.nl, prof.code_rate)
    plan = ops.tb_plan(tbs, T * g.cw_symbols * g.qm, g.qm, g.nl)
    rng = np.random.default_rng(32)
    payload = synth.random_payload(rng, 2, plan.tbs)
    want_grid, want_coded = synth.tx_grid(cfg, plan, payload, slots_per_tb=T, want_coded=True)
    tx = GpuTransmitter(cfg, plan, max_tbs=2, slots_per_tb=T)
    grid, coded = tx.encode(torch.from_numpy(payload).cuda(), want_coded=True)
    torch.cuda.synchronize()
    tx.close()
    assert np.array_equal(coded.cpu().numpy(), want_coded)
    assert np.array_equal(grid.cpu().numpy().view(np.uint32), want_grid.view(np.uint32))


@pytest.mark.parametrize("prof", [P.FULL_CARRIER_4RX, P.SISO_52PRB], ids=lambda p: p.name)
def test_gpu_channel_matches_numpy(gpu_stream, prof):
    """Noiseless: the device channel + OFDM modulation against numpy on the
    same gains (fp32 against fp64: NMSE <= -80 dB), and the int16 Q8.8
    output is the rounded float output."""
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(33)
    S = 2
    payload = synth.random_payload(rng, S, plan.tbs)
    spec = synth.ChannelSpec(taps=16, delay_spread_s=300e-9, doppler_hz=300.0, rx_corr=0.5)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    gains, delays = device_channel(gen, cfg, spec, S)
    tx = GpuTransmitter(cfg, plan, max_tbs=S)
    grid = tx.encode(torch.from_numpy(payload).cuda())
    iq = tx.channel(grid, gains, delays, float("inf"), seed=1, iq_format=abi.IQ_COMPLEX64)
    iq16 = tx.channel(grid, gains, delays, float("inf"), seed=1, iq_format=abi.IQ_INT16_Q8_8)
    torch.cuda.synchronize()
    tx.close()
    gn, dn, gr = gains.cpu().numpy().astype(np.complex128), delays.cpu().numpy().astype(np.float64), grid.cpu().numpy()
    freqs = (np.arange(g.used) - g.used // 2) * spec.scs_hz
    got = iq.cpu().numpy()
    for s in range(S):
        steer = np.exp(-2j * np.pi * freqs[None, :] * dn[s][:, None])
        H = np.einsum("rltk,tu->rlku", gn[s], steer)
        y = np.einsum("rlku,lku->rku", H, gr[s].astype(np.complex128))
        want = synth.ofdm_modulate(cfg, y.astype(np.complex64)).astype(np.complex128)
        have = got[s, ..., 0] + 1j * got[s, ..., 1]
        nmse = 10 * np.log10(np.sum(np.abs(have - want) ** 2) / np.sum(np.abs(want) ** 2))
        print(f"{prof.name} slot {s}: IQ NMSE vs numpy {nmse:.1f} dB", flush=True)
        assert nmse <= -80.0
    q = np.clip(np.rint(got * 256.0), -32768, 32767).astype(np.int16)
    assert np.array_equal(iq16.cpu().numpy(), q)


def test_gpu_channel_noise(gpu_stream):
    """A zero grid gives pure noise: variance 10^(-snr/10) per complex sample
    (within 1 % over 1.7 M samples), zero mean, uncorrelated re / im, the
    same seed the same noise and another seed other noise."""
    prof = P.FULL_CARRIER_4RX
    cfg, g, plan = plan_for(prof)
    tx = GpuTransmitter(cfg, plan, max_tbs=1)
    grid = torch.zeros((7, g.nl, 14, g.used), dtype=torch.complex64, device="cuda")
    gains = torch.zeros((7, g.nr, g.nl, 1, 14), dtype=torch.complex64, device="cuda")
    delays = torch.zeros((7, 1), dtype=torch.float32, device="cuda")
    a = tx.channel(grid, gains, delays, 20.0, seed=11)
    b = tx.channel(grid, gains, delays, 20.0, seed=11)
    c = tx.channel(grid, gains, delays, 20.0, seed=12)
    torch.cuda.synchronize()
    tx.close()
    x = a.cpu().numpy().astype(np.float64)
    z = x[..., 0] + 1j * x[..., 1]
    var = np.mean(np.abs(z) ** 2)
    print(f"noise: {z.size} samples, variance {var:.6f} (want 0.01), mean {np.abs(z.mean()):.2e}", flush=True)
    assert abs(var / 0.01 - 1.0) < 0.01
    assert abs(z.mean()) < 5e-4
    assert abs(np.mean(x[..., 0] * x[..., 1])) < 5e-5
    assert abs(np.mean(x[..., 0] ** 2) / np.mean(x[..., 1] ** 2) - 1.0) < 0.01
    assert torch.equal(a, b) and not torch.equal(a, c)


def test_gpu_generated_slots_decode(gpu_stream):
    """16 full-carrier slots generated on the device (42 dB, 300 ns / 100 Hz
    channel) decode through the batched chain to the payload."""
    from puschrx.chain import ChainOptions, NativeChain
    prof = P.FULL_CARRIER_4RX
    cfg, g, plan = plan_for(prof)
    S = 16
    spec = synth.ChannelSpec(taps=16, delay_spread_s=300e-9, doppler_hz=100.0, rx_corr=0.5)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    payload = torch.randint(0, 2, (S, plan.tbs), dtype=torch.uint8, device="cuda", generator=gen)
    tx = GpuTransmitter(cfg, plan, max_tbs=S)
    gains, delays = device_channel(gen, cfg, spec, S)
    iq = tx.channel(tx.encode(payload), gains, delays, 42.0, seed=77)
    opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, iq_format=abi.IQ_COMPLEX64)
    chain = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=False)
    chain.enqueue(iq, S)
    ok = chain.fetch("tb_ok", torch.uint8, S)
    bits = chain.fetch("bits", torch.int8, S * plan.C * plan.K)
    torch.cuda.synchronize()
    chain.close()
    tx.close()
    per_cb = plan.Kprime - plan.crc_cb
    b = bits.cpu().numpy().reshape(S, plan.C, plan.K)[:, :, :per_cb].reshape(S, -1)[:, : plan.tbs].astype(np.uint8)
    print(f"device-generated slots: tb_ok {int(ok.sum())}/{S}", flush=True)
    assert int(ok.sum()) == S
    assert np.array_equal(b, payload.cpu().numpy())
