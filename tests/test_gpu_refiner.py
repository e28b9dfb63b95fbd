"""Learned CE refiner on the GPU: forward output, loss and every parameter
gradient against the CPU oracle (oracle/or_refiner.c with the device's bf16
rounding points), and training on device-synthesised slots."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle_lib import load_oracle, ptr
from puschrx import profiles as P
from puschrx.refiner import Refiner, TrainSpec
from test_refiner_oracle import CIN, COUT, refiner_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oracle():
    return load_oracle()


GRAD_CASES = [(1, 40, None, None), (2, 70, None, None), (1, 70, [1, 2, 4, 8, 16, 1], None),
              (2, 150, [2, 1, 16, 3, 5, 16], None), (2, 90, [1, 2, 4, 8, 16, 1], [0.3, 0.02])]


@pytest.mark.parametrize("S,W,dil,scale", GRAD_CASES,
                         ids=["w40", "w70", "w70_dil1-16", "w150_dil_mixed", "w90_dil_scaled"])
def test_refiner_gradients_match_oracle(oracle, gpu_stream, S, W, dil, scale):
    rng = np.random.default_rng(11 + W)
    cfg = P.FULL_CARRIER_4RX.config()
    params = refiner_params(oracle, rng)
    x = torch.from_numpy(rng.normal(0, 1, (S, 14, W, 32)).astype(np.float32)).to(torch.bfloat16)
    xf = x.float().numpy()
    t = (xf + rng.normal(0, 0.3, xf.shape)).astype(np.float32)
    ref = Refiner(cfg, max_slots=S, params=params, dilation=dil)
    sc = None if scale is None else np.asarray(scale, np.float32)
    if sc is not None:
        ref.set_slot_scale(sc)
    print(f"refiner gradients S={S} W={W} dilation {ref.dilation} residual scale {scale}", flush=True)
    grad, loss = ref.gradients(x.cuda(), torch.from_numpy(t).cuda(), S, W)
    n = S * 14 * W * 32
    out = ref.fetch("out", torch.float32, n)
    torch.cuda.synchronize()
    grad = grad.cpu().numpy().astype(np.float64)
    want_g = np.zeros(params.size)
    want_l = np.zeros(1)
    want_o = np.zeros(n, np.float32)
    d = np.asarray(ref.dilation, np.int32)
    assert oracle.or_refiner_grad_full(ptr(params), ptr(np.ascontiguousarray(xf)), ptr(t), S, W, 1, ptr(d),
                                       None if sc is None else ptr(sc), ptr(want_g), ptr(want_l), ptr(want_o)) == 0
    # fp32 accumulation on the device against double in the oracle flips an
    # occasional bf16 rounding of an activation (one bf16 ulp, 2^-8), which
    # the later layers carry: agreement is to ~1e-4 in norm, not bit-exact.
    o = out.cpu().numpy()
    assert np.linalg.norm(o - want_o) <= 1e-3 * np.linalg.norm(want_o)
    assert float(loss.item()) == pytest.approx(want_l[0], rel=1e-4)
    for i in range(6):
        off = oracle.or_refiner_param_offset(i)
        nw = 9 * COUT[i] * CIN[i]
        for a, b, what in ((off, off + nw, "weights"), (off + nw, off + nw + COUT[i], "bias")):
            err = np.linalg.norm(grad[a:b] - want_g[a:b]) / np.linalg.norm(want_g[a:b])
            assert err < 1e-2, (i, what, err)
    ref.close()


def test_refiner_training_beats_ls_interpolation(gpu_stream):
    """Training on slots of one channel (300 ns, 100 Hz, Rx correlation 0.5,
    34 dB) teaches the network something its linear-interpolation input
    does not do; evaluated on a held-out batch of the same channel."""
    cfg = P.FULL_CARRIER_4RX.config()
    spec = TrainSpec(snr_db=(34.0, 34.0), delay_spread_s=(300e-9, 300e-9), doppler_hz=(100.0, 100.0),
                     rx_corr=(0.5, 0.5))
    ref = Refiner(cfg, max_slots=8, seed=3)

    def held_out():
        ref.synth(np.random.default_rng(99), 4, spec, seed=99)
        return ref.evaluate(4)

    before = held_out()
    print("before", before, flush=True)
    ref.train(np.random.default_rng(5), steps=800, slots=8, spec=spec, lr=3e-4, log_every=100)
    after = held_out()
    print("after", after, flush=True)
    assert abs(after["ls_interp"] - before["ls_interp"]) < 1e-6
    assert after["refiner"] < before["refiner"] - 1.0
    assert after["refiner"] < after["ls_interp"] - 0.4
    ref.close()


def test_refiner_lmmse_input_mode(gpu_stream):
    """Input mode 1: the batch input is the LMMSE estimate of the synthesised
    slots (the chain's filter: 300 ns, 100 Hz, rank 384) rounded to bf16,
    and a fresh network (zero last layer) returns the estimate itself (the
    residual is added to it in fp32, not to its bf16 copy)."""
    from puschrx import ops

    cfg = P.FULL_CARRIER_4RX.config()
    spec = TrainSpec(snr_db=(34.0, 34.0), delay_spread_s=(300e-9, 300e-9), doppler_hz=(100.0, 100.0),
                     rx_corr=(0.5, 0.5))
    filt = ops.LmmseFilter(cfg, tau_s=300e-9, fd_hz=100.0, scs_hz=30e3, max_rank=384)
    ref = Refiner(cfg, max_slots=4, seed=5, input_mode=1, estimator=filt)
    ref.synth(np.random.default_rng(21), 4, spec, seed=21)
    res = ref.evaluate(4, lmmse=filt)
    print(res, flush=True)
    n = 4 * 14 * ref.g.used * 32
    out = ref.fetch("out", torch.float32, n)
    assert torch.equal(out, ref.lmmse_grid(4, filt))
    # bf16 rounding of the estimate costs well under a dB at 34 dB.
    assert res["lmmse"] - 0.05 <= res["input"] <= res["lmmse"] + 1.0
    assert res["input"] < -38.0
    ref.close()
    filt.close()


def test_refiner_against_lmmse(gpu_stream):
    """The committed refiner against the method without it -- LMMSE with the
    chain's prior (300 ns, 100 Hz, rank 384) -- on held-out slots (seeds used
    neither in training nor in its evaluation):

      * NMSE on draws of the training family (delay spread 30-500 ns,
        Doppler 0-500 Hz, Rx correlation 0-0.7) at 38 dB, where the fixed
        prior fits only some slots: the refiner is better (recorded at
        training time: 1.2 dB);
      * NMSE on the workload channel (300 ns, 100 Hz, 0.5) at 38 dB, where the
        prior is exact: the refiner is worse by at most 1.5 dB (recorded:
        0.4-0.5 dB);
      * BLER of the batched chain with and without the refiner on 16
        device-generated workload slots at 40 dB: no better, no more than
        4 of 16 TBs worse (recorded on 64 slots: 0.016 for both)."""
    from puschrx import abi, ops
    from puschrx.chain import ChainOptions, NativeChain
    from puschrx.refiner import WEIGHTS
    from puschrx.txgpu import GpuTransmitter, device_channel_family

    cfg = P.FULL_CARRIER_4RX.config()
    g = ops.Geometry(cfg)
    prof = P.FULL_CARRIER_4RX
    plan = ops.tb_plan(ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate), g.cw_symbols * g.qm, g.qm,
                       g.nl)
    meta = Refiner.load_meta(WEIGHTS)
    assert meta["input_mode"] == 1 and meta["residual_scale"] == 1
    filt = ops.LmmseFilter(cfg, tau_s=300e-9, fd_hz=100.0, scs_hz=30e3, max_rank=384)
    ref = Refiner.from_file(cfg, 16, WEIGHTS, estimator=filt)
    spec = TrainSpec(snr_db=(0.0, 46.0))
    ref.synth(np.random.default_rng(4711), 8, spec, seed=4711, snr_db=38.0)
    fam = ref.evaluate(8, lmmse=filt)
    ref.synth(np.random.default_rng(4712), 8, spec, seed=4712, snr_db=38.0, delay_spread=300e-9, doppler=100.0,
              rx_corr=0.5)
    wl = ref.evaluate(8, lmmse=filt)
    print(f"NMSE at 38 dB, training family: refiner {fam['refiner']:.2f} dB, LMMSE {fam['lmmse']:.2f} dB "
          f"(gain {fam['lmmse'] - fam['refiner']:+.2f} dB)", flush=True)
    print(f"NMSE at 38 dB, workload channel: refiner {wl['refiner']:.2f} dB, LMMSE {wl['lmmse']:.2f} dB "
          f"(gain {wl['lmmse'] - wl['refiner']:+.2f} dB)", flush=True)
    assert fam["refiner"] < fam["lmmse"] - 0.3
    assert wl["refiner"] < wl["lmmse"] + 1.5

    S = 16
    opts = ChainOptions(tau_s=300e-9, fd_hz=100.0, ce_rank=384, iq_format=abi.IQ_COMPLEX64)
    chain = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=False)
    tx = GpuTransmitter(cfg, plan, max_tbs=S)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(4713)
    payload = torch.randint(0, 2, (S, plan.tbs), dtype=torch.uint8, device="cuda", generator=gen)
    gains, delays = device_channel_family(gen, cfg, S, np.full(S, 300e-9), np.full(S, 100.0), np.full(S, 0.5))
    iq = tx.channel(tx.encode(payload), gains, delays, 40.0, seed=4714)
    bler = {}
    for leg in ("lmmse", "refiner"):
        chain.set_refiner(ref if leg == "refiner" else None, 0.0, 60.0)
        chain.enqueue(iq, S)
        ok = chain.fetch("tb_ok", torch.uint8, S)
        torch.cuda.synchronize()
        bler[leg] = 1.0 - float(ok.float().mean())
    chain.set_refiner(None)
    print(f"BLER at 40 dB, workload channel, {S} slots: refiner {bler['refiner']:.3f}, LMMSE {bler['lmmse']:.3f}",
          flush=True)
    assert bler["refiner"] <= bler["lmmse"] + 4.0 / S
    chain.close()
    tx.close()
    ref.close()
    filt.close()


ENQUEUE_CASES = [("full_273prb", None, 3, [1, 0, 1]), ("narrow_20prb_pad512", 20, 2, None)]


@pytest.mark.parametrize("name,num_rb,S,keep", ENQUEUE_CASES, ids=[c[0] for c in ENQUEUE_CASES])
def test_refiner_enqueue_grid_matches_forward(oracle, gpu_stream, name, num_rb, S, keep):
    """Inference in input mode 1 (airan_ce_refiner_EnqueueWithNoise: the
    estimate grid in, the refined grid written in place by the last layer)
    against the fp32 route whose output the oracle pins (the forward of
    airan_ce_refiner_Gradients on the same input, rounded to fp16 into the
    grid): bit-exact for the kept slots, zero beyond the used subcarriers,
    untouched for the others.  The input is bf16-exact so that both routes'
    residual bases are the same; the noise variances have exact square
    roots (the residual scales 0.5, 0.25, 0.125)."""
    import dataclasses

    from puschrx import abi, ops

    prof = P.FULL_CARRIER_4RX
    if num_rb is not None:
        prof = dataclasses.replace(prof, num_rb=num_rb, fft=1024, padded_subcarriers=512)
    cfg = prof.config()
    rng = np.random.default_rng(77 + S)
    params = refiner_params(oracle, rng)
    ref = Refiner(cfg, max_slots=S, params=params, input_mode=1, dilation=[1, 2, 4, 8, 16, 1], residual_scale=1)
    g = ref.g
    used, pad = g.used, g.pad
    print(f"refiner enqueue {name}: S={S} used={used} pad={pad} keep={keep}", flush=True)
    def exact_grid():
        # [S][16 links][14][pad], values exact in bf16 and in fp16.
        v = rng.normal(0, 0.5, (S, 16, 14, pad)).astype(np.float32)
        v[np.abs(v) < 2.0 ** -10] = 0.0
        return torch.from_numpy(v).to(torch.bfloat16).float()

    grid_re, grid_im = exact_grid(), exact_grid()
    in_re, in_im = grid_re.half().cuda(), grid_im.half().cuda()
    h_re, h_im = in_re.clone(), in_im.clone()
    scales = np.array([0.5, 0.25, 0.125][:S], np.float32)
    noise = torch.from_numpy(np.repeat((scales ** 2)[:, None], 4, 1)).half().cuda()
    keep_t = None if keep is None else torch.tensor(keep, dtype=torch.uint8, device="cuda")
    ops._check("refiner EnqueueWithNoise", ref.lib.airan_ce_refiner_EnqueueWithNoise(
        ref.handle, C_vp(h_re), C_vp(h_im), C_vp(noise), S, None if keep_t is None else C_vp(keep_t),
        C_vp(h_re), C_vp(h_im), C_vp_int(ops.stream_ptr())))

    # The fp32 route on the same input: x [S][14][used][32], channel 2 link + {re, im}.
    x = torch.stack([grid_re[..., :used], grid_im[..., :used]], -1)      # [S][16][14][used][2]
    x = x.permute(0, 2, 3, 1, 4).reshape(S, 14, used, 32).to(torch.bfloat16).contiguous().cuda()
    ref.set_slot_scale(scales)
    ref.gradients(x, torch.zeros(x.shape, dtype=torch.float32, device="cuda"), S, used)
    out = ref.fetch("out", torch.float32, S * 14 * used * 32).reshape(S, 14, used, 16, 2)
    want = torch.zeros((2, S, 16, 14, pad), dtype=torch.float16, device="cuda")
    want[..., :used] = out.permute(4, 0, 3, 1, 2).half()
    torch.cuda.synchronize()
    for s in range(S):
        if keep is not None and keep[s] == 0:
            assert torch.equal(h_re[s], in_re[s]) and torch.equal(h_im[s], in_im[s]), f"slot {s} not kept but written"
            continue
        assert torch.equal(h_re[s], want[0, s]), (s, (h_re[s].float() - want[0, s].float()).abs().max().item())
        assert torch.equal(h_im[s], want[1, s]), (s, (h_im[s].float() - want[1, s].float()).abs().max().item())
    # The network changed the estimate (the check above is not the identity).
    assert not torch.equal(h_re[..., :used], in_re[..., :used])
    ref.close()


def C_vp(t):
    import ctypes as C
    return C.c_void_p(t.data_ptr())


def C_vp_int(p):
    import ctypes as C
    return C.c_void_p(p)
