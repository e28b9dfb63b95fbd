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
