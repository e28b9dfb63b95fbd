"""Ragged multi-grant batches on the GPU (include/airan_pusch_grants.h):
grants of mixed PRB counts, ranks, modulation orders, code rates and FFT
sizes, each its own cell, decoded in one batch with one launch per stage,
equal bit for bit to the single-grant chain run on each grant."""
import numpy as np
import pytest

from puschrx import abi, synth
from puschrx import grants as G
from puschrx.chain import ChainOptions, NativeChain

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

OPTS = dict(iq_format=abi.IQ_COMPLEX64, ldpc_iterations=8, early_stop=True, ce_rank=384, tau_s=300e-9, fd_hz=100.0)


def _slots(rng, compiled, snr_lo=24.0, snr_hi=36.0):
    iqs, payloads = [], []
    for rt, cfg, plan in compiled:
        payload = synth.random_payload(rng, 1, plan.tbs)
        grid = synth.tx_grid(cfg, plan, payload)
        spec = synth.ChannelSpec(taps=int(rng.integers(8, 17)), delay_spread_s=float(rng.uniform(50e-9, 300e-9)),
                                 doppler_hz=10.0, rx_corr=0.5, snr_db=float(rng.uniform(snr_lo, snr_hi)))
        H = synth.channel(rng, cfg, spec, 1)
        iqs.append(synth.receive(rng, cfg, grid, H, spec.snr_db, iq_format=abi.IQ_COMPLEX64)[0])
        payloads.append(payload[0])
    return iqs, payloads


def test_ragged_batch_matches_single_grant_chain(gpu_stream):
    rng = np.random.default_rng(0xC4)
    grants = G.config4_grants(rng, 12)
    # Two grants of one shape, so that one CE application serves both.
    grants.append(G.Grant(name="twin", rank=grants[0].rank, qm=grants[0].qm, num_rb=grants[0].num_rb,
                          fft_size=grants[0].fft_size, code_rate=grants[0].code_rate, rnti=777,
                          padded_subcarriers=grants[0].padded_subcarriers, slot_number=3))
    compiled = G.compile_grants(G.Radio(), grants)
    iqs, payloads = _slots(rng, compiled)
    print(f"{len(grants)} grants: PRB {[g.num_rb for g in grants]}, ranks {[g.rank for g in grants]}, "
          f"Qm {[g.qm for g in grants]}, FFT {[g.fft_size for g in grants]}", flush=True)
    batch = G.GrantBatch(compiled, OPTS, detect_tile=128)
    assert batch.num_shapes == len({(g.num_rb, g.rank) for g in grants})
    iq_dev = batch.pack_iq(iqs)
    for rep in range(2):
        print(f"ragged batch replay {rep}", flush=True)
        batch.enqueue(iq_dev)
        torch.cuda.synchronize()
    tb_ok = batch.fetch("tb_ok", torch.uint8, batch.n).cpu().numpy()
    ncb = batch.num_cb
    cb_ok = batch.fetch("cb_ok", torch.uint8, ncb).cpu().numpy()
    iters = batch.fetch("iterations", torch.uint8, ncb).cpu().numpy()
    ok_any = 0
    for i, ((rt, cfg, plan), lo) in enumerate(zip(compiled, batch.layout)):
        print(f"grant {i}: single-grant chain", flush=True)
        opts = ChainOptions(tau_s=OPTS["tau_s"], fd_hz=OPTS["fd_hz"], ce_rank=OPTS["ce_rank"],
                            iq_format=abi.IQ_COMPLEX64, ldpc_iterations=8, early_stop=True)
        one = NativeChain(cfg, plan, max_slots=1, opts=opts, use_graph=False)
        one.enqueue(torch.from_numpy(iqs[i]).cuda(), 1)
        g = one.g
        want = {"bits": one.fetch("bits", torch.int8, plan.C * plan.K), "cw": one.fetch("cw_llr", torch.int16,
                                                                                       g.qm * g.cw_stride),
                "tb_ok": one.fetch("tb_ok", torch.uint8, 1), "cb_ok": one.fetch("cb_ok", torch.uint8, plan.C),
                "it": one.fetch("iterations", torch.uint8, plan.C)}
        got_bits = batch.fetch("bits", torch.int8, plan.C * plan.K, lo.bits_offset)
        got_cw = batch.fetch("cw_llr", torch.int16, g.qm * g.cw_stride, lo.cw_offset)
        torch.cuda.synchronize()
        want = {k: v.cpu().numpy() for k, v in want.items()}
        one.close()
        assert np.array_equal(got_cw.cpu().numpy(), want["cw"]), i
        assert np.array_equal(got_bits.cpu().numpy(), want["bits"]), i
        assert np.array_equal(cb_ok[lo.first_cb: lo.first_cb + plan.C], want["cb_ok"]), i
        assert np.array_equal(iters[lo.first_cb: lo.first_cb + plan.C], want["it"]), i
        assert int(tb_ok[i]) == int(want["tb_ok"][0]), i
        if tb_ok[i]:
            per_cb = plan.Kprime - plan.crc_cb
            b = got_bits.cpu().numpy().reshape(plan.C, plan.K)[:, :per_cb].reshape(-1)[: plan.tbs]
            assert np.array_equal(b, payloads[i]), i
            ok_any += 1
    print(f"ragged batch: {int(tb_ok.sum())}/{len(grants)} TBs pass their CRC")
    assert ok_any > 0
    batch.close()


def test_ragged_detect_tiles_agree(gpu_stream):
    """The autotuned detector tile changes the launch shape, not the result."""
    rng = np.random.default_rng(0xC5)
    compiled = G.compile_grants(G.Radio(), G.config4_grants(rng, 6))
    iqs, _ = _slots(rng, compiled)
    batch = G.GrantBatch(compiled, OPTS, detect_tile=256)
    iq_dev = batch.pack_iq(iqs)
    res = []
    last = batch.layout[-1]
    n_bits = int(last.bits_offset + last.num_cb * last.K)
    for tile in (256, 64, 128):
        print(f"detector tile {tile}", flush=True)
        batch.set_detect_tile(tile)
        batch.enqueue(iq_dev)
        res.append((batch.fetch("bits", torch.int8, n_bits).cpu().numpy(),
                    batch.fetch("tb_ok", torch.uint8, batch.n).cpu().numpy()))
        torch.cuda.synchronize()
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0]) and np.array_equal(r[1], res[0][1])
    batch.close()
