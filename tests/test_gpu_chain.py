"""End-to-end receive chain on the GPU: slots from the synthetic transmitter
through a fading channel (or a noiseless identity loopback) are decoded by the
op-by-op chain with passing CRCs and the transmitted payload."""
import numpy as np
import pytest

from puschrx import abi, ops, synth
from puschrx import profiles as P
from puschrx.chain import ChainOptions, OpChain

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def plan_for(prof):
    cfg = prof.config()
    g = ops.Geometry(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    return cfg, g, ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl)


def tb_bits(plan, bits):
    b = bits.cpu().numpy().reshape(plan.C, plan.K).astype(np.uint8)
    per_cb = plan.Kprime - plan.crc_cb
    return np.concatenate([b[r, :per_cb] for r in range(plan.C)])[: plan.tbs]


CASES = [
    (P.SISO_52PRB, 30.0),
    (P.MID_4RX_RANK2, 30.0),
    (P.FULL_CARRIER_4RX, 42.0),
]


@pytest.mark.parametrize("prof,snr", CASES, ids=[c[0].name for c in CASES])
def test_chain_decodes_fading_slots(gpu_stream, prof, snr):
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(21)
    S = 2
    payload = synth.random_payload(rng, S, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    spec = synth.ChannelSpec(taps=8, delay_spread_s=100e-9, doppler_hz=10.0, rx_corr=0.3, snr_db=snr)
    H = synth.channel(rng, cfg, spec, S)
    iq = synth.receive(rng, cfg, grid, H, snr)
    chain = OpChain(cfg, plan, ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz))
    for s in range(S):
        print(f"slot {s} of {prof.name}", flush=True)
        r = chain.run_slot(torch.from_numpy(iq[s]).cuda())
        torch.cuda.synchronize()
        assert int(r["tb_ok"].item()) == 1
        assert np.array_equal(tb_bits(plan, r["bits"]), payload[s])
    chain.close()


@pytest.mark.parametrize("prof", [P.SISO_52PRB, P.FULL_CARRIER_4RX], ids=lambda p: p.name)
def test_chain_noiseless_identity_loopback(gpu_stream, prof):
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(4)
    payload = synth.random_payload(rng, 1, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    iq = synth.receive(rng, cfg, grid, synth.identity_channel(cfg, 1), snr_db=60.0)
    chain = OpChain(cfg, plan, ChainOptions(tau_s=100e-9, fd_hz=0.0))
    r = chain.run_slot(torch.from_numpy(iq[0]).cuda(), keep=True)
    torch.cuda.synchronize()
    # Near-noiseless identity channel: the equalised symbols are the grid's.
    xr = r["xhat_re"].cpu().numpy().astype(np.float32).reshape(16, 14, g.pad)
    xi = r["xhat_im"].cpu().numpy().astype(np.float32).reshape(16, 14, g.pad)
    x = (xr + 1j * xi)[: g.nl, :, : g.used][:, g.data_sym]
    want = grid[0][:, g.data_sym]
    assert np.mean(np.abs(x - want) ** 2) < 1e-3
    assert int(r["tb_ok"].item()) == 1
    assert np.array_equal(tb_bits(plan, r["bits"]), payload[0])
    chain.close()


NATIVE_CASES = [
    (P.SISO_52PRB, 30.0),
    (P.MID_4RX_RANK2, 30.0),
    (P.FULL_CARRIER_4RX, 42.0),
]


@pytest.mark.parametrize("prof,snr", NATIVE_CASES, ids=[c[0].name for c in NATIVE_CASES])
def test_native_chain_matches_op_chain(gpu_stream, prof, snr):
    """The batched chain (fused detect/demap/descramble, graph replay)
    reproduces the operator-by-operator chain bit for bit on every slot,
    across two graph replays with different slots in the same input buffer."""
    from puschrx.chain import NativeChain
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(31)
    S = 3
    spec = synth.ChannelSpec(taps=8, delay_spread_s=100e-9, doppler_hz=10.0, rx_corr=0.3, snr_db=snr)
    opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, iq_format=abi.IQ_COMPLEX64)
    op = OpChain(cfg, plan, opts)
    nat = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True)
    iq_dev = torch.empty((S, g.nr, g.samples, 2), dtype=torch.float32, device="cuda")
    plane = 14 * g.pad
    for replay in range(2):
        print(f"{prof.name} replay {replay}", flush=True)
        payload = synth.random_payload(rng, S, plan.tbs)
        grid = synth.tx_grid(cfg, plan, payload)
        H = synth.channel(rng, cfg, spec, S)
        iq = synth.receive(rng, cfg, grid, H, snr, iq_format=abi.IQ_COMPLEX64)
        iq_dev.copy_(torch.from_numpy(iq))
        nat.enqueue(iq_dev, S)
        torch.cuda.synchronize()
        out = {
            "rx_grid_re": nat.fetch("rx_grid_re", torch.float16, S * g.nr * plane),
            "h_ls_re": nat.fetch("h_ls_re", torch.float16, S * g.nr * g.nl * g.ndmrs * ops.Geometry(cfg).pilot_pad),
            "noise_ls": nat.fetch("noise_ls", torch.float16, S * g.nr),
            "noise_ce": nat.fetch("noise_ce", torch.float16, S * g.nr),
            "h_re": nat.fetch("h_re", torch.float16, S * g.nr * g.nl * plane),
            "cw_llr": nat.fetch("cw_llr", torch.int16, S * g.qm * g.cw_stride),
            "bits": nat.fetch("bits", torch.int8, S * plan.C * plan.K),
            "iterations": nat.fetch("iterations", torch.uint8, S * plan.C),
            "cb_ok": nat.fetch("cb_ok", torch.uint8, S * plan.C),
            "tb_ok": nat.fetch("tb_ok", torch.uint8, S),
        }
        # The decoder reads the codeword LLRs directly when it can (no
        # rate-dematched buffer); otherwise that buffer is compared too.
        fused = not nat.outputs().ldpc_llr
        if not fused:
            out["ldpc_llr"] = nat.fetch("ldpc_llr", torch.int16, S * plan.C * 68 * plan.Z)
        torch.cuda.synchronize()
        nat_np = {k: v.cpu().numpy() for k, v in out.items()}
        for s in range(S):
            r = op.run_slot(iq_dev[s].contiguous(), keep=True)
            torch.cuda.synchronize()
            def part(name, n):
                return nat_np[name].reshape(S, -1)[s][:n]
            assert np.array_equal(part("rx_grid_re", g.nr * plane).view(np.uint16),
                                  r["rx_re"].cpu().numpy().view(np.uint16))
            assert np.array_equal(part("h_ls_re", r["h_ls_re"].numel()).view(np.uint16),
                                  r["h_ls_re"].cpu().numpy().view(np.uint16))
            assert np.array_equal(part("noise_ls", g.nr).view(np.uint16), r["noise"].cpu().numpy().view(np.uint16))
            assert np.array_equal(part("noise_ce", g.nr).view(np.uint16),
                                  r["noise_ce"].cpu().numpy().view(np.uint16))
            h_op = r["h_re"].cpu().numpy().reshape(g.nr, 16, plane)[:, : g.nl]
            h_nat = part("h_re", g.nr * g.nl * plane).reshape(g.nr, g.nl, plane)
            assert np.array_equal(h_nat.view(np.uint16), h_op.view(np.uint16))
            assert np.array_equal(part("cw_llr", g.qm * g.cw_stride), r["cw_nr"].cpu().numpy())
            if not fused:
                assert np.array_equal(part("ldpc_llr", plan.C * 68 * plan.Z), r["ldpc_llr"].cpu().numpy())
            assert np.array_equal(part("bits", plan.C * plan.K), r["bits"].cpu().numpy())
            assert np.array_equal(part("iterations", plan.C), r["iterations"].cpu().numpy())
            assert np.array_equal(part("cb_ok", plan.C), r["cb_ok"].cpu().numpy())
            assert int(nat_np["tb_ok"][s]) == int(r["tb_ok"].item()) == 1
            assert np.array_equal(tb_bits(plan, torch.from_numpy(part("bits", plan.C * plan.K))), payload[s])
    nat.close()
    op.close()


FUSED_LOAD_CASES = [(prof, rv) for prof in (P.SISO_52PRB, P.MID_4RX_RANK2, P.FULL_CARRIER_4RX) for rv in (0, 1, 2, 3)]


@pytest.mark.parametrize("prof,rv", FUSED_LOAD_CASES, ids=[f"{c[0].name}-rv{c[1]}" for c in FUSED_LOAD_CASES])
def test_fused_rate_dematch_load_matches_operator(gpu_stream, prof, rv):
    """Where at most 8 rows are active (the full carrier at rv 0 here), the
    batched chain's decoder de-rate-matches the codeword LLRs while loading
    each block (chunked fast path, plane ends and chunks across the filler
    one by one); the operator chain runs the rate-dematch operator and the
    decoder on its output.  One decoding iteration without early stop leaves
    the decoded bits a function of the loaded LLR values, so equal bits on
    noisy slots check the load position by position.  The other cases (rv
    1-3 activate more rows) run the batched chain's separate rate dematch
    against the operator chain under the same settings."""
    from puschrx.chain import NativeChain
    cfg = prof.config()
    g = ops.Geometry(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    plan = ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl, rv=rv)
    rng = np.random.default_rng(100 + rv)
    S = 2
    snr = {"SISO_52PRB": 12.0, "MID_4RX_RANK2": 20.0}.get(prof.name, 30.0)
    spec = synth.ChannelSpec(taps=8, delay_spread_s=100e-9, doppler_hz=10.0, rx_corr=0.3, snr_db=snr)
    opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, iq_format=abi.IQ_COMPLEX64,
                        ldpc_iterations=1, early_stop=False)
    op = OpChain(cfg, plan, opts)
    nat = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True)
    print(f"{prof.name} rv {rv}: k0 {plan.k0}, E {plan.E_low}/{plan.E_high}, fused {not nat.outputs().ldpc_llr}",
          flush=True)
    payload = synth.random_payload(rng, S, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    H = synth.channel(rng, cfg, spec, S)
    iq = synth.receive(rng, cfg, grid, H, snr, iq_format=abi.IQ_COMPLEX64)
    iq_dev = torch.from_numpy(iq).cuda()
    nat.enqueue(iq_dev, S)
    bits = nat.fetch("bits", torch.int8, S * plan.C * plan.K).cpu().numpy().reshape(S, -1)
    for s in range(S):
        r = op.run_slot(iq_dev[s].contiguous())
        want = r["bits"].cpu().numpy()
        assert np.array_equal(bits[s], want), (s, int((bits[s] != want).sum()))
        print(f"  slot {s}: {int(r['cb_ok'].sum().item())}/{plan.C} code blocks pass after one iteration",
              flush=True)
    nat.close()
    op.close()


CRC_CASES = [
    (P.SISO_52PRB, (6.0, 12.0, 16.0, 20.0, 40.0)),
    (P.FULL_CARRIER_4RX, (24.0, 30.0, 33.0, 36.0, 46.0)),
]


@pytest.mark.parametrize("prof,snrs", CRC_CASES, ids=[c[0].name for c in CRC_CASES])
def test_native_chain_crc_flags_match_crc_operator(gpu_stream, prof, snrs):
    """The decoder's CRC epilogue (code-block CRC24B and the transport block
    CRC from each block's hard bits) gives the same flags as the standalone
    CRC operator on the same decoded bits, on slots at SNRs where some code
    blocks fail and some pass."""
    from puschrx.chain import NativeChain
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(41)
    S = len(snrs)
    spec = synth.ChannelSpec(taps=8, delay_spread_s=100e-9, doppler_hz=10.0, rx_corr=0.3)
    payload = synth.random_payload(rng, S, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    H = synth.channel(rng, cfg, spec, S)
    iq = np.concatenate([synth.receive(rng, cfg, grid[s:s + 1], H[s:s + 1], snr, iq_format=abi.IQ_COMPLEX64)
                         for s, snr in enumerate(snrs)])
    opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, iq_format=abi.IQ_COMPLEX64)
    nat = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=False)
    print(f"{prof.name} snrs {snrs}", flush=True)
    nat.enqueue(torch.from_numpy(iq).cuda(), S)
    torch.cuda.synchronize()
    bits = nat.fetch("bits", torch.int8, S * plan.C * plan.K)
    cb_nat = nat.fetch("cb_ok", torch.uint8, S * plan.C).cpu().numpy()
    tb_nat = nat.fetch("tb_ok", torch.uint8, S).cpu().numpy()
    cb_op, tb_op, _ = ops.crc_check(plan, S, bits)
    torch.cuda.synchronize()
    cb_op = cb_op.cpu().numpy().reshape(-1)
    tb_op = tb_op.cpu().numpy().reshape(-1)
    assert np.array_equal(cb_nat, cb_op)
    assert np.array_equal(tb_nat, tb_op)
    # Both outcomes occur, so the test sees failing and passing checks.
    assert 0 < int(cb_nat.sum()) < cb_nat.size
    assert 0 < int(tb_nat.sum()) < S
    nat.close()


def test_native_chains_on_two_streams_match_serial(gpu_stream):
    """Two chain instances replaying their graphs on two streams at once (the
    decoder's work queues, CRC tables and scratch are per instance) decode
    every batch exactly as one instance decoding the batches in turn."""
    from puschrx.chain import NativeChain
    prof = P.FULL_CARRIER_4RX
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(43)
    S = 4
    spec = synth.ChannelSpec(taps=16, delay_spread_s=300e-9, doppler_hz=100.0, rx_corr=0.5)
    opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, ce_rank=384,
                        iq_format=abi.IQ_COMPLEX64)
    batches = []
    for snr in (36.0, 40.0):
        payload = synth.random_payload(rng, S, plan.tbs)
        grid = synth.tx_grid(cfg, plan, payload)
        H = synth.channel(rng, cfg, spec, S)
        batches.append(torch.from_numpy(synth.receive(rng, cfg, grid, H, snr, iq_format=abi.IQ_COMPLEX64)).cuda())
    names = (("bits", torch.int8, S * plan.C * plan.K), ("cb_ok", torch.uint8, S * plan.C),
             ("tb_ok", torch.uint8, S), ("iterations", torch.uint8, S * plan.C))

    def outputs(chain):
        out = [chain.fetch(n, dt, cnt) for n, dt, cnt in names]
        torch.cuda.synchronize()
        return [o.cpu().numpy() for o in out]

    ref_chain = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True)
    want = []
    for iq in batches:
        ref_chain.enqueue(iq, S)
        want.append(outputs(ref_chain))
    ref_chain.close()

    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    chains = []
    for st in streams:
        with torch.cuda.stream(st):
            chains.append(NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True))
    torch.cuda.synchronize()
    for rep in range(3):
        print(f"two streams, replay {rep}", flush=True)
        for j in range(2):
            with torch.cuda.stream(streams[j]):
                chains[j].enqueue(batches[j], S)
        torch.cuda.synchronize()
        for j in range(2):
            with torch.cuda.stream(streams[j]):
                got = outputs(chains[j])
            for n, a, b in zip(names, got, want[j]):
                assert np.array_equal(a, b), (rep, j, n[0])
    for c in chains:
        c.close()


def test_native_chain_per_slot_dmrs(gpu_stream):
    """20 consecutive slots with slot numbers 0..19, each carrying the DMRS
    of its own number (c_init depends on it, mimo_dmrs_gen.cpp:38-45): the
    batched chain decodes every slot and matches the op chain run per slot
    with that slot's number.  Negative control: enqueued with the config's
    slot number for all, every slot but slot 0 fails its CRC.  The slot
    numbers are read when the graph replays: rewriting the buffer in place
    switches between the two outcomes."""
    from puschrx.chain import NativeChain
    prof = P.MID_4RX_RANK2
    cfg, g, plan = plan_for(prof)
    assert cfg.slot_number == 0
    rng = np.random.default_rng(51)
    S = 20
    sn = np.arange(S, dtype=np.uint16)
    spec = synth.ChannelSpec(taps=8, delay_spread_s=100e-9, doppler_hz=10.0, rx_corr=0.3, snr_db=30.0)
    payload = synth.random_payload(rng, S, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload, slot_numbers=sn)
    H = synth.channel(rng, cfg, spec, S)
    iq = synth.receive(rng, cfg, grid, H, spec.snr_db, iq_format=abi.IQ_COMPLEX64)
    opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, iq_format=abi.IQ_COMPLEX64)
    nat = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True)
    iq_dev = torch.from_numpy(iq).cuda()
    sn_dev = torch.from_numpy(sn.astype(np.int16)).cuda()

    def run():
        nat.enqueue(iq_dev, S, sn_dev)
        out = (nat.fetch("bits", torch.int8, S * plan.C * plan.K), nat.fetch("tb_ok", torch.uint8, S),
               nat.fetch("cb_ok", torch.uint8, S * plan.C), nat.fetch("iterations", torch.uint8, S * plan.C))
        torch.cuda.synchronize()
        return [o.cpu().numpy() for o in out]

    print("per-slot DMRS: slot numbers 0..19", flush=True)
    bits, tb_ok, cb_ok, iters = run()
    assert tb_ok.tolist() == [1] * S
    op = OpChain(cfg, plan, opts)
    for s in range(S):
        r = op.run_slot(iq_dev[s].contiguous(), slot_number=s)
        torch.cuda.synchronize()
        assert np.array_equal(bits.reshape(S, -1)[s], r["bits"].cpu().numpy()), s
        assert np.array_equal(iters.reshape(S, -1)[s], r["iterations"].cpu().numpy()), s
        assert np.array_equal(cb_ok.reshape(S, -1)[s], r["cb_ok"].cpu().numpy()), s
        assert np.array_equal(tb_bits(plan, torch.from_numpy(bits.reshape(S, -1)[s])), payload[s]), s
    op.close()
    print("per-slot DMRS: wrong slot numbers (all 0), same graph", flush=True)
    sn_dev.zero_()
    bits_bad, tb_bad, _, _ = run()
    assert tb_bad[0] == 1 and np.array_equal(bits_bad.reshape(S, -1)[0], bits.reshape(S, -1)[0])
    for s in range(1, S):
        # Against the wrong pilots the estimate is noise-like, and no slot
        # decodes its payload.  A slot whose estimate collapses to ~0 gets
        # all-zero LLRs and decodes to the all-zero codeword, whose CRC24A
        # remainder is zero (no initial inversion in TS 38.212 5.1): the only
        # way such a slot can pass its CRC.
        b = bits_bad.reshape(S, -1)[s]
        assert not np.array_equal(tb_bits(plan, torch.from_numpy(b)), payload[s]), s
        if tb_bad[s]:
            assert not b.any(), s
    print("per-slot DMRS: slot numbers restored in place", flush=True)
    sn_dev.copy_(torch.from_numpy(sn.astype(np.int16)).cuda())
    bits2, tb2, _, _ = run()
    assert tb2.tolist() == [1] * S and np.array_equal(bits2, bits)
    # The plain Enqueue uses the config's slot number: the same as the control.
    nat.enqueue(iq_dev, S)
    tb3 = nat.fetch("tb_ok", torch.uint8, S)
    torch.cuda.synchronize()
    assert np.array_equal(tb3.cpu().numpy(), tb_bad)
    nat.close()


def _or_plan(plan):
    from oracle_lib import OrTbPlan
    p = OrTbPlan()
    for name, _ in OrTbPlan._fields_:
        if hasattr(plan, name):
            setattr(p, name, getattr(plan, name))
    return p


def test_reference_profile_23_slot_tb(oracle, gpu_stream):
    """The reference's own workload end to end (chain/pusch_mimo_rx_chain,
    rx_chain.cpp:463-664): the fd16x64 profile (64 Rx, rank 4, 133 PRB, FFT
    2048, 256QAM, int16 IQ), one transport block over 23 slots with the
    reference's fixed plan (C = 143 code blocks of K = 8448, no CRC), the
    reference LDPC arithmetic (int8 messages, 3 iterations, no early stop),
    through the reference's flat channel H[rx, l] = exp(j 2 pi rx l / 64)
    (chain/pusch_mimo_loopback/rank2_23slot.py:93-103) with AWGN, slot numbers
    0..22 (each slot's own DMRS).

    Checked: the batched chain equals the op-by-op chain at the descrambled
    codeword LLRs, the rate-dematched decoder input and the decoded bits; the
    integer tail equals the oracle (rate dematch of the 23 slots' LLRs, then
    the REF-mode decoder) bit for bit; and the bits are the payload (BER 0,
    as the reference's loopback evidence)."""
    from puschrx.chain import NativeChain
    from oracle_lib import cref, ptr  # noqa: F401
    import ctypes as C
    prof = P.REFERENCE_FD16X64
    cfg = prof.config()
    g = ops.Geometry(cfg)
    T = 23
    plan = ops.reference_tb_plan(cfg, T)
    assert (plan.C, plan.K, plan.Z, plan.crc_tb, plan.G) == (143, 8448, 384, 0, T * g.cw_symbols * g.qm)
    rng = np.random.default_rng(0x523236)
    sn = np.arange(T, dtype=np.uint16)
    payload = synth.random_payload(rng, 1, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload, slots_per_tb=T, slot_numbers=sn)
    rx = np.arange(g.nr)[:, None]
    lay = np.arange(g.nl)[None, :]
    hflat = np.exp(2j * np.pi * rx * lay / g.nr).astype(np.complex64)
    H = np.broadcast_to(hflat[None, :, :, None, None], (T, g.nr, g.nl, 14, g.used))
    snr = 10.0
    iq = synth.receive(rng, cfg, grid, H, snr)  # int16 Q8.8
    opts = ChainOptions(ldpc_iterations=3, ldpc_mode=abi.LDPC_MODE_REF, early_stop=False, tau_s=100e-9,
                        fd_hz=0.0, iq_format=abi.IQ_INT16_Q8_8)
    iq_dev = torch.from_numpy(iq).cuda()
    sn_dev = torch.from_numpy(sn.astype(np.int16)).cuda()
    print("reference profile: native chain, 23 slots", flush=True)
    nat = NativeChain(cfg, plan, max_slots=T, opts=opts, use_graph=True)
    assert nat.slots_per_tb == T
    nat.enqueue(iq_dev, T, sn_dev)
    assert not nat.outputs().tb_ok  # the reference plan carries no CRC
    cw_nat = nat.fetch("cw_llr", torch.int16, T * g.qm * g.cw_stride)
    lam_nat = nat.fetch("ldpc_llr", torch.int16, plan.C * 68 * plan.Z)
    bits_nat = nat.fetch("bits", torch.int8, plan.C * plan.K)
    it_nat = nat.fetch("iterations", torch.uint8, plan.C)
    torch.cuda.synchronize()
    cw_nat, lam_nat = cw_nat.cpu().numpy(), lam_nat.cpu().numpy()
    bits_nat, it_nat = bits_nat.cpu().numpy(), it_nat.cpu().numpy()
    nat.close()
    assert np.all(it_nat == 3)
    assert np.array_equal(bits_nat.reshape(plan.C, plan.K).reshape(-1), payload[0].astype(np.int8))

    print("reference profile: op-by-op chain, 23 slots", flush=True)
    op = OpChain(cfg, plan, opts)
    r = op.run_tb(iq_dev, slot_numbers=sn, keep=True)
    torch.cuda.synchronize()
    op.close()
    assert np.array_equal(cw_nat, r["cw_nr"].cpu().numpy())
    assert np.array_equal(lam_nat, r["ldpc_llr"].cpu().numpy())
    assert np.array_equal(bits_nat, r["bits"].cpu().numpy())

    print("reference profile: oracle rate dematch + REF decoder", flush=True)
    op_plan = _or_plan(plan)
    want_lam = np.zeros(plan.C * 68 * plan.Z, np.int16)
    oracle.or_rate_dematch(C.byref(op_plan), g.nl, g.qm, g.cw_symbols, g.cw_stride, T, ptr(cw_nat), ptr(want_lam))
    assert np.array_equal(lam_nat, want_lam)
    want_bits = np.zeros(plan.C * plan.K, np.int8)
    want_it = np.zeros(plan.C, np.uint8)
    lam = want_lam.copy()
    oracle.or_ldpc_decode(plan.Z, plan.C, ptr(lam), abi.LDPC_MODE_REF, 3, ops.ldpc_active_rows(plan), plan.K, 0,
                          ptr(want_bits), ptr(want_it))
    assert np.array_equal(bits_nat, want_bits)
    assert np.array_equal(it_nat, want_it)


PRECISIONS = [(abi.DFT_FP16, abi.GRAM_FP16), (abi.DFT_BF16, abi.GRAM_BF16), (abi.DFT_FP32, abi.GRAM_FP16)]


@pytest.mark.parametrize("dft,gram", PRECISIONS, ids=["fp16", "bf16", "fp32dft_fp16gram"])
def test_native_chain_mixed_precision(oracle, gpu_stream, dft, gram):
    """The chain with the OFDM DFT and the detector's Gram/matched filter on
    the matrix cores (BF16 / FP16 operands, FP32 accumulation) against the
    FP32 chain on the same slots: the received grid within the DFT's stated
    NMSE of the FP32 FFT's, the detector's outputs (through the LLRs) nearly
    identical, and every slot decoded with a passing CRC at 42 dB; the FP16
    Gram on the FP32 grid changes at most a few LLRs by one step."""
    from puschrx.chain import NativeChain
    prof = P.FULL_CARRIER_4RX
    cfg, g, plan = plan_for(prof)
    rng = np.random.default_rng(61)
    S = 4
    spec = synth.ChannelSpec(taps=16, delay_spread_s=300e-9, doppler_hz=100.0, rx_corr=0.5, snr_db=42.0)
    payload = synth.random_payload(rng, S, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    H = synth.channel(rng, cfg, spec, S)
    iq = torch.from_numpy(synth.receive(rng, cfg, grid, H, spec.snr_db, iq_format=abi.IQ_COMPLEX64)).cuda()
    out = {}
    for name, (d, gm) in (("fp32", (abi.DFT_FP32, abi.GRAM_FP32)), ("mixed", (dft, gram))):
        opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, iq_format=abi.IQ_COMPLEX64,
                            dft_precision=d, gram_precision=gm)
        nat = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True)
        nat.enqueue(iq, S)
        plane = 14 * g.pad
        r = {"rx_re": nat.fetch("rx_grid_re", torch.float16, S * g.nr * plane),
             "rx_im": nat.fetch("rx_grid_im", torch.float16, S * g.nr * plane),
             "cw": nat.fetch("cw_llr", torch.int16, S * g.qm * g.cw_stride),
             "bits": nat.fetch("bits", torch.int8, S * plan.C * plan.K),
             "tb_ok": nat.fetch("tb_ok", torch.uint8, S)}
        torch.cuda.synchronize()
        out[name] = {k: v.cpu().numpy() for k, v in r.items()}
        nat.close()
    a, b = out["fp32"], out["mixed"]
    ya = a["rx_re"].astype(np.float64) + 1j * a["rx_im"].astype(np.float64)
    yb = b["rx_re"].astype(np.float64) + 1j * b["rx_im"].astype(np.float64)
    err = 10 * np.log10(np.sum(np.abs(yb - ya) ** 2) / np.sum(np.abs(ya) ** 2))
    llr_diff = float(np.mean(a["cw"] != b["cw"]))
    sign_flip = float(np.mean(np.sign(a["cw"]) != np.sign(b["cw"])))
    print(f"dft {dft} gram {gram}: grid NMSE vs FP32 {err:.1f} dB, LLRs changed {llr_diff:.4f}, "
          f"signs flipped {sign_flip:.5f}, tb_ok {b['tb_ok'].tolist()}")
    if dft == abi.DFT_FP32:
        assert np.array_equal(a["rx_re"].view(np.uint16), b["rx_re"].view(np.uint16))
        assert llr_diff < 0.01
    else:
        assert err <= {abi.DFT_FP16: -60.0, abi.DFT_BF16: -40.0}[dft]
    assert b["tb_ok"].tolist() == [1] * S
    assert np.array_equal(a["bits"], b["bits"])


def test_config2_stated_point(gpu_stream):
    """The mid-size plain MIMO workload at its stated point: 4 Rx, rank 2 on
    CDM ports 0/1, 106 PRB, 64QAM R ~ 0.6, 12-tap 300 ns channel with 50 Hz
    Jakes Doppler and Rx correlation 0.5, AWGN 18 dB (per Rx and layer), 32
    complex64 slots through the batched chain with a matched CE prior.  The
    BLER is printed beside the FP32 CPU chain's on the same slots; every
    passing TB carries the transmitted payload, and the GPU chain loses no
    more than one TB against the CPU chain."""
    from oracle_lib import CpuChain, OrCeParams, host_cpu, load_oracle, or_tb_plan
    from puschrx.chain import NativeChain
    prof = P.MID_4RX_RANK2
    cfg, g, plan = plan_for(prof)
    S = 32
    spec = synth.ChannelSpec(taps=12, delay_spread_s=300e-9, doppler_hz=50.0, rx_corr=0.5, snr_db=18.0)
    rng = np.random.default_rng(0xC0F2)
    payload = synth.random_payload(rng, S, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    iq = synth.receive(rng, cfg, grid, synth.channel(rng, cfg, spec, S), spec.snr_db, iq_format=abi.IQ_COMPLEX64)
    opts = ChainOptions(tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz, iq_format=abi.IQ_COMPLEX64)
    print(f"config 2: {S} slots at {spec.snr_db} dB, TBS {plan.tbs}, C = {plan.C}", flush=True)
    chain = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=False)
    chain.enqueue(torch.from_numpy(iq).cuda(), S)
    tb_ok = chain.fetch("tb_ok", torch.uint8, S)
    bits = chain.fetch("bits", torch.int8, S * plan.C * plan.K)
    torch.cuda.synchronize()
    chain.close()
    ok = tb_ok.cpu().numpy()
    b = bits.cpu().numpy().reshape(S, plan.C, plan.K).astype(np.uint8)
    per_cb = plan.Kprime - plan.crc_cb
    for s in range(S):
        if ok[s]:
            assert np.array_equal(b[s, :, :per_cb].reshape(-1)[: plan.tbs], payload[s])
    gpu_bler = 1.0 - float(ok.mean())
    msg = f"config 2 at 18 dB: GPU chain BLER {gpu_bler:.3f} over {S} slots"
    if host_cpu()["isa"] is not None:
        op = or_tb_plan(load_oracle(), plan.tbs, plan.G, g.qm, g.nl)
        cc = CpuChain(cfg, op, OrCeParams(spec.delay_spread_s, spec.doppler_hz, 30e3), 8)
        try:
            _, cpu_ok = cc.run(iq, 1)
        finally:
            cc.close()
        msg += f", FP32 CPU chain BLER {1.0 - float(cpu_ok.mean()):.3f}"
        assert int(ok.sum()) >= int(cpu_ok.sum()) - 1
    print(msg, flush=True)
