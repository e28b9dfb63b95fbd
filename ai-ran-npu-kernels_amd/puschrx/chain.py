"""Op-by-op PUSCH receive chain over the C-ABI operators, in the order of the
reference's chain/pusch_mimo_rx_chain/rx_chain.cpp:440-640:

  OFDM demod (+RE demap) -> DMRS gen -> DMRS LS -> LMMSE CE -> detector
  I/O pack -> LMMSE detection -> QAM demap -> layer demap -> descramble ->
  rate dematch -> LDPC decode -> CRC.

A transport block spans one slot (run_slot) or T consecutive slots
(run_tb).  Every stage is a separate launch with its buffers allocated per
call; the native batched chain (airan_pusch_chain_*)
is the production path and must reproduce this composition bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import abi, ops


@dataclass
class ChainOptions:
    ldpc_iterations: int = 8
    ldpc_mode: int = abi.LDPC_MODE_FP16
    early_stop: bool = True
    tau_s: float = 300e-9
    fd_hz: float = 100.0
    scs_hz: float = 30e3
    ce_rank: int = 384
    iq_format: int = abi.IQ_INT16_Q8_8
    receiver: int = abi.CHAIN_RX_LMMSE   # CHAIN_RX_LS_ZF: SISO LS + linear interpolation, zero forcing
    dft_precision: int = abi.DFT_FP32    # DFT_BF16 / DFT_FP16: the OFDM DFT on the matrix cores
    gram_precision: int = abi.GRAM_FP32  # GRAM_BF16 / GRAM_FP16: the detector's Gram on the matrix cores


class OpChain:
    def __init__(self, cfg, plan, opts: ChainOptions | None = None):
        self.cfg = cfg
        self.plan = plan
        self.opts = opts or ChainOptions()
        self.g = ops.Geometry(cfg)
        self.ls_zf = self.opts.receiver == abi.CHAIN_RX_LS_ZF
        self.filter = None if self.ls_zf else ops.LmmseFilter(cfg, tau_s=self.opts.tau_s, fd_hz=self.opts.fd_hz,
                                                              scs_hz=self.opts.scs_hz, max_rank=self.opts.ce_rank)
        self._dmrs = {cfg.slot_number: ops.dmrs_gen(cfg)}
        self.active_rows = ops.ldpc_active_rows(plan)
        self.no_eff_floor = float(abi.load_library().airan_pusch_NoEffFloor(cfg.qm))

    def dmrs(self, slot_number: int):
        """DMRS reference of a slot number (the rest of the config fixed)."""
        if slot_number not in self._dmrs:
            c = type(self.cfg).from_buffer_copy(self.cfg)
            c.slot_number = slot_number
            self._dmrs[slot_number] = ops.dmrs_gen(c)
        return self._dmrs[slot_number]

    def front(self, iq: torch.Tensor, slot_number: int | None = None, out: dict | None = None):
        """Slot-local stages of one slot: OFDM demod (+RE demap) -> DMRS LS ->
        LMMSE CE -> detector I/O pack -> detection -> QAM demap -> layer
        demap (receiver LS_ZF: SISO LS + linear interpolation and zero
        forcing with n0 = the DMRS-LS noise estimate in place of CE and
        detection).  Returns the codeword LLRs before descrambling
        [Qm, codeword_stride] int16; `out` (when given) receives the
        intermediates."""
        cfg, g, o = self.cfg, self.g, self.opts
        dmrs_re, dmrs_im = self.dmrs(cfg.slot_number if slot_number is None else slot_number)
        rx_re, rx_im = ops.ofdm_demod(cfg, iq, g.nr, o.iq_format, fused_redemap=True)
        h_re, h_im, sc, cnt, noise = ops.dmrs_ls(cfg, rx_re, rx_im, dmrs_re, dmrs_im)
        if self.ls_zf:
            hg_re, hg_im = ops.ls_siso(cfg, rx_re, rx_im, dmrs_re, dmrs_im)
            noise_ce = noise
            n0 = noise[:1].expand(14 * g.pad).contiguous()
            xr, xi, ne = ops.zf_siso(rx_re, rx_im, hg_re, hg_im, n0)
        else:
            hg_re, hg_im, _, noise_ce = ops.ce_lmmse(cfg, self.filter, h_re, h_im, sc, cnt, noise)
            hrm_r, hrm_i, yv_r, yv_i, no = ops.detect_io_pack(cfg, rx_re, rx_im, hg_re, hg_im, noise_ce)
            xr, xi, ne = ops.detect_lmmse(cfg, hrm_r, hrm_i, yv_r, yv_i, no)
        # Keep the demapper's integer arithmetic in range (airan_pusch_NoEffFloor).
        ne_q = torch.clamp_min(ne, self.no_eff_floor)
        llr = ops.qam_demod(cfg, xr, xi, ne_q)
        cw = ops.layer_demap(cfg, llr)
        if out is not None:
            out.update(rx_re=rx_re, rx_im=rx_im, h_ls_re=h_re, h_ls_im=h_im, noise=noise, noise_ce=noise_ce,
                       h_re=hg_re, h_im=hg_im, xhat_re=xr, xhat_im=xi, no_eff=ne, layer_llr=llr, cw_llr=cw)
        return cw

    def tail(self, cw: torch.Tensor, num_slots: int, restart_per_slot: bool = False, out: dict | None = None):
        """Per-TB stages over the codeword LLRs of the TB's `num_slots`
        slots ([num_slots, Qm, codeword_stride], contiguous): descramble (the
        Gold sequence runs on across the slots, descramble_mimo.cpp:176-184),
        rate dematch (code blocks may span slots), LDPC decode and, when the
        plan carries CRCs, the CRC check."""
        plan, o = self.plan, self.opts
        cw_nr = ops.descramble(self.cfg, cw, num_slots, restart_per_slot=restart_per_slot)
        lam = ops.rate_dematch(self.cfg, cw_nr, num_slots, plan)
        bits, _, iters = ops.ldpc_decode(lam, plan.C, plan.Z, plan.K, o.ldpc_iterations, self.active_rows,
                                         o.ldpc_mode, o.early_stop)
        res = dict(bits=bits, iterations=iters)
        if plan.crc_tb:
            cb_ok, tb_ok, _ = ops.crc_check(plan, 1, bits)
            res.update(cb_ok=cb_ok, tb_ok=tb_ok)
        if out is not None:
            out.update(cw_nr=cw_nr, ldpc_llr=lam)
        return res

    def run_slot(self, iq: torch.Tensor, keep: bool = False, slot_number: int | None = None):
        """iq: one slot, [NR, samples, 2] on the device.  Returns a dict with
        bits [C*K] int8, cb_ok, tb_ok and (keep=True) the intermediates.
        slot_number: the slot's number (its DMRS), default the config's."""
        out = {}
        cw = self.front(iq, slot_number, out if keep else None)
        out.update(self.tail(cw, 1, restart_per_slot=True, out=out if keep else None))
        if not keep:
            out = {k: out[k] for k in ("bits", "cb_ok", "tb_ok", "iterations") if k in out}
        return out

    def run_tb(self, iq: torch.Tensor, slot_numbers=None, keep: bool = False):
        """One transport block over T = iq.shape[0] consecutive slots
        (iq [T, NR, samples, 2]): the slot-local stages per slot, then the
        TB stages over the T slots' codeword LLRs.  slot_numbers: each
        slot's number (default the config's)."""
        T = iq.shape[0]
        g = self.g
        cw = torch.empty((T, g.qm, g.cw_stride), dtype=torch.int16, device=iq.device)
        for t in range(T):
            sn = None if slot_numbers is None else int(slot_numbers[t])
            cw[t].copy_(self.front(iq[t].contiguous(), sn).view(g.qm, g.cw_stride))
        out = {"cw_llr": cw} if keep else {}
        out.update(self.tail(cw, T, out=out if keep else None))
        return out

    def close(self):
        if self.filter is not None:
            self.filter.close()


_HIP = None


def _hip():
    """libamdhip64 for device-to-device copies out of chain-owned buffers."""
    global _HIP
    if _HIP is None:
        import ctypes as C
        lib = C.CDLL("libamdhip64.so")
        lib.hipMemcpyAsync.restype = C.c_int
        lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _HIP = lib
    return _HIP


class NativeChain:
    """The batched chain of include/airan_pusch_chain.h (slots per launch,
    HIP-graph replay)."""

    def __init__(self, cfg, plan, max_slots: int, opts: ChainOptions | None = None, use_graph: bool = True):
        import ctypes as C
        self._C = C
        self.lib = abi.load_library()
        self.cfg, self.plan = cfg, plan
        self.opts = opts or ChainOptions()
        self.g = ops.Geometry(cfg)
        o = abi.AiranPuschChainOptionsV1()
        o.abi_version, o.struct_size = abi.ABI_VERSION, C.sizeof(o)
        o.max_slots = max_slots
        o.iq_format = self.opts.iq_format
        o.ldpc_iterations = self.opts.ldpc_iterations
        o.ldpc_mode = self.opts.ldpc_mode
        o.early_stop = 1 if self.opts.early_stop else 0
        o.ce_max_rank = self.opts.ce_rank
        o.use_graph = 1 if use_graph else 0
        o.tau_s, o.fd_hz, o.scs_hz, o.eig_floor = self.opts.tau_s, self.opts.fd_hz, self.opts.scs_hz, 1e-12
        o.receiver = self.opts.receiver
        o.dft_precision = self.opts.dft_precision
        o.gram_precision = self.opts.gram_precision
        self.handle = C.c_void_p()
        ops._check("chain Create", self.lib.airan_pusch_chain_Create(C.byref(cfg), C.byref(plan), C.byref(o),
                                                                    C.c_void_p(ops.stream_ptr()),
                                                                    C.byref(self.handle)))

    def enqueue(self, iq: torch.Tensor, num_slots: int, slot_numbers: torch.Tensor | None = None):
        """slot_numbers: device int16 [num_slots] (uint16 values), the slot
        number of each slot; None uses the config's for every slot.  With a
        HIP graph the tensor's storage is captured: rewrite it in place."""
        C = self._C
        if slot_numbers is None:
            rc = self.lib.airan_pusch_chain_Enqueue(self.handle, C.c_void_p(iq.data_ptr()), num_slots,
                                                    C.c_void_p(ops.stream_ptr()))
        else:
            if slot_numbers.dtype != torch.int16 or not slot_numbers.is_cuda or slot_numbers.numel() < num_slots:
                raise ValueError("slot_numbers: a device int16 tensor with one entry per slot")
            rc = self.lib.airan_pusch_chain_EnqueueSlots(self.handle, C.c_void_p(iq.data_ptr()), num_slots,
                                                         C.c_void_p(slot_numbers.data_ptr()),
                                                         C.c_void_p(ops.stream_ptr()))
        ops._check("chain Enqueue", rc)

    @property
    def slots_per_tb(self) -> int:
        return int(self.lib.airan_pusch_chain_SlotsPerTb(self.handle))

    def outputs(self) -> abi.AiranPuschChainOutputsV1:
        o = abi.AiranPuschChainOutputsV1()
        o.abi_version, o.struct_size = abi.ABI_VERSION, self._C.sizeof(o)
        ops._check("chain Outputs", self.lib.airan_pusch_chain_Outputs(self.handle, self._C.byref(o)))
        return o

    def fetch(self, name: str, dtype, count: int) -> torch.Tensor:
        """Copy `count` elements of output buffer `name` into a new tensor."""
        src = getattr(self.outputs(), name)
        if not src:
            raise ValueError(f"chain output {name} is not produced for this plan")
        out = torch.empty(count, dtype=dtype, device="cuda")
        nbytes = count * out.element_size()
        rc = _hip().hipMemcpyAsync(out.data_ptr(), src, nbytes, 3, ops.stream_ptr())
        if rc != 0:
            raise ops.AiranError("hipMemcpyAsync", rc)
        return out

    def set_refiner(self, refiner, snr_min_db: float = -100.0, snr_max_db: float = 100.0):
        """Attach a puschrx.refiner.Refiner (None detaches): slots whose SNR
        estimate lies in [snr_min_db, snr_max_db] use its channel estimate."""
        h = refiner.handle if refiner is not None else self._C.c_void_p()
        ops._check("chain SetRefiner", self.lib.airan_pusch_chain_SetRefiner(self.handle, h, snr_min_db, snr_max_db))
        self._refiner = refiner

    def device_bytes(self) -> int:
        return int(self.lib.airan_pusch_chain_DeviceBytes(self.handle))

    def set_timing(self, enabled: bool):
        ops._check("chain SetTiming", self.lib.airan_pusch_chain_SetTiming(self.handle, 1 if enabled else 0))

    def stage_times(self) -> dict:
        """Milliseconds per stage of the last Enqueue (timing on)."""
        n = len(abi.CHAIN_STAGES)
        ms = (self._C.c_float * n)()
        ops._check("chain StageTimes", self.lib.airan_pusch_chain_StageTimes(self.handle, ms, n))
        return dict(zip(abi.CHAIN_STAGES, (float(v) for v in ms)))

    def close(self):
        if self.handle:
            self.lib.airan_pusch_chain_Destroy(self.handle)
            self.handle = self._C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
