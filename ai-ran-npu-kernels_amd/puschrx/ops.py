"""Per-operator host calls: build the OpArgs struct, enqueue on torch's current
stream, raise on a non-OK status.  One function per reference operator
interface (see include/airan_pusch_rx.h for the reference file:line of each).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import abi
from .abi import load_library


class AiranError(RuntimeError):
    def __init__(self, what: str, status: int):
        super().__init__(f"{what} failed with status {status}")
        self.status = status


def _check(what: str, st: int) -> None:
    if st != abi.OK:
        raise AiranError(what, st)


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream or 0


def _dp(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


def _hdr(cls, **kw):
    a = cls()
    a.abi_version = abi.ABI_VERSION
    a.struct_size = C.sizeof(cls)
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def layout(cfg: abi.PuschMimoConfig) -> abi.PuschMimoLayout:
    lib = load_library()
    lay = abi.PuschMimoLayout()
    _check("DerivePuschMimoLayout", lib.airan_DerivePuschMimoLayout(C.byref(cfg), C.byref(lay)))
    return lay


class Geometry:
    """Host copy of the derived geometry (same rules as csrc/common.h)."""

    def __init__(self, cfg: abi.PuschMimoConfig):
        self.nr = cfg.num_rx_antennas
        self.nl = cfg.num_layers
        self.qm = cfg.qm
        self.fft = cfg.fft_size
        self.used = cfg.used_subcarriers
        self.pad = cfg.padded_subcarriers
        self.dmrs_sym = [s for s in range(14) if (cfg.dmrs_symbol_mask >> s) & 1]
        self.data_sym = [s for s in range(14) if not (cfg.dmrs_symbol_mask >> s) & 1]
        self.ndmrs = len(self.dmrs_sym)
        self.ndata = len(self.data_sym)
        a = lambda v, m: (v + m - 1) // m * m  # noqa: E731
        self.n_dmrs_re = self.used // 2
        self.ref_pad = a(self.n_dmrs_re, 128)
        self.pilot_pad = a(self.n_dmrs_re, 64)
        self.n_occ = self.used // 4
        self.data_re = self.ndata * self.used
        self.data_stride = a(self.data_re, 64)
        self.qam_row = a(self.used, 64)
        self.qam_plane = self.ndata * self.qam_row
        self.cw_symbols = self.nl * self.data_re
        self.cw_stride = self.nl * self.data_stride
        self.cp_short = self.fft * 9 // 128
        self.cp_long = self.cp_short + self.fft // 64
        self.samples = 14 * self.fft + 13 * self.cp_short + self.cp_long
        self.n_re = 14 * self.pad


def _dev(n, dtype, device):
    return torch.empty(n, dtype=dtype, device=device)


# ------------------------------------------------------------------ RE demap
def re_demap(cfg, fft_re: torch.Tensor, fft_im: torch.Tensor):
    lib = load_library()
    g = Geometry(cfg)
    dev = fft_re.device
    lay = layout(cfg)
    meta = abi.ReDemapKernelMetadata()
    _check("re_demap BuildCurrentProfile",
           lib.airan_re_demap_batch_BuildCurrentProfile(C.byref(cfg), C.byref(abi.PuschMimoLayout()), C.byref(meta)))
    idx = (C.c_uint32 * g.pad)()
    _check("re_demap BuildGatherIndex", lib.airan_re_demap_batch_BuildGatherIndex(C.byref(cfg), idx, g.pad))
    idx_d = torch.tensor(list(idx), dtype=torch.int64).to(torch.int32).to(dev)
    meta_d = torch.frombuffer(bytearray(bytes(meta)), dtype=torch.uint8).to(dev)
    rre = _dev(g.nr * 14 * g.pad, torch.float16, dev)
    rim = _dev(g.nr * 14 * g.pad, torch.float16, dev)
    a = _hdr(abi.ReDemapBatchOpArgsV1, fft_grid_re=_dp(fft_re), fft_grid_im=_dp(fft_im), rx_grid_re=_dp(rre),
             rx_grid_im=_dp(rim), config=C.pointer(cfg), layout=C.pointer(lay), stream=C.c_void_p(stream_ptr()))
    _check("re_demap Enqueue", lib.airan_re_demap_batch_Enqueue(C.byref(a), _dp(idx_d), idx_d.numel() * 4,
                                                                _dp(meta_d), 128, _dp(meta_d), 128))
    return rre, rim


# ------------------------------------------------------------------ DMRS gen
def dmrs_gen(cfg, device="cuda"):
    lib = load_library()
    g = Geometry(cfg)
    lay = layout(cfg)
    rt = C.c_void_p()
    _check("dmrs CreateRuntime", lib.airan_mimo_dmrs_gen_CreateRuntime(C.byref(rt)))
    re = _dev(g.nl * g.ndmrs * g.ref_pad, torch.float16, device)
    im = _dev(g.nl * g.ndmrs * g.ref_pad, torch.float16, device)
    a = _hdr(abi.MimoDmrsGenOpArgsV1, dmrs_re=_dp(re), dmrs_im=_dp(im), config=C.pointer(cfg),
             layout=C.pointer(lay), stream=C.c_void_p(stream_ptr()))
    try:
        _check("dmrs Enqueue", lib.airan_mimo_dmrs_gen_Enqueue(rt, C.byref(a)))
        torch.cuda.current_stream().synchronize()
    finally:
        lib.airan_mimo_dmrs_gen_DestroyRuntime(rt)
    return re, im


# ------------------------------------------------------------------- DMRS LS
def dmrs_ls(cfg, rx_re, rx_im, ref_re, ref_im):
    lib = load_library()
    g = Geometry(cfg)
    dev = rx_re.device
    lay = layout(cfg)
    n = g.nr * g.nl * g.ndmrs * g.pilot_pad
    hr, hi = _dev(n, torch.float16, dev), _dev(n, torch.float16, dev)
    sc = _dev(g.nl * g.ndmrs * g.pilot_pad, torch.int16, dev)
    cnt = _dev(16, torch.int16, dev)
    noise = _dev(g.nr, torch.float16, dev)
    a = _hdr(abi.MimoDmrsLsOpArgsV1, rx_grid_re=_dp(rx_re), rx_grid_im=_dp(rx_im), dmrs_ref_re=_dp(ref_re),
             dmrs_ref_im=_dp(ref_im), h_ls_re=_dp(hr), h_ls_im=_dp(hi), pilot_sc=_dp(sc), pilot_count=_dp(cnt),
             noise_var_rx=_dp(noise), config=C.pointer(cfg), layout=C.pointer(lay),
             stream=C.c_void_p(stream_ptr()))
    _check("dmrs_ls Enqueue", lib.airan_mimo_dmrs_ls_Enqueue(C.byref(a)))
    return hr, hi, sc, cnt, noise


# ------------------------------------------------------------------ QAM demap
def qam_demod(cfg, x_re, x_im, no_eff):
    lib = load_library()
    g = Geometry(cfg)
    dev = x_re.device
    lay = layout(cfg)
    t = abi.QamDemodBatchTilingData()
    _check("qam BuildCurrentProfile",
           lib.airan_qam256_demod_batch_BuildCurrentProfile(C.byref(cfg), C.byref(abi.PuschMimoLayout()),
                                                            C.byref(t)))
    t_d = torch.frombuffer(bytearray(bytes(t)), dtype=torch.uint8).to(dev)
    out = _dev(g.nl * g.qm * g.qam_plane, torch.int16, dev)
    a = _hdr(abi.QamDemod256BatchOpArgsV1, x_re=_dp(x_re), x_im=_dp(x_im), no_eff=_dp(no_eff),
             layer_llr=_dp(out), config=C.pointer(cfg), layout=C.pointer(lay), stream=C.c_void_p(stream_ptr()))
    _check("qam Enqueue", lib.airan_qam256_demod_batch_Enqueue(C.byref(a), _dp(t_d), 128, _dp(t_d), 128))
    return out


# --------------------------------------------------------------- layer demap
def layer_demap(cfg, layer_llr, out=None):
    lib = load_library()
    g = Geometry(cfg)
    lay = layout(cfg)
    if out is None:
        out = _dev(g.qm * g.cw_stride, torch.int16, layer_llr.device)
    a = _hdr(abi.LayerDemapOpArgsV1, layer_llr=_dp(layer_llr), cw_llr=_dp(out), config=C.pointer(cfg),
             layout=C.pointer(lay), stream=C.c_void_p(stream_ptr()))
    _check("layer_demap Enqueue", lib.airan_layer_demap_Enqueue(C.byref(a)))
    return out


# ---------------------------------------------------------------- descramble
def descramble(cfg, llr_qam, num_slots: int, restart_per_slot: bool = False, sign=None):
    lib = load_library()
    g = Geometry(cfg)
    lay = layout(cfg)
    out = _dev(num_slots * g.qm * g.cw_stride, torch.int16, llr_qam.device)
    a = _hdr(abi.DescrambleMimoOpArgsV1, llr_qam=_dp(llr_qam), llr_nr=_dp(out), config=C.pointer(cfg),
             layout=C.pointer(lay), num_slots=num_slots, stream=C.c_void_p(stream_ptr()),
             flags=abi.DESCR_RESTART_PER_SLOT if restart_per_slot else 0)
    sbytes = 0 if sign is None else sign.numel() * 2
    _check("descramble Enqueue", lib.airan_descramble_mimo_Enqueue(C.byref(a), _dp(sign), sbytes, None, 0, None, 0))
    return out


# ---------------------------------------------------------------- TB plan
def tb_plan(tbs: int, G: int, qm: int, layers: int, rv: int = 0) -> abi.AiranTbPlanV1:
    lib = load_library()
    p = abi.AiranTbPlanV1()
    _check("BuildTbPlan", lib.airan_pusch_BuildTbPlan(tbs, G, qm, layers, rv, C.byref(p)))
    return p


def reference_tb_plan(cfg, num_slots: int) -> abi.AiranTbPlanV1:
    lib = load_library()
    p = abi.AiranTbPlanV1()
    _check("ReferenceTbPlan", lib.airan_pusch_ReferenceTbPlan(C.byref(cfg), num_slots, C.byref(p)))
    return p


def tbs(n_prb: int, n_re_per_prb: int, qm: int, layers: int, rate: float) -> int:
    return int(load_library().airan_pusch_Tbs(n_prb, n_re_per_prb, qm, layers, rate))


# -------------------------------------------------------------- rate dematch
def rate_dematch_profile(cfg, num_slots: int, plan: abi.AiranTbPlanV1):
    lib = load_library()
    meta = abi.RateDematchKernelMetadata()
    desc = (abi.RateMatchDescriptor * plan.C)()
    _check("rate BuildProfileForTb",
           lib.airan_rate_dematch_mimo_BuildProfileForTb(C.byref(cfg), num_slots, C.byref(plan),
                                                         C.byref(abi.PuschMimoLayout()), C.byref(meta), desc,
                                                         plan.C))
    return meta, desc


def rate_dematch(cfg, cw_llr, num_slots: int, plan: abi.AiranTbPlanV1, out=None):
    lib = load_library()
    lay = layout(cfg)
    dev = cw_llr.device
    meta, desc = rate_dematch_profile(cfg, num_slots, plan)
    desc_d = torch.frombuffer(bytearray(bytes(desc)), dtype=torch.uint8).to(dev)
    meta_d = torch.frombuffer(bytearray(bytes(meta)), dtype=torch.uint8).to(dev)
    if out is None:
        out = _dev(plan.C * 68 * plan.Z, torch.int16, dev)
    a = _hdr(abi.RateDematchMimoOpArgsV1, cw_llr=_dp(cw_llr), ldpc_llr=_dp(out), config=C.pointer(cfg),
             layout=C.pointer(lay), num_slots=num_slots, stream=C.c_void_p(stream_ptr()))
    _check("rate Enqueue", lib.airan_rate_dematch_mimo_Enqueue(C.byref(a), C.byref(meta), _dp(desc_d),
                                                               desc_d.numel(), None, 0, _dp(meta_d),
                                                               meta_d.numel()))
    return out


# ---------------------------------------------------------------------- LDPC
def ldpc_decode(lam_in, num_cb: int, z: int, k: int, iterations: int, active_rows: int = 46,
                mode: int = abi.LDPC_MODE_FP16, early_stop: bool = False, want_lam: bool = False):
    lib = load_library()
    dev = lam_in.device
    bits = _dev(num_cb * k, torch.int8, dev)
    lam_out = _dev(num_cb * 68 * z, torch.int16, dev) if want_lam else None
    iters = _dev(num_cb, torch.uint8, dev)
    a = _hdr(abi.LdpcDecodeOpArgsV1, lam_in=_dp(lam_in), bits=_dp(bits), lam_out=_dp(lam_out),
             iterations_out=_dp(iters), num_cb=num_cb, z=z, k=k, max_iterations=iterations,
             active_rows=active_rows, mode=mode, early_stop=1 if early_stop else 0,
             stream=C.c_void_p(stream_ptr()))
    nws = lib.airan_ldpc_decode_WorkspaceBytes(num_cb, z, active_rows)
    ws = _dev(max(nws, 8), torch.uint8, dev)
    _check("ldpc Enqueue", lib.airan_ldpc_decode_Enqueue(C.byref(a), _dp(ws), ws.numel()))
    return bits, lam_out, iters


def ldpc_active_rows(plan: abi.AiranTbPlanV1) -> int:
    return int(load_library().airan_ldpc_ActiveRows(C.byref(plan)))


# ----------------------------------------------------------------------- CRC
def crc_check(plan: abi.AiranTbPlanV1, num_tbs: int, bits, want_bytes: bool = False):
    lib = load_library()
    dev = bits.device
    cb_ok = _dev(num_tbs * plan.C, torch.uint8, dev)
    tb_ok = _dev(num_tbs, torch.uint8, dev)
    nbytes = (plan.tbs + 7) // 8
    tb_bytes = _dev(num_tbs * nbytes, torch.uint8, dev) if want_bytes else None
    ws = _dev(max(int(lib.airan_crc_check_WorkspaceBytes(num_tbs)), 8), torch.uint8, dev)
    _check("crc Enqueue", lib.airan_crc_check_Enqueue(C.byref(plan), num_tbs, _dp(bits), _dp(cb_ok), _dp(tb_ok),
                                                      _dp(tb_bytes), _dp(ws), ws.numel(),
                                                      C.c_void_p(stream_ptr())))
    return cb_ok, tb_ok, tb_bytes


# ---------------------------------------------------------------- OFDM demod
def ofdm_demod(cfg, iq, batch: int, iq_format: int = abi.IQ_INT16_Q8_8, fused_redemap: bool = True,
               precision: int = 0):
    """CP removal + unitary DFT of `batch` antenna streams (one slot each).
    fused_redemap: return the used-subcarrier grid [batch,14,padded];
    otherwise the stage-4 planes [batch,14,fft].  precision: abi.DFT_FP32
    (FFT) or DFT_BF16 / DFT_FP16 (two DFT products on the matrix cores)."""
    lib = load_library()
    g = Geometry(cfg)
    dev = iq.device
    n = batch * 14 * (g.pad if fused_redemap else g.fft)
    ore, oim = _dev(n, torch.float16, dev), _dev(n, torch.float16, dev)
    a = _hdr(abi.OfdmDemodBatchOpArgsV1, iq=_dp(iq), iq_format=iq_format, batch=batch, precision=precision,
             config=C.pointer(cfg), stream=C.c_void_p(stream_ptr()))
    if fused_redemap:
        _check("ofdm_demod_redemap Enqueue", lib.airan_ofdm_demod_redemap_Enqueue(C.byref(a), _dp(ore), _dp(oim)))
    else:
        a.out_re, a.out_im = _dp(ore), _dp(oim)
        _check("ofdm_demod Enqueue", lib.airan_ofdm_demod_batch_Enqueue(C.byref(a)))
    return ore, oim


# ------------------------------------------------------------------ CE LMMSE
class LmmseFilter:
    """Owns a device filter built by airan_channel_est_lmmse_CreateFilter."""

    def __init__(self, cfg, tau_s: float = 100e-9, fd_hz: float = 0.0, scs_hz: float = 30e3,
                 max_rank: int = 512, eig_floor: float = 1e-12):
        self._lib = load_library()
        self.cfg = cfg
        p = _hdr(abi.LmmseFilterParamsV1, max_rank=max_rank, tau_s=tau_s, fd_hz=fd_hz, scs_hz=scs_hz,
                 eig_floor=eig_floor)
        self.params = p
        self.handle = C.c_void_p()
        _check("ce CreateFilter", self._lib.airan_channel_est_lmmse_CreateFilter(
            C.byref(cfg), C.byref(p), C.c_void_p(stream_ptr()), C.byref(self.handle)))

    def rank(self, layer: int) -> int:
        return int(self._lib.airan_channel_est_lmmse_FilterRank(self.handle, layer))

    def close(self):
        if self.handle:
            self._lib.airan_channel_est_lmmse_DestroyFilter(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ce_lmmse(cfg, filt: LmmseFilter, h_ls_re, h_ls_im, pilot_sc, pilot_count, noise_var_rx, out=None,
             workspace=None):
    """LMMSE channel estimate [NR,16,14,padded] fp16 planes (layers >= L
    untouched).  Returns (h_re, h_im, workspace, noise_est [NR] fp16)."""
    lib = load_library()
    g = Geometry(cfg)
    dev = h_ls_re.device
    lay = layout(cfg)
    cnt_h = (C.c_uint16 * 16)()
    sc_h = (C.c_uint16 * (g.nl * g.ndmrs * g.pilot_pad))()
    for l in range(g.nl):
        comb = (cfg.dmrs_ports[l] - 1000) // 2
        occ = sum((cfg.dmrs_ports[q] - 1000) // 2 == comb for q in range(g.nl)) == 2
        n = g.n_occ if occ else g.n_dmrs_re
        for d in range(g.ndmrs):
            cnt_h[l * g.ndmrs + d] = n
            for p in range(n):
                sc_h[(l * g.ndmrs + d) * g.pilot_pad + p] = comb + (1 + 4 * p if occ else 2 * p)
    model = abi.LmmseWeightModelV1()
    _check("ce BuildWeightModel", lib.airan_channel_est_lmmse_BuildWeightModel(C.byref(cfg), C.byref(lay), cnt_h,
                                                                               sc_h, C.byref(model)))
    if out is None:
        out = torch.zeros(g.nr * 16 * 14 * g.pad * 2, dtype=torch.float16, device=dev)
    hr, hi = out[: g.nr * 16 * 14 * g.pad], out[g.nr * 16 * 14 * g.pad:]
    a = _hdr(abi.ChannelEstLmmseOpArgsV1, h_ls_re=_dp(h_ls_re), h_ls_im=_dp(h_ls_im), pilot_count=_dp(pilot_count),
             pilot_sc=_dp(pilot_sc), pilot_count_host=cnt_h, pilot_sc_host=sc_h, weight_model=C.pointer(model),
             h_grid_re=_dp(hr), h_grid_im=_dp(hi), config=C.pointer(cfg), layout=C.pointer(lay),
             stream=C.c_void_p(stream_ptr()))
    nws = int(lib.airan_channel_est_lmmse_WorkspaceBytes(C.byref(cfg)))
    if workspace is None:
        workspace = _dev(nws, torch.uint8, dev)
    noise_est = _dev(g.nr, torch.float16, dev)
    _check("ce Enqueue", lib.airan_channel_est_lmmse_Enqueue(C.byref(a), filt.handle, _dp(noise_var_rx),
                                                             _dp(noise_est), _dp(workspace), workspace.numel()))
    return hr, hi, workspace, noise_est


# ----------------------------------------------------------------- detection
def layer_plan(cfg) -> abi.MimoDetectLayerPlan:
    p = _hdr(abi.MimoDetectLayerPlan, num_rx_antennas=cfg.num_rx_antennas, layer_capacity=16,
             total_layers=cfg.num_layers, num_allocations=1)
    p.allocations[0].pusch_index = 0
    p.allocations[0].layer_offset = 0
    p.allocations[0].num_layers = cfg.num_layers
    p.allocations[0].num_tx_ports = cfg.num_tx_ports
    return p


def detect_io_pack(cfg, rx_re, rx_im, h_re, h_im, noise_var_rx):
    lib = load_library()
    g = Geometry(cfg)
    dev = rx_re.device
    plan = layer_plan(cfg)
    meta = abi.MimoDetectIoPackKernelMetadata()
    _check("io_pack BuildCurrentProfile",
           lib.airan_mimo_detect_io_pack_BuildCurrentProfile(C.byref(cfg), 1, C.byref(plan), C.byref(meta)))
    n_re = g.n_re
    hrm_r, hrm_i = _dev(n_re * g.nr * 16, torch.float16, dev), _dev(n_re * g.nr * 16, torch.float16, dev)
    yv_r, yv_i = _dev(n_re // 8 * g.nr * 16, torch.float16, dev), _dev(n_re // 8 * g.nr * 16, torch.float16, dev)
    no = _dev(n_re, torch.float16, dev)
    a = _hdr(abi.MimoDetectIoPackOpArgsV1, rx_grid_re=_dp(rx_re), rx_grid_im=_dp(rx_im), h_grid_re=_dp(h_re),
             h_grid_im=_dp(h_im), noise_var_rx=_dp(noise_var_rx), hrm_re=_dp(hrm_r), hrm_im=_dp(hrm_i),
             yvpad_re=_dp(yv_r), yvpad_im=_dp(yv_i), no=_dp(no), configs=C.pointer(cfg), num_configs=1,
             layer_plan=C.pointer(plan), stream=C.c_void_p(stream_ptr()))
    meta_d = torch.frombuffer(bytearray(bytes(meta)), dtype=torch.uint8).to(dev)
    _check("io_pack Enqueue", lib.airan_mimo_detect_io_pack_Enqueue(C.byref(a), _dp(meta_d), meta_d.numel()))
    return hrm_r, hrm_i, yv_r, yv_i, no


def ls_siso(cfg, rx_re, rx_im, ref_re, ref_im, batch: int = 1, ref_per_slot: bool = False):
    """SISO LS + linear interpolation: rx [batch,14,padded] fp16 planes ->
    H [batch,14,padded] fp16 planes (airan_channel_est_ls_siso_Enqueue)."""
    lib = load_library()
    g = Geometry(cfg)
    dev = rx_re.device
    n = batch * 14 * g.pad
    hr, hi = _dev(n, torch.float16, dev), _dev(n, torch.float16, dev)
    a = _hdr(abi.ChannelEstLsSisoOpArgsV1, rx_grid_re=_dp(rx_re), rx_grid_im=_dp(rx_im), dmrs_ref_re=_dp(ref_re),
             dmrs_ref_im=_dp(ref_im), h_re=_dp(hr), h_im=_dp(hi), config=C.pointer(cfg), batch=batch,
             ref_per_slot=1 if ref_per_slot else 0, stream=C.c_void_p(stream_ptr()))
    _check("ls_siso Enqueue", lib.airan_channel_est_ls_siso_Enqueue(C.byref(a)))
    return hr, hi


def zf_siso(y_re, y_im, h_re, h_im, n0):
    """Single-tap ZF over n = y_re.numel() REs (airan_equalize_zf_siso_Enqueue)."""
    lib = load_library()
    n = y_re.numel()
    dev = y_re.device
    xr, xi, ne = _dev(n, torch.float16, dev), _dev(n, torch.float16, dev), _dev(n, torch.float16, dev)
    a = _hdr(abi.EqualizeZfSisoOpArgsV1, y_re=_dp(y_re), y_im=_dp(y_im), h_re=_dp(h_re), h_im=_dp(h_im),
             n0=_dp(n0), x_re=_dp(xr), x_im=_dp(xi), no_eff=_dp(ne), n=n, stream=C.c_void_p(stream_ptr()))
    _check("zf_siso Enqueue", lib.airan_equalize_zf_siso_Enqueue(C.byref(a)))
    return xr, xi, ne


def detect_lmmse(cfg, hrm_r, hrm_i, yv_r, yv_i, no):
    lib = load_library()
    g = Geometry(cfg)
    dev = hrm_r.device
    plan = layer_plan(cfg)
    n = 16 * g.n_re
    xr, xi, ne = _dev(n, torch.float16, dev), _dev(n, torch.float16, dev), _dev(n, torch.float16, dev)
    a = _hdr(abi.MimoDetectLmmseOpArgsV1, hrm_re=_dp(hrm_r), hrm_im=_dp(hrm_i), yvpad_re=_dp(yv_r),
             yvpad_im=_dp(yv_i), no=_dp(no), xhat_re=_dp(xr), xhat_im=_dp(xi), no_eff=_dp(ne),
             config=C.pointer(cfg), layer_plan=C.pointer(plan), stream=C.c_void_p(stream_ptr()))
    _check("detect Enqueue", lib.airan_mimo_detect_lmmse_Enqueue(C.byref(a)))
    return xr, xi, ne
