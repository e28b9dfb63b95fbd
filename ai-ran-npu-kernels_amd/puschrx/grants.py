"""Grant lists -> 192-byte runtime configs, and ragged multi-grant batches.

The compiler restates the rules of the reference's control plane
(common/mimo_config_compiler.py:79-252: a radio profile plus a PUSCH grant
become one strictly validated PuschMimoRuntimeConfig) for a list of grants:
rank within the radio's supported ranks, the symbol allocation inside one
14-symbol slot, DMRS symbols strictly increasing inside the allocation, the
nr_4port DMRS port policy (`auto` -> {1000}, {1000, 1002}, {1000..1002},
{1000..1003}; explicit ports must equal it), precoding `auto` -> codebook for
rank 3 else bypass (TPMI 6 / 0, PRG = the grant's PRBs), the transport and
scheduler integer ranges, and the receiver capacity; every packed config is
then checked by the library's airan_ValidateMimoRuntimeConfig (the rules of
common/pusch_mimo_runtime_config.cpp:23-170).

`config4_grants` draws the ragged workload (96 grants per slot: PRB
uniform in 4..273, rank 1..4 on 4 Rx, 16/64/256QAM, code rate 0.3..0.9, each
grant its own cell with the smallest FFT that holds it at <= 85 % occupancy).
`GrantBatch` runs a list of grants through the library's ragged batch
(include/airan_pusch_grants.h).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi, ops
from .profiles import Profile, align_up

STANDARD_DMRS_PORTS = {1: [1000], 2: [1000, 1002], 3: [1000, 1001, 1002], 4: [1000, 1001, 1002, 1003]}


class GrantError(ValueError):
    pass


@dataclass
class Radio:
    """The static radio profile (common/radio_profiles/*.json, the fields the
    compiler reads)."""
    name: str = "fd4x4"
    rx_antennas: int = 4
    tx_antennas: int = 4
    supported_ranks: tuple = (1, 2, 3, 4)
    max_rx_antennas: int = 4


@dataclass
class Grant:
    """One PUSCH grant (common/pusch_grants/*.json, flattened)."""
    name: str
    rank: int
    qm: int
    num_rb: int
    fft_size: int
    code_rate: float
    slot_number: int = 0
    rnti: int = 12345
    data_scrambling_id: int = 321
    dmrs_scrambling_id: int = 321
    codeword_index: int = 0
    num_symbols: int = 14
    start_symbol: int = 0
    rb_start: int = 0
    dmrs_symbols: tuple = (2, 11)
    ports: object = "auto"
    n_scid: int = 0
    additional_position: int = 0
    mapping_type: int = 0
    precoding_mode: str = "auto"
    padded_subcarriers: int | None = None
    extra: dict = field(default_factory=dict)


def _int(v, where, lo, hi):
    if isinstance(v, bool) or not isinstance(v, (int, np.integer)) or not lo <= int(v) <= hi:
        raise GrantError(f"{where} must be an integer in [{lo},{hi}]")
    return int(v)


def compile_grant(radio: Radio, g: Grant) -> abi.PuschMimoRuntimeConfig:
    """One grant -> its runtime config (mimo_config_compiler.py:79-252)."""
    rank = _int(g.rank, "grant.rank", 1, 4)
    if rank not in radio.supported_ranks:
        raise GrantError(f"Rank{rank} is not supported by radio {radio.name}")
    nsym = _int(g.num_symbols, "grant.waveform.num_symbols", 1, 14)
    start = _int(g.start_symbol, "grant.waveform.start_symbol", 0, 13)
    if start + nsym > 14:
        raise GrantError("grant symbol allocation exceeds one 14-symbol slot")
    dmrs = [_int(s, "grant.dmrs.symbols[]", 0, 13) for s in g.dmrs_symbols]
    if not dmrs or dmrs != sorted(set(dmrs)):
        raise GrantError("grant.dmrs.symbols must be unique and strictly increasing")
    if any(s < start or s >= start + nsym for s in dmrs):
        raise GrantError("DMRS symbol lies outside the PUSCH symbol allocation")
    ports = STANDARD_DMRS_PORTS[rank] if g.ports == "auto" else list(g.ports)
    if len(ports) != rank or len(set(ports)) != rank:
        raise GrantError(f"grant.dmrs.ports must contain exactly {rank} unique ports")
    if ports != STANDARD_DMRS_PORTS[rank]:
        raise GrantError(f"installed nr_4port policy requires Rank{rank} DMRS ports {STANDARD_DMRS_PORTS[rank]}")
    if g.precoding_mode not in ("auto", "bypass", "codebook", "non_codebook"):
        raise GrantError("unsupported grant.precoding.mode")
    _int(g.slot_number, "grant.slot_number", 0, 1023)
    _int(g.rnti, "grant.transport.rnti", 1, 65535)
    _int(g.data_scrambling_id, "grant.transport.data_scrambling_id", 0, 1023)
    _int(g.dmrs_scrambling_id, "grant.dmrs.scrambling_id", 0, 65535)
    _int(g.codeword_index, "grant.transport.codeword_index", 0, 1)
    _int(g.additional_position, "grant.dmrs.additional_position", 0, 3)
    _int(g.mapping_type, "grant.dmrs.mapping_type", 0, 1)
    _int(g.n_scid, "grant.dmrs.n_scid", 0, 1)
    if radio.max_rx_antennas < radio.rx_antennas:
        raise GrantError("radio receiver capacity is below physical RX antennas")
    if g.qm not in (2, 4, 6, 8):
        raise GrantError("grant.waveform.qm must be 2, 4, 6 or 8")
    prof = Profile(name=g.name, rx=radio.rx_antennas, rank=rank, num_rb=g.num_rb, fft=g.fft_size, qm=g.qm,
                   dmrs_symbols=tuple(dmrs), rnti=g.rnti, data_scrambling_id=g.data_scrambling_id,
                   dmrs_scrambling_id=g.dmrs_scrambling_id, slot_number=g.slot_number, n_scid=g.n_scid,
                   padded_subcarriers=g.padded_subcarriers, code_rate=g.code_rate, ports=tuple(ports))
    rt = prof.runtime_config()
    why = C.create_string_buffer(256)
    st = abi.load_library().airan_ValidateMimoRuntimeConfig(C.byref(rt), why, 256)
    if st != 0:
        raise GrantError(f"runtime config of {g.name} rejected ({st}): {why.value.decode()}")
    return rt


def compile_grants(radio: Radio, grants: list) -> list:
    """Grant list -> (runtime config, operator config, TB plan) per grant."""
    lib = abi.load_library()
    out = []
    for g in grants:
        rt = compile_grant(radio, g)
        cfg = abi.PuschMimoConfig()
        why = C.create_string_buffer(256)
        st = lib.airan_DerivePuschMimoConfig(C.byref(rt), radio.rx_antennas, C.byref(cfg), why, 256, g.slot_number)
        if st != 0:
            raise GrantError(f"{g.name}: {why.value.decode()}")
        geo = ops.Geometry(cfg)
        plan = ops.tb_plan(ops.tbs(g.num_rb, 12 * geo.ndata, g.qm, rank_of(cfg), g.code_rate),
                           geo.cw_symbols * geo.qm, geo.qm, geo.nl)
        out.append((rt, cfg, plan))
    return out


def rank_of(cfg) -> int:
    return int(cfg.num_layers)


def fft_for(num_rb: int, occupancy: float = 0.85) -> int:
    """Smallest power-of-two FFT (128..4096) holding 12 * num_rb subcarriers
    at <= `occupancy` of its bins."""
    used = 12 * num_rb
    n = 128
    while n < 4096 and used > occupancy * n:
        n *= 2
    return n


def config4_grants(rng: np.random.Generator, n: int = 96, slot_number: int = 0, prb=(4, 273), ranks=(1, 2, 3, 4),
                   qms=(4, 6, 8), rate=(0.3, 0.9)) -> list:
    """Ragged grants of one slot: PRB uniform in prb, rank uniform in ranks,
    Qm uniform in qms, code rate uniform in rate; each its own cell."""
    out = []
    for i in range(n):
        nrb = int(rng.integers(prb[0], prb[1] + 1))
        fft = fft_for(nrb)
        out.append(Grant(name=f"g{i:03d}", rank=int(rng.choice(ranks)), qm=int(rng.choice(qms)), num_rb=nrb,
                         fft_size=fft, code_rate=float(rng.uniform(*rate)), slot_number=slot_number,
                         rnti=int(rng.integers(1, 65536)), data_scrambling_id=int(rng.integers(0, 1024)),
                         dmrs_scrambling_id=int(rng.integers(0, 1024)),
                         padded_subcarriers=min(align_up(12 * nrb, 128), fft)))
    return out


class GrantBatch:
    """A ragged batch of grants on the device (airan_pusch_grants_*)."""

    def __init__(self, compiled: list, opts=None, detect_tile: int = 256, use_graph: bool = True):
        import torch
        self._torch = torch
        self.lib = abi.load_library()
        self.n = len(compiled)
        self.cfgs = (abi.PuschMimoConfig * self.n)(*[c for _, c, _ in compiled])
        self.plans = (abi.AiranTbPlanV1 * self.n)(*[p for _, _, p in compiled])
        o = abi.AiranGrantBatchOptionsV1()
        o.abi_version, o.struct_size = abi.ABI_VERSION, C.sizeof(o)
        opts = opts or {}
        o.iq_format = opts.get("iq_format", abi.IQ_COMPLEX64)
        o.ldpc_iterations = opts.get("ldpc_iterations", 8)
        o.early_stop = 1 if opts.get("early_stop", True) else 0
        o.ce_max_rank = opts.get("ce_rank", 384)
        o.use_graph = 1 if use_graph else 0
        o.tau_s, o.fd_hz, o.scs_hz = opts.get("tau_s", 300e-9), opts.get("fd_hz", 100.0), opts.get("scs_hz", 30e3)
        o.detect_tile = detect_tile
        self.options = o
        self.handle = C.c_void_p()
        ops._check("grants Create", self.lib.airan_pusch_grants_Create(self.cfgs, self.plans, self.n, C.byref(o),
                                                                      C.c_void_p(ops.stream_ptr()),
                                                                      C.byref(self.handle)))
        lay = (abi.AiranGrantLayoutV1 * self.n)()
        total = C.c_uint64()
        ops._check("grants Layout", self.lib.airan_pusch_grants_Layout(self.handle, lay, self.n, C.byref(total)))
        self.layout = list(lay)
        self.iq_samples = int(total.value)

    def pack_iq(self, iqs: list):
        """Per-grant IQ arrays [NR, samples, 2] -> one device buffer."""
        torch = self._torch
        dt = np.float32 if self.options.iq_format == abi.IQ_COMPLEX64 else np.int16
        host = np.zeros((self.iq_samples, 2), dt)
        for lo, x in zip(self.layout, iqs):
            host[lo.iq_offset: lo.iq_offset + lo.iq_samples] = np.asarray(x, dt).reshape(-1, 2)
        return torch.from_numpy(host).cuda()

    def enqueue(self, iq_dev):
        ops._check("grants Enqueue", self.lib.airan_pusch_grants_Enqueue(self.handle, C.c_void_p(iq_dev.data_ptr()),
                                                                        C.c_void_p(ops.stream_ptr())))

    def outputs(self) -> abi.AiranGrantBatchOutputsV1:
        o = abi.AiranGrantBatchOutputsV1()
        o.abi_version, o.struct_size = abi.ABI_VERSION, C.sizeof(o)
        ops._check("grants Outputs", self.lib.airan_pusch_grants_Outputs(self.handle, C.byref(o)))
        return o

    def fetch(self, name: str, dtype, count: int, offset: int = 0):
        from .chain import _hip
        torch = self._torch
        src = getattr(self.outputs(), name)
        out = torch.empty(count, dtype=dtype, device="cuda")
        es = out.element_size()
        rc = _hip().hipMemcpyAsync(out.data_ptr(), src + offset * es, count * es, 3, ops.stream_ptr())
        if rc != 0:
            raise ops.AiranError("hipMemcpyAsync", rc)
        return out

    def set_detect_tile(self, tile: int):
        ops._check("grants SetDetectTile", self.lib.airan_pusch_grants_SetDetectTile(self.handle, tile))

    @property
    def num_shapes(self) -> int:
        return int(self.lib.airan_pusch_grants_NumShapes(self.handle))

    @property
    def num_cb(self) -> int:
        return int(self.outputs().num_cb)

    def close(self):
        if self.handle:
            self.lib.airan_pusch_grants_Destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
