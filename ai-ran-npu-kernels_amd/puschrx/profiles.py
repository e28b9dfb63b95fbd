"""Runtime-config construction for the profiles this library serves.

Plays the role of the reference's control-plane compiler
(common/mimo_config_compiler.py:223-252 -> ResolvedPlan.to_runtime_config,
common/mimo_profile_compiler.py:51-103): a radio profile plus a grant become
one 192-byte PuschMimoRuntimeConfig, from which every operator derives its
PuschMimoConfig with airan_DerivePuschMimoConfig.
"""
from __future__ import annotations

import ctypes as C
import hashlib
from dataclasses import dataclass, field

from . import abi

# Ports per rank (common/mimo_config_compiler.py:26-31): FD-OCC pairs on comb 0/1.
PORTS_BY_RANK = {1: (1000,), 2: (1000, 1002), 3: (1000, 1001, 1002), 4: (1000, 1001, 1002, 1003)}


def align_up(v: int, a: int) -> int:
    return (v + a - 1) // a * a


@dataclass
class Profile:
    """A radio profile plus one PUSCH grant (the inputs of the reference compiler)."""
    name: str
    rx: int
    rank: int
    num_rb: int
    fft: int
    qm: int
    dmrs_symbols: tuple = (2, 11)
    tx: int | None = None
    rnti: int = 12345
    data_scrambling_id: int = 321
    dmrs_scrambling_id: int = 321
    slot_number: int = 0
    n_scid: int = 0
    num_cdm_groups_without_data: int = 2
    data_re_alignment: int = 64
    qam_row_alignment: int = 64
    padded_subcarriers: int | None = None
    channel_gain: float = 0.16
    awgn_std_int16: float = 80.0
    code_rate: float | None = None      # None: the reference's fixed 143-CB TB
    layer_bucket: int | None = None     # detector storage width when padded
    ports: tuple | None = None
    extra: dict = field(default_factory=dict)

    @property
    def used(self) -> int:
        return 12 * self.num_rb

    @property
    def padded(self) -> int:
        if self.padded_subcarriers is not None:
            return self.padded_subcarriers
        return align_up(self.used, 128)

    def port_list(self) -> tuple:
        return self.ports if self.ports is not None else PORTS_BY_RANK[self.rank]

    def runtime_config(self) -> abi.PuschMimoRuntimeConfig:
        """Pack the 192-byte runtime config (mimo_profile_compiler.py:51-103)."""
        c = abi.PuschMimoRuntimeConfig()
        c.abi_version = abi.RUNTIME_ABI_VERSION
        c.struct_size = C.sizeof(c)
        c.flags = 1  # profile hash present
        digest = hashlib.sha256(repr((self.name, self.rx, self.rank, self.num_rb, self.fft, self.qm,
                                      tuple(self.dmrs_symbols))).encode()).digest()
        for i, b in enumerate(digest):
            c.profile_sha256[i] = b
        # Rank 3 is carried on 4 logical ports with codebook precoding
        # (common/tests/test_mimo_config_compiler.py:22: rank 3 -> 4 ports).
        tx_ports = 4 if self.rank == 3 else self.rank
        tx = self.tx if self.tx is not None else tx_ports
        layer_bucket = self.layer_bucket if self.layer_bucket is not None else self.rank
        if layer_bucket != self.rank:
            c.flags |= 2  # capacity adapter required
        c.architecture = 1
        c.num_tx_antennas = tx
        c.num_tx_rf_chains = tx
        c.num_rx_antennas = self.rx
        c.num_rx_rf_chains = self.rx
        c.num_layers = self.rank
        c.num_tx_ports = tx_ports
        c.rx_bucket = self.rx
        c.layer_bucket = layer_bucket
        c.max_rx_antennas = self.rx
        c.max_layers = 4
        c.qm = self.qm
        c.num_symbols = 14
        c.fft_size = self.fft
        c.num_rb = self.num_rb
        c.rb_start = 0
        c.used_subcarriers = self.used
        c.padded_subcarriers = self.padded
        mask = 0
        for s in self.dmrs_symbols:
            mask |= 1 << s
        c.num_dmrs_symbols = len(self.dmrs_symbols)
        c.dmrs_symbol_mask = mask
        ports = self.port_list()
        c.dmrs_port_count = len(ports)
        for i, p in enumerate(ports):
            c.dmrs_ports[i] = p
        c.dmrs_type = 1
        c.dmrs_length = 1
        c.num_cdm_groups_without_data = self.num_cdm_groups_without_data
        c.n_scid = self.n_scid
        c.precoding_mode = 0 if tx_ports == self.rank else 1
        c.tpmi = 0
        c.prg_size_rb = self.num_rb
        c.channel_gain = self.channel_gain
        c.awgn_std_int16 = self.awgn_std_int16
        data_re = (14 - len(self.dmrs_symbols)) * self.used
        c.data_re_per_layer = data_re
        c.data_stride_per_layer = align_up(data_re, self.data_re_alignment)
        c.qam_row_stride = align_up(self.used, self.qam_row_alignment)
        c.grid_re_per_port = 14 * self.padded
        c.slot_number = self.slot_number
        c.dmrs_scrambling_id = self.dmrs_scrambling_id
        c.data_scrambling_id = self.data_scrambling_id
        c.rnti = self.rnti
        c.start_symbol = 0
        c.num_allocated_symbols = 14
        c.dmrs_additional_position = 0
        c.mapping_type = 0
        c.codeword_index = 0
        return c

    def config(self, slot_number: int = 0, lib=None) -> abi.PuschMimoConfig:
        """Operator ABI v1 config derived through the library."""
        lib = lib or abi.load_library()
        rt = self.runtime_config()
        out = abi.PuschMimoConfig()
        why = C.create_string_buffer(256)
        st = lib.airan_DerivePuschMimoConfig(C.byref(rt), self.rx, C.byref(out), why, 256, slot_number)
        if st != 0:
            raise ValueError(f"runtime config rejected ({st}): {why.value.decode()}")
        return out


# The reference's fixed profile: fd16x64 radio, full_band grant, rank 4
# (common/radio_profiles/fd16x64.json, common/pusch_grants/full_band.json,
# chain/pusch_mimo_rx_chain/rx_chain.cpp:64-67,232-237).
REFERENCE_FD16X64 = Profile(name="fd16x64_full_band", rx=64, rank=4, num_rb=133, fft=2048, qm=8,
                            tx=16, padded_subcarriers=1664, layer_bucket=16)

# Headline profile: 4 Rx, rank 4, 273 PRB, FFT 4096, 256QAM, R ~ 0.85
# (MCS 25 of TS 38.214 Table 5.1.3.1-2: R = 873/1024).
FULL_CARRIER_4RX = Profile(name="fd4x4_273prb", rx=4, rank=4, num_rb=273, fft=4096, qm=8,
                           padded_subcarriers=3328, code_rate=873 / 1024)

# Mid-size plain MIMO profile: 4 Rx, rank 2 on CDM ports 0/1 (one comb,
# FD-OCC2), 106 PRB, FFT 2048, 64QAM, R ~ 0.6.
MID_4RX_RANK2 = Profile(name="fd4x2_106prb", rx=4, rank=2, num_rb=106, fft=2048, qm=6,
                        padded_subcarriers=1280, code_rate=0.6, ports=(1000, 1001))

# Small SISO profile: 1 Rx, rank 1, 52 PRB, FFT 1024, 256QAM, one DMRS symbol.
SISO_52PRB = Profile(name="fd1x1_52prb", rx=1, rank=1, num_rb=52, fft=1024, qm=8, dmrs_symbols=(2,),
                     padded_subcarriers=640, code_rate=0.75)

# The plain SISO workload: 1 Tx / 1 Rx, 52 PRB in a 1024-point FFT, one DMRS
# symbol (2), 256QAM at R ~ 0.75, data scrambling c_init = 0x1234 * 2^15 + 0
# (rnti 0x1234, n_ID 0); received by LS + linear interpolation and ZF.
SISO_CONFIG1 = Profile(name="fd1x1_52prb_ls_zf", rx=1, rank=1, num_rb=52, fft=1024, qm=8, dmrs_symbols=(2,),
                       padded_subcarriers=640, code_rate=0.75, rnti=0x1234, data_scrambling_id=0)

# SISO on the reference's grid (133 PRB, FFT 2048, DMRS {2, 11}): the geometry
# of channel_est_ls_siso / equalize_zf_siso (1596 used, 1664 padded).
SISO_133PRB = Profile(name="fd1x1_133prb", rx=1, rank=1, num_rb=133, fft=2048, qm=8,
                      padded_subcarriers=1664, code_rate=0.75)
