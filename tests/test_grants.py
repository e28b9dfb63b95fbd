"""Grant-list compiler (puschrx/grants.py): the reference control plane's
rules (common/mimo_config_compiler.py:79-252) on each grant, the packed
192-byte configs accepted by the library's validator, and the random
ragged workload's ranges."""
import ctypes as C
import dataclasses
import struct

import numpy as np
import pytest

from puschrx import abi
from puschrx import grants as G
from puschrx import profiles as P


def _full_band():
    # common/pusch_grants/full_band.json on the fd16x64 radio.
    return G.Grant(name="full_band", rank=4, qm=8, num_rb=133, fft_size=2048, code_rate=None, rnti=12345,
                   data_scrambling_id=321, dmrs_scrambling_id=321, padded_subcarriers=1664)


def test_full_band_grant_known_bytes():
    """The reference grant full_band.json compiles to the scheduler bytes
    common/tests/test_mimo_config_compiler.py:35-36 asserts."""
    radio = G.Radio(name="fd16x64", rx_antennas=64, tx_antennas=16, max_rx_antennas=64)
    rt = G.compile_grant(radio, _full_band())
    b = bytes(memoryview(rt))
    assert len(b) == 192
    assert struct.unpack_from("<10H", b, 152) == (0, 321, 321, 12345, 0, 14, 0, 0, 0, 0)
    # The same bytes (except the profile hash) as the fixed reference profile.
    ref = bytes(memoryview(P.REFERENCE_FD16X64.runtime_config()))
    assert b[64:] == ref[64:]


BAD = [
    ("rank", 5, "grant.rank"),
    ("num_symbols", 15, "num_symbols"),
    ("dmrs_symbols", (11, 2), "strictly increasing"),
    ("ports", [1000, 1001], "nr_4port"),
    ("rnti", 0, "rnti"),
    ("data_scrambling_id", 1024, "data_scrambling_id"),
    ("slot_number", 1024, "slot_number"),
    ("n_scid", 2, "n_scid"),
    ("precoding_mode", "magic", "precoding"),
    ("qm", 5, "qm"),
]


@pytest.mark.parametrize("field,value,msg", BAD, ids=[b[0] for b in BAD])
def test_grant_rules_reject(field, value, msg):
    g = dataclasses.replace(G.Grant(name="g", rank=2, qm=6, num_rb=50, fft_size=1024, code_rate=0.5), **{field: value})
    with pytest.raises(G.GrantError, match=msg):
        G.compile_grant(G.Radio(), g)


def test_rank_outside_radio_and_dmrs_outside_allocation():
    with pytest.raises(G.GrantError, match="not supported"):
        G.compile_grant(G.Radio(supported_ranks=(1, 2)), G.Grant(name="g", rank=3, qm=4, num_rb=20, fft_size=512,
                                                                 code_rate=0.5))
    with pytest.raises(G.GrantError, match="outside"):
        G.compile_grant(G.Radio(), G.Grant(name="g", rank=1, qm=4, num_rb=20, fft_size=512, code_rate=0.5,
                                           start_symbol=3, num_symbols=11))


def test_config4_grants_are_valid_and_ragged():
    rng = np.random.default_rng(7)
    gl = G.config4_grants(rng, 96)
    comp = G.compile_grants(G.Radio(), gl)
    assert len(comp) == 96
    lib = abi.load_library()
    for g, (rt, cfg, plan) in zip(gl, comp):
        assert 4 <= g.num_rb <= 273 and g.rank in (1, 2, 3, 4) and g.qm in (4, 6, 8) and 0.3 <= g.code_rate <= 0.9
        assert 12 * g.num_rb <= 0.85 * g.fft_size or g.fft_size == 4096
        assert g.fft_size == 128 or 12 * g.num_rb > 0.85 * g.fft_size / 2
        assert lib.airan_ValidateMimoRuntimeConfig(C.byref(rt), None, 0) == 0
        assert cfg.num_layers == g.rank and list(cfg.dmrs_ports)[: g.rank] == G.STANDARD_DMRS_PORTS[g.rank]
        assert plan.C >= 1 and plan.crc_tb in (16, 24) and plan.G == cfg.num_layers * g.qm * (14 - 2) * 12 * g.num_rb
    assert len({g.fft_size for g in gl}) >= 4 and len({(g.num_rb, g.rank) for g in gl}) > 48
