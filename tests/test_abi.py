"""CPU checks of the C ABI: the library loads without a GPU, exports every
entry point that include/*.h declares, its ctypes mirror matches the C struct
layouts, and the host-side contract functions behave like the reference's."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from puschrx import abi, ops
from puschrx import profiles as P

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(airan_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if getattr(lib, n, None) is None]
    assert not missing, missing
    assert set(abi.PROTOTYPES) <= set(declared_functions())


STRUCTS = ["PuschMimoConfig", "MimoDetectLayerPlan", "PuschMimoLayout", "PuschMimoRuntimeConfig",
           "OfdmDemodBatchOpArgsV1", "ReDemapBatchOpArgsV1", "ReDemapKernelMetadata", "MimoDmrsGenOpArgsV1",
           "MimoDmrsGenKernelMetadata", "MimoDmrsLsOpArgsV1", "MimoDmrsLsKernelMetadata", "LmmseWeightModelV1",
           "ChannelEstLmmseOpArgsV1", "LmmseFilterParamsV1", "MimoDetectIoPackOpArgsV1",
           "MimoDetectIoPackKernelMetadata", "MimoDetectLmmseOpArgsV1", "QamDemod256BatchOpArgsV1",
           "QamDemodBatchTilingData", "LayerDemapOpArgsV1", "LayerDemapKernelMetadata", "DescrambleMimoOpArgsV1",
           "DescrambleKernelMetadata", "AiranTbPlanV1", "RateMatchDescriptor", "RateDematchMimoOpArgsV1",
           "RateDematchKernelMetadata", "LdpcDecodeOpArgsV1"]


def test_ctypes_layouts_match_c(tmp_path):
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "airan_pusch_rx.h"', "int main(void){"]
    for s in STRUCTS:
        src.append(f'printf("{s} %zu\\n", sizeof({s}));')
    src.append("return 0;}")
    (tmp_path / "s.c").write_text("\n".join(src))
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(tmp_path / "s.c"), "-o", str(tmp_path / "s")],
                   check=True)
    out = subprocess.run([str(tmp_path / "s")], capture_output=True, text=True, check=True).stdout
    sizes = dict(line.split() for line in out.strip().splitlines())
    for s in STRUCTS:
        assert C.sizeof(getattr(abi, s)) == int(sizes[s]), s
    # the four reference-shared ABI structs (common/pusch_mimo_types.h:204-217)
    assert C.sizeof(abi.PuschMimoRuntimeConfig) == 192
    assert abi.PuschMimoRuntimeConfig.slot_number.offset == 152
    assert abi.PuschMimoRuntimeConfig.channel_gain.offset == 128
    assert abi.PuschMimoRuntimeConfig.dmrs_ports.offset == 82


def test_runtime_config_file_and_env(lib, tmp_path, monkeypatch):
    rt = P.REFERENCE_FD16X64.runtime_config()
    path = tmp_path / "fd16x64_rank4.bin"
    path.write_bytes(bytes(memoryview(rt)))
    out = abi.PuschMimoRuntimeConfig()
    why = C.create_string_buffer(128)
    assert lib.airan_LoadMimoRuntimeConfig(str(path).encode(), C.byref(out), why, 128) == 0
    assert bytes(memoryview(out)) == bytes(memoryview(rt))
    monkeypatch.setenv("PUSCH_MIMO_RUNTIME_CONFIG", str(path))
    en = C.c_int32(0)
    assert lib.airan_LoadMimoRuntimeConfigFromEnv(4, C.byref(out), C.byref(en), why, 128) == 0 and en.value == 1
    assert lib.airan_LoadMimoRuntimeConfigFromEnv(3, C.byref(out), C.byref(en), why, 128) == 4
    path.write_bytes(bytes(memoryview(rt))[:100])
    assert lib.airan_LoadMimoRuntimeConfig(str(path).encode(), C.byref(out), why, 128) == 5
    monkeypatch.delenv("PUSCH_MIMO_RUNTIME_CONFIG")
    assert lib.airan_LoadMimoRuntimeConfigFromEnv(4, C.byref(out), C.byref(en), why, 128) == 0 and en.value == 0


def test_product_config_contract_matches_oracle(lib, oracle):
    from oracle_lib import cref
    for prof in (P.REFERENCE_FD16X64, P.FULL_CARRIER_4RX, P.SISO_52PRB):
        rt = prof.runtime_config()
        for field, value in [(None, None), ("rnti", 0), ("qm", 5), ("num_dmrs_symbols", 5), ("flags", 8)]:
            r = abi.PuschMimoRuntimeConfig.from_buffer_copy(bytes(memoryview(rt)))
            if field:
                setattr(r, field, value)
            w1, w2 = C.create_string_buffer(128), C.create_string_buffer(128)
            a = lib.airan_ValidateMimoRuntimeConfig(C.byref(r), w1, 128)
            b = oracle.or_validate_runtime_config(cref(r), w2, 128)
            assert a == b and w1.value == w2.value


def test_tbs_and_plan_match_oracle(lib, oracle):
    from oracle_lib import or_tb_plan
    rng = np.random.default_rng(0)
    for _ in range(200):
        prb = int(rng.integers(4, 274))
        qm = int(rng.choice([2, 4, 6, 8]))
        layers = int(rng.integers(1, 5))
        rate = float(rng.uniform(0.3, 0.93))
        t = ops.tbs(prb, 144, qm, layers, rate)
        assert t == oracle.or_tbs(prb, 144, qm, layers, rate)
        G = prb * 144 * qm * layers
        p = abi.AiranTbPlanV1()
        st = lib.airan_pusch_BuildTbPlan(t, G, qm, layers, 0, C.byref(p))
        q = or_tb_plan(oracle, t, G, qm, layers) if st == 0 else None
        if q is not None:
            assert (p.C, p.K, p.Kprime, p.Z, p.filler, p.E_low, p.E_high, p.n_low, p.k0) == \
                (q.C, q.K, q.Kprime, q.Z, q.filler, q.E_low, q.E_high, q.n_low, q.k0)
            assert lib.airan_ldpc_ActiveRows(C.byref(p)) == oracle.or_ldpc_active_rows(p.Z, p.Kprime, p.E_high,
                                                                                       p.k0, p.Ncb)


def test_reference_plan_distribution(lib):
    cfg = P.Profile(name="r1", rx=64, rank=1, num_rb=133, fft=2048, qm=8, tx=16, padded_subcarriers=1664).config()
    p = ops.reference_tb_plan(cfg, 23)
    # rate_dematch_ref.py:71-76
    assert (p.C, p.n_low, p.E_low, p.E_high) == (143, 87, 24640, 24648)


def test_ops_reject_bad_arguments(lib):
    cfg = P.FULL_CARRIER_4RX.config()
    lay = ops.layout(cfg)
    a = abi.LayerDemapOpArgsV1()
    a.abi_version, a.struct_size = 1, C.sizeof(a)
    a.layer_llr = a.cw_llr = 0x1000  # aliasing
    a.config, a.layout, a.stream = C.pointer(cfg), C.pointer(lay), 0x1
    assert lib.airan_layer_demap_ValidateOpArgs(C.byref(a)) == abi.INVALID_ARGUMENT
    a.cw_llr = 0x2000
    assert lib.airan_layer_demap_ValidateOpArgs(C.byref(a)) == abi.OK
    bad = abi.PuschMimoLayout.from_buffer_copy(bytes(memoryview(lay)))
    bad.data_stride += 64
    a.layout = C.pointer(bad)
    assert lib.airan_layer_demap_ValidateOpArgs(C.byref(a)) == abi.LAYOUT_MISMATCH
    a.layout = C.pointer(lay)
    a.abi_version = 2
    assert lib.airan_layer_demap_ValidateOpArgs(C.byref(a)) == abi.INVALID_ARGUMENT
    cfg2 = abi.PuschMimoConfig.from_buffer_copy(bytes(memoryview(cfg)))
    cfg2.num_rb = 300
    a.abi_version, a.config = 1, C.pointer(cfg2)
    assert lib.airan_layer_demap_ValidateOpArgs(C.byref(a)) == abi.UNSUPPORTED_PROFILE


@pytest.mark.parametrize("prof", [P.FULL_CARRIER_4RX, P.REFERENCE_FD16X64], ids=lambda p: p.name)
def test_compiled_cpp_caller(lib, tmp_path, prof):
    """tests/abi_caller.cpp, compiled with g++ against include/*.h and linked
    to the library (as a host C++ chain would bind it), drives the host-side
    contract functions; its numbers equal the ctypes reading of the same
    library."""
    exe = tmp_path / "abi_caller"
    libdir = ROOT / "ai-ran-npu-kernels_amd" / "lib"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "abi_caller.cpp"), f"-L{libdir}", "-lpuschrx_hip",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    rt = prof.runtime_config()
    path = tmp_path / "runtime.bin"
    path.write_bytes(bytes(rt))
    out = subprocess.run([str(exe), str(path), str(prof.rx)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {}
    for line in out.stdout.splitlines():
        k, _, v = line.partition(" ")
        vals[k] = v
    assert vals["result"] == "OK"
    cfg = prof.config()
    lay = ops.layout(cfg)
    assert int(vals["data_stride"]) == lay.data_stride and int(vals["codeword_stride"]) == lay.codeword_stride
    g = ops.Geometry(cfg)
    tbs = ops.tbs(cfg.num_rb, 12 * g.ndata, g.qm, g.nl, 873 / 1024)
    plan = ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl)
    assert (int(vals["tbs"]), int(vals["C"]), int(vals["Z"]), int(vals["E_low"]), int(vals["E_high"])) == \
        (plan.tbs, plan.C, plan.Z, plan.E_low, plan.E_high)
    assert int(vals["active_rows"]) == lib.airan_ldpc_ActiveRows(C.byref(plan))
    assert int(vals["bad_rnti"].split()[0]) != 0
