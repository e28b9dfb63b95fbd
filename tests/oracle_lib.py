"""ctypes access to the CPU oracle (oracle/lib/liboracle.so) and to the
reference host sources built into oracle/_ref/libref.so.  Test-only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle" / "lib" / "liboracle.so"
REF = ROOT / "oracle" / "_ref" / "libref.so"

_ORACLE = None


class OrTbPlan(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("tbs", "C", "K", "Kprime", "Z", "filler", "G", "E_low", "E_high",
                                          "n_low", "N", "Ncb", "k0", "bg", "crc_tb", "crc_cb")]


class OrCeParams(C.Structure):
    _fields_ = [("tau_s", C.c_double), ("fd_hz", C.c_double), ("scs_hz", C.c_double)]


def load_oracle():
    global _ORACLE
    if _ORACLE is not None:
        return _ORACLE
    if not ORACLE.exists():
        raise RuntimeError(f"{ORACLE} missing: run __graft_entry__.build()")
    lib = C.CDLL(str(ORACLE))
    vp, u16p, u32, u64 = C.c_void_p, C.POINTER(C.c_uint16), C.c_uint32, C.c_uint64
    lib.or_f2h.restype = C.c_uint16
    lib.or_f2h.argtypes = [C.c_float]
    lib.or_h2f.restype = C.c_float
    lib.or_h2f.argtypes = [C.c_uint16]
    lib.or_gold.argtypes = [u32, u64, u64, vp]
    lib.or_dmrs_cinit.restype = u32
    lib.or_dmrs_cinit.argtypes = [vp, C.c_int]
    lib.or_validate_runtime_config.argtypes = [vp, C.c_char_p, C.c_int]
    lib.or_derive_config.argtypes = [vp, C.c_uint16, C.c_uint16, vp, C.c_char_p, C.c_int]
    lib.or_re_demap_gather_index.argtypes = [vp, vp]
    lib.or_re_demap.argtypes = [vp, vp, vp, vp, vp]
    lib.or_dmrs_gen.argtypes = [vp, vp, vp]
    lib.or_dmrs_ls.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.or_ce_lmmse.argtypes = [vp, vp, vp, vp, vp]
    lib.or_ce_noise_rel.argtypes = [vp, vp, C.c_double, vp]
    lib.or_detect_lmmse.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.or_ls_siso.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.or_zf_siso.argtypes = [C.c_size_t, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.or_qam_axis_llr.argtypes = [C.c_int, C.c_float, C.c_float, vp]
    lib.or_qam_axis_llr_bruteforce.argtypes = [C.c_int, C.c_float, C.c_float, vp]
    lib.or_qam_demod.argtypes = [vp, vp, vp, vp, vp]
    lib.or_layer_demap.argtypes = [vp, vp, vp]
    lib.or_descramble.argtypes = [vp, u32, u64, vp, vp]
    lib.or_tbs.restype = u32
    lib.or_tbs.argtypes = [u32, u32, u32, u32, C.c_double]
    lib.or_tb_plan.argtypes = [u32, u32, u32, u32, u32, C.POINTER(OrTbPlan)]
    lib.or_rate_dematch.argtypes = [C.POINTER(OrTbPlan), u32, u32, u32, u32, u32, vp, vp]
    lib.or_ldpc_active_rows.argtypes = [u32, u32, u32, u32, u32]
    lib.or_ldpc_decode.argtypes = [u32, u32, vp, C.c_int, C.c_int, C.c_int, u32, C.c_int, vp, vp]
    lib.or_ldpc_decode_posterior.argtypes = [u32, u32, vp, C.c_int, C.c_int, C.c_int, u32, C.c_int, vp, vp, vp]
    lib.or_ldpc_encode.argtypes = [u32, vp, vp]
    lib.or_ldpc_syndrome_ok.argtypes = [u32, vp]
    lib.or_crc.restype = u32
    lib.or_crc.argtypes = [C.c_int, vp, u64]
    lib.or_ofdm_demod.argtypes = [vp, vp, vp]
    lib.or_chain_slot.argtypes = [vp, C.POINTER(OrTbPlan), C.POINTER(OrCeParams), C.c_int, vp, vp, vp]
    lib.or_num_threads.restype = C.c_int
    lib.or_refiner_param_offset.restype = C.c_size_t
    lib.or_refiner_param_offset.argtypes = [C.c_int]
    lib.or_refiner_grad.restype = C.c_int
    lib.or_refiner_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
    lib.or_refiner_grad_dil.restype = C.c_int
    lib.or_refiner_grad_dil.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
    lib.or_refiner_grad_full.restype = C.c_int
    lib.or_refiner_grad_full.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.or_bf16_round.restype = C.c_double
    lib.or_bf16_round.argtypes = [C.c_double]
    _ORACLE = lib
    return lib


def load_ref():
    if not REF.exists():
        return None
    lib = C.CDLL(str(REF))
    vp = C.c_void_p
    lib.ref_validate_runtime_config.argtypes = [vp, C.c_char_p, C.c_int]
    lib.ref_derive_config.argtypes = [vp, C.c_uint, C.c_uint, vp]
    lib.ref_re_demap_gather_index.argtypes = [vp, C.c_size_t]
    lib.ref_re_demap_reference.argtypes = [vp, vp, C.c_uint, vp, vp, vp]
    lib.ref_re_demap_profile.argtypes = [vp, vp]
    lib.ref_layer_demap_reference.argtypes = [vp, vp, vp]
    lib.ref_dmrs_gen_profile.argtypes = [vp, vp, vp]
    lib.ref_dmrs_gold_basis.argtypes = [vp, vp]
    lib.ref_dmrs_ls_profile.argtypes = [vp, vp, vp, vp]
    lib.ref_dmrs_ls_describe.argtypes = [vp, C.c_uint, C.c_uint, vp]
    lib.ref_dmrs_ls_validate_natural.argtypes = [vp, vp, C.c_uint, C.c_uint]
    lib.ref_dmrs_ls_validate_798.argtypes = [vp, C.c_uint, C.c_uint]
    lib.ref_pilot_sc_hash.restype = C.c_ulonglong
    lib.ref_pilot_sc_hash.argtypes = [vp, C.c_uint]
    return lib


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def cref(s):
    return C.cast(C.pointer(s), C.c_void_p)


class OracleGeom:
    """Python copy of OrGeom (oracle/oracle.h) for shapes in tests."""

    def __init__(self, cfg):
        a = lambda v, m: (v + m - 1) // m * m  # noqa: E731
        self.nr, self.nl, self.qm = cfg.num_rx_antennas, cfg.num_layers, cfg.qm
        self.fft, self.used, self.pad = cfg.fft_size, cfg.used_subcarriers, cfg.padded_subcarriers
        self.dmrs_sym = [s for s in range(14) if (cfg.dmrs_symbol_mask >> s) & 1]
        self.data_sym = [s for s in range(14) if not (cfg.dmrs_symbol_mask >> s) & 1]
        self.ndmrs, self.ndata = len(self.dmrs_sym), len(self.data_sym)
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


def or_tb_plan(oracle, tbs, G, qm, layers, rv=0) -> OrTbPlan:
    p = OrTbPlan()
    st = oracle.or_tb_plan(tbs, G, qm, layers, rv, C.byref(p))
    assert st == 0, st
    return p


def gold(oracle, cinit: int, start: int, count: int) -> np.ndarray:
    c = np.zeros(count, np.uint8)
    oracle.or_gold(cinit, start, count, ptr(c))
    return c


# ---------------------------------------------------------------- CPU chain
# oracle/lib/libcpuchain_{v3,v4}.so: the FP32 C++ chain timed as the CPU
# baseline (one build per x86-64 level; the highest the host supports runs).
_V4_FLAGS = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
_V3_FLAGS = {"avx2", "fma", "bmi2", "movbe"}


def host_cpu() -> dict:
    """CPU model, feature level and logical CPU count of this host."""
    model, flags = "unknown", set()
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    import os
    level = "v4" if _V4_FLAGS <= flags else ("v3" if _V3_FLAGS <= flags else None)
    return {"model": model, "isa": level, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


class CpuChain:
    """One grant of the FP32 CPU chain: the CE eigenpairs come from LAPACK
    (numpy.linalg.eigh, FP64) once, as the GPU chain takes them from
    rocSOLVER; the top `rank` eigenpairs above 1e-12 of the largest are kept."""

    def __init__(self, cfg, op: OrTbPlan, prm: OrCeParams, ldpc_iters: int = 8, rank: int = 384, isa=None):
        info = host_cpu()
        isa = isa or info["isa"]
        if isa is None:
            raise RuntimeError("host CPU lacks x86-64-v3 (AVX2 / FMA)")
        path = ROOT / "oracle" / "lib" / f"libcpuchain_{isa}.so"
        if not path.exists():
            raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
        lib = C.CDLL(str(path))
        vp = C.c_void_p
        lib.cpu_chain_create.argtypes = [vp, vp, vp, C.c_int, C.POINTER(vp)]
        lib.cpu_chain_num_models.argtypes = [vp]
        lib.cpu_chain_model_rpp.argtypes = [vp, C.c_int, vp]
        lib.cpu_chain_set_eigen.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        lib.cpu_chain_run.argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, vp]
        lib.cpu_chain_destroy.argtypes = [vp]
        lib.cpu_chain_max_threads.restype = C.c_int
        self.lib, self.isa, self.op = lib, isa, op
        self._keep = (cfg, op, prm)
        self.h = vp()
        if lib.cpu_chain_create(C.byref(cfg), C.byref(op), C.byref(prm), ldpc_iters, C.byref(self.h)) != 0:
            raise RuntimeError("cpu_chain_create failed")
        cache = {}
        for q in range(lib.cpu_chain_num_models(self.h)):
            n = lib.cpu_chain_model_rpp(self.h, q, None)
            r = np.zeros((n, n, 2))
            lib.cpu_chain_model_rpp(self.h, q, r.ctypes.data)
            key = r.tobytes()
            if key not in cache:
                lam, v = np.linalg.eigh(r[..., 0] + 1j * r[..., 1])
                lam, v = lam[::-1], v[:, ::-1]
                kept = max(1, int(min(rank, np.count_nonzero(lam >= 1e-12 * lam[0]))))
                cache[key] = (np.ascontiguousarray(lam[:kept]),
                              np.ascontiguousarray(v[:, :kept].T).view(np.float64).copy())
            lam, vec = cache[key]
            if lib.cpu_chain_set_eigen(self.h, q, len(lam), lam.ctypes.data, vec.ctypes.data) != 0:
                raise RuntimeError("cpu_chain_set_eigen failed")

    def max_threads(self) -> int:
        return int(self.lib.cpu_chain_max_threads())

    def run(self, iq: np.ndarray, threads: int):
        """iq complex64 as float32 [slots, NR, samples, 2] -> (bits int8
        [slots, C, K], tb_ok uint8 [slots])."""
        iq = np.ascontiguousarray(iq, dtype=np.float32)
        n = iq.shape[0]
        bits = np.zeros((n, self.op.C, self.op.K), np.int8)
        ok = np.zeros(n, np.uint8)
        if self.lib.cpu_chain_run(self.h, iq.ctypes.data, n, None, threads, bits.ctypes.data, ok.ctypes.data) != 0:
            raise RuntimeError("cpu_chain_run failed")
        return bits, ok

    def close(self):
        if self.h:
            self.lib.cpu_chain_destroy(self.h)
            self.h = C.c_void_p()
