"""ctypes mirror of include/airan_pusch_types.h and include/airan_pusch_rx.h.

Struct layouts restate the reference ABI (common/pusch_mimo_types.h:48-188 and
the per-operator OpArgs structs cited in airan_pusch_rx.h); a test compares
every sizeof/offsetof against the C compiler.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parents[1]
LIB_PATH = PKG_DIR / "lib" / "libpuschrx_hip.so"

u8, u16, u32, u64, i16, i32 = C.c_uint8, C.c_uint16, C.c_uint32, C.c_uint64, C.c_int16, C.c_int32
vp = C.c_void_p

ABI_VERSION = 1
RUNTIME_ABI_VERSION = 2

OK = 0
INVALID_ARGUMENT = -1
UNSUPPORTED_PROFILE = -2
LAYOUT_MISMATCH = -3
RESOURCE_TOO_SMALL = -4
LAUNCH_FAILED = -5

IQ_INT16_Q8_8 = 0
IQ_COMPLEX64 = 1
DFT_FP32 = 0
DFT_BF16 = 1
DFT_FP16 = 2
LDPC_MODE_REF = 0
LDPC_MODE_FP16 = 1
OBS_COMB2 = 0
OBS_FD_OCC2 = 1
DESCR_RESTART_PER_SLOT = 1


class PuschMimoConfig(C.Structure):
    _fields_ = [
        ("abi_version", u16), ("struct_size", u16), ("flags", u32),
        ("num_layers", u16), ("num_tx_ports", u16), ("num_rx_antennas", u16), ("qm", u16),
        ("num_symbols", u16), ("fft_size", u16), ("num_rb", u16), ("rb_start", u16),
        ("slot_number", u16), ("start_symbol", u8), ("num_allocated_symbols", u8),
        ("used_subcarriers", u16), ("padded_subcarriers", u16),
        ("dmrs_symbol_mask", u16), ("dmrs_ports", u16 * 4), ("dmrs_scrambling_id", u16),
        ("data_scrambling_id", u16), ("rnti", u16), ("dmrs_type", u8), ("dmrs_length", u8),
        ("dmrs_additional_position", u8), ("mapping_type", u8),
        ("num_cdm_groups_without_data", u8), ("n_scid", u8), ("codeword_index", u8),
        ("transform_precoding", u8), ("codebook_enabled", u8), ("tpmi", u16),
        ("prg_size_rb", u16), ("reserved", u32 * 8),
    ]


class MimoDetectAllocation(C.Structure):
    _fields_ = [("pusch_index", u16), ("layer_offset", u16), ("num_layers", u16), ("num_tx_ports", u16)]


class MimoDetectLayerPlan(C.Structure):
    _fields_ = [
        ("abi_version", u16), ("struct_size", u16), ("num_rx_antennas", u16),
        ("layer_capacity", u16), ("total_layers", u16), ("num_allocations", u16),
        ("allocations", MimoDetectAllocation * 16), ("reserved", u32 * 8),
    ]


class PuschMimoLayout(C.Structure):
    _fields_ = [
        ("num_dmrs_symbols", u16), ("num_data_symbols", u16), ("num_data_re", u32),
        ("data_stride", u32), ("codeword_symbols", u32), ("codeword_stride", u32),
    ]


class PuschMimoRuntimeConfig(C.Structure):
    _fields_ = [
        ("abi_version", u16), ("struct_size", u16), ("flags", u32), ("profile_sha256", u8 * 32),
        ("architecture", u16), ("num_tx_antennas", u16), ("num_tx_rf_chains", u16),
        ("num_rx_antennas", u16), ("num_rx_rf_chains", u16), ("num_layers", u16),
        ("num_tx_ports", u16), ("rx_bucket", u16), ("layer_bucket", u16),
        ("max_rx_antennas", u16), ("max_layers", u16),
        ("qm", u16), ("num_symbols", u16), ("fft_size", u16), ("num_rb", u16), ("rb_start", u16),
        ("used_subcarriers", u16), ("padded_subcarriers", u16),
        ("num_dmrs_symbols", u16), ("dmrs_symbol_mask", u16), ("dmrs_port_count", u16),
        ("dmrs_ports", u16 * 16),
        ("dmrs_type", u16), ("dmrs_length", u16), ("num_cdm_groups_without_data", u16),
        ("n_scid", u16), ("precoding_mode", u16), ("tpmi", u16), ("prg_size_rb", u16),
        ("channel_gain", C.c_float), ("awgn_std_int16", C.c_float),
        ("data_re_per_layer", u32), ("data_stride_per_layer", u32), ("qam_row_stride", u32),
        ("grid_re_per_port", u32),
        ("slot_number", u16), ("dmrs_scrambling_id", u16), ("data_scrambling_id", u16),
        ("rnti", u16), ("start_symbol", u16), ("num_allocated_symbols", u16),
        ("dmrs_additional_position", u16), ("mapping_type", u16), ("codeword_index", u16),
        ("reserved16", u16), ("reserved", u32 * 5),
    ]


def _args(*fields):
    return [("abi_version", u16), ("struct_size", u16)] + list(fields)


class OfdmDemodBatchOpArgsV1(C.Structure):
    _fields_ = _args(("iq", vp), ("iq_format", u32), ("batch", u32), ("out_re", vp), ("out_im", vp),
                     ("precision", u32), ("reserved", u32), ("config", C.POINTER(PuschMimoConfig)),
                     ("stream", vp))


class ReDemapBatchOpArgsV1(C.Structure):
    _fields_ = _args(("fft_grid_re", vp), ("fft_grid_im", vp), ("rx_grid_re", vp), ("rx_grid_im", vp),
                     ("config", C.POINTER(PuschMimoConfig)), ("layout", C.POINTER(PuschMimoLayout)),
                     ("stream", vp))


class ReDemapKernelMetadata(C.Structure):
    _fields_ = [("magic", u32), ("num_rx_antennas", u32), ("num_symbols", u32), ("fft_size", u32),
                ("used_subcarriers", u32), ("padded_subcarriers", u32), ("block_dim", u32),
                ("symbols_per_tile", u32), ("reserved", u32 * 24)]


class MimoDmrsGenOpArgsV1(C.Structure):
    _fields_ = _args(("dmrs_re", vp), ("dmrs_im", vp), ("config", C.POINTER(PuschMimoConfig)),
                     ("layout", C.POINTER(PuschMimoLayout)), ("stream", vp))


class MimoDmrsGenKernelMetadata(C.Structure):
    _fields_ = [("magic", u32), ("num_layers", u32), ("num_dmrs_symbols", u32),
                ("dmrs_symbols", u32 * 4), ("port_indices", u32 * 4), ("comb_delta", u32 * 4),
                ("wf_odd_negative", u32 * 4), ("wt_negative", u32 * 8),
                ("output_layer_stride", u32), ("output_dmrs_stride", u32), ("reserved", u32 * 3)]


class MimoDmrsPortOccSemantics(C.Structure):
    _fields_ = [("port_index", u32), ("comb_delta", u32), ("wf_odd_negative", u32),
                ("wt_negative", u32 * 2)]


class MimoDmrsLsOpArgsV1(C.Structure):
    _fields_ = _args(("rx_grid_re", vp), ("rx_grid_im", vp), ("dmrs_ref_re", vp), ("dmrs_ref_im", vp),
                     ("h_ls_re", vp), ("h_ls_im", vp), ("pilot_sc", vp), ("pilot_count", vp),
                     ("noise_var_rx", vp), ("config", C.POINTER(PuschMimoConfig)),
                     ("layout", C.POINTER(PuschMimoLayout)), ("stream", vp))


class MimoDmrsLsKernelMetadata(C.Structure):
    _fields_ = [("magic", u32), ("num_rx", u32), ("num_layers", u32), ("num_dmrs_symbols", u32),
                ("num_symbols", u32), ("used_subcarriers", u32), ("padded_subcarriers", u32),
                ("dmrs_ref_stride", u32), ("pilot_stride", u32), ("block_dim", u32),
                ("dmrs_symbols", u32 * 4), ("ports", u32 * 4), ("comb_delta", u32 * 4),
                ("observation_model", u32 * 4), ("reserved", u32 * 6)]


class LmmseWeightModelV1(C.Structure):
    _fields_ = [("magic", u32), ("abi_version", u16), ("struct_size", u16), ("num_layers", u16),
                ("num_dmrs_symbols", u16), ("rank", u16), ("pilot_stride", u16),
                ("pilot_count", (u16 * 2) * 4), ("pilot_sc_hash", (u64 * 2) * 4), ("reserved", u32 * 8)]


class ChannelEstLmmseOpArgsV1(C.Structure):
    _fields_ = _args(("h_ls_re", vp), ("h_ls_im", vp), ("pilot_count", vp), ("pilot_sc", vp),
                     ("pilot_count_host", C.POINTER(u16)), ("pilot_sc_host", C.POINTER(u16)),
                     ("weight_model", C.POINTER(LmmseWeightModelV1)), ("h_grid_re", vp),
                     ("h_grid_im", vp), ("config", C.POINTER(PuschMimoConfig)),
                     ("layout", C.POINTER(PuschMimoLayout)), ("stream", vp))


class LmmseFilterParamsV1(C.Structure):
    _fields_ = _args(("max_rank", u32), ("tau_s", C.c_double), ("fd_hz", C.c_double),
                     ("scs_hz", C.c_double), ("eig_floor", C.c_double))


class MimoDetectIoPackOpArgsV1(C.Structure):
    _fields_ = _args(("rx_grid_re", vp), ("rx_grid_im", vp), ("h_grid_re", vp), ("h_grid_im", vp),
                     ("noise_var_rx", vp), ("hrm_re", vp), ("hrm_im", vp), ("yvpad_re", vp),
                     ("yvpad_im", vp), ("no", vp), ("configs", C.POINTER(PuschMimoConfig)),
                     ("num_configs", u16), ("reserved", u16),
                     ("layer_plan", C.POINTER(MimoDetectLayerPlan)), ("stream", vp))


class MimoDetectIoPackKernelMetadata(C.Structure):
    _fields_ = [("magic", u32), ("active_layers", u32), ("num_rx_antennas", u32),
                ("layer_capacity", u32), ("n_re", u32), ("abi_version", u32), ("reserved", u32 * 2)]


class MimoDetectLmmseOpArgsV1(C.Structure):
    _fields_ = _args(("hrm_re", vp), ("hrm_im", vp), ("yvpad_re", vp), ("yvpad_im", vp), ("no", vp),
                     ("xhat_re", vp), ("xhat_im", vp), ("no_eff", vp),
                     ("config", C.POINTER(PuschMimoConfig)),
                     ("layer_plan", C.POINTER(MimoDetectLayerPlan)), ("stream", vp))


class QamDemod256BatchOpArgsV1(C.Structure):
    _fields_ = _args(("x_re", vp), ("x_im", vp), ("no_eff", vp), ("layer_llr", vp),
                     ("config", C.POINTER(PuschMimoConfig)), ("layout", C.POINTER(PuschMimoLayout)),
                     ("stream", vp))


class QamDemodBatchTilingData(C.Structure):
    _fields_ = [("n_sc_used", i32), ("n_sc_pad", i32), ("n_data_symbols", i32), ("n_data_re", i32),
                ("llr_layer_stride", i32), ("block_dim", i32), ("llr_clip", i32),
                ("y_scaled_clip", i32), ("num_layers", i32), ("grid_layer_stride", i32),
                ("physical_data_symbols", i32 * 14), ("group_size", i32), ("reserved", i32 * 7)]


class LayerDemapOpArgsV1(C.Structure):
    _fields_ = _args(("layer_llr", vp), ("cw_llr", vp), ("config", C.POINTER(PuschMimoConfig)),
                     ("layout", C.POINTER(PuschMimoLayout)), ("stream", vp))


class LayerDemapKernelMetadata(C.Structure):
    _fields_ = [("magic", u32), ("num_layers", u32), ("qm", u32), ("num_data_re", u32),
                ("data_stride", u32), ("codeword_symbols", u32), ("codeword_stride", u32),
                ("input_symbol_stride", u32), ("symbols_per_group", u32), ("group_data_re", u32),
                ("gather_chunk", u32), ("gather_index_elems", u32), ("reserved", u32 * 20)]


class DescrambleMimoOpArgsV1(C.Structure):
    _fields_ = _args(("llr_qam", vp), ("llr_nr", vp), ("config", C.POINTER(PuschMimoConfig)),
                     ("layout", C.POINTER(PuschMimoLayout)), ("num_slots", u32), ("stream", vp),
                     ("flags", u32), ("reserved", u32))


class DescrambleKernelMetadata(C.Structure):
    _fields_ = [("magic", u32), ("num_slots", u32), ("num_layers", u32), ("qm", u32),
                ("num_data_re", u32), ("data_stride", u32), ("codeword_symbols", u32),
                ("codeword_stride", u32), ("tile_elems", u32), ("max_slots", u32), ("reserved", u32 * 22)]


class AiranTbPlanV1(C.Structure):
    _fields_ = [("tbs", u32), ("crc_tb", u32), ("crc_cb", u32), ("C", u32), ("K", u32),
                ("Kprime", u32), ("Z", u32), ("filler", u32), ("N", u32), ("Ncb", u32), ("k0", u32),
                ("rv", u32), ("G", u32), ("E_low", u32), ("E_high", u32), ("n_low", u32),
                ("reserved", u32 * 4)]


class RateMatchDescriptor(C.Structure):
    _fields_ = [("e", u32), ("k0", u32), ("ncb", u32), ("cw_symbol_offset", u32)]


class RateDematchMimoOpArgsV1(C.Structure):
    _fields_ = _args(("cw_llr", vp), ("ldpc_llr", vp), ("config", C.POINTER(PuschMimoConfig)),
                     ("layout", C.POINTER(PuschMimoLayout)), ("num_slots", u32), ("stream", vp))


class RateDematchKernelMetadata(C.Structure):
    _fields_ = [("magic", u32), ("num_slots", u32), ("num_layers", u32), ("qm", u32),
                ("num_data_re", u32), ("codeword_symbols", u32), ("codeword_stride", u32),
                ("total_coded_bits", u32), ("code_blocks", u32), ("ldpc_n", u32), ("n_2z", u32),
                ("ncb", u32), ("llr_scale", u32), ("llr_clip", u32), ("tile_elems", u32),
                ("max_slots", u32), ("descriptor_words", u32), ("z", u32), ("kprime", u32),
                ("filler", u32), ("reserved", u32 * 12)]


class LdpcDecodeOpArgsV1(C.Structure):
    _fields_ = _args(("lam_in", vp), ("bits", vp), ("lam_out", vp), ("iterations_out", vp),
                     ("num_cb", u32), ("z", u32), ("k", u32), ("max_iterations", u32),
                     ("active_rows", u32), ("mode", u32), ("early_stop", u32), ("reserved", u32),
                     ("stream", vp))


class ChannelEstLsSisoOpArgsV1(C.Structure):
    _fields_ = _args(("rx_grid_re", vp), ("rx_grid_im", vp), ("dmrs_ref_re", vp), ("dmrs_ref_im", vp),
                     ("h_re", vp), ("h_im", vp), ("config", C.POINTER(PuschMimoConfig)), ("batch", u32),
                     ("ref_per_slot", u32), ("stream", vp))


class EqualizeZfSisoOpArgsV1(C.Structure):
    _fields_ = _args(("y_re", vp), ("y_im", vp), ("h_re", vp), ("h_im", vp), ("n0", vp), ("x_re", vp),
                     ("x_im", vp), ("no_eff", vp), ("n", u32), ("reserved", u32), ("stream", vp))


class AiranPuschChainOptionsV1(C.Structure):
    _fields_ = _args(("max_slots", u32), ("iq_format", u32), ("ldpc_iterations", u32), ("ldpc_mode", u32),
                     ("early_stop", u32), ("ce_max_rank", u32), ("use_graph", u32), ("tau_s", C.c_double),
                     ("fd_hz", C.c_double), ("scs_hz", C.c_double), ("eig_floor", C.c_double),
                     ("receiver", u32), ("dft_precision", u32), ("gram_precision", u32),
                     ("reserved", u32 * 5))


CHAIN_RX_LMMSE = 0
CHAIN_RX_LS_ZF = 1
GRAM_FP32 = 0
GRAM_BF16 = 1
GRAM_FP16 = 2


class AiranGrantBatchOptionsV1(C.Structure):
    _fields_ = _args(("iq_format", u32), ("ldpc_iterations", u32), ("early_stop", u32), ("ce_max_rank", u32),
                     ("use_graph", u32), ("tau_s", C.c_double), ("fd_hz", C.c_double), ("scs_hz", C.c_double),
                     ("detect_tile", u32), ("reserved", u32 * 7))


class AiranGrantLayoutV1(C.Structure):
    _fields_ = [("iq_offset", C.c_uint64), ("iq_samples", C.c_uint64), ("bits_offset", C.c_uint64),
                ("cw_offset", C.c_uint64), ("first_cb", u32), ("num_cb", u32), ("K", u32), ("fft", u32)]


class AiranGrantBatchOutputsV1(C.Structure):
    _fields_ = _args(("num_grants", u32), ("bits", vp), ("tb_ok", vp), ("cb_ok", vp), ("iterations", vp),
                     ("cw_llr", vp), ("num_cb", C.c_uint64))


class AiranPuschChainOutputsV1(C.Structure):
    _fields_ = _args(("num_slots", u32), ("bits", vp), ("tb_ok", vp), ("cb_ok", vp), ("iterations", vp),
                     ("rx_grid_re", vp), ("rx_grid_im", vp), ("h_ls_re", vp), ("h_ls_im", vp),
                     ("noise_ls", vp), ("noise_ce", vp), ("h_re", vp), ("h_im", vp), ("cw_llr", vp),
                     ("ldpc_llr", vp))


P = C.POINTER

# name -> (restype, argtypes).  Every entry is an exported symbol of the
# library; tests/test_abi.py checks the list against include/*.h.
class AiranCeRefinerBatchV1(C.Structure):
    _fields_ = [("abi_version", C.c_uint16), ("struct_size", C.c_uint16), ("x", C.c_void_p),
                ("target", C.c_void_p), ("out", C.c_void_p), ("h_ls_re", C.c_void_p), ("h_ls_im", C.c_void_p),
                ("noise_ls", C.c_void_p)]


PROTOTYPES = {
    "airan_InitDevice": (i32, []),
    "airan_ValidateMimoRuntimeConfig": (i32, [P(PuschMimoRuntimeConfig), C.c_char_p, C.c_size_t]),
    "airan_LoadMimoRuntimeConfig": (i32, [C.c_char_p, P(PuschMimoRuntimeConfig), C.c_char_p, C.c_size_t]),
    "airan_LoadMimoRuntimeConfigFromEnv": (i32, [u16, P(PuschMimoRuntimeConfig), P(i32), C.c_char_p, C.c_size_t]),
    "airan_DerivePuschMimoConfig": (i32, [P(PuschMimoRuntimeConfig), u16, P(PuschMimoConfig), C.c_char_p,
                                          C.c_size_t, u16]),
    "airan_DerivePuschMimoLayout": (i32, [P(PuschMimoConfig), P(PuschMimoLayout)]),
    "airan_ofdm_demod_batch_ValidateOpArgs": (i32, [P(OfdmDemodBatchOpArgsV1)]),
    "airan_ofdm_demod_batch_Enqueue": (i32, [P(OfdmDemodBatchOpArgsV1)]),
    "airan_ofdm_demod_redemap_Enqueue": (i32, [P(OfdmDemodBatchOpArgsV1), vp, vp]),
    "airan_re_demap_batch_BuildCurrentProfile": (i32, [P(PuschMimoConfig), P(PuschMimoLayout),
                                                       P(ReDemapKernelMetadata)]),
    "airan_re_demap_batch_ValidateOpArgs": (i32, [P(ReDemapBatchOpArgsV1)]),
    "airan_re_demap_batch_BuildGatherIndex": (i32, [P(PuschMimoConfig), P(u32), C.c_size_t]),
    "airan_re_demap_batch_Enqueue": (i32, [P(ReDemapBatchOpArgsV1), vp, C.c_size_t, vp, C.c_size_t, vp,
                                           C.c_size_t]),
    "airan_mimo_dmrs_gen_GetCurrentPortOccSemantics": (i32, [u16, P(MimoDmrsPortOccSemantics)]),
    "airan_mimo_dmrs_gen_BuildCurrentProfile": (i32, [P(PuschMimoConfig), P(PuschMimoLayout),
                                                      P(MimoDmrsGenKernelMetadata), P(i32)]),
    "airan_mimo_dmrs_gen_ValidateOpArgs": (i32, [P(MimoDmrsGenOpArgsV1)]),
    "airan_mimo_dmrs_gen_CreateRuntime": (i32, [P(vp)]),
    "airan_mimo_dmrs_gen_Enqueue": (i32, [vp, P(MimoDmrsGenOpArgsV1)]),
    "airan_mimo_dmrs_gen_DestroyRuntime": (None, [vp]),
    "airan_mimo_dmrs_ls_BuildCurrentProfile": (i32, [P(PuschMimoConfig), P(PuschMimoLayout),
                                                     P(MimoDmrsLsKernelMetadata), P(u16), P(u16)]),
    "airan_mimo_dmrs_ls_ValidateOpArgs": (i32, [P(MimoDmrsLsOpArgsV1)]),
    "airan_mimo_dmrs_ls_Enqueue": (i32, [P(MimoDmrsLsOpArgsV1)]),
    "airan_channel_est_lmmse_PilotScHash": (u64, [P(u16), u32]),
    "airan_channel_est_lmmse_BuildWeightModel": (i32, [P(PuschMimoConfig), P(PuschMimoLayout), P(u16), P(u16),
                                                       P(LmmseWeightModelV1)]),
    "airan_channel_est_lmmse_ValidateObservationModel": (i32, [P(PuschMimoConfig), P(PuschMimoLayout), P(u16),
                                                               P(u16), P(LmmseWeightModelV1)]),
    "airan_channel_est_lmmse_ValidateOpArgs": (i32, [P(ChannelEstLmmseOpArgsV1)]),
    "airan_channel_est_lmmse_CreateFilter": (i32, [P(PuschMimoConfig), P(LmmseFilterParamsV1), vp, P(vp)]),
    "airan_channel_est_lmmse_DestroyFilter": (None, [vp]),
    "airan_channel_est_lmmse_FilterRank": (u32, [vp, u32]),
    "airan_channel_est_lmmse_WorkspaceBytes": (C.c_size_t, [P(PuschMimoConfig)]),
    "airan_channel_est_lmmse_Enqueue": (i32, [P(ChannelEstLmmseOpArgsV1), vp, vp, vp, vp, C.c_size_t]),
    "airan_mimo_detect_io_pack_BuildCurrentProfile": (i32, [P(PuschMimoConfig), C.c_size_t,
                                                            P(MimoDetectLayerPlan),
                                                            P(MimoDetectIoPackKernelMetadata)]),
    "airan_mimo_detect_io_pack_ValidateOpArgs": (i32, [P(MimoDetectIoPackOpArgsV1)]),
    "airan_mimo_detect_io_pack_Enqueue": (i32, [P(MimoDetectIoPackOpArgsV1), vp, C.c_size_t]),
    "airan_mimo_detect_lmmse_ValidateOpArgs": (i32, [P(MimoDetectLmmseOpArgsV1)]),
    "airan_mimo_detect_lmmse_Enqueue": (i32, [P(MimoDetectLmmseOpArgsV1)]),
    "airan_qam256_demod_batch_BuildCurrentProfile": (i32, [P(PuschMimoConfig), P(PuschMimoLayout),
                                                           P(QamDemodBatchTilingData)]),
    "airan_qam256_demod_batch_ValidateOpArgs": (i32, [P(QamDemod256BatchOpArgsV1)]),
    "airan_qam256_demod_batch_Enqueue": (i32, [P(QamDemod256BatchOpArgsV1), vp, C.c_size_t, vp, C.c_size_t]),
    "airan_layer_demap_BuildCurrentProfile": (i32, [P(PuschMimoConfig), P(PuschMimoLayout),
                                                    P(LayerDemapKernelMetadata)]),
    "airan_layer_demap_ValidateOpArgs": (i32, [P(LayerDemapOpArgsV1)]),
    "airan_layer_demap_Enqueue": (i32, [P(LayerDemapOpArgsV1)]),
    "airan_descramble_mimo_BuildCurrentProfile": (i32, [P(PuschMimoConfig), u32, P(PuschMimoLayout),
                                                        P(DescrambleKernelMetadata)]),
    "airan_descramble_mimo_ValidateOpArgs": (i32, [P(DescrambleMimoOpArgsV1)]),
    "airan_descramble_mimo_BuildGoldSign": (i32, [P(PuschMimoConfig), P(PuschMimoLayout), u32, P(i16),
                                                  C.c_size_t]),
    "airan_descramble_mimo_Enqueue": (i32, [P(DescrambleMimoOpArgsV1), vp, C.c_size_t, vp, C.c_size_t, vp,
                                            C.c_size_t]),
    "airan_pusch_Tbs": (u32, [u32, u32, u32, u32, C.c_double]),
    "airan_pusch_BuildTbPlan": (i32, [u32, u32, u32, u32, u32, P(AiranTbPlanV1)]),
    "airan_pusch_ReferenceTbPlan": (i32, [P(PuschMimoConfig), u32, P(AiranTbPlanV1)]),
    "airan_rate_dematch_mimo_BuildCurrentProfile": (i32, [P(PuschMimoConfig), u32, P(PuschMimoLayout),
                                                          P(RateDematchKernelMetadata), P(RateMatchDescriptor),
                                                          C.c_size_t]),
    "airan_rate_dematch_mimo_BuildProfileForTb": (i32, [P(PuschMimoConfig), u32, P(AiranTbPlanV1),
                                                        P(PuschMimoLayout), P(RateDematchKernelMetadata),
                                                        P(RateMatchDescriptor), C.c_size_t]),
    "airan_rate_dematch_mimo_ValidateOpArgs": (i32, [P(RateDematchMimoOpArgsV1)]),
    "airan_rate_dematch_mimo_Enqueue": (i32, [P(RateDematchMimoOpArgsV1), P(RateDematchKernelMetadata), vp,
                                              C.c_size_t, vp, C.c_size_t, vp, C.c_size_t]),
    "airan_channel_est_ls_siso_ValidateOpArgs": (i32, [P(ChannelEstLsSisoOpArgsV1)]),
    "airan_channel_est_ls_siso_Enqueue": (i32, [P(ChannelEstLsSisoOpArgsV1)]),
    "airan_equalize_zf_siso_ValidateOpArgs": (i32, [P(EqualizeZfSisoOpArgsV1)]),
    "airan_equalize_zf_siso_Enqueue": (i32, [P(EqualizeZfSisoOpArgsV1)]),
    "airan_ldpc_decode_ValidateOpArgs": (i32, [P(LdpcDecodeOpArgsV1)]),
    "airan_ldpc_decode_WorkspaceBytes": (C.c_size_t, [u32, u32, u32]),
    "airan_ldpc_decode_Enqueue": (i32, [P(LdpcDecodeOpArgsV1), vp, C.c_size_t]),
    "airan_ldpc_decode_EnqueueNoAlloc": (i32, [P(LdpcDecodeOpArgsV1)]),
    "airan_ldpc_decode_NoAllocMaxWorkspaceBytes": (C.c_size_t, []),
    "airan_ldpc_ActiveRows": (u32, [P(AiranTbPlanV1)]),
    "airan_crc_check_WorkspaceBytes": (C.c_size_t, [u32]),
    "airan_crc_check_Enqueue": (i32, [P(AiranTbPlanV1), u32, vp, vp, vp, vp, vp, C.c_size_t, vp]),
    "airan_pusch_NoEffFloor": (C.c_float, [u32]),
    # host references (bit-exact, no GPU)
    "airan_re_demap_batch_Reference": (i32, [P(PuschMimoConfig), vp, vp, u32, vp, vp, vp]),
    "airan_layer_demap_ReferenceDemap": (i32, [vp, P(PuschMimoConfig), P(PuschMimoLayout), vp]),
    "airan_descramble_mimo_ReferenceDescramble": (i32, [vp, vp, P(PuschMimoConfig), P(PuschMimoLayout), u32, vp]),
    "airan_rate_dematch_mimo_ReferenceRateDematch": (i32, [vp, P(PuschMimoConfig), P(PuschMimoLayout), u32,
                                                           P(RateMatchDescriptor), C.c_size_t, vp]),
    "airan_rate_dematch_mimo_ReferenceRateDematchForTb": (i32, [vp, P(PuschMimoConfig), P(PuschMimoLayout), u32,
                                                                P(AiranTbPlanV1), P(RateMatchDescriptor),
                                                                C.c_size_t, vp]),
    "airan_mimo_dmrs_ls_DescribeObservationModels": (i32, [vp, vp, u32, u32, vp]),
    "airan_mimo_dmrs_ls_ValidateNaturalLmmseContract": (i32, [vp, vp, vp, u32, u32]),
    "airan_mimo_dmrs_ls_ValidateLmmse798Compatibility": (i32, [vp, vp, u32, u32]),
    # synthetic transmitter (include/airan_pusch_tx.h)
    "airan_pusch_tx_EncodeTbs": (i32, [P(PuschMimoConfig), P(AiranTbPlanV1), u32, u32, vp, vp, vp]),
    "airan_pusch_tx_EncodeTbsSlots": (i32, [P(PuschMimoConfig), P(AiranTbPlanV1), u32, u32, vp, vp, vp, vp]),
    "airan_pusch_txgpu_Create": (i32, [P(PuschMimoConfig), P(AiranTbPlanV1), u32, u32, P(vp)]),
    "airan_pusch_txgpu_Destroy": (None, [vp]),
    "airan_pusch_txgpu_EncodeTbs": (i32, [vp, vp, u32, vp, vp, vp, vp]),
    "airan_pusch_txgpu_Channel": (i32, [vp, vp, u32, vp, vp, u32, C.c_float, C.c_float, C.c_uint64, u32, vp, vp]),
    "airan_ldpc_bg1_Encode": (i32, [u32, vp, vp]),
    "airan_crc_Compute": (u32, [u32, vp, C.c_uint64]),
    # batched chain (include/airan_pusch_chain.h)
    "airan_pusch_chain_Create": (i32, [P(PuschMimoConfig), P(AiranTbPlanV1), P(AiranPuschChainOptionsV1), vp,
                                       P(vp)]),
    "airan_pusch_chain_Destroy": (None, [vp]),
    "airan_pusch_chain_Enqueue": (i32, [vp, vp, u32, vp]),
    "airan_pusch_chain_EnqueueSlots": (i32, [vp, vp, u32, vp, vp]),
    "airan_pusch_chain_SlotsPerTb": (u32, [vp]),
    "airan_pusch_chain_Outputs": (i32, [vp, P(AiranPuschChainOutputsV1)]),
    "airan_pusch_chain_DeviceBytes": (C.c_size_t, [vp]),
    "airan_pusch_chain_SetTiming": (i32, [vp, u32]),
    "airan_pusch_chain_StageTimes": (i32, [vp, P(C.c_float), u32]),
    "airan_pusch_chain_SetRefiner": (i32, [vp, vp, C.c_float, C.c_float]),
    # ragged multi-grant batches (include/airan_pusch_grants.h)
    "airan_pusch_grants_Create": (i32, [P(PuschMimoConfig), P(AiranTbPlanV1), u32, P(AiranGrantBatchOptionsV1), vp,
                                        P(vp)]),
    "airan_pusch_grants_Destroy": (None, [vp]),
    "airan_pusch_grants_Layout": (i32, [vp, P(AiranGrantLayoutV1), u32, P(C.c_uint64)]),
    "airan_pusch_grants_Enqueue": (i32, [vp, vp, vp]),
    "airan_pusch_grants_Outputs": (i32, [vp, P(AiranGrantBatchOutputsV1)]),
    "airan_pusch_grants_SetDetectTile": (i32, [vp, u32]),
    "airan_pusch_grants_NumShapes": (u32, [vp]),
    "airan_pusch_grants_FilterCacheSize": (u32, []),
    "airan_pusch_grants_ClearFilterCache": (None, []),
    # learned CE refiner (include/airan_ce_refiner.h)
    "airan_ce_refiner_Create": (i32, [P(PuschMimoConfig), u32, vp, C.c_uint64, vp, P(vp)]),
    "airan_ce_refiner_Destroy": (None, [vp]),
    "airan_ce_refiner_GetParams": (i32, [vp, vp]),
    "airan_ce_refiner_SetParams": (i32, [vp, vp]),
    "airan_ce_refiner_Enqueue": (i32, [vp, vp, vp, u32, vp, vp, vp, vp]),
    "airan_ce_refiner_SetInput": (i32, [vp, u32, vp]),
    "airan_ce_refiner_SetDilation": (i32, [vp, vp, u32]),
    "airan_ce_refiner_EnqueueWithNoise": (i32, [vp, vp, vp, vp, u32, vp, vp, vp, vp]),
    "airan_ce_refiner_SetResidualScale": (i32, [vp, u32]),
    "airan_ce_refiner_SetSlotScale": (i32, [vp, vp, u32]),
    "airan_ce_refiner_GetDilation": (i32, [vp, vp, u32]),
    "airan_ce_refiner_SynthBatch": (i32, [vp, vp, vp, u32, vp, u32, C.c_float, C.c_uint64, vp]),
    "airan_ce_refiner_Batch": (i32, [vp, P(AiranCeRefinerBatchV1)]),
    "airan_ce_refiner_TrainStep": (i32, [vp, u32, C.c_float, vp, vp]),
    "airan_ce_refiner_Forward": (i32, [vp, u32, vp]),
    "airan_ce_refiner_Gradients": (i32, [vp, vp, vp, u32, u32, vp, vp, vp]),
}

REFINER_PARAMS = 184672

CHAIN_STAGES = ("ofdm_demod", "dmrs_ls", "ce_lmmse", "detect_demap", "rate_dematch", "ldpc_decode", "crc_check")

_LIB = None


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    """Load libpuschrx_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build() "
                           "(the HIP path has no fallback)")
    lib = C.CDLL(str(p))
    for name, (res, argt) in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = argt
    if path is None:
        _LIB = lib
    return lib
