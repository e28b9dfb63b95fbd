/*
 * airan_pusch_rx.h -- C ABI of libpuschrx_hip.so, the gfx950 PUSCH MIMO
 * receive chain.
 *
 * One entry point per reference operator interface, with the reference's
 * argument structs, Status values and ownership rules: the caller owns every
 * device buffer and passes a hipStream_t as `void *stream`; no entry point
 * synchronises the stream; argument structs carry abi_version/struct_size and
 * are validated fail-closed before anything is enqueued; input/output
 * aliasing is rejected.  Geometry is taken from PuschMimoConfig instead of
 * being fixed at compile time, so the reference profile (64 Rx, 133 PRB,
 * FFT 2048, 256QAM, 23-slot TB) and the 4 Rx / 273 PRB / FFT 4096 profiles
 * use the same entry points.
 *
 * Derived geometry (identical in every operator, see
 * airan_DerivePuschMimoLayout):
 *   D        = popcount(dmrs_symbol_mask), data symbols = the other symbols
 *   num_data_re = (14 - D) * used, data_stride = align64(num_data_re)
 *   codeword_symbols = L * num_data_re, codeword_stride = L * data_stride
 *   qam_row_stride = align64(used), qam_plane = (14 - D) * qam_row_stride
 *   dmrs_ref_pad = align128(used / 2), pilot_pad = align64(used / 2)
 *   OFDM stage-4 factorisation fft = P * 64 (P = fft / 64)
 */
#ifndef AIRAN_PUSCH_RX_H
#define AIRAN_PUSCH_RX_H

#include <stddef.h>
#include <stdint.h>

#include "airan_pusch_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ============================================================ device init */
/* Loads the per-device constant tables (Gold jump-ahead matrices, FFT
 * twiddles, LDPC base graph) and takes the device address of the decoder's
 * no-alloc message array for the current device.  Called implicitly by the
 * first Enqueue on a device; call it explicitly before capturing a stream
 * into a HIP graph. */
int32_t airan_InitDevice(void);

/* ======================================================= runtime config */
/* common/pusch_mimo_runtime_config.h:19-37.  `why` receives the reason text
 * (may be NULL).  Return values are the AIRAN_RTCFG_* codes. */
int32_t airan_ValidateMimoRuntimeConfig(const PuschMimoRuntimeConfig *config, char *why,
                                        size_t why_len);
int32_t airan_LoadMimoRuntimeConfig(const char *path, PuschMimoRuntimeConfig *config,
                                    char *why, size_t why_len);
/* Reads the path from PUSCH_MIMO_RUNTIME_CONFIG; *enabled = 0 when unset. */
int32_t airan_LoadMimoRuntimeConfigFromEnv(uint16_t expected_rank,
                                           PuschMimoRuntimeConfig *config,
                                           int32_t *enabled, char *why, size_t why_len);
int32_t airan_DerivePuschMimoConfig(const PuschMimoRuntimeConfig *runtime,
                                    uint16_t operator_rx_capacity, PuschMimoConfig *config,
                                    char *why, size_t why_len, uint16_t slot_number);
/* Layout every operator derives from a config (no reference counterpart as a
 * single function: each reference operator fills it inside
 * BuildCurrentProfile, e.g. re_demap_batch.cpp:62-67). */
int32_t airan_DerivePuschMimoLayout(const PuschMimoConfig *config, PuschMimoLayout *layout);

/* ============================================================ OFDM demod */
/* Replaces the raw launch ofdm_demod_batch_kernel(iq, w32r, w32i, w64r, w64i,
 * twr, twi, out_re, out_im, batch, ws, tiling)
 * (chain/pusch_mimo_rx_chain/rx_chain.cpp:471-475; kernel
 * ofdm/ofdm_demod_batch/ofdm_demod_kernel.cpp:889).  The DFT factors the
 * reference loads from files are generated inside the library.
 * Input: `batch` antenna streams of one slot each (15 * fft samples).
 * Output: stage-4 layout fp16 planes [batch,14,P,64], bin k at
 * [(k % P) * 64 + k / P], unitary DFT, no fftshift. */
#define AIRAN_IQ_INT16_Q8_8 0   /* int16 [batch, samples, 2], value / 256 */
#define AIRAN_IQ_COMPLEX64 1    /* float [batch, samples, 2] */
#define AIRAN_DFT_FP32 0        /* fp32 radix-16 FFT */
/* The reference's form: two dense DFT products (W_P x then x W_64, N = 64 P)
 * on the matrix cores with FP32 accumulation, samples block-scaled per OFDM
 * symbol by a power of two and rounded to the 16-bit type, the twiddled
 * intermediate rounded again; FFT 1024, 2048 and 4096 only. */
#define AIRAN_DFT_BF16 1
#define AIRAN_DFT_FP16 2
typedef struct OfdmDemodBatchOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *iq;
    uint32_t iq_format;
    uint32_t batch;             /* number of antenna streams */
    void *out_re;
    void *out_im;
    uint32_t precision;         /* AIRAN_DFT_* */
    uint32_t reserved;
    const PuschMimoConfig *config;
    void *stream;
} OfdmDemodBatchOpArgsV1;
int32_t airan_ofdm_demod_batch_ValidateOpArgs(const OfdmDemodBatchOpArgsV1 *args);
int32_t airan_ofdm_demod_batch_Enqueue(const OfdmDemodBatchOpArgsV1 *args);
/* Fused OFDM demod + RE demap: writes the natural used-subcarrier grid
 * [batch,14,padded] fp16 planes directly (zero tail), i.e. the output of
 * re_demap_batch applied to the stage-4 output above. */
int32_t airan_ofdm_demod_redemap_Enqueue(const OfdmDemodBatchOpArgsV1 *args, void *rx_grid_re,
                                         void *rx_grid_im);

/* ============================================================== RE demap */
/* mapping/re_demap_batch/re_demap_batch.h:30-94 */
typedef struct ReDemapBatchOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *fft_grid_re;    /* fp16 [NR,14,fft] stage-4 order */
    const void *fft_grid_im;
    void *rx_grid_re;           /* fp16 [NR,14,padded] natural order */
    void *rx_grid_im;
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    void *stream;
} ReDemapBatchOpArgsV1;
typedef struct ReDemapKernelMetadata {
    uint32_t magic;             /* "RDB1" */
    uint32_t num_rx_antennas;
    uint32_t num_symbols;
    uint32_t fft_size;
    uint32_t used_subcarriers;
    uint32_t padded_subcarriers;
    uint32_t block_dim;
    uint32_t symbols_per_tile;
    uint32_t reserved[24];
} ReDemapKernelMetadata;
int32_t airan_re_demap_batch_BuildCurrentProfile(const PuschMimoConfig *config,
                                                 PuschMimoLayout *layout,
                                                 ReDemapKernelMetadata *metadata);
int32_t airan_re_demap_batch_ValidateOpArgs(const ReDemapBatchOpArgsV1 *args);
/* re_demap_batch.cpp:105-123; takes the config because the geometry is not
 * fixed.  index_elems >= padded_subcarriers; entries are fp16 byte offsets. */
int32_t airan_re_demap_batch_BuildGatherIndex(const PuschMimoConfig *config, uint32_t *index,
                                              size_t index_elems);
/* mapping/re_demap_batch/re_demap_batch.h:69-76 (Reference): bit-exact host
 * gather of the used subcarriers through the byte-offset table, zero tail.
 * fft grids [NR][14][fft], rx grids [NR][14][padded] (fp16 bits); the
 * config argument supplies the geometry the reference fixes at compile time. */
int32_t airan_re_demap_batch_Reference(const PuschMimoConfig *config, const uint16_t *fft_grid_re,
                                       const uint16_t *fft_grid_im, uint32_t num_rx_antennas,
                                       const uint32_t *gather_index, uint16_t *rx_grid_re,
                                       uint16_t *rx_grid_im);
int32_t airan_re_demap_batch_Enqueue(const ReDemapBatchOpArgsV1 *args, const void *gather_index,
                                     size_t gather_index_bytes, void *workspace,
                                     size_t workspace_bytes, const void *tiling,
                                     size_t tiling_bytes);

/* ============================================================== DMRS gen */
/* mimo/mimo_dmrs_gen/mimo_dmrs_gen.h:50-121.  Output [L,D,dmrs_ref_pad] fp16
 * planes; the Gold sequence is generated on the device from c_init. */
typedef struct MimoDmrsGenOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    void *dmrs_re;
    void *dmrs_im;
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    void *stream;
} MimoDmrsGenOpArgsV1;
typedef struct MimoDmrsGenKernelMetadata {
    uint32_t magic;             /* "MDG1" */
    uint32_t num_layers;
    uint32_t num_dmrs_symbols;
    uint32_t dmrs_symbols[4];
    uint32_t port_indices[4];
    uint32_t comb_delta[4];
    uint32_t wf_odd_negative[4];
    uint32_t wt_negative[8];
    uint32_t output_layer_stride;
    uint32_t output_dmrs_stride;
    uint32_t reserved[3];
} MimoDmrsGenKernelMetadata;
typedef struct MimoDmrsPortOccSemantics {
    uint32_t port_index;
    uint32_t comb_delta;
    uint32_t wf_odd_negative;
    uint32_t wt_negative[2];
} MimoDmrsPortOccSemantics;
typedef struct MimoDmrsGenRuntimeV1 MimoDmrsGenRuntimeV1;
int32_t airan_mimo_dmrs_gen_GetCurrentPortOccSemantics(uint16_t port,
                                                       MimoDmrsPortOccSemantics *semantics);
int32_t airan_mimo_dmrs_gen_BuildCurrentProfile(const PuschMimoConfig *config,
                                                PuschMimoLayout *layout,
                                                MimoDmrsGenKernelMetadata *metadata,
                                                int32_t *cinit /* [16] */);
int32_t airan_mimo_dmrs_gen_ValidateOpArgs(const MimoDmrsGenOpArgsV1 *args);
int32_t airan_mimo_dmrs_gen_CreateRuntime(MimoDmrsGenRuntimeV1 **runtime);
int32_t airan_mimo_dmrs_gen_Enqueue(MimoDmrsGenRuntimeV1 *runtime, const MimoDmrsGenOpArgsV1 *args);
void airan_mimo_dmrs_gen_DestroyRuntime(MimoDmrsGenRuntimeV1 *runtime);

/* ================================================================ DMRS LS */
/* channel_est/mimo_dmrs_ls/mimo_dmrs_ls.h:70-160.  The reference launches
 * mimo_dmrs_ls_kernel raw (rx_chain.cpp:503-506); here Enqueue takes the
 * same buffers through the struct. */
typedef struct MimoDmrsLsOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *rx_grid_re;     /* fp16 [NR,14,padded] */
    const void *rx_grid_im;
    const void *dmrs_ref_re;    /* fp16 [L,D,dmrs_ref_pad] */
    const void *dmrs_ref_im;
    void *h_ls_re;              /* fp16 [NR,L,D,pilot_pad] */
    void *h_ls_im;
    void *pilot_sc;             /* uint16 [L,D,pilot_pad] */
    void *pilot_count;          /* uint16 [16] */
    void *noise_var_rx;         /* fp16 [NR] */
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    void *stream;
} MimoDmrsLsOpArgsV1;
typedef struct MimoDmrsLsKernelMetadata {
    uint32_t magic;             /* "MLS1" */
    uint32_t num_rx;
    uint32_t num_layers;
    uint32_t num_dmrs_symbols;
    uint32_t num_symbols;
    uint32_t used_subcarriers;
    uint32_t padded_subcarriers;
    uint32_t dmrs_ref_stride;
    uint32_t pilot_stride;
    uint32_t block_dim;
    uint32_t dmrs_symbols[4];
    uint32_t ports[4];
    uint32_t comb_delta[4];
    uint32_t observation_model[4];  /* 0 = comb (used/2), 1 = FD-OCC2 (used/4) */
    uint32_t reserved[6];
} MimoDmrsLsKernelMetadata;
int32_t airan_mimo_dmrs_ls_BuildCurrentProfile(const PuschMimoConfig *config,
                                               PuschMimoLayout *layout,
                                               MimoDmrsLsKernelMetadata *metadata,
                                               uint16_t *pilot_count /* [16] */,
                                               uint16_t *pilot_sc /* [L,D,pilot_pad] */);
int32_t airan_mimo_dmrs_ls_ValidateOpArgs(const MimoDmrsLsOpArgsV1 *args);
int32_t airan_mimo_dmrs_ls_Enqueue(const MimoDmrsLsOpArgsV1 *args);
/* channel_est/mimo_dmrs_ls/mimo_dmrs_ls.h:118-142: the per-layer observation
 * contract of the DMRS-LS outputs.  pilot_count [L][D], pilot_sc
 * [L][D][pilot_pad] as BuildCurrentProfile writes them.  The reference's
 * pilot counts are those of its 133-PRB profile (798 comb / 399 FD-OCC2
 * observations, stride 832); here they follow `config` (used / 2 and
 * used / 4, stride pilot_pad), or the reference's values for config NULL.
 * Models: AIRAN_OBS_COMB2 / AIRAN_OBS_FD_OCC2 (the reference's COMB2_798 /
 * FD_OCC2_399, same values). */
#define AIRAN_OBS_COMB2 0u
#define AIRAN_OBS_FD_OCC2 1u
int32_t airan_mimo_dmrs_ls_DescribeObservationModels(const PuschMimoConfig *config, const uint16_t *pilot_count,
                                                     uint32_t num_layers, uint32_t num_dmrs_symbols,
                                                     uint32_t *models);
int32_t airan_mimo_dmrs_ls_ValidateNaturalLmmseContract(const PuschMimoConfig *config, const uint16_t *pilot_count,
                                                        const uint16_t *pilot_sc, uint32_t num_layers,
                                                        uint32_t num_dmrs_symbols);
/* AIRAN_LAYOUT_MISMATCH (-3, the reference's LMMSE_OBSERVATION_MODEL_MISMATCH)
 * unless every layer is on the comb model. */
int32_t airan_mimo_dmrs_ls_ValidateLmmse798Compatibility(const PuschMimoConfig *config, const uint16_t *pilot_count,
                                                         uint32_t num_layers, uint32_t num_dmrs_symbols);

/* ============================================================= CE LMMSE */
/* channel_est/channel_est_lmmse_mimo/channel_est_lmmse.h:127-171.  The
 * reference applies fixed offline rank-96 factors; here an LmmseFilter is
 * built per grant geometry and prior (FP64 eigendecomposition on the
 * device) and applied with the per-slot noise estimate. */
typedef struct LmmseWeightModelV1 {
    uint32_t magic;             /* "CEW1" */
    uint16_t abi_version;
    uint16_t struct_size;
    uint16_t num_layers;
    uint16_t num_dmrs_symbols;
    uint16_t rank;
    uint16_t pilot_stride;
    uint16_t pilot_count[4][2];
    uint64_t pilot_sc_hash[4][2];
    uint32_t reserved[8];
} LmmseWeightModelV1;
typedef struct ChannelEstLmmseOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *h_ls_re;        /* fp16 [NR,L,D,pilot_pad] */
    const void *h_ls_im;
    const void *pilot_count;
    const void *pilot_sc;
    const uint16_t *pilot_count_host;
    const uint16_t *pilot_sc_host;
    const LmmseWeightModelV1 *weight_model;
    void *h_grid_re;            /* fp16 [NR,16,14,padded]; only layers < L written */
    void *h_grid_im;
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    void *stream;
} ChannelEstLmmseOpArgsV1;
typedef struct LmmseFilterParamsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    uint32_t max_rank;          /* eigenpairs kept (rounded up to 16), <= 512 */
    double tau_s;               /* RMS delay spread of the exponential prior */
    double fd_hz;               /* Jakes Doppler of the time prior */
    double scs_hz;              /* subcarrier spacing */
    double eig_floor;           /* drop eigenvalues below eig_floor * max */
} LmmseFilterParamsV1;
typedef struct LmmseFilter LmmseFilter;
uint64_t airan_channel_est_lmmse_PilotScHash(const uint16_t *pilot_sc, uint32_t count);
int32_t airan_channel_est_lmmse_BuildWeightModel(const PuschMimoConfig *config,
                                                 const PuschMimoLayout *layout,
                                                 const uint16_t *pilot_count,
                                                 const uint16_t *pilot_sc,
                                                 LmmseWeightModelV1 *model);
int32_t airan_channel_est_lmmse_ValidateObservationModel(const PuschMimoConfig *config,
                                                         const PuschMimoLayout *layout,
                                                         const uint16_t *pilot_count,
                                                         const uint16_t *pilot_sc,
                                                         const LmmseWeightModelV1 *model);
int32_t airan_channel_est_lmmse_ValidateOpArgs(const ChannelEstLmmseOpArgsV1 *args);
/* Builds the filter on the device; synchronises `stream` before returning. */
int32_t airan_channel_est_lmmse_CreateFilter(const PuschMimoConfig *config,
                                             const LmmseFilterParamsV1 *params, void *stream,
                                             LmmseFilter **filter);
void airan_channel_est_lmmse_DestroyFilter(LmmseFilter *filter);
/* Eigenpairs kept for `layer` (top of R_pp, at most params->max_rank). */
uint32_t airan_channel_est_lmmse_FilterRank(const LmmseFilter *filter, uint32_t layer);
/* The noise variance the filter uses is estimated from the energy of h_ls
 * outside the kept eigenspace (n - kept dimensions per layer and DMRS
 * symbol); noise_var_rx (fp16 [NR] from mimo_dmrs_ls, whose adjacent-pilot
 * differences also pick up the channel's frequency selectivity) is the
 * fallback when the filter keeps every eigenvector.  noise_out (fp16 [NR],
 * may be NULL) receives the per-Rx estimate used, for the detector.
 * workspace >= WorkspaceBytes. */
size_t airan_channel_est_lmmse_WorkspaceBytes(const PuschMimoConfig *config);
int32_t airan_channel_est_lmmse_Enqueue(const ChannelEstLmmseOpArgsV1 *args,
                                        const LmmseFilter *filter, const void *noise_var_rx,
                                        void *noise_out, void *workspace, size_t workspace_bytes);

/* ======================================================= detector IO pack */
/* mimo/mimo_detect_io_pack/mimo_detect_io_pack.h:69-121 */
typedef struct MimoDetectIoPackOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *rx_grid_re;     /* fp16 [NR,14,padded] */
    const void *rx_grid_im;
    const void *h_grid_re;      /* fp16 [NR,16,14,padded] */
    const void *h_grid_im;
    const void *noise_var_rx;   /* fp16 [NR] */
    void *hrm_re;               /* fp16 [N_RE,NR,16] */
    void *hrm_im;
    void *yvpad_re;             /* fp16 [N_RE/8,NR,16] grouped right-hand sides */
    void *yvpad_im;
    void *no;                   /* fp16 [N_RE] */
    const PuschMimoConfig *configs;
    uint16_t num_configs;
    uint16_t reserved;
    const MimoDetectLayerPlan *layer_plan;
    void *stream;
} MimoDetectIoPackOpArgsV1;
typedef struct MimoDetectIoPackKernelMetadata {
    uint32_t magic;             /* "MIP1" */
    uint32_t active_layers;
    uint32_t num_rx_antennas;
    uint32_t layer_capacity;
    uint32_t n_re;
    uint32_t abi_version;
    uint32_t reserved[2];
} MimoDetectIoPackKernelMetadata;
int32_t airan_mimo_detect_io_pack_BuildCurrentProfile(const PuschMimoConfig *configs,
                                                      size_t num_configs,
                                                      const MimoDetectLayerPlan *layer_plan,
                                                      MimoDetectIoPackKernelMetadata *metadata);
int32_t airan_mimo_detect_io_pack_ValidateOpArgs(const MimoDetectIoPackOpArgsV1 *args);
int32_t airan_mimo_detect_io_pack_Enqueue(const MimoDetectIoPackOpArgsV1 *args,
                                          const void *metadata, size_t metadata_bytes);

/* ============================================================== detector */
/* Replaces the raw launch mimo_detect_bri_kernel(hrm, y, no, ..., xhat,
 * no_eff, yv, ws, tiling) (rx_chain.cpp:557-562): unbiased LMMSE,
 * A = H^H H + no I, xhat = (A^-1 H^H y) / d, d = max(1 - no Re diag(A^-1),
 * 1e-3), no_eff = (1 - d) / d (mimo_detect_bri_ref.py:72-86), solved exactly
 * per RE in FP32 (the reference's block Richardson iteration needs
 * NR >= 8 min(B, L), mimo_detect_bri.h:93-101, which 4 Rx violates). */
typedef struct MimoDetectLmmseOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *hrm_re;         /* fp16 [N_RE,NR,16] */
    const void *hrm_im;
    const void *yvpad_re;       /* fp16 [N_RE/8,NR,16] */
    const void *yvpad_im;
    const void *no;             /* fp16 [N_RE] */
    void *xhat_re;              /* fp16 [16,N_RE] */
    void *xhat_im;
    void *no_eff;               /* fp16 [16,N_RE] */
    const PuschMimoConfig *config;
    const MimoDetectLayerPlan *layer_plan;
    void *stream;
} MimoDetectLmmseOpArgsV1;
int32_t airan_mimo_detect_lmmse_ValidateOpArgs(const MimoDetectLmmseOpArgsV1 *args);
int32_t airan_mimo_detect_lmmse_Enqueue(const MimoDetectLmmseOpArgsV1 *args);

/* ============================================================ QAM demap */
/* mapping/qam_demod_256_mimo_batch/qam256_demod_batch.h:47-100, for
 * Qm in {2,4,6,8}.  Output [L,Qm,14-D,qam_row_stride] int16 (Q11.5, clip
 * +-2560, positive => bit 0), row tail zero. */
typedef struct QamDemod256BatchOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *x_re;           /* fp16 [>=L,14,padded] */
    const void *x_im;
    const void *no_eff;
    void *layer_llr;
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    void *stream;
} QamDemod256BatchOpArgsV1;
typedef struct QamDemodBatchTilingData {
    int32_t n_sc_used;
    int32_t n_sc_pad;
    int32_t n_data_symbols;
    int32_t n_data_re;
    int32_t llr_layer_stride;
    int32_t block_dim;
    int32_t llr_clip;
    int32_t y_scaled_clip;
    int32_t num_layers;
    int32_t grid_layer_stride;
    int32_t physical_data_symbols[14];
    int32_t group_size;
    int32_t reserved[7];
} QamDemodBatchTilingData;
int32_t airan_qam256_demod_batch_BuildCurrentProfile(const PuschMimoConfig *config,
                                                     PuschMimoLayout *layout,
                                                     QamDemodBatchTilingData *tiling);
int32_t airan_qam256_demod_batch_ValidateOpArgs(const QamDemod256BatchOpArgsV1 *args);
int32_t airan_qam256_demod_batch_Enqueue(const QamDemod256BatchOpArgsV1 *args, void *workspace,
                                         size_t workspace_bytes, const void *tiling,
                                         size_t tiling_bytes);

/* =========================================================== layer demap */
/* mapping/layer_demap_mimo/layer_demap.h:63-116 */
typedef struct LayerDemapOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *layer_llr;      /* int16 [L,Qm,14-D,qam_row_stride] */
    void *cw_llr;               /* int16 [Qm,codeword_stride] */
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    void *stream;
} LayerDemapOpArgsV1;
typedef struct LayerDemapKernelMetadata {
    uint32_t magic;             /* "LDM1" */
    uint32_t num_layers;
    uint32_t qm;
    uint32_t num_data_re;
    uint32_t data_stride;
    uint32_t codeword_symbols;
    uint32_t codeword_stride;
    uint32_t input_symbol_stride;
    uint32_t symbols_per_group;
    uint32_t group_data_re;
    uint32_t gather_chunk;
    uint32_t gather_index_elems;
    uint32_t reserved[20];
} LayerDemapKernelMetadata;
int32_t airan_layer_demap_BuildCurrentProfile(const PuschMimoConfig *config,
                                              PuschMimoLayout *layout,
                                              LayerDemapKernelMetadata *metadata);
int32_t airan_layer_demap_ValidateOpArgs(const LayerDemapOpArgsV1 *args);
int32_t airan_layer_demap_Enqueue(const LayerDemapOpArgsV1 *args);
/* mapping/layer_demap_mimo/layer_demap.h:104-109 (ReferenceDemap): bit-exact
 * host layer demap, layer LLRs [L][Qm][14-D][qam_row_stride] ->
 * codeword [Qm][codeword_stride], tails zero; AIRAN_LAYOUT_MISMATCH when
 * `layout` is not the config's. */
int32_t airan_layer_demap_ReferenceDemap(const int16_t *layer_llr, const PuschMimoConfig *config,
                                         const PuschMimoLayout *layout, int16_t *cw_llr);

/* ============================================================ descramble */
/* fec/descramble_mimo/descramble_mimo.h:48-113.  sign may be NULL: the Gold
 * sequence is then generated on the device (the default); a host-built sign
 * table (BuildGoldSign) is still accepted.  Extension field `flags` (present
 * when struct_size covers it): AIRAN_DESCR_RESTART_PER_SLOT restarts the
 * Gold sequence in every slot (one TB per slot, TS 38.211 6.3.1.1) instead of
 * continuing it across the slot batch (the reference's multi-slot TB). */
#define AIRAN_DESCR_RESTART_PER_SLOT 1u
typedef struct DescrambleMimoOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *llr_qam;        /* int16 [slots,Qm,codeword_stride] */
    void *llr_nr;
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    uint32_t num_slots;
    void *stream;
    uint32_t flags;
    uint32_t reserved;
} DescrambleMimoOpArgsV1;
typedef struct DescrambleKernelMetadata {
    uint32_t magic;             /* "DMM1" */
    uint32_t num_slots;
    uint32_t num_layers;
    uint32_t qm;
    uint32_t num_data_re;
    uint32_t data_stride;
    uint32_t codeword_symbols;
    uint32_t codeword_stride;
    uint32_t tile_elems;
    uint32_t max_slots;
    uint32_t reserved[22];
} DescrambleKernelMetadata;
int32_t airan_descramble_mimo_BuildCurrentProfile(const PuschMimoConfig *config,
                                                  uint32_t num_slots, PuschMimoLayout *layout,
                                                  DescrambleKernelMetadata *metadata);
int32_t airan_descramble_mimo_ValidateOpArgs(const DescrambleMimoOpArgsV1 *args);
int32_t airan_descramble_mimo_BuildGoldSign(const PuschMimoConfig *config,
                                            const PuschMimoLayout *layout, uint32_t num_slots,
                                            int16_t *sign, size_t sign_elems);
int32_t airan_descramble_mimo_Enqueue(const DescrambleMimoOpArgsV1 *args, const void *sign,
                                      size_t sign_bytes, void *workspace,
                                      size_t workspace_bytes, const void *tiling,
                                      size_t tiling_bytes);
/* fec/descramble_mimo/descramble_mimo.h:92-98 (ReferenceDescramble): bit-exact
 * host descramble of num_slots slots [slots][Qm][codeword_stride], one Gold
 * sequence continuing across the slots; sign is BuildGoldSign's table, or
 * NULL to generate the same sequence here.  Tails zero. */
int32_t airan_descramble_mimo_ReferenceDescramble(const int16_t *llr_qam, const int16_t *sign,
                                                  const PuschMimoConfig *config,
                                                  const PuschMimoLayout *layout, uint32_t num_slots,
                                                  int16_t *llr_nr);

/* ========================================================= TB / CB plan */
/* TS 38.214 5.1.3.2 and TS 38.212 5.1-5.4.  The reference has no TB layer:
 * its profile fixes C = 143 code blocks of K = 8448 without CRC or filler
 * (rate_dematch_mimo.h:24-31, ldpc_decode.h:26-33). */
typedef struct AiranTbPlanV1 {
    uint32_t tbs;               /* transport block size (bits, without CRC) */
    uint32_t crc_tb;            /* 24 (CRC24A) or 16 */
    uint32_t crc_cb;            /* 24 (CRC24B) when C > 1 else 0 */
    uint32_t C;                 /* code blocks */
    uint32_t K;                 /* 22 * Z */
    uint32_t Kprime;            /* bits per code block incl. CB CRC, excl. filler */
    uint32_t Z;
    uint32_t filler;
    uint32_t N;                 /* 66 * Z */
    uint32_t Ncb;
    uint32_t k0;
    uint32_t rv;
    uint32_t G;                 /* coded bits of the TB */
    uint32_t E_low, E_high, n_low; /* first n_low CBs use E_low */
    uint32_t reserved[4];
} AiranTbPlanV1;
/* TBS from (PRB, data REs per PRB, Qm, layers, code rate). */
uint32_t airan_pusch_Tbs(uint32_t n_prb, uint32_t n_re_per_prb, uint32_t qm, uint32_t layers,
                         double code_rate);
int32_t airan_pusch_BuildTbPlan(uint32_t tbs, uint32_t G, uint32_t qm, uint32_t layers,
                                uint32_t rv, AiranTbPlanV1 *plan);
/* The reference's fixed 23-slot profile: C = 143, K' = K = 8448, Z = 384,
 * no CRC, rv0 (rate_dematch_mimo.cpp:90-112). */
int32_t airan_pusch_ReferenceTbPlan(const PuschMimoConfig *config, uint32_t num_slots,
                                    AiranTbPlanV1 *plan);
/* Smallest post-detection noise variance the QAM demapper's integer
 * arithmetic represents: below it u = rint(y fp16(32 D / n)) saturates at
 * +-1000 for the outer levels and t = rint(fp16(32 D^2 / n)) overflows, so
 * the magnitude-bit LLRs collapse (qam_demod_256_mimo_batch/scripts/
 * gen_data.py:20-27 at n -> 0).  The chain clamps no_eff to this floor
 * before demapping: 0.04 (2^(Qm/2) - 1) D^2, i.e. |y| up to 1.25x the
 * outermost level stays below |u| = 1000.  No counterpart in the reference,
 * which leaves high-SNR operating points to the saturated arithmetic. */
float airan_pusch_NoEffFloor(uint32_t qm);

/* ========================================================= rate dematch */
/* fec/rate_dematch_mimo/rate_dematch_mimo.h:58-137, generalised with filler
 * bits (LLR +5120) and puncturing (E < Ncb). */
typedef struct RateMatchDescriptor {
    uint32_t e;
    uint32_t k0;
    uint32_t ncb;
    uint32_t cw_symbol_offset;
} RateMatchDescriptor;
typedef struct RateDematchMimoOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *cw_llr;         /* int16 [slots,Qm,codeword_stride] */
    void *ldpc_llr;             /* int16 [C,68Z] */
    const PuschMimoConfig *config;
    const PuschMimoLayout *layout;
    uint32_t num_slots;
    void *stream;
} RateDematchMimoOpArgsV1;
typedef struct RateDematchKernelMetadata {
    uint32_t magic;             /* "RDM1" */
    uint32_t num_slots;
    uint32_t num_layers;
    uint32_t qm;
    uint32_t num_data_re;
    uint32_t codeword_symbols;
    uint32_t codeword_stride;
    uint32_t total_coded_bits;
    uint32_t code_blocks;
    uint32_t ldpc_n;
    uint32_t n_2z;
    uint32_t ncb;
    uint32_t llr_scale;
    uint32_t llr_clip;
    uint32_t tile_elems;
    uint32_t max_slots;
    uint32_t descriptor_words;
    uint32_t z;                 /* extension: lifting size */
    uint32_t kprime;            /* extension: K' (filler starts at K'-2Z) */
    uint32_t filler;            /* extension: filler bits per CB */
    uint32_t reserved[12];
} RateDematchKernelMetadata;
/* Reference entry (fixed C = 143 profile). */
int32_t airan_rate_dematch_mimo_BuildCurrentProfile(const PuschMimoConfig *config,
                                                    uint32_t num_slots, PuschMimoLayout *layout,
                                                    RateDematchKernelMetadata *metadata,
                                                    RateMatchDescriptor *descriptors,
                                                    size_t descriptor_count);
/* General entry: descriptors from a TB plan. */
int32_t airan_rate_dematch_mimo_BuildProfileForTb(const PuschMimoConfig *config,
                                                  uint32_t num_slots, const AiranTbPlanV1 *plan,
                                                  PuschMimoLayout *layout,
                                                  RateDematchKernelMetadata *metadata,
                                                  RateMatchDescriptor *descriptors,
                                                  size_t descriptor_count);
int32_t airan_rate_dematch_mimo_ValidateOpArgs(const RateDematchMimoOpArgsV1 *args);
/* descriptors/tiling are device copies of the BuildProfile outputs; the
 * host copy of the metadata is also required (grid sizing). */
int32_t airan_rate_dematch_mimo_Enqueue(const RateDematchMimoOpArgsV1 *args,
                                        const RateDematchKernelMetadata *metadata_host,
                                        const void *descriptors, size_t descriptor_bytes,
                                        void *workspace, size_t workspace_bytes,
                                        const void *tiling, size_t tiling_bytes);
/* fec/rate_dematch_mimo/rate_dematch_mimo.h:112-120 (ReferenceRateDematch):
 * bit-exact host rate dematch of the reference's fixed plan (descriptors as
 * BuildCurrentProfile gives them, else AIRAN_LAYOUT_MISMATCH) into
 * ldpc_llr [C][68 Z]. */
int32_t airan_rate_dematch_mimo_ReferenceRateDematch(const int16_t *cw_llr, const PuschMimoConfig *config,
                                                     const PuschMimoLayout *layout, uint32_t num_slots,
                                                     const RateMatchDescriptor *descriptors,
                                                     size_t descriptor_count, int16_t *ldpc_llr);
/* The same for a TB plan (filler bits +5120, puncturing, the descriptors of
 * BuildProfileForTb); no counterpart in the reference, whose plan is fixed. */
int32_t airan_rate_dematch_mimo_ReferenceRateDematchForTb(const int16_t *cw_llr, const PuschMimoConfig *config,
                                                          const PuschMimoLayout *layout, uint32_t num_slots,
                                                          const AiranTbPlanV1 *plan,
                                                          const RateMatchDescriptor *descriptors,
                                                          size_t descriptor_count, int16_t *ldpc_llr);

/* ============================================ SISO LS + linear interpolation */
/* Replaces the raw launch channel_est_ls_kernel(y_re, y_im, x_ref_re,
 * x_ref_im, out_h_re, out_h_im, weave_idx, ws, tiling)
 * (channel_est/channel_est_ls_siso/channel_est_ls_kernel.cpp; algorithm
 * operator.yaml:82-93, golden scripts/channel_est_ref.py:97-121).  One Rx,
 * one layer on port 1000 (comb 0), per slot, all arithmetic in fp16 in the
 * reference's order:
 *   h_ls[d][p]   = y[d_sym][2p] * conj(x[d][p])   (re: yr xr + yi xi,
 *                  im: yi xr - yr xi, every product and sum rounded),
 *                  h_ls[d][p] = 0 for p >= used / 2 (zero-extended grid)
 *   H_d[2p]      = h_ls[d][p] + 0
 *   H_d[2p + 1]  = (h_ls[d][p] + h_ls[d][p + 1]) * 0.5
 *   H[sym]       = H_0 * fp16((b - sym) / (b - a)) + H_1 * fp16((sym - a) / (b - a))
 *                  for DMRS symbols a < b (the reference's {2, 11}: (11 - sym) / 9
 *                  and (sym - 2) / 9, extrapolated outside [a, b]); H_0 on every
 *                  symbol when the slot has one DMRS symbol.
 * Subcarriers >= used are written as zero.  dmrs_ref: one reference
 * [D][dmrs_ref_pad] for every slot, or one per slot (ref_per_slot). */
typedef struct ChannelEstLsSisoOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *rx_grid_re;     /* fp16 [batch,14,padded] */
    const void *rx_grid_im;
    const void *dmrs_ref_re;    /* fp16 [1 or batch][D,dmrs_ref_pad] (airan_mimo_dmrs_gen layout, L = 1) */
    const void *dmrs_ref_im;
    void *h_re;                 /* fp16 [batch,14,padded] */
    void *h_im;
    const PuschMimoConfig *config;
    uint32_t batch;             /* slots */
    uint32_t ref_per_slot;      /* 0: one DMRS reference for every slot */
    void *stream;
} ChannelEstLsSisoOpArgsV1;
int32_t airan_channel_est_ls_siso_ValidateOpArgs(const ChannelEstLsSisoOpArgsV1 *args);
int32_t airan_channel_est_ls_siso_Enqueue(const ChannelEstLsSisoOpArgsV1 *args);

/* ============================================================= SISO ZF */
/* Replaces the raw launch equalize_kernel(y_re, y_im, h_re, h_im, n0, x_re,
 * x_im, no_eff, ws, tiling) (equalization/equalize_zf_siso/
 * equalize_kernel.cpp; operator.yaml:72-84): per RE, from fp16 inputs in fp32,
 *   g = max(hr^2 + hi^2, 1e-4), x = ((hr yr + hi yi) / g, (hr yi - hi yr) / g),
 *   no_eff = n0 / g, outputs rounded to fp16 (IEEE division). */
#define AIRAN_ZF_G_FLOOR 1.0e-4f
typedef struct EqualizeZfSisoOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *y_re;           /* fp16 [n] */
    const void *y_im;
    const void *h_re;           /* fp16 [n] */
    const void *h_im;
    const void *n0;             /* fp16 [n], pre-equalisation noise variance */
    void *x_re;                 /* fp16 [n] */
    void *x_im;
    void *no_eff;               /* fp16 [n] */
    uint32_t n;                 /* resource elements (e.g. batch * 14 * padded) */
    uint32_t reserved;
    void *stream;
} EqualizeZfSisoOpArgsV1;
int32_t airan_equalize_zf_siso_ValidateOpArgs(const EqualizeZfSisoOpArgsV1 *args);
int32_t airan_equalize_zf_siso_Enqueue(const EqualizeZfSisoOpArgsV1 *args);

/* ================================================================= LDPC */
/* Replaces the raw launch ldpc_decode_kernel(lam_in, packed_bc, packed_s,
 * degrees, edge_offsets, prev_msg, bits, lam_out, scratch)
 * (rx_chain.cpp:654-657; kernel fec/ldpc_decode/ldpc_decode_kernel.cpp:315).
 * BG1 layered normalized min-sum, alpha = 0.75, rows in order. */
#define AIRAN_LDPC_MODE_REF 0   /* fp16 posterior in Q8.8 units, int8 messages (reference arithmetic) */
#define AIRAN_LDPC_MODE_FP16 1  /* fp16 posterior in LLR units (input / 256), fp16 messages */
typedef struct LdpcDecodeOpArgsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *lam_in;         /* int16 [num_cb,68Z] Q8.8 */
    void *bits;                 /* int8 [num_cb,K]: 0/1 */
    void *lam_out;              /* int16 [num_cb,68Z] clamped posterior, or NULL */
    void *iterations_out;       /* uint8 [num_cb] iterations run, or NULL */
    uint32_t num_cb;
    uint32_t z;
    uint32_t k;                 /* bits written per CB (<= 22Z) */
    uint32_t max_iterations;
    uint32_t active_rows;       /* 4..46; rows past the last transmitted parity
                                   column carry no information (MODE_FP16) */
    uint32_t mode;              /* AIRAN_LDPC_MODE_* */
    uint32_t early_stop;        /* stop a CB once its active rows are satisfied */
    uint32_t reserved;
    void *stream;
} LdpcDecodeOpArgsV1;
int32_t airan_ldpc_decode_ValidateOpArgs(const LdpcDecodeOpArgsV1 *args);
size_t airan_ldpc_decode_WorkspaceBytes(uint32_t num_cb, uint32_t z, uint32_t active_rows);
int32_t airan_ldpc_decode_Enqueue(const LdpcDecodeOpArgsV1 *args, void *workspace,
                                  size_t workspace_bytes);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Decodes exactly as airan_ldpc_decode_Enqueue, with the
 * decoder's message state (WorkspaceBytes) in a fixed device array of the
 * library instead of a caller workspace, at most
 * NoAllocMaxWorkspaceBytes() = 32 MiB (the reference's 143-block transport
 * block at 46 rows needs 20.2 MB); larger requests return
 * AIRAN_RESOURCE_TOO_SMALL.  Allocates nothing: the array's device address
 * is taken once per device by airan_InitDevice (called here if it has not
 * run on the current device; call it first to keep this entry point to the
 * decoder launch alone).  The array is shared by every call on one GPU and
 * keeps the last call's contents (the decoder clears the state it reads),
 * so calls that overlap on one GPU (concurrent streams or threads) race:
 * serialise them. */
size_t airan_ldpc_decode_NoAllocMaxWorkspaceBytes(void);
int32_t airan_ldpc_decode_EnqueueNoAlloc(const LdpcDecodeOpArgsV1 *args);
/* Active rows for a TB plan (see LdpcDecodeOpArgsV1.active_rows). */
uint32_t airan_ldpc_ActiveRows(const AiranTbPlanV1 *plan);

/* ================================================================== CRC */
/* CB CRC24B and TB CRC24A/CRC16 check on decoded bits (TS 38.212 5.1).  The
 * reference receiver has no CRC (fec/ldpc_decode/operator.yaml:49).
 * bits: int8 [num_tbs*C, K]; cb_ok uint8 [num_tbs*C] (1 when C == 1);
 * tb_ok uint8 [num_tbs]; tb_bytes (optional) uint8 packed payload bytes
 * [num_tbs, ceil(tbs/8)], MSB first; workspace >= WorkspaceBytes(num_tbs),
 * 16-byte aligned for the matrix-core path.  plan->C <= 256 (every TS 38.214
 * TB size needs at most 152 code blocks). */
size_t airan_crc_check_WorkspaceBytes(uint32_t num_tbs);
int32_t airan_crc_check_Enqueue(const AiranTbPlanV1 *plan, uint32_t num_tbs, const void *bits,
                                void *cb_ok, void *tb_ok, void *tb_bytes, void *workspace,
                                size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* AIRAN_PUSCH_RX_H */
