/*
 * airan_pusch_types.h -- C-compatible restatement of the shared PUSCH MIMO ABI.
 *
 * Byte layouts are identical to the reference so that a 192-byte runtime
 * config file, a PuschMimoConfig and a PuschMimoLayout produced by either side
 * are interchangeable:
 *   PuschMimoConfig          common/pusch_mimo_types.h:48-86   (ABI v1, 96 B)
 *   MimoDetectAllocation     common/pusch_mimo_types.h:89-94
 *   MimoDetectLayerPlan      common/pusch_mimo_types.h:100-109
 *   PuschMimoLayout          common/pusch_mimo_types.h:111-118
 *   PuschMimoRuntimeConfig   common/pusch_mimo_types.h:129-188 (ABI v2, 192 B)
 * Offsets are asserted below exactly as common/pusch_mimo_types.h:204-217 does.
 *
 * Complex tensors at operator boundaries are separate real/imaginary planes
 * (fp16 unless an operator states otherwise); grids are [outer, symbol,
 * subcarrier] with the subcarrier axis contiguous.
 */
#ifndef AIRAN_PUSCH_TYPES_H
#define AIRAN_PUSCH_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AIRAN_PUSCH_MIMO_ABI_VERSION 1
#define AIRAN_PUSCH_MIMO_RUNTIME_ABI_VERSION 2
#define AIRAN_PUSCH_MIMO_MAX_TX_ANTENNAS 16
#define AIRAN_PUSCH_MIMO_MAX_RX_ANTENNAS 64
#define AIRAN_PUSCH_MIMO_MAX_PUSCH_LAYERS 4
#define AIRAN_PUSCH_MIMO_MAX_DETECT_LAYERS 16
#define AIRAN_PUSCH_MIMO_MAX_DMRS_PORTS 16
#define AIRAN_PUSCH_MIMO_MAX_ALLOCATIONS 16

#define AIRAN_MIMO_ARCH_FULL_DIGITAL 1
#define AIRAN_MIMO_ARCH_HYBRID 2
#define AIRAN_MIMO_ARCH_ANALOG 3

#define AIRAN_MIMO_PRECODING_BYPASS 0
#define AIRAN_MIMO_PRECODING_CODEBOOK 1
#define AIRAN_MIMO_PRECODING_NON_CODEBOOK 2

#define AIRAN_RUNTIME_FLAG_PROFILE_HASH (1u << 0)
#define AIRAN_RUNTIME_FLAG_ADAPTER_REQUIRED (1u << 1)
#define AIRAN_RUNTIME_KNOWN_FLAGS \
    (AIRAN_RUNTIME_FLAG_PROFILE_HASH | AIRAN_RUNTIME_FLAG_ADAPTER_REQUIRED)

typedef struct PuschMimoConfig {
    uint16_t abi_version;
    uint16_t struct_size;
    uint32_t flags;

    uint16_t num_layers;
    uint16_t num_tx_ports;
    uint16_t num_rx_antennas;
    uint16_t qm;

    uint16_t num_symbols;
    uint16_t fft_size;
    uint16_t num_rb;
    uint16_t rb_start;
    uint16_t slot_number;
    uint8_t start_symbol;
    uint8_t num_allocated_symbols;
    uint16_t used_subcarriers;
    uint16_t padded_subcarriers;

    uint16_t dmrs_symbol_mask;
    uint16_t dmrs_ports[4];
    uint16_t dmrs_scrambling_id;
    uint16_t data_scrambling_id;
    uint16_t rnti;
    uint8_t dmrs_type;
    uint8_t dmrs_length;
    uint8_t dmrs_additional_position;
    uint8_t mapping_type;
    uint8_t num_cdm_groups_without_data;
    uint8_t n_scid;
    uint8_t codeword_index;

    uint8_t transform_precoding;
    uint8_t codebook_enabled;
    uint16_t tpmi;
    uint16_t prg_size_rb;
    uint32_t reserved[8];
} PuschMimoConfig;

typedef struct MimoDetectAllocation {
    uint16_t pusch_index;
    uint16_t layer_offset;
    uint16_t num_layers;
    uint16_t num_tx_ports;
} MimoDetectAllocation;

typedef struct MimoDetectLayerPlan {
    uint16_t abi_version;
    uint16_t struct_size;
    uint16_t num_rx_antennas;
    uint16_t layer_capacity;
    uint16_t total_layers;
    uint16_t num_allocations;
    MimoDetectAllocation allocations[AIRAN_PUSCH_MIMO_MAX_ALLOCATIONS];
    uint32_t reserved[8];
} MimoDetectLayerPlan;

typedef struct PuschMimoLayout {
    uint16_t num_dmrs_symbols;
    uint16_t num_data_symbols;
    uint32_t num_data_re;
    uint32_t data_stride;
    uint32_t codeword_symbols;
    uint32_t codeword_stride;
} PuschMimoLayout;

typedef struct PuschMimoRuntimeConfig {
    uint16_t abi_version;
    uint16_t struct_size;
    uint32_t flags;
    uint8_t profile_sha256[32];

    uint16_t architecture;
    uint16_t num_tx_antennas;
    uint16_t num_tx_rf_chains;
    uint16_t num_rx_antennas;
    uint16_t num_rx_rf_chains;
    uint16_t num_layers;
    uint16_t num_tx_ports;
    uint16_t rx_bucket;
    uint16_t layer_bucket;
    uint16_t max_rx_antennas;
    uint16_t max_layers;

    uint16_t qm;
    uint16_t num_symbols;
    uint16_t fft_size;
    uint16_t num_rb;
    uint16_t rb_start;
    uint16_t used_subcarriers;
    uint16_t padded_subcarriers;

    uint16_t num_dmrs_symbols;
    uint16_t dmrs_symbol_mask;
    uint16_t dmrs_port_count;
    uint16_t dmrs_ports[AIRAN_PUSCH_MIMO_MAX_DMRS_PORTS];

    uint16_t dmrs_type;
    uint16_t dmrs_length;
    uint16_t num_cdm_groups_without_data;
    uint16_t n_scid;
    uint16_t precoding_mode;
    uint16_t tpmi;
    uint16_t prg_size_rb;
    float channel_gain;
    float awgn_std_int16;

    uint32_t data_re_per_layer;
    uint32_t data_stride_per_layer;
    uint32_t qam_row_stride;
    uint32_t grid_re_per_port;

    uint16_t slot_number;
    uint16_t dmrs_scrambling_id;
    uint16_t data_scrambling_id;
    uint16_t rnti;
    uint16_t start_symbol;
    uint16_t num_allocated_symbols;
    uint16_t dmrs_additional_position;
    uint16_t mapping_type;
    uint16_t codeword_index;
    uint16_t reserved16;
    uint32_t reserved[5];
} PuschMimoRuntimeConfig;

#ifdef __cplusplus
#define AIRAN_STATIC_ASSERT(cond, msg) static_assert(cond, msg)
#else
#define AIRAN_STATIC_ASSERT(cond, msg) _Static_assert(cond, msg)
#endif

AIRAN_STATIC_ASSERT(sizeof(PuschMimoConfig) == 96, "PuschMimoConfig ABI size changed");
AIRAN_STATIC_ASSERT(sizeof(PuschMimoRuntimeConfig) == 192, "runtime config ABI size changed");
AIRAN_STATIC_ASSERT(offsetof(PuschMimoRuntimeConfig, profile_sha256) == 8, "hash offset");
AIRAN_STATIC_ASSERT(offsetof(PuschMimoRuntimeConfig, architecture) == 40, "topology offset");
AIRAN_STATIC_ASSERT(offsetof(PuschMimoRuntimeConfig, dmrs_ports) == 82, "DMRS offset");
AIRAN_STATIC_ASSERT(offsetof(PuschMimoRuntimeConfig, channel_gain) == 128, "channel offset");
AIRAN_STATIC_ASSERT(offsetof(PuschMimoRuntimeConfig, data_re_per_layer) == 136, "geometry offset");
AIRAN_STATIC_ASSERT(offsetof(PuschMimoRuntimeConfig, slot_number) == 152, "scheduler offset");
AIRAN_STATIC_ASSERT(sizeof(PuschMimoLayout) == 20, "layout size");
AIRAN_STATIC_ASSERT(sizeof(MimoDetectLayerPlan) == 172, "layer plan size");

/* Status of the runtime-config functions (common/pusch_mimo_runtime_config.h:10-17). */
#define AIRAN_RTCFG_SUCCESS 0
#define AIRAN_RTCFG_NULL_ARGUMENT 1
#define AIRAN_RTCFG_ABI_MISMATCH 2
#define AIRAN_RTCFG_UNSUPPORTED_VALUE 3
#define AIRAN_RTCFG_INCONSISTENT_SHAPE 4
#define AIRAN_RTCFG_IO_ERROR 5

/* Operator Status values shared by every operator entry point
 * (e.g. mapping/re_demap_batch/re_demap_batch.h:18-25). */
#define AIRAN_OK 0
#define AIRAN_INVALID_ARGUMENT (-1)
#define AIRAN_UNSUPPORTED_PROFILE (-2)
#define AIRAN_LAYOUT_MISMATCH (-3)
#define AIRAN_RESOURCE_TOO_SMALL (-4)
#define AIRAN_LAUNCH_FAILED (-5)

#ifdef __cplusplus
}
#endif

#endif /* AIRAN_PUSCH_TYPES_H */
