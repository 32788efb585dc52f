/*
 * hgap.h -- C ABI of the MI355X (gfx950) Huffman encoder/decoder with gap arrays.
 *
 * Stream format (bit-identical to the reference; reference paths relative to its root):
 *   u64 k; k x {u8 symbol, u8 length} in code-assignment order (descending count, equal counts
 *   in descending byte value); u32 n; u32 payload_words W; u32 gap_elements G;
 *   u32 gap[ceil(G/8)]; u32 payload[W]                       (encoder/src/huff.cpp:111-125)
 *   payload: codewords concatenated MSB-first in 32-bit words, zero tail;
 *   gap element s: number of bits by which the codeword that starts in 128-bit segment s spills
 *   into segment s+1 (0 when none does); 8 elements per u32, low nibble first
 *   (encoder/src/encoder.cu:307-312,371-377).
 *
 * Device entry points are asynchronous on the given HIP stream (a hipStream_t passed as void*;
 * NULL = the default stream), take caller-owned device buffers and an explicit workspace, and
 * never allocate.  Every function returns an hgap_status; nothing calls exit().
 * Device pointers (inputs, payload, gap, outputs, workspace, tables) must be 16-byte aligned.
 */
#ifndef HGAP_H
#define HGAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *hgap_stream_t; /* hipStream_t */

typedef enum hgap_status {
    HGAP_OK = 0,
    HGAP_ERR_INVALID_ARG = 1, /* null pointer, misaligned pointer, bad enum */
    HGAP_ERR_SIZE = 2,        /* sizes disagree, or exceed the format's 32-bit fields */
    HGAP_ERR_CODEBOOK = 3,    /* lengths > 16, not prefix-free, not monotone, symbol missing */
    HGAP_ERR_WORKSPACE = 4,   /* workspace smaller than the *_workspace_bytes() figure */
    HGAP_ERR_HIP = 5,         /* a HIP runtime call failed */
    HGAP_ERR_FORMAT = 6,      /* stream truncated or header inconsistent */
    HGAP_ERR_CAPACITY = 7,    /* caller buffer too small */
    HGAP_ERR_UNSUPPORTED = 8  /* a configuration the kernels do not handle */
} hgap_status;

#define HGAP_MAX_CODE_LEN 16
#define HGAP_SEGMENT_BITS 128
#define HGAP_GAP_PER_WORD 8

/* Codebook: the reference's Codetable[256] (encoder/include/constants.hpp:40-43) plus the
 * code-assignment order the stream header lists. */
typedef struct hgap_codebook {
    uint32_t symbol_count;     /* k, 0..256 */
    uint8_t order_symbol[256]; /* code-assignment (file) order */
    uint8_t order_length[256];
    uint32_t code[256];        /* by byte value; valid where length != 0 */
    uint32_t length[256];      /* by byte value; 0 = byte absent */
} hgap_codebook;

/* Output sizes: encoder/src/symbols.cpp:45-55 and encoder/src/huff.cpp:84-88. */
typedef struct hgap_sizes {
    uint64_t bits;          /* sum of count x length */
    uint32_t payload_words; /* ceil(bits / 32)  = W */
    uint32_t gap_elements;  /* ceil(bits / 128) = G */
    uint32_t gap_words;     /* ceil(G / 8) */
    uint32_t reserved;
} hgap_sizes;

/* Decode table image (built on the host, copied to device memory by the caller).  Entry layout
 * is internal to this library; only the sizes below are needed to place it. */
typedef struct hgap_dectable_info {
    uint32_t prefix_bits;   /* first-level index width */
    uint32_t l1_entries;    /* 1 << prefix_bits */
    uint32_t l2_entries;    /* second-level entries (0 when one level covers all codes) */
    uint32_t bytes;         /* image size: 4 * (l1_entries + l2_entries + 64); the last 64
                             * words hold the code length of each byte value */
    uint32_t min_length;    /* shortest code length */
    uint32_t max_length;    /* longest code length */
    uint32_t max_symbols_per_segment; /* bound on codewords starting in one 128-bit segment */
    uint32_t symbol_count;
} hgap_dectable_info;

/* Launch shape of the device steps.  hgap_default_config() gives the tuned defaults; the
 * autotuner (huffman_coding_gap_arrays_amd/tuner.py) sweeps these fields. */
typedef struct hgap_launch_config {
    uint32_t enc_threads;          /* encoder workgroup size: 256, 512 or 1024 */
    uint32_t enc_bytes_per_thread; /* input bytes per encoder thread: 16, 32 or 64 */
    uint32_t dec_threads;          /* decoder workgroup size = segments per tile: 256/512/1024 */
    uint32_t lookback_window;      /* predecessor tiles read per look-back round: 1..256 */
    uint32_t blocks_per_cu;        /* persistent workgroups per CU; 0 = as many as fit */
    uint32_t prefix_bits;          /* decode-table first-level width; 0 = library default */
} hgap_launch_config;

const char *hgap_status_string(int status);
void hgap_default_config(hgap_launch_config *cfg);
/* hgap_default_config with the chained encoder's default shape (hgap_encode_ranges,
 * hgap_batch_encode_ranges); the per-tile look-back encoder's default stays
 * hgap_default_config's.  No counterpart in the reference (its launch shape is fixed). */
void hgap_default_chained_config(hgap_launch_config *cfg);
/* Device (ordinal) properties used by the launchers; -1 when no usable device. */
int hgap_device_cu_count(void);

/* ---- host-side steps of the path ---------------------------------------------------------- */

/* store_symbols + qsort + boundary_PM + code assignment, bit-exact with
 * encoder/src/symbols.cpp:29-43, encoder/src/huff.cpp:17-21,78-81 and
 * encoder/src/package_merge.cpp:107-182.  A single distinct byte gets length 1, code 0; an
 * all-zero histogram gives k = 0 (the reference cannot run on an empty input). */
hgap_status hgap_codebook_build(const uint32_t hist[256], hgap_codebook *cb);

/* Rebuilds the codebook from a stream header's (symbol, length) list using the canonical rule
 * of decoder/src/get_table.cpp:17-20; checks lengths and prefix-freeness. */
hgap_status hgap_codebook_from_order(const uint8_t *order_symbol, const uint8_t *order_length,
                                     uint32_t k, hgap_codebook *cb);

/* get_outputfilesize + size arithmetic (encoder/src/symbols.cpp:45-55, huff.cpp:84-88). */
hgap_status hgap_encode_sizes(const hgap_codebook *cb, const uint32_t hist[256],
                              hgap_sizes *sizes);

/* Decode table; prefix_bits = 0 picks the default split.  `image` may be NULL to query info. */
hgap_status hgap_dectable_build(const hgap_codebook *cb, uint32_t prefix_bits,
                                hgap_dectable_info *info, void *image, size_t image_cap);

/* Stream header (encoder/src/huff.cpp:111-123; decoder/src/huff.cpp:27-39,71-74). */
uint64_t hgap_stream_bytes(uint32_t k, uint32_t payload_words, uint32_t gap_elements);
uint64_t hgap_stream_header_bytes(uint32_t k);
hgap_status hgap_stream_write_header(const hgap_codebook *cb, uint32_t n,
                                     const hgap_sizes *sizes, uint8_t *dst, size_t cap,
                                     size_t *written);
/* Parses a header; reports where the gap array and the payload start inside the stream. */
hgap_status hgap_stream_parse_header(const uint8_t *src, size_t size, hgap_codebook *cb,
                                     uint32_t *n, hgap_sizes *sizes, uint64_t *gap_offset,
                                     uint64_t *payload_offset);

/* ---- device steps (asynchronous on `stream`) ---------------------------------------------- */

/* Byte histogram: d_hist[256] (u32) receives the counts of d_in[0..n).  Replaces
 * histgram() (encoder/src/encoder.cu:118-140) without its per-call allocation. */
hgap_status hgap_histogram(const void *d_in, uint64_t n, uint32_t *d_hist,
                           hgap_stream_t stream);

/* Encoder workspace: tile counter + one 64-bit look-back status word per input tile (sized
 * for the smallest tile any configuration uses). */
size_t hgap_encode_workspace_bytes(uint64_t n);
/* Writes payload[W] and gap[ceil(G/8)] for d_in[0..n) with codebook `cb`; sizes must come from
 * hgap_encode_sizes for the same input.  Replaces encode() (encoder/src/encoder.cu:382-457):
 * no host copies, no memset of the payload, the caller keeps ownership of d_in.  Launch shape:
 * hgap_default_config's, with smaller workgroups when its LDS image does not fit. */
hgap_status hgap_encode(const void *d_in, uint64_t n, const hgap_codebook *cb,
                        const hgap_sizes *sizes, uint32_t *d_payload, uint32_t *d_gap,
                        void *d_workspace, size_t workspace_bytes, hgap_stream_t stream);

/* hgap_encode with an explicit launch shape (NULL = defaults); HGAP_ERR_UNSUPPORTED when the
 * shape does not fit the device (e.g. its LDS image for this codebook's longest code). */
hgap_status hgap_encode_cfg(const void *d_in, uint64_t n, const hgap_codebook *cb,
                            const hgap_sizes *sizes, uint32_t *d_payload, uint32_t *d_gap,
                            void *d_workspace, size_t workspace_bytes,
                            const hgap_launch_config *cfg, hgap_stream_t stream);

/* The histogram and encode pair of the reference's pipeline (histgram() then encode(),
 * encoder/src/huff.cpp:69-99) with per-range counts handed from one to the other.  Range r is
 * input bytes [r * HGAP_RANGE_BYTES, (r + 1) * HGAP_RANGE_BYTES), the last one ragged.
 * hgap_histogram_ranges writes d_hist[256] as hgap_histogram does, plus d_range_hist[256 * r
 * + b] = count of byte b in range r (hgap_range_count(n) x 256 u32).  hgap_encode_ranges
 * writes the same payload and gap array as hgap_encode; it takes d_range_hist from
 * hgap_histogram_ranges over the same input (other counts give a wrong stream, never an access
 * outside the caller's buffers), derives every range's bit offset from them before encoding,
 * and so needs no tile to wait for another.  Workspace: hgap_encode_workspace_bytes(n).
 * cfg NULL = the chained encoder's own default shape (256 threads x 32 bytes), or 512 x 32
 * when its resident workgroups divide the ranges clearly more evenly (a workgroup encodes
 * whole ranges); inputs under HGAP_CHAINED_MIN_BYTES then take hgap_encode instead (too few
 * ranges to fill the device; the counts go unused).  A non-NULL cfg always runs the chained
 * encoder with that shape. */
#define HGAP_RANGE_BYTES (256u * 1024u)
#define HGAP_CHAINED_MIN_BYTES (64ull << 20)
uint32_t hgap_range_count(uint64_t n);
hgap_status hgap_histogram_ranges(const void *d_in, uint64_t n, uint32_t *d_hist,
                                  uint32_t *d_range_hist, hgap_stream_t stream);
hgap_status hgap_encode_ranges(const void *d_in, uint64_t n, const hgap_codebook *cb,
                               const hgap_sizes *sizes, const uint32_t *d_range_hist,
                               uint32_t *d_payload, uint32_t *d_gap, void *d_workspace,
                               size_t workspace_bytes, const hgap_launch_config *cfg,
                               hgap_stream_t stream);

/* Decoder workspace: tile counter + one 64-bit look-back status word per tile of segments. */
size_t hgap_decode_workspace_bytes(uint32_t gap_elements);
/* Decodes n bytes into d_out from payload[W] + gap array using a table image already in device
 * memory (d_table, described by `info`).  Replaces decoder_l1_l2
 * (decoder/src/decoder.cu:732-815) minus its host copies and timing loop. */
hgap_status hgap_decode(const uint32_t *d_payload, uint32_t payload_words, const uint32_t *d_gap,
                        uint32_t gap_elements, const void *d_table,
                        const hgap_dectable_info *info, void *d_out, uint64_t n,
                        void *d_workspace, size_t workspace_bytes, hgap_stream_t stream);

/* hgap_decode with an explicit launch shape (NULL = defaults). */
hgap_status hgap_decode_cfg(const uint32_t *d_payload, uint32_t payload_words,
                            const uint32_t *d_gap, uint32_t gap_elements, const void *d_table,
                            const hgap_dectable_info *info, void *d_out, uint64_t n,
                            void *d_workspace, size_t workspace_bytes,
                            const hgap_launch_config *cfg, hgap_stream_t stream);

/* ---- batched steps: many independent buffers per launch ----------------------------------- */
/* Job descriptors live in device memory (the caller copies the arrays filled by the plan
 * functions), so a captured graph replays the same launches.  Jobs of one call share one launch
 * shape; a job's look-back status words are [base, base + num_tiles) of the call's workspace. */

typedef struct hgap_hist_job {
    const void *in;  /* device bytes, 16-byte aligned */
    uint64_t n;
    uint32_t *hist;  /* device u32[256], overwritten */
    uint32_t range_base; /* set by hgap_batch_plan_histogram: the job's first row of range
                          * counts (hgap_batch_histogram_ranges) */
    uint32_t reserved;
    uint64_t base;   /* set by hgap_batch_plan_histogram: offset in the concatenated batch */
} hgap_hist_job;

typedef struct hgap_enc_job {
    const void *in;  /* device bytes, 16-byte aligned */
    uint64_t n;
    uint32_t *payload;      /* device, payload_words words */
    uint32_t *gap;          /* device, ceil(gap_elements / 8) words (zeroed by the call) */
    uint32_t payload_words;
    uint32_t gap_words;
    uint32_t table;         /* index of this job's 256-word code table (hgap_encode_table) */
    uint32_t max_len;       /* longest code length of its codebook */
    uint32_t base;          /* set by hgap_batch_plan_encode(_ranges) */
    uint32_t num_tiles;     /* set by hgap_batch_plan_encode(_ranges) */
    uint32_t range_base;    /* hgap_batch_encode_ranges: the job's first row of range counts */
    uint32_t reserved;
} hgap_enc_job;

typedef struct hgap_dec_job {
    const uint32_t *payload; /* device */
    const uint32_t *gap;     /* device */
    void *out;               /* device, n bytes */
    uint64_t n;
    const uint32_t *table;   /* device decode-table image (hgap_dectable_build) */
    uint32_t payload_words;
    uint32_t gap_elements;
    uint32_t prefix_bits;    /* from the table's hgap_dectable_info */
    uint32_t l1_entries;
    uint32_t table_words;    /* words of the table image: hgap_dectable_info.bytes / 4 */
    uint32_t max_symbols;    /* max_symbols_per_segment */
    uint32_t max_length;
    uint32_t base;           /* set by hgap_batch_plan_decode */
    uint32_t num_tiles;      /* set by hgap_batch_plan_decode */
    uint32_t reserved;
} hgap_dec_job;

/* Sets each job's base and range_base; returns the total byte count of the batch. */
uint64_t hgap_batch_plan_histogram(hgap_hist_job *jobs, uint32_t count);
hgap_status hgap_batch_histogram(const hgap_hist_job *d_jobs, uint32_t count, uint64_t total,
                                 hgap_stream_t stream);
/* hgap_batch_histogram plus per-range counts: row range_base + r of d_range_hist (u32[256]
 * each) counts range r of the job (HGAP_RANGE_BYTES each, the last one ragged).
 * total_ranges = the sum of hgap_range_count(n) over the jobs. */
hgap_status hgap_batch_histogram_ranges(const hgap_hist_job *d_jobs, uint32_t count,
                                        uint32_t total_ranges, uint32_t *d_range_hist,
                                        hgap_stream_t stream);

/* Codebooks of `count` histograms (count x 256 u32) on `threads` host threads (0 = all). */
hgap_status hgap_codebook_build_batch(const uint32_t *hists, uint32_t count, hgap_codebook *cbs,
                                      uint32_t threads);
/* The encoder's per-byte code table ((code << 5) | length) of one codebook. */
hgap_status hgap_encode_table(const hgap_codebook *cb, uint32_t table[256]);
/* Per buffer of a batch: codebook, sizes and (when enc_tables is not NULL) the encoder's code
 * table (count x 256 u32), on `threads` host threads (0 = all cores). */
hgap_status hgap_batch_prepare(const uint32_t *hists, uint32_t count, hgap_codebook *cbs,
                               hgap_sizes *sizes, uint32_t *enc_tables, uint32_t threads);
/* Decode tables of `count` codebooks, first-level widths prefix_bits[i] (NULL = defaults).
 * With arena == NULL only the infos are filled (their `bytes` size the images); otherwise
 * image i is written at arena + offsets[i] (4-byte aligned, within arena_bytes). */
hgap_status hgap_dectable_build_batch(const hgap_codebook *cbs, uint32_t count,
                                      const uint32_t *prefix_bits, hgap_dectable_info *infos,
                                      void *arena, const uint64_t *offsets, size_t arena_bytes,
                                      uint32_t threads);

/* Sets base/num_tiles for the tile shape in `cfg` (NULL = default); returns total tiles. */
uint32_t hgap_batch_plan_encode(hgap_enc_job *jobs, uint32_t count,
                                const hgap_launch_config *cfg);
uint32_t hgap_batch_plan_decode(hgap_dec_job *jobs, uint32_t count,
                                const hgap_launch_config *cfg);
/* Workspace of hgap_batch_encode (look-back status words of `total_tiles` tiles). */
size_t hgap_batch_workspace_bytes(uint32_t total_tiles);
/* Workspace of hgap_batch_decode for `total_tiles` tiles of the shape in `cfg` (NULL =
 * default): per-segment codeword counts, per-tile offsets and the offsets scan's status words
 * (about 1.1 bytes per 128-bit segment).  0 for an invalid shape. */
size_t hgap_batch_decode_workspace_bytes(uint32_t total_tiles, const hgap_launch_config *cfg);
/* max_len / max_table_words: maxima over the jobs (they size LDS).  max_symbols: the codewords
 * per segment the decoder's LDS image is sized for (the jobs' max_symbols_per_segment, or less:
 * tiles with more codewords than fit take a slower path, with the same output). */
hgap_status hgap_batch_encode(const hgap_enc_job *d_jobs, uint32_t count, uint32_t total_tiles,
                              uint32_t max_len, const uint32_t *d_tables, void *d_workspace,
                              size_t workspace_bytes, const hgap_launch_config *cfg,
                              hgap_stream_t stream);
/* The chained batched encode (hgap_encode_ranges for every job): hgap_batch_plan_encode_ranges
 * sets base (the job's first range in this call) and num_tiles and returns the call's total
 * ranges; each job's range_base points at its rows of d_range_hist, from
 * hgap_batch_histogram_ranges over the same inputs.  The same streams as hgap_batch_encode;
 * a range's first tile finds its offset without waiting on another workgroup.  Workspace:
 * hgap_batch_encode_ranges_workspace_bytes(total_ranges) (a status word and a job index per
 * range). */
uint32_t hgap_batch_plan_encode_ranges(hgap_enc_job *jobs, uint32_t count,
                                       const hgap_launch_config *cfg);
size_t hgap_batch_encode_ranges_workspace_bytes(uint32_t total_ranges);
hgap_status hgap_batch_encode_ranges(const hgap_enc_job *d_jobs, uint32_t count,
                                     uint32_t total_ranges, uint32_t max_len,
                                     const uint32_t *d_tables, const uint32_t *d_range_hist,
                                     void *d_workspace, size_t workspace_bytes,
                                     const hgap_launch_config *cfg, hgap_stream_t stream);
hgap_status hgap_batch_decode(const hgap_dec_job *d_jobs, uint32_t count, uint32_t total_tiles,
                              uint32_t max_table_words, uint32_t max_symbols, void *d_workspace,
                              size_t workspace_bytes, const hgap_launch_config *cfg,
                              hgap_stream_t stream);

/* HGAP_OK when the launch shape `cfg` fits the device (LDS) for an encoder whose longest code
 * is `max_len` (0 = skip the encoder check) and a decode table described by `info` (NULL = skip
 * the decoder check), for both the single-buffer and the batched kernels; HGAP_ERR_UNSUPPORTED
 * when it does not.  No counterpart in the reference (its launch shape is fixed). */
hgap_status hgap_config_supported(const hgap_launch_config *cfg, uint32_t max_len,
                                  const hgap_dectable_info *info);
/* LDS bytes of the decoder's per-tile output image for `dec_threads` segments per tile, a
 * table image of `table_words` words and segments of at most `max_symbols` codewords (0: the
 * shape does not fit).  Sized for occupancy; tiles with more symbols than fit go through it
 * in windows (slower, same output).  No counterpart in the reference. */
uint32_t hgap_decode_image_bytes(uint32_t dec_threads, uint32_t table_words,
                                 uint32_t max_symbols);

/* ---- launch-shape predictor network ---------------------------------------------------------
 * A two-layer perceptron y = W2 relu(W1 x + b1) + b2 in fp32, trained by full-batch Adam on a
 * masked mean squared error.  Parameters are one flat device array: W1 [hidden x in] row-major,
 * b1 [hidden], W2 [out x hidden] row-major, b2 [out].  No counterpart in the reference. */
uint32_t hgap_mlp_param_count(uint32_t in, uint32_t hidden, uint32_t out);
/* d_x [n x in] -> d_h [n x hidden] (post-ReLU), d_y [n x out]. */
hgap_status hgap_mlp_forward(const float *d_x, uint32_t n, uint32_t in, uint32_t hidden,
                             uint32_t out, const float *d_params, float *d_h, float *d_y,
                             hgap_stream_t stream);
/* Gradient of L = sum(mask * (y - target)^2) / mask_sum with respect to the parameters, from
 * the forward pass's d_h and d_y; d_dy [n x out] and d_dh [n x hidden] are scratch outputs. */
hgap_status hgap_mlp_backward(const float *d_x, const float *d_h, const float *d_y,
                              const float *d_target, const float *d_mask, uint32_t n, uint32_t in,
                              uint32_t hidden, uint32_t out, float mask_sum,
                              const float *d_params, float *d_dy, float *d_dh, float *d_grad,
                              hgap_stream_t stream);
/* One Adam update (step counts from 1) of `count` parameters with moments d_m, d_v. */
hgap_status hgap_mlp_adam_step(float *d_params, const float *d_grad, float *d_m, float *d_v,
                               uint32_t count, uint32_t step, float lr, float beta1, float beta2,
                               float eps, hgap_stream_t stream);

/* Synthetic data for tests and benchmarks: byte i = perm[rank], rank = first r with
 * (splitmix64(seed + (start+i+1)*0x9E3779B97F4A7C15) >> 11) < cdf[r]  (cdf[255] = 2^53). */
hgap_status hgap_synth_fill(void *d_out, uint64_t n, uint64_t start, uint64_t seed,
                            const uint64_t cdf[256], const uint8_t perm[256],
                            hgap_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* HGAP_H */
