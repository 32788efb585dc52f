/*
 * hgap_oracle.h -- CPU restatement of the reference Huffman-with-gap-arrays path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (huffman_coding_gap_arrays_amd/) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).
 *
 * Parity pins (what checks each function; see tests/test_oracle_cpu.py):
 *   orc_histogram         closed form (numpy bincount)                         pinned
 *   orc_sort_symbols      rule (stable count-ascending, byte-ascending ties)     pinned by rule +
 *                                                                               golden vectors
 *   orc_boundary_pm       golden vectors produced by the compiled reference
 *                         (oracle/_ref), plus Kraft/optimality/monotonicity     pinned
 *   orc_assign_codes      golden vectors from the compiled reference + prefix-free check  pinned
 *   orc_sizes             closed form                                           pinned
 *   orc_encode            closed form: concatenation of codewords, gap definition,
 *                         brute-force bit writer in numpy                        pinned
 *   orc_decode_segments   round trip to the original input                     pinned
 *   orc_stream_write/read byte layout checked field by field                     pinned
 *   orc_synth_fill        numpy restatement of the counter-based generator       pinned
 *   orc_*_mt              word-for-word equal to the serial functions            pinned
 */
#ifndef HGAP_ORACLE_H
#define HGAP_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEN 16      /* MAX_CODEWORD_LENGTH, encoder/include/constants.hpp:5 */
#define ORC_SEGMENT 128     /* SEGMENTSIZE, encoder/include/constants.hpp:12 */
#define ORC_GAP_PER_WORD 8  /* GAP_ELEMENTS_NUM, encoder/include/constants.hpp:10 */

/* flags reported by orc_boundary_pm: situations where the reference reads memory it never
 * wrote (package_merge.cpp:56-60 with the leaf list exhausted) or writes past its per-level
 * node list (package_merge.cpp:113).  The restatement treats the unwritten weight as 0. */
#define ORC_PM_UNINIT_LEFT 1
#define ORC_PM_LIST_OVERFLOW 2

void orc_histogram(const uint8_t *in, uint64_t n, uint64_t hist[256]);

int orc_sort_symbols(const uint64_t hist[256], uint8_t sym[256], uint64_t cnt[256]);

int orc_boundary_pm(const uint64_t *cnt, int k, uint8_t *len, int *flags);

void orc_assign_codes(const uint8_t *sym, const uint8_t *len, int k,
                      uint32_t code_by_byte[256], uint32_t len_by_byte[256],
                      uint8_t *order_sym, uint8_t *order_len);

int orc_build_codebook(const uint64_t hist[256], uint32_t code_by_byte[256],
                       uint32_t len_by_byte[256], uint8_t order_sym[256],
                       uint8_t order_len[256], int *flags);

void orc_sizes(const uint64_t hist[256], const uint32_t len_by_byte[256], uint64_t *bits,
               uint64_t *payload_words, uint64_t *gap_elements, uint64_t *gap_words);

int orc_encode(const uint8_t *in, uint64_t n, const uint32_t code_by_byte[256],
               const uint32_t len_by_byte[256], uint32_t *payload, uint64_t payload_words,
               uint32_t *gap, uint64_t gap_elements);

int64_t orc_decode_segments(const uint32_t *payload, uint64_t payload_words, const uint32_t *gap,
                            uint64_t gap_elements, const uint8_t *order_sym,
                            const uint8_t *order_len, int k, uint8_t *out, uint64_t n,
                            uint32_t *seg_counts);

int64_t orc_stream_size(int k, uint64_t payload_words, uint64_t gap_elements);
int64_t orc_stream_write(uint8_t *dst, uint64_t cap, int k, const uint8_t *order_sym,
                         const uint8_t *order_len, uint32_t n, uint32_t payload_words,
                         uint32_t gap_elements, const uint32_t *gap, const uint32_t *payload);
int orc_stream_parse(const uint8_t *src, uint64_t size, int *k, uint8_t *order_sym,
                     uint8_t *order_len, uint32_t *n, uint32_t *payload_words,
                     uint32_t *gap_elements, uint64_t *gap_offset, uint64_t *payload_offset);

/* All-cores (OpenMP) variants with outputs identical to the serial functions above
 * (hgap_oracle_mt.c); threads <= 0 means every available core. */
void orc_histogram_mt(const uint8_t *in, uint64_t n, uint64_t hist[256], int threads);
int orc_encode_mt(const uint8_t *in, uint64_t n, const uint32_t code_by_byte[256],
                  const uint32_t len_by_byte[256], uint32_t *payload, uint64_t payload_words,
                  uint32_t *gap, uint64_t gap_elements, int threads);
int64_t orc_decode_segments_mt(const uint32_t *payload, uint64_t payload_words,
                               const uint32_t *gap, uint64_t gap_elements,
                               const uint8_t *order_sym, const uint8_t *order_len, int k,
                               uint8_t *out, uint64_t n, int threads);

uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i);
void orc_synth_fill(uint8_t *out, uint64_t n, uint64_t start, uint64_t seed,
                    const uint64_t cdf[256], const uint8_t perm[256]);

#ifdef __cplusplus
}
#endif
#endif
