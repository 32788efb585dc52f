// hgap_internal.h -- constants shared by the host code and the HIP kernels.
#ifndef HGAP_INTERNAL_H
#define HGAP_INTERNAL_H

#include <cstdint>

namespace hgap {

// First-level width of the decode table when the caller does not pick one: 2^11 entries
// (8 KiB of LDS) cover every code of length <= 11 in a single lookup.
constexpr uint32_t kDefaultPrefixBits = 11;

// Default launch shapes (override per call with hgap_launch_config).  Encoder tile = threads x
// bytes per thread; decoder tile = threads segments of 128 bits.  Bigger tiles mean fewer
// look-back participants; the look-back window is the number of predecessor status words read
// per round.
// (tuned on the 1 GiB Zipf(1.1) buffer with huffman_coding_gap_arrays_amd/tuner.py)
constexpr uint32_t kDefaultEncThreads = 512;
constexpr uint32_t kDefaultEncBytesPerThread = 64;
// The chained encoder's default shape (hgap_encode_ranges without one): 8 KiB tiles, 32 per
// range.  On the 1 GiB Zipf(1.1) buffer 0.57 ms, against 0.62-0.63 ms for 512 x 64, 512 x 32
// and 256 x 64.
constexpr uint32_t kDefaultChainedThreads = 256;
constexpr uint32_t kDefaultChainedBytesPerThread = 32;
constexpr uint32_t kDefaultDecThreads = 256;
constexpr uint32_t kDefaultLookbackWindow = 64;
// smallest tiles any configuration uses (they size the look-back workspaces)
constexpr uint32_t kMinEncTileBytes = 256 * 16;
constexpr uint32_t kMinDecThreads = 256;

// Decode table image: first level, second level, then this many words holding one byte per
// symbol value: that symbol's code length (0 for bytes absent from the codebook).
constexpr uint32_t kDecLengthWords = 64;

// Histogram: 256 threads, one private bin array per lane index (64 copies of 256 bins).
constexpr int kHistThreads = 256;

// Code lengths for counts w[0..k-1] sorted ascending (boundary package-merge with the
// reference's tie rules); defined in hgap_host.cpp.
void package_merge_lengths(const uint64_t *w, int k, uint8_t *len);

}  // namespace hgap

#endif
