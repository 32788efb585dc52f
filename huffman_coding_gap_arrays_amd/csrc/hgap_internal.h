// hgap_internal.h -- constants shared by the host code and the HIP kernels.
#ifndef HGAP_INTERNAL_H
#define HGAP_INTERNAL_H

#include <cstdint>

namespace hgap {

// First-level width of the decode table when the caller does not pick one: 2^11 entries
// (8 KiB of LDS) cover every code of length <= 11 in a single lookup.
constexpr uint32_t kDefaultPrefixBits = 11;

// Encoder tiling: 256 threads x 32 input bytes = 8 KiB of input per tile.
constexpr int kEncThreads = 256;
constexpr int kEncBytesPerThread = 32;
constexpr int kEncTileBytes = kEncThreads * kEncBytesPerThread;

// Decoder tiling: 256 threads, one 128-bit segment per thread.
constexpr int kDecThreads = 256;

// Histogram: 256 threads, one private bin array per lane index (64 copies of 256 bins).
constexpr int kHistThreads = 256;

// Code lengths for counts w[0..k-1] sorted ascending (boundary package-merge with the
// reference's tie rules); defined in hgap_host.cpp.
void package_merge_lengths(const uint64_t *w, int k, uint8_t *len);

}  // namespace hgap

#endif
