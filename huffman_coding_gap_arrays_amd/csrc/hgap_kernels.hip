// hgap_kernels.hip -- gfx950 kernels and launchers of the Huffman-with-gap-arrays path.
//
//   hist_kernel     byte histogram; one private 256-bin array per lane index in LDS, merged once
//                   per workgroup into the global histogram
//   hist_range_kernel
//                   the same, plus the counts of every 256 KiB range of the input
//   enc_kernel      persistent encoder: per-thread code lengths, workgroup scan, single-pass
//                   decoupled look-back across tiles (64-bit status words in HBM), bit-packing
//                   into an LDS tile with gap nibbles emitted on the fly, 128-bit stores.
//                   Chained form (after enc_range_prep_kernel): a workgroup encodes a range's
//                   tiles in order from the range's offset, which the range counts give in
//                   advance, so no tile waits on another workgroup
//   dec_count_kernel, dec_offsets_kernel, dec_write_kernel
//                   decoder in three launches: per-segment codeword counts (one lane per
//                   128-bit segment from its gap offset, two-level table in LDS), per-tile output
//                   offsets, then a second decode of each segment into an LDS image of the
//                   tile's output and 128-bit stores
//   hist_batch_kernel, hist_range_batch_kernel, enc_batch_kernel, enc_range_batch_kernel,
//   dec_count_batch_kernel, dec_write_batch_kernel
//                   the same steps over many buffers (each with its own codebook or table) in
//                   one launch per step
//
// Reference citations are file:line relative to the reference repository root.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <type_traits>

#include "hgap.h"
#include "hgap_internal.h"

namespace hgap {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------------
// wave64 helpers

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Inclusive prefix sum over the 64 lanes of a wave: four row_shr steps build sums inside each
// 16-lane row, row_bcast:15 / row_bcast:31 carry row totals into the rows above.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x) {
    x += __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Block-wide exclusive scan over THREADS threads (THREADS / 64 waves); returns the exclusive
// prefix and the block total.  `s_wave` holds one word per wave.  Contains one barrier.
template <int THREADS>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t x, uint32_t *s_wave,
                                                   uint32_t &total) {
    constexpr int W = THREADS / 64;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_sum(x);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t v = s_wave[i];
        tot += v;
        before += (uint32_t)i < wave ? v : 0u;
    }
    total = tot;
    return before + incl - x;
}

// ---------------------------------------------------------------------------------------------
// Decoupled look-back (single pass) over per-tile 64-bit status words in HBM.
// Status word: 0 = nothing published; kFlagA | aggregate; kFlagP | inclusive prefix.  Flag and
// value travel in one 64-bit word, so relaxed agent-scope atomics are enough.  Tile ids come
// from an atomic counter, so every predecessor is resident or finished (forward progress).
// Each round reads a window of 64 * wpl predecessors (wpl <= 4 per lane: lane l reads tiles
// end - wpl*l - j, j < wpl) and stops at the nearest inclusive prefix.  A wide window lets
// the prefix frontier keep up when many tiles are in flight: the tile rate a look-back can
// sustain is about (window / round-trip latency).
// Called by all 64 lanes of one wave; returns the exclusive prefix of `tile`.
constexpr uint64_t kFlagA = 1ull << 62;
constexpr uint64_t kFlagP = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;
constexpr uint32_t kMaxWindowPerLane = 4;

__device__ __forceinline__ uint64_t status_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// `published`: the tile's aggregate is already in its status word (stored by an earlier launch).
__device__ uint64_t look_back(uint64_t *status, uint32_t tile, uint64_t aggregate,
                              uint32_t wpl, bool published = false) {
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) status_store(&status[0], kFlagP | aggregate);
        return 0;
    }
    if (lane == 0 && !published) status_store(&status[tile], kFlagA | aggregate);
    uint64_t excl = 0;
    int64_t end = (int64_t)tile - 1;
    while (true) {
        const int64_t first = end - (int64_t)(lane * wpl);
        uint64_t v[kMaxWindowPerLane];
#pragma unroll
        for (uint32_t j = 0; j < kMaxWindowPerLane; j++) {
            const int64_t idx = first - (int64_t)j;
            // outside the window: an empty aggregate; before tile 0: an empty inclusive prefix
            v[j] = j >= wpl ? kFlagA : (idx >= 0 ? status_load(&status[idx]) : kFlagP);
        }
        while (true) {
            bool missing = false;
#pragma unroll
            for (uint32_t j = 0; j < kMaxWindowPerLane; j++) missing |= v[j] == 0;
            if (__ballot(missing) == 0) break;
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (uint32_t j = 0; j < kMaxWindowPerLane; j++)
                if (v[j] == 0) v[j] = status_load(&status[first - (int64_t)j]);
        }
        // this lane's sum up to and including its nearest inclusive prefix (or all of it)
        uint64_t sum = 0;
        bool seen_p = false;
#pragma unroll
        for (uint32_t j = 0; j < kMaxWindowPerLane; j++) {
            if (!seen_p) sum += v[j] & kValueMask;
            seen_p |= (v[j] & kFlagP) != 0;
        }
        const uint64_t pmask = __ballot(seen_p);
        if (pmask != 0) {
            const uint32_t fl = (uint32_t)__builtin_ctzll(pmask);  // lane of nearest prefix
            excl += wave_sum_u64(lane <= fl ? sum : 0ull);
            break;
        }
        excl += wave_sum_u64(sum);
        end -= 64 * (int64_t)wpl;
    }
    if (lane == 0) status_store(&status[tile], kFlagP | (excl + aggregate));
    return excl;
}

// ---------------------------------------------------------------------------------------------
// Where an encoder tile claims the workgroup's next tile: after its own look-back, so tiles are
// claimed in the order they finish looking back (a tile never waits on a predecessor that has
// not started), and the next tile's input load overlaps this tile's packing and stores.
struct NextTile {
    uint32_t *counter;  // ticket counter; nullptr = the caller claims tiles itself
    uint32_t *s_ticket; // LDS word broadcasting the ticket
    uint32_t tile;      // out: the claimed ticket
};

// zero fill (the histogram, the gap array and the look-back workspaces)
__global__ void fill_zero_kernel(uint32_t *a, uint64_t na, uint32_t *b, uint64_t nb) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += stride) {
        if (i < na)
            a[i] = 0u;
        else
            b[i - na] = 0u;
    }
}

// ---------------------------------------------------------------------------------------------
// Histogram (replaces cu_histgram + mergeHistogram, encoder/src/encoder.cu:33-116).
// Every lane index owns one 256-bin array in LDS (shared by the waves of the workgroup, so
// updates are LDS atomics, but lanes of one wave never collide on an address).  Bin b of lane
// l's copy sits at word 64 b + l: whatever bytes the lanes count, they hit distinct banks.
// The single-buffer kernel runs 1024-lane workgroups (two per CU next to 64 KiB of bins) so
// that enough 16-byte loads are in flight to cover HBM latency.
constexpr int kHistCopies = 64;
constexpr int kHistThreadsSingle = 1024;

__global__ __launch_bounds__(kHistThreadsSingle) void hist_kernel(const uint8_t *__restrict__ in,
                                                                  uint64_t n,
                                                                  uint32_t *__restrict__ hist) {
    constexpr int kHistThreads = kHistThreadsSingle;
    __shared__ uint32_t s_bins[kHistCopies * 256];
    for (int i = threadIdx.x; i < kHistCopies * 256; i += kHistThreads) s_bins[i] = 0u;
    __syncthreads();
    // bin b of lane l's copy is word 64 b + l: within each 32-lane LDS pass every lane hits
    // its own bank, whatever bytes the lanes count
    uint32_t *mine = s_bins + lane_id();
    const uint64_t vecs = n >> 4;
    const u32x4 *in4 = reinterpret_cast<const u32x4 *>(in);
    const uint64_t stride = (uint64_t)gridDim.x * kHistThreads;
    uint64_t v = (uint64_t)blockIdx.x * kHistThreads + threadIdx.x;
    // 4 independent 16-byte loads in flight per thread
    for (; v + 3 * stride < vecs; v += 4 * stride) {
        u32x4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = __builtin_nontemporal_load(&in4[v + u * stride]);
#pragma unroll
        for (int u = 0; u < 4; u++) {
#pragma unroll
            for (int j = 0; j < 16; j++)
                atomicAdd(&mine[((q[u][j >> 2] >> (8 * (j & 3))) & 0xFFu) << 6], 1u);
        }
    }
    for (; v < vecs; v += stride) {
        const u32x4 q = __builtin_nontemporal_load(&in4[v]);
#pragma unroll
        for (int j = 0; j < 16; j++)
            atomicAdd(&mine[((q[j >> 2] >> (8 * (j & 3))) & 0xFFu) << 6], 1u);
    }
    // tail bytes (n % 16) in the first workgroup
    if (blockIdx.x == 0) {
        const uint64_t t = (vecs << 4) + threadIdx.x;
        if (t < n) atomicAdd(&mine[(uint32_t)in[t] << 6], 1u);
    }
    __syncthreads();
    // thread b < 256 owns bin b: sum of the 64 copies (read in a rotated order so that the
    // lanes of a pass hit different banks), then one global atomic per bin per workgroup
    if (threadIdx.x < 256) {
        uint32_t sum = 0;
        for (int c = 0; c < kHistCopies; c++)
            sum += s_bins[(threadIdx.x << 6) + ((c + threadIdx.x) & 63u)];
        if (sum) atomicAdd(&hist[threadIdx.x], sum);
    }
}

// The same histogram, plus one histogram per range of `range_bytes` consecutive input bytes
// (range r = bytes [r * range_bytes, (r + 1) * range_bytes), the last one ragged).  A workgroup
// counts whole ranges (ranges blockIdx.x, blockIdx.x + gridDim.x, ...); after each one,
// thread b < 256 sums bin b over the 64 copies, clears them as it reads, and stores the range's
// count; it keeps a running total of bin b for the one global atomic at the end.
// Counts bytes [lo, hi) of `in` (lo a multiple of 16) into the lane copies `mine`: 16-byte
// loads, four in flight per thread, then the last hi % 16 bytes one per thread.
template <int THREADS>
__device__ __forceinline__ void hist_count_range(const uint8_t *in, uint64_t lo, uint64_t hi,
                                                 uint32_t *mine) {
    const u32x4 *in4 = reinterpret_cast<const u32x4 *>(in);
    const uint64_t v1 = hi >> 4;
    uint64_t v = (lo >> 4) + threadIdx.x;
    for (; v + 3 * THREADS < v1; v += 4 * THREADS) {
        u32x4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = __builtin_nontemporal_load(&in4[v + u * THREADS]);
#pragma unroll
        for (int u = 0; u < 4; u++) {
#pragma unroll
            for (int j = 0; j < 16; j++)
                atomicAdd(&mine[((q[u][j >> 2] >> (8 * (j & 3))) & 0xFFu) << 6], 1u);
        }
    }
    for (; v < v1; v += THREADS) {
        const u32x4 q = __builtin_nontemporal_load(&in4[v]);
#pragma unroll
        for (int j = 0; j < 16; j++)
            atomicAdd(&mine[((q[j >> 2] >> (8 * (j & 3))) & 0xFFu) << 6], 1u);
    }
    const uint64_t t = (v1 << 4) + threadIdx.x;
    if (t < hi) atomicAdd(&mine[(uint32_t)in[t] << 6], 1u);
}

// After a barrier, thread b < 256: the sum of bin b over the 64 copies, which it clears (read
// in a rotated order so that the lanes of a pass hit different banks).
__device__ __forceinline__ uint32_t hist_take_bin(uint32_t *s_bins) {
    uint32_t sum = 0;
    for (int c = 0; c < kHistCopies; c++) {
        const uint32_t i = (threadIdx.x << 6) + ((c + threadIdx.x) & 63u);
        sum += s_bins[i];
        s_bins[i] = 0u;
    }
    return sum;
}

__global__ __launch_bounds__(kHistThreadsSingle) void hist_range_kernel(
    const uint8_t *__restrict__ in, uint64_t n, uint32_t *__restrict__ hist,
    uint32_t *__restrict__ range_hist, uint32_t ranges, uint32_t range_bytes) {
    constexpr int kThreads = kHistThreadsSingle;
    __shared__ uint32_t s_bins[kHistCopies * 256];
    for (int i = threadIdx.x; i < kHistCopies * 256; i += kThreads) s_bins[i] = 0u;
    __syncthreads();
    uint32_t *mine = s_bins + lane_id();
    uint32_t total = 0;
    for (uint32_t r = blockIdx.x; r < ranges; r += gridDim.x) {
        const uint64_t lo = (uint64_t)r * range_bytes;
        hist_count_range<kThreads>(in, lo, lo + range_bytes < n ? lo + range_bytes : n, mine);
        __syncthreads();
        if (threadIdx.x < 256) {
            const uint32_t sum = hist_take_bin(s_bins);
            range_hist[256ull * r + threadIdx.x] = sum;
            total += sum;
        }
        __syncthreads();
    }
    if (threadIdx.x < 256 && total) atomicAdd(&hist[threadIdx.x], total);
}

// ---------------------------------------------------------------------------------------------
// Encoder (replaces cuencoder + cu_get_gaparray, encoder/src/encoder.cu:142-379).
//
// Tile = THREADS threads x K consecutive input bytes.  Per tile:
//   1. each thread looks up its K codes (code << 5 | length) in an LDS copy of the codebook
//      and sums the lengths (its input bytes were loaded into registers while the previous
//      tile was processed);
//   2. workgroup scan of the bit counts, then decoupled look-back for the tile's global bit
//      offset E;
//   3. each thread packs its codes (two at a time) MSB-first into 32-bit words of a zeroed LDS
//      tile image whose word 0 is global word (E/32) rounded down to a multiple of 4, with LDS
//      atomic OR (a thread's first and last words may be shared with its neighbours).  When a
//      completed word closes a 128-bit segment, the bits of the straddling code past the
//      boundary (< 16) are that segment's gap nibble (encoder.cu:307-312), stored as one byte
//      per segment in LDS by the thread that completes the word;
//   4. the tile owns global words [E/32, (E+T)/32) (the last tile also the final partial word);
//      its first word also holds the last E%32 bits of the previous tile, which are rebuilt
//      here from the previous tile's last input bytes instead of waiting for that tile;
//      owned words are written with 128-bit stores;
//   5. spill bytes are packed eight to a gap word (nibble i = segment 8g + i); gap words whose
//      eight segment boundaries all lie inside the tile are stored; the (at most two) gap words
//      shared with neighbouring tiles are OR-ed into the pre-zeroed array.
struct EncTable {
    uint32_t e[256];  // (code << 5) | length, by byte value
};

struct EncBuffer {
    const uint8_t *in;
    uint64_t n;
    uint32_t *payload;
    uint32_t payload_words;
    uint32_t gap_words;
    uint32_t *gap;
    uint64_t *status;  // one status word per tile of this buffer
    uint32_t num_tiles;
    uint32_t reserved;
};

// LDS words of the output image for a tile of `tile_bytes` inputs whose codes are at most
// `max_len` bits, and of its gap-word image
__host__ __device__ constexpr uint32_t enc_out_words(uint32_t tile_bytes, uint32_t max_len) {
    return (tile_bytes * max_len + 31) / 32 + 8;
}
__host__ __device__ constexpr uint32_t enc_gap_words(uint32_t tile_bytes, uint32_t max_len) {
    // one spill byte per segment the tile's bits touch, offset by up to 7 bytes
    return (tile_bytes * max_len) / 512 + 8;
}

// A thread's K input bytes of a tile, loaded into registers one tile ahead so their HBM latency
// overlaps the previous tile's work; nb = how many of them lie inside the buffer.
template <int K>
struct EncStage {
    u32x4 q[K / 16];
    int nb;
};

template <int THREADS, int K>
__device__ __forceinline__ EncStage<K> enc_fetch(const EncBuffer &buf, uint32_t tile) {
    EncStage<K> st;
#pragma unroll
    for (int v = 0; v < K / 16; v++) st.q[v] = u32x4{0u, 0u, 0u, 0u};
    st.nb = 0;
    if (tile >= buf.num_tiles) return st;
    const uint64_t mine = (uint64_t)tile * (THREADS * K) + (uint64_t)threadIdx.x * K;
    if (mine >= buf.n) return st;
    st.nb = buf.n - mine >= (uint64_t)K ? K : (int)(buf.n - mine);
    if (st.nb == K) {
        const u32x4 *p = reinterpret_cast<const u32x4 *>(buf.in + mine);
#pragma unroll
        for (int v = 0; v < K / 16; v++) st.q[v] = __builtin_nontemporal_load(&p[v]);
    } else {
#pragma unroll
        for (int j = 0; j < K; j++)
            if (j < st.nb)
                st.q[j >> 4][(j >> 2) & 3] |= (uint32_t)buf.in[mine + j] << (8 * (j & 3));
    }
    return st;
}

// byte j of a thread's staged input
template <int K>
__device__ __forceinline__ uint32_t enc_byte(const EncStage<K> &st, int j) {
    return (st.q[j >> 4][(j >> 2) & 3] >> (8 * (j & 3))) & 0xFFu;
}

// How a chained encoder finds E.  The input is cut into ranges of whole tiles; a workgroup
// encodes a range's tiles in order.  The first tile of a range looks back over per-range status
// words whose aggregates (the ranges' bit counts, from per-range histograms) were all stored
// before the launch, so it never waits on another workgroup; it also claims the workgroup's
// next range.  A later tile's E is the previous tile's E plus its bit count, and the previous
// tile's last partial word (`carry`) supplies the bits of E's word before E, so the tile does
// not rebuild them from input bytes.
struct EncChain {
    uint32_t range;            // range of the current tile
    uint32_t next_range;       // claimed by the current range's first tile
    uint32_t ranges;
    uint32_t tiles_per_range;
    uint32_t end;              // one past the current range's last tile
    bool first;                // the current tile is its range's first
    uint64_t E;                // in: E of the current tile (when !first); out: the next tile's
    uint32_t carry;            // in: bits of E's word before E (when !first); out: the next's
};

// PREFETCH: load the next tile's input bytes into `ahead` once the next tile is known.
template <int THREADS, int K, bool CHAINED, bool PREFETCH>
__device__ __forceinline__ void enc_tile(const EncBuffer &buf, uint32_t tile,
                                         const EncStage<K> &st, const uint32_t *s_tab,
                                         uint32_t *s_out, uint32_t *s_gap, uint32_t *s_wave,
                                         uint64_t *s_excl, uint32_t wpl, NextTile &next,
                                         EncChain &ch, EncStage<K> &ahead) {
    constexpr uint32_t TILE = THREADS * K;
    const uint32_t lane = lane_id();
    const uint8_t *in = buf.in;
    const int nb = st.nb;

    // 1. codes and bit count; entries past the buffer's end are 0 (length 0: they append
    // nothing)
    const uint64_t tile_start = (uint64_t)tile * TILE;
    uint32_t e[K];
    uint32_t bits = 0;
    if (nb == K) {
#pragma unroll
        for (int j = 0; j < K; j++) {
            e[j] = s_tab[enc_byte<K>(st, j)];
            bits += e[j] & 31u;
        }
    } else {
#pragma unroll
        for (int j = 0; j < K; j++) {
            e[j] = j < nb ? s_tab[enc_byte<K>(st, j)] : 0u;
            bits += e[j] & 31u;
        }
    }

    // 2. tile scan, zero the used part of the LDS images (a chained tile that knows E puts the
    // carried bits into its first word instead), look-back
    uint32_t tile_bits;
    const uint32_t excl = block_excl_sum<THREADS>(bits, s_wave, tile_bits);
    const uint32_t used_words = (tile_bits >> 5) + 8, used_gap = (tile_bits >> 9) + 5;
    const bool known = CHAINED && !ch.first;
    const uint32_t i0 = known ? (uint32_t)((ch.E >> 5) & 3u) : 0xFFFFFFFFu;
    const uint32_t carry = known ? ch.carry : 0u;
    for (uint32_t i = threadIdx.x; i < used_words; i += THREADS) s_out[i] = i == i0 ? carry : 0u;
    for (uint32_t i = threadIdx.x; i < used_gap; i += THREADS) s_gap[i] = 0u;
    if (!known && threadIdx.x < 64) {
        uint64_t E;
        if (CHAINED) {
            const uint64_t agg = status_load(&buf.status[ch.range]) & kValueMask;
            E = look_back(buf.status, ch.range, agg, wpl, true);
        } else {
            E = look_back(buf.status, tile, tile_bits, wpl);
        }
        if (threadIdx.x == 0) {
            *s_excl = E;
            if (next.counter) *next.s_ticket = atomicAdd(next.counter, 1u);
        }
    }
    __syncthreads();
    const uint64_t E = known ? ch.E : *s_excl;
    if (CHAINED) {
        if (ch.first) ch.next_range = *next.s_ticket;
        next.tile = tile + 1 < ch.end ? tile + 1
                    : (ch.next_range < ch.ranges ? ch.next_range * ch.tiles_per_range
                                                 : buf.num_tiles);
    } else if (next.counter) {
        next.tile = *next.s_ticket;
    }
    // the next tile's input bytes into registers now, so that their HBM latency overlaps this
    // tile's packing and stores.  (Copying them HBM->LDS with global_load_lds instead made the
    // compiler wait for the whole copy before this tile's first LDS access.)
    if (PREFETCH) ahead = enc_fetch<THREADS, K>(buf, next.tile);
    const uint64_t base_w = (E >> 5) & ~3ull;  // global word of s_out[0], 4 x global segment E/128
    const uint64_t gap_base = E >> 10;         // global gap word of bytes 0..7 of s_gap
    // one spill byte per segment: byte (E/128 mod 8) + i of s_gap is segment E/128 + i
    uint8_t *s_spill = reinterpret_cast<uint8_t *>(s_gap) + ((E >> 7) & 7u);

    // 3. pack, two codes at a time: a pair is at most 32 bits, so appending it to at most 31
    // pending bits completes at most one word
    if (nb > 0) {
        const uint64_t B = E + excl;
        uint32_t w = (uint32_t)((B >> 5) - base_w);
        int nacc = (int)(B & 31);
        uint64_t acc = 0;
#pragma unroll
        for (int j = 0; j < K; j += 2) {
            const uint32_t l1 = e[j] & 31u, l2 = e[j + 1] & 31u;
            const uint32_t pair = ((e[j] >> 5) << l2) | (e[j + 1] >> 5);
            acc = (acc << (l1 + l2)) | pair;
            nacc += (int)(l1 + l2);
            if (nacc >= 32) {
                nacc -= 32;
                // the zeroed image takes an OR for every word: only a thread's first word is
                // shared with its predecessor, and OR needs no branch on which one it is
                atomicOr(&s_out[w], (uint32_t)(acc >> nacc));
                if ((w & 3u) == 3u) {
                    // word w closes a segment (its thread is the one holding the segment's last
                    // bit); nacc bits lie past the boundary: all of them belong to the second
                    // code if it straddles, else the first code straddles by nacc - l2 (0 when
                    // it ends exactly on the boundary)
                    s_spill[w >> 2] =
                        (uint8_t)((uint32_t)nacc >= l2 ? (uint32_t)nacc - l2 : (uint32_t)nacc);
                }
                w++;
            }
        }
        if (nacc > 0) atomicOr(&s_out[w], (uint32_t)(acc << (32 - nacc)));
    }
    __syncthreads();

    // 4a. first word: add the previous tile's last E%32 bits (a chained tile that knows E has
    // them already)
    const uint32_t r = (uint32_t)(E & 31);
    if (!known && r != 0) {
        if (threadIdx.x < 64) {
            uint32_t l = 0, c = 0;
            if (lane < 32) {
                const uint32_t ee = s_tab[in[tile_start - 1 - lane]];
                l = ee & 31u;
                c = ee >> 5;
            }
            const uint32_t cum = wave_incl_sum(l);
            const uint32_t after = cum - l;  // bits between this code's end and the tile start
            const uint32_t part = (lane < 32 && after < r) ? (c << (32 - (r - after))) : 0u;
            const uint32_t prev_bits = wave_or_u32(part);
            if (threadIdx.x == 0) s_out[(E >> 5) - base_w] |= prev_bits;
        }
        __syncthreads();
    }
    const uint64_t T = tile_bits;
    if (CHAINED) {  // the word holding the tile's end is complete up to the end: carry it
        const uint64_t end_bit = E + T;
        ch.carry = (end_bit & 31u) ? s_out[(end_bit >> 5) - base_w] : 0u;
        ch.E = end_bit;
    }

    // 4b. payload words
    const uint64_t own_lo = E >> 5;
    uint64_t own_hi = (tile + 1 == buf.num_tiles) ? ((E + T + 31) >> 5) : ((E + T) >> 5);
    if (own_hi > buf.payload_words) own_hi = buf.payload_words;
    if (own_hi > own_lo) {
        const uint32_t chunks = (uint32_t)((own_hi - base_w + 3) >> 2);
        for (uint32_t c = threadIdx.x; c < chunks; c += THREADS) {
            const uint64_t g = base_w + 4ull * c;
            if (g >= own_lo && g + 4 <= own_hi) {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(&s_out[4 * c]);
                __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(buf.payload + g));
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (g + i >= own_lo && g + i < own_hi) buf.payload[g + i] = s_out[4 * c + i];
            }
        }
    }

    // 5. gap words: eight spill bytes -> eight nibbles, low nibble first
    if (T > 0) {
        const uint64_t g_hi = (E + T - 1) >> 10;
        for (uint64_t g = gap_base + threadIdx.x; g <= g_hi; g += THREADS) {
            if (g >= buf.gap_words) break;
            const uint32_t lo = s_gap[2 * (g - gap_base)], hi = s_gap[2 * (g - gap_base) + 1];
            // bytes b0..b3 (each < 16) -> (b0 | b1 << 4) in byte 0 and (b2 | b3 << 4) in byte 2
            const uint32_t tl = lo | (lo >> 4), th = hi | (hi >> 4);
            const uint32_t val = __builtin_amdgcn_perm(th, tl, 0x06040200u);
            const bool inside = (g * 1024 + 128 > E) && ((g + 1) * 1024 <= E + T);
            if (inside)
                buf.gap[g] = val;
            else if (val)
                atomicOr(&buf.gap[g], val);
        }
    }
    __syncthreads();
}

// Waves per SIMD enc_kernel asks the compiler for.  Shapes of up to 512 threads x 32 bytes fit
// in 80 VGPRs (512 x 32 with a 12-byte spill), so six waves share a SIMD: the chained 256 x 32
// encoder ran 8% faster than at its natural 88 VGPRs (five waves), 512 x 32 16% faster.
// 64-byte shapes would spill heavily and keep the compiler's choice.  (The batched encoders
// already fit in 80 VGPRs at 256 x 32.)
template <int THREADS, int K>
constexpr int enc_waves_per_eu() {
    return THREADS <= 512 && K <= 32 ? 6 : 1;
}

// The next tile's input is loaded into registers one tile ahead (claimed right after the
// current tile's look-back).  CHAINED: the counter hands out ranges of `tiles_per_range` tiles
// (buf.status has one word per range, prepared by enc_range_prep_kernel); otherwise single
// tiles.
template <int THREADS, int K, bool CHAINED>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(enc_waves_per_eu<THREADS, K>())))
void enc_kernel(EncBuffer buf, EncTable table, uint32_t *__restrict__ counter,
                uint32_t out_words, uint32_t wpl, uint32_t tiles_per_range) {
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    __shared__ uint32_t s_tab[256];
    __shared__ uint32_t s_wave[THREADS / 64];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_tile;
    uint32_t *s_out = s_dyn;
    uint32_t *s_gap = s_dyn + out_words;
    for (int i = threadIdx.x; i < 256; i += THREADS) s_tab[i] = table.e[i];
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    EncChain ch{};
    uint32_t first_tile = s_tile;
    if (CHAINED) {
        ch.ranges = (buf.num_tiles + tiles_per_range - 1) / tiles_per_range;
        ch.tiles_per_range = tiles_per_range;
        ch.next_range = s_tile;
        first_tile = s_tile < ch.ranges ? s_tile * tiles_per_range : buf.num_tiles;
    }
    NextTile next{counter, &s_tile, first_tile};
    EncStage<K> ahead = enc_fetch<THREADS, K>(buf, next.tile);
    if (CHAINED) {
        while (ch.next_range < ch.ranges) {
            ch.range = ch.next_range;
            const uint32_t lo = ch.range * tiles_per_range;
            ch.end = lo + tiles_per_range < buf.num_tiles ? lo + tiles_per_range : buf.num_tiles;
            for (uint32_t cur = lo; cur < ch.end; cur++) {
                ch.first = cur == lo;
                const EncStage<K> st = ahead;
                enc_tile<THREADS, K, true, true>(buf, cur, st, s_tab, s_out, s_gap, s_wave,
                                                 &s_excl, wpl, next, ch, ahead);
            }
        }
        return;
    }
    while (next.tile < buf.num_tiles) {
        const uint32_t cur = next.tile;
        const EncStage<K> st = ahead;
        enc_tile<THREADS, K, false, true>(buf, cur, st, s_tab, s_out, s_gap, s_wave, &s_excl,
                                          wpl, next, ch, ahead);
    }
}

// Before a chained encode: zero the gap array and the range counter, and store every range's
// bit count (its histogram dotted with the code lengths) in its status word as a look-back
// aggregate (range 0's as its inclusive prefix).  One wave per range; lane l reads bins
// 4l..4l+3.
__global__ __launch_bounds__(256) void enc_range_prep_kernel(
    const uint32_t *__restrict__ range_hist, uint32_t ranges, EncTable table,
    uint64_t *__restrict__ status, uint32_t *__restrict__ counter, uint32_t *__restrict__ gap,
    uint32_t gap_words) {
    const uint32_t lane = lane_id();
    uint32_t len[4];
#pragma unroll
    for (int j = 0; j < 4; j++) len[j] = table.e[4 * lane + j] & 31u;
    const uint32_t waves = gridDim.x * 4u;
    for (uint32_t r = blockIdx.x * 4u + (threadIdx.x >> 6); r < ranges; r += waves) {
        const u32x4 h = *reinterpret_cast<const u32x4 *>(range_hist + 256ull * r + 4 * lane);
        const uint64_t bits = wave_sum_u64((uint64_t)h[0] * len[0] + (uint64_t)h[1] * len[1] +
                                           (uint64_t)h[2] * len[2] + (uint64_t)h[3] * len[3]);
        if (lane == 0) status_store(&status[r], (r == 0 ? kFlagP : kFlagA) | (bits & kValueMask));
    }
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < gap_words; i += stride) gap[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

// ---------------------------------------------------------------------------------------------
// Decoder (replaces gpu_dec_l1_l2, decoder/src/decoder.cu:454-730).
//
// Tile = THREADS consecutive 128-bit segments, one per thread.  Segment s starts at bit
// 128*s + (gap element s-1) and decodes every codeword that starts inside it (and inside the
// payload), as in decoder.cu:505-569.  A lane holds its segment's four payload words and the
// word after them in registers (loaded one tile ahead).  Three launches:
//   1. dec_count_kernel: walk 1, each lane counts its segment's codewords (64-bit bit buffer,
//      one- or two-level table in LDS, up to two codewords per lookup) and stores the count
//      (one byte per segment) and each wave's sum;
//   2. dec_offsets_kernel: each tile's output offset O, a scan of the wave sums;
//   3. dec_write_kernel: a workgroup scan of the tile's counts, then walk 2, each lane
//      decoding its segment again straight into an LDS image of the tile's output (image byte
//      j = output byte (O & ~15) + j).  A step stores two bytes whatever the entry holds, so a
//      lane leaves one stray byte on its successor's first byte, which the successor writes
//      again after its own walk (see dec_write_tile).  Owned output bytes [O, O + count) leave
//      with 128-bit stores (byte stores at the two shared 16-byte edges); bytes at or past n
//      (decoded zero padding) are dropped.
// The image region holds `image_bytes` bytes (chosen by the launcher for occupancy); a tile
// whose symbols do not fit goes through it in windows.
struct DecBuffer {
    const uint32_t *payload;
    uint32_t payload_words;
    uint32_t gap_elements;
    const uint32_t *gap;
    uint8_t *out;
    uint64_t n;
    uint64_t *status;  // unused by the three-launch decoder
    uint32_t num_tiles;
    uint32_t last_limit;  // payload bits in the last segment (dec_last_limit)
};

// Bits of the stream's last segment that lie inside the payload (1..128); every other segment
// has all 128.
__host__ __device__ inline uint32_t dec_last_limit(uint32_t payload_words, uint32_t gap_elements) {
    if (gap_elements == 0) return 0u;
    const uint64_t bits = 32ull * payload_words;
    const uint64_t before = 128ull * (gap_elements - 1);
    if (bits <= before) return 0u;
    return bits - before < 128 ? (uint32_t)(bits - before) : 128u;
}

struct DecTable {
    uint32_t prefix_bits;
    uint32_t l1_entries;
};

template <bool SINGLE>
__device__ __forceinline__ uint32_t dec_lookup(const uint32_t *s_tab, uint32_t hi, uint32_t P,
                                               uint32_t l1_entries) {
    uint32_t ent = s_tab[hi >> (32 - P)];
    if (!SINGLE && (ent >> 31)) {
        const uint32_t width = (ent >> 24) & 31u;
        const uint32_t idx2 = (hi << P) >> (32 - width);
        ent = s_tab[l1_entries + (ent & 0xFFFFFFu) + idx2];
    }
    return ent;
}

// A lane's share of a tile's global inputs: its segment's 4 payload words, the word after them
// (lane 63 of each wave only; the other lanes take it from their successor lane), and the gap
// word holding its start offset.  Loaded into registers one tile ahead, so their HBM latency
// overlaps the previous tile's work.
struct DecStage {
    u32x4 q;
    uint32_t extra;
    uint32_t gapw;
};

template <int THREADS>
__device__ __forceinline__ DecStage dec_fetch(const DecBuffer &buf, uint32_t tile) {
    DecStage st;
    st.q = u32x4{0u, 0u, 0u, 0u};
    st.extra = 0u;
    st.gapw = 0u;
    if (tile >= buf.num_tiles) return st;
    const uint64_t wi = 4ull * tile * THREADS + 4ull * threadIdx.x;
    if (wi + 4 <= buf.payload_words) {
        st.q = *reinterpret_cast<const u32x4 *>(buf.payload + wi);
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) st.q[i] = wi + i < buf.payload_words ? buf.payload[wi + i] : 0u;
    }
    if (lane_id() == 63 && wi + 4 < buf.payload_words) st.extra = buf.payload[wi + 4];
    const uint32_t seg = tile * THREADS + threadIdx.x;
    if (seg > 0 && seg < buf.gap_elements) st.gapw = buf.gap[(seg - 1) >> 3];
    return st;
}

// One lane's segment: payload words w0..w4 (w4 = first word of the next segment), bit limit and
// start offset.  Bit positions are relative to w0; `bb` holds bits [p, fill) MSB-aligned, `nx`
// the 32 bits after `fill`, then w3/w4 queue behind it.  Every step starts a codeword at
// p < limit and takes the entry's bits and symbols as they are.  A two-codeword entry whose
// second codeword starts at or past `limit` belongs to the next segment; p then reaches
// `limit`, so only a segment's last step can be such an entry.  A step consumes at most 16
// bits, so after a refill (>= 32 bits buffered) two steps can run before the next.
struct DecSeg {
    uint32_t w0, w1, w2, w3, w4;
    uint32_t p0, limit;
};

// The lane's segment from its staged words: called by all 64 lanes of the wave (the word after
// the segment comes from the successor lane; lane 63 loaded it).  Segments past the stream get
// limit 0 (no codewords).
__device__ __forceinline__ DecSeg dec_segment(const DecBuffer &buf, const DecStage &st,
                                              uint32_t seg) {
    DecSeg g;
    g.w0 = st.q[0];
    g.w1 = st.q[1];
    g.w2 = st.q[2];
    g.w3 = st.q[3];
    g.w4 = __builtin_amdgcn_ds_bpermute(4 * ((lane_id() + 1) & 63u), (int)st.q[0]);
    if (lane_id() == 63) g.w4 = st.extra;
    g.p0 = seg > 0 ? (st.gapw >> (((seg - 1) & 7u) * 4u)) & 15u : 0u;
    g.limit = seg + 1 < buf.gap_elements ? 128u
              : (seg + 1 == buf.gap_elements ? buf.last_limit : 0u);
    return g;
}

// the lane's first symbol (its segment's first codeword)
template <bool SINGLE>
__device__ __forceinline__ uint8_t dec_first(const DecSeg &g, const uint32_t *s_tab, uint32_t P,
                                             uint32_t l1) {
    const uint64_t bb = ((((uint64_t)g.w0) << 32) | g.w1) << g.p0;
    return (uint8_t)dec_lookup<SINGLE>(s_tab, (uint32_t)(bb >> 32), P, l1);
}

// Field access of direct table entries, as stored (sym1 | bits << 8 | sym2 << 16 | symbols << 24)
// or rotated right by 8 (bits | sym2 << 8 | symbols << 16 | sym1 << 24, the count walk's copy of
// a one-level table: the 64-bit shift then takes the entry itself as its amount).
struct EntStored {
    static __device__ __forceinline__ uint32_t bits(uint32_t e) { return (e >> 8) & 0xFFu; }
    static __device__ __forceinline__ uint32_t symbols(uint32_t e) { return e >> 24; }
    static __device__ __forceinline__ uint32_t sym2(uint32_t e) { return (e >> 16) & 0xFFu; }
};
struct EntRotated {
    static __device__ __forceinline__ uint32_t bits(uint32_t e) { return e & 0xFFu; }
    static __device__ __forceinline__ uint32_t symbols(uint32_t e) { return (e >> 16) & 0xFFu; }
    static __device__ __forceinline__ uint32_t sym2(uint32_t e) { return (e >> 8) & 0xFFu; }
};

// Bit buffer refill, shared by both walks: `q` = fill - 32, so a refill is due when p > q and
// shifts the next word in by p - q.
#define HGAP_REFILL()                                                                         \
    if (p > q) {                                                                              \
        bb |= ((uint64_t)nx) << (p - q);                                                      \
        q += 32u;                                                                             \
        nx = nq;                                                                              \
        nq = nq2;                                                                             \
    }

// walk 1: the number of codewords starting in [p0, limit)
template <bool SINGLE, typename E>
__device__ __forceinline__ uint32_t dec_count(const DecSeg &g, const uint32_t *s_tab,
                                              const uint8_t *s_len, uint32_t P, uint32_t l1) {
    uint32_t p = g.p0, q = 32, nx = g.w2, nq = g.w3, nq2 = g.w4;
    uint64_t bb = ((((uint64_t)g.w0) << 32) | g.w1) << p;
    uint32_t cnt = 0, last = 0;
#define HGAP_COUNT_STEP()                                                                     \
    {                                                                                         \
        const uint32_t ent = dec_lookup<SINGLE>(s_tab, (uint32_t)(bb >> 32), P, l1);            \
        cnt += E::symbols(ent);                                                               \
        p += E::bits(ent);                                                                    \
        bb <<= E::bits(ent) & 63u;                                                            \
        last = ent;                                                                           \
    }
    while (p < g.limit) {
        HGAP_COUNT_STEP()
        if (p < g.limit) HGAP_COUNT_STEP()
        HGAP_REFILL()
    }
#undef HGAP_COUNT_STEP
    // undo the last step's second codeword when it starts at p - len2 >= limit
    if (E::symbols(last) == 2u && p - s_len[E::sym2(last)] >= g.limit) cnt--;
    return cnt;
}

// walk 2: the same steps, storing the entry's two output bytes per step at the writer (LDS
// image or global output).  The byte a step stores past the lane's count is overwritten by its
// next step, except after the last step: one stray byte just past the lane's range.
template <bool SINGLE, typename Writer>
__device__ __forceinline__ void dec_write(const DecSeg &g, const uint32_t *s_tab, uint32_t P,
                                          uint32_t l1, Writer wp) {
    uint32_t p = g.p0, q = 32, nx = g.w2, nq = g.w3, nq2 = g.w4;
    uint64_t bb = ((((uint64_t)g.w0) << 32) | g.w1) << p;
#define HGAP_WRITE_STEP()                                                                     \
    {                                                                                         \
        const uint32_t ent = dec_lookup<SINGLE>(s_tab, (uint32_t)(bb >> 32), P, l1);            \
        wp.put2(ent);                                                                         \
        wp.advance(EntStored::symbols(ent));                                                  \
        p += EntStored::bits(ent);                                                            \
        bb <<= EntStored::bits(ent) & 63u;                                                    \
    }
    while (p < g.limit) {
        HGAP_WRITE_STEP()
        if (p < g.limit) HGAP_WRITE_STEP()
        HGAP_REFILL()
    }
#undef HGAP_WRITE_STEP
}
#undef HGAP_REFILL

// LDS writer for walk 2: two byte stores, of the entry's bytes 0 and 2 (single-byte stores
// take them from the data register without shifts; a 16-bit store at an odd address is much
// slower on gfx950 LDS).
struct DecLdsOut {
    uint8_t *p;
    __device__ __forceinline__ void put2(uint32_t ent) {
        p[0] = (uint8_t)ent;
        p[1] = (uint8_t)(ent >> 16);
    }
    __device__ __forceinline__ void advance(uint32_t d) { p += d; }
};

// LDS writer for walk 2 of a tile written through windows of the image region: only bytes
// whose position x (relative to the window's first image byte) lies in [0, span) are stored.
struct DecLdsWindow {
    uint8_t *img;
    int32_t x;
    uint32_t span;
    __device__ __forceinline__ void put2(uint32_t ent) {
        if ((uint32_t)x < span) img[x] = (uint8_t)ent;
        if ((uint32_t)(x + 1) < span) img[x + 1] = (uint8_t)(ent >> 16);
    }
    __device__ __forceinline__ void advance(uint32_t d) { x += (int32_t)d; }
};

// LDS bytes of the decoder's image region, from the device's LDS budget: the region that lets
// 32 / (THREADS / 64) workgroups share a CU next to the table (or fewer when the table is
// large), capped at what the largest possible tile needs (16 + THREADS * max_symbols + 32,
// rounded to 16) and at least 16 * THREADS + 48 bytes; 0 when even that does not fit.
__host__ __device__ constexpr uint32_t dec_image_need(uint32_t threads, uint32_t max_symbols) {
    return (16u + threads * max_symbols + 32u + 15u) & ~15u;
}

// ---------------------------------------------------------------------------------------------
// Single-buffer decode in three launches, with no waiting between tiles:
//   dec_count_kernel    walk 1 of every segment; its codeword count, one byte per segment
//   dec_offsets_kernel  each tile's symbol total (from per-wave sums the count kernel stores
//                       beside the bytes), scanned into the tile's output offset (blocks of
//                       1024 tiles chained by the look-back)
//   dec_write_kernel    per tile: scan of its segment counts, walk 2 into an LDS image aligned
//                       like the output (image byte j = output byte (O & ~15) + j), aligned
//                       128-bit stores (a tile larger than the image region goes through it in
//                       windows)
// The payload is read twice; in exchange no tile waits on another, and the persistent
// workgroups take tiles blockIdx.x, blockIdx.x + gridDim.x, ... without a ticket, loading the
// next tile's inputs into registers one tile ahead.
template <int THREADS, bool SINGLE>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void dec_count_kernel(DecBuffer buf, DecTable t, const uint32_t *__restrict__ table,
                      uint32_t table_words, uint8_t *__restrict__ seg_counts,
                      uint32_t *__restrict__ wave_sums) {
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    uint32_t *s_tab = s_dyn;
    const uint8_t *s_len = reinterpret_cast<const uint8_t *>(s_tab + table_words - kDecLengthWords);
    // one-level tables are stored rotated (EntRotated); the length block is copied as is
    using E = typename std::conditional<SINGLE, EntRotated, EntStored>::type;
    const uint32_t entries = table_words - kDecLengthWords;
    for (uint32_t i = threadIdx.x; i < table_words; i += THREADS) {
        const uint32_t e = table[i];
        s_tab[i] = SINGLE && i < entries ? __builtin_amdgcn_alignbit(e, e, 8) : e;
    }
    __syncthreads();
    uint32_t tile = blockIdx.x;
    DecStage st = dec_fetch<THREADS>(buf, tile);
    while (tile < buf.num_tiles) {
        const uint32_t nt = tile + gridDim.x;
        const DecStage sn = dec_fetch<THREADS>(buf, nt);
        const uint32_t seg = tile * THREADS + threadIdx.x;
        const DecSeg g = dec_segment(buf, st, seg);
        const uint32_t cnt =
            seg < buf.gap_elements
                ? dec_count<SINGLE, E>(g, s_tab, s_len, t.prefix_bits, t.l1_entries)
                : 0u;
        seg_counts[seg] = (uint8_t)cnt;  // <= 128 codewords start in a segment
        const uint32_t wsum = wave_incl_sum(cnt);
        if (lane_id() == 63) wave_sums[seg >> 6] = wsum;
        st = sn;
        tile = nt;
    }
}

constexpr uint32_t kOffsetThreads = 256;
constexpr uint32_t kOffsetTilesPerThread = 4;
constexpr uint32_t kOffsetTiles = kOffsetThreads * kOffsetTilesPerThread;  // per block

// Tile t's output offset: the sum of the wave sums of tiles before t.  A block covers 1024
// tiles (4 consecutive per thread); blocks are chained by the look-back in claim order.
template <int SEGS>
__global__ __launch_bounds__(kOffsetThreads) void dec_offsets_kernel(
    const uint32_t *__restrict__ wave_sums, uint32_t tiles, uint64_t *__restrict__ tile_offsets,
    uint64_t *__restrict__ status, uint32_t *__restrict__ counter, uint32_t wpl) {
    constexpr uint32_t W = SEGS / 64;  // wave sums per tile (4, 8 or 16)
    __shared__ uint32_t s_wave[kOffsetThreads / 64];
    __shared__ uint32_t s_blk;
    __shared__ uint64_t s_base;
    if (threadIdx.x == 0) s_blk = atomicAdd(counter, 1u);  // blocks in claim order
    __syncthreads();
    const uint32_t blk = s_blk;
    const uint32_t tile0 = blk * kOffsetTiles + threadIdx.x * kOffsetTilesPerThread;
    uint32_t sums[kOffsetTilesPerThread];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kOffsetTilesPerThread; k++) {
        uint32_t v = 0;
        if (tile0 + k < tiles) {
            const u32x4 *p = reinterpret_cast<const u32x4 *>(wave_sums + (uint64_t)(tile0 + k) * W);
#pragma unroll
            for (uint32_t i = 0; i < W / 4; i++) {
                const u32x4 x = p[i];
                v += x[0] + x[1] + x[2] + x[3];
            }
        }
        sums[k] = v;
        mine += v;
    }
    uint32_t total;  // <= 1024 tiles x 1024 segments x 128 codewords = 2^27
    const uint32_t excl = block_excl_sum<kOffsetThreads>(mine, s_wave, total);
    if (threadIdx.x < 64) {
        const uint64_t before = look_back(status, blk, total, wpl);
        if (threadIdx.x == 0) s_base = before;
    }
    __syncthreads();
    uint64_t o = s_base + excl;
#pragma unroll
    for (uint32_t k = 0; k < kOffsetTilesPerThread; k++) {
        if (tile0 + k < tiles) tile_offsets[tile0 + k] = o;
        o += sums[k];
    }
}

// Stores image bytes [j0, j1) of a tile (j0 a multiple of 16; the image region holds them from
// its start) to output bytes base + [j0, j1), base = O & ~15, keeping only [O, O + tile_syms)
// below n: 16-byte stores, byte stores in chunks that are not wholly inside.
template <int THREADS>
__device__ __forceinline__ void dec_copy_out(const DecBuffer &buf, const uint8_t *s_img,
                                             uint64_t O, uint32_t tile_syms, uint32_t j0,
                                             uint32_t j1) {
    const uint64_t base = O & ~15ull;
    uint64_t lo = base + j0 > O ? base + j0 : O;
    uint64_t hi = O + tile_syms;
    if (hi > base + j1) hi = base + j1;
    if (hi > buf.n) hi = buf.n;
    const uint32_t chunks = (j1 - j0 + 15u) >> 4;
    for (uint32_t c = threadIdx.x; c < chunks; c += THREADS) {
        const u32x4 v = *reinterpret_cast<const u32x4 *>(s_img + 16 * c);
        const uint64_t gpos = base + j0 + 16ull * c;
        if (gpos >= lo && gpos + 16 <= hi) {
            __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(buf.out + gpos));
        } else {
            for (int i = 0; i < 16; i++)
                if (gpos + i >= lo && gpos + i < hi)
                    buf.out[gpos + i] = (uint8_t)(v[i >> 2] >> (8 * (i & 3)));
        }
    }
}

// One tile of the write kernels: the lane's segment `seg` (staged in `st`) has `cnt` codewords;
// the tile's output starts at buf.out + O.  Ends with no barrier: the caller's next LDS image
// write must follow a barrier that every wave reaches after this tile's copy-out.
template <int THREADS, bool SINGLE>
__device__ __forceinline__ void dec_write_tile(const DecBuffer &buf, DecTable t,
                                               const uint32_t *s_tab, uint8_t *s_img,
                                               uint32_t image_bytes, uint32_t *s_wave,
                                               const DecStage &st, uint32_t seg, uint32_t cnt,
                                               uint64_t O) {
    const uint32_t P = t.prefix_bits;
    const DecSeg g = dec_segment(buf, st, seg);
    uint32_t tile_syms;
    const uint32_t excl = block_excl_sum<THREADS>(cnt, s_wave, tile_syms);
    const uint32_t lead = (uint32_t)(O & 15u);
    const uint32_t a = lead + excl;  // this lane's first image byte
    const uint32_t total = lead + tile_syms;
    if (total + 17u <= image_bytes) {  // uniform
        // walk 2 into the image.  Every segment of the stream has a codeword, so the stray
        // byte on a lane's first byte comes from the lane before it.  One wave's LDS stores
        // land in program order: a lane writing its first symbol again after its walk lands
        // after its predecessor's stray byte when both are in this wave, and lane 63 writes
        // its successor's first symbol (decoded from the next segment's first word and gap
        // nibble) after its own stray byte, for the lane 0 of the next wave.
        if (cnt) {
            dec_write<SINGLE>(g, s_tab, P, t.l1_entries, DecLdsOut{s_img + a});
            s_img[a] = dec_first<SINGLE>(g, s_tab, P, t.l1_entries);
            if (lane_id() == 63u) {
                // element `seg` shares lane 63's gap word (seg % 8 == 7)
                DecSeg nx{};
                nx.w0 = g.w4;
                nx.p0 = (st.gapw >> ((seg & 7u) * 4u)) & 15u;
                s_img[a + cnt] = dec_first<SINGLE>(nx, s_tab, P, t.l1_entries);
            }
        }
        __syncthreads();
        dec_copy_out<THREADS>(buf, s_img, O, tile_syms, 0u, total);
    } else {
        // more symbols than the region holds: windows of `span` image bytes, each written by
        // the lanes whose bytes meet it (bytes outside the window are not stored), then stored
        const uint32_t span = image_bytes & ~15u;
        for (uint32_t j0 = 0; j0 < total; j0 += span) {  // uniform
            if (cnt && a < j0 + span && a + cnt > j0)
                dec_write<SINGLE>(g, s_tab, P, t.l1_entries,
                                  DecLdsWindow{s_img, (int32_t)(a - j0), span});
            __syncthreads();
            if (cnt && a >= j0 && a < j0 + span)
                s_img[a - j0] = dec_first<SINGLE>(g, s_tab, P, t.l1_entries);
            __syncthreads();
            dec_copy_out<THREADS>(buf, s_img, O, tile_syms, j0, std::min(total, j0 + span));
            __syncthreads();  // the next window overwrites the region
        }
    }
}

template <int THREADS, bool SINGLE>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void dec_write_kernel(DecBuffer buf, DecTable t, const uint32_t *__restrict__ table,
                      uint32_t table_words, uint32_t image_offset, uint32_t image_bytes,
                      const uint8_t *__restrict__ seg_counts,
                      const uint64_t *__restrict__ tile_offsets) {
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    __shared__ uint32_t s_wave[THREADS / 64];
    uint32_t *s_tab = s_dyn;
    uint8_t *s_img = reinterpret_cast<uint8_t *>(s_dyn + image_offset);  // 16-byte aligned
    for (uint32_t i = threadIdx.x; i < table_words; i += THREADS) s_tab[i] = table[i];
    __syncthreads();
    uint32_t tile = blockIdx.x;
    DecStage st = dec_fetch<THREADS>(buf, tile);
    uint32_t cnt = tile < buf.num_tiles ? seg_counts[tile * THREADS + threadIdx.x] : 0u;
    while (tile < buf.num_tiles) {
        const uint32_t nt = tile + gridDim.x;
        const DecStage sn = dec_fetch<THREADS>(buf, nt);
        const uint32_t cn = nt < buf.num_tiles ? seg_counts[nt * THREADS + threadIdx.x] : 0u;
        dec_write_tile<THREADS, SINGLE>(buf, t, s_tab, s_img, image_bytes, s_wave, st,
                                        tile * THREADS + threadIdx.x, cnt, tile_offsets[tile]);
        st = sn;
        cnt = cn;
        tile = nt;
    }
}

// ---------------------------------------------------------------------------------------------
// Batched steps: many independent buffers, each with its own codebook or decode table, in one
// persistent launch per step.  Job descriptors live in device memory (so a captured graph
// replays the same launches).  Encode/decode jobs own look-back status words
// [tile_base, tile_base + num_tiles); tiles are claimed in job order from one counter, so every
// predecessor tile of a job is resident or finished.

// job index owning global unit `g` (units = tiles or bytes), jobs sorted by base
template <typename J>
__device__ __forceinline__ uint32_t find_job(const J *jobs, uint32_t count, uint64_t g) {
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if ((uint64_t)jobs[mid].base <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct HistJobDev {
    const uint8_t *in;
    uint64_t n;
    uint32_t *hist;
    uint32_t range_base;  // first row of this job's per-range counts
    uint32_t reserved;
    uint64_t base;  // first byte of this job in the batch's concatenation
};

// zero every job's 256 bins (one workgroup per job, strided)
__global__ void hist_batch_zero_kernel(const HistJobDev *jobs, uint32_t count) {
    for (uint32_t j = blockIdx.x; j < count; j += gridDim.x) jobs[j].hist[threadIdx.x] = 0u;
}

// Each workgroup counts one contiguous slice of the concatenated jobs; at every job boundary
// inside its slice it flushes its 64 LDS bin copies into that job's histogram.
__global__ __launch_bounds__(kHistThreads) void hist_batch_kernel(const HistJobDev *jobs,
                                                                  uint32_t count,
                                                                  uint64_t total) {
    __shared__ uint32_t s_bins[kHistCopies * 256];
    uint32_t *mine = s_bins + lane_id();  // bin b of this lane's copy: mine[64 b]
    for (int i = threadIdx.x; i < kHistCopies * 256; i += kHistThreads) s_bins[i] = 0u;
    const uint64_t per = (total + gridDim.x - 1) / gridDim.x;
    const uint64_t a = (uint64_t)blockIdx.x * per;
    const uint64_t b = a + per < total ? a + per : total;
    __syncthreads();
    if (a >= b) return;
    uint32_t j = find_job(jobs, count, a);
    for (uint64_t pos = a; pos < b;) {
        const HistJobDev job = jobs[j];
        const uint64_t lo = pos - job.base;
        const uint64_t hi = (b < job.base + job.n ? b : job.base + job.n) - job.base;
        // [lo, hi) of this job: bytes up to the first 16-byte boundary, vectors, tail bytes
        const uint64_t vlo = (lo + 15) & ~15ull, vhi = hi & ~15ull;
        if (vlo >= vhi) {
            for (uint64_t i = lo + threadIdx.x; i < hi; i += kHistThreads)
                atomicAdd(&mine[(uint32_t)job.in[i] << 6], 1u);
        } else {
            for (uint64_t i = lo + threadIdx.x; i < vlo; i += kHistThreads)
                atomicAdd(&mine[(uint32_t)job.in[i] << 6], 1u);
            for (uint64_t i = vhi + threadIdx.x; i < hi; i += kHistThreads)
                atomicAdd(&mine[(uint32_t)job.in[i] << 6], 1u);
            const u32x4 *in4 = reinterpret_cast<const u32x4 *>(job.in);
            for (uint64_t v = (vlo >> 4) + threadIdx.x; v < (vhi >> 4); v += kHistThreads) {
                const u32x4 q = __builtin_nontemporal_load(&in4[v]);
#pragma unroll
                for (int k = 0; k < 16; k++)
                    atomicAdd(&mine[((q[k >> 2] >> (8 * (k & 3))) & 0xFFu) << 6], 1u);
            }
        }
        __syncthreads();
        uint32_t sum = 0;
        for (int c = 0; c < kHistCopies; c++) {
            const uint32_t w = (threadIdx.x << 6) + ((c + threadIdx.x) & 63u);
            sum += s_bins[w];
            s_bins[w] = 0u;
        }
        if (sum) atomicAdd(&job.hist[threadIdx.x], sum);
        __syncthreads();
        pos = job.base + hi;
        j++;
    }
}

// Per-range counts of a batch (hist_range_kernel over many buffers): workgroup b takes the
// contiguous ranges [b*C, (b+1)*C) of the batch (range_base order) and, after each one, stores
// its counts and adds them to the job's histogram (the job's histogram words were zeroed).
__global__ __launch_bounds__(kHistThreadsSingle) void hist_range_batch_kernel(
    const HistJobDev *__restrict__ jobs, uint32_t count, uint32_t total_ranges,
    uint32_t range_bytes, uint32_t *__restrict__ range_hist) {
    constexpr int kThreads = kHistThreadsSingle;
    __shared__ uint32_t s_bins[kHistCopies * 256];
    for (int i = threadIdx.x; i < kHistCopies * 256; i += kThreads) s_bins[i] = 0u;
    __syncthreads();
    uint32_t *mine = s_bins + lane_id();
    const uint32_t chunk = (total_ranges + gridDim.x - 1) / gridDim.x;
    const uint32_t u0 = blockIdx.x * chunk;
    const uint32_t u1 = u0 + chunk < total_ranges ? u0 + chunk : total_ranges;
    if (u0 >= u1) return;  // uniform
    // the last job whose first range is at or before u0 (jobs without bytes own no ranges)
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].range_base <= u0) lo = mid; else hi = mid - 1;
    }
    uint32_t j = lo;
    for (uint32_t u = u0; u < u1; u++) {
        while (j + 1 < count && jobs[j + 1].range_base <= u) j++;
        const HistJobDev job = jobs[j];
        const uint64_t a = (uint64_t)(u - job.range_base) * range_bytes;
        if (a < job.n)  // always, when total_ranges matches the jobs
            hist_count_range<kThreads>(job.in, a,
                                       a + range_bytes < job.n ? a + range_bytes : job.n, mine);
        __syncthreads();
        if (threadIdx.x < 256) {
            const uint32_t sum = hist_take_bin(s_bins);
            range_hist[256ull * u + threadIdx.x] = sum;
            if (sum) atomicAdd(&job.hist[threadIdx.x], sum);
        }
        __syncthreads();
    }
}

struct EncJobDev {
    const uint8_t *in;
    uint64_t n;
    uint32_t *payload;
    uint32_t *gap;
    uint32_t payload_words;
    uint32_t gap_words;
    uint32_t table;    // index of this job's 256-word code table
    uint32_t max_len;
    uint32_t base;     // first tile (chained: first range) and look-back status word of this job
    uint32_t num_tiles;
    uint32_t range_base;  // chained: first row of this job's per-range counts
    uint32_t reserved;
};

// zero every job's gap words (one workgroup per job, strided)
template <typename J>
__global__ void gap_batch_zero_kernel(const J *jobs, uint32_t count) {
    for (uint32_t j = blockIdx.x; j < count; j += gridDim.x) {
        uint32_t *g = jobs[j].gap;
        for (uint32_t i = threadIdx.x; i < jobs[j].gap_words; i += blockDim.x) g[i] = 0u;
    }
}

template <int THREADS, int K>
__global__ __launch_bounds__(THREADS) void enc_batch_kernel(
    const EncJobDev *__restrict__ jobs, uint32_t count, uint32_t total_tiles,
    const uint32_t *__restrict__ tables, uint32_t *__restrict__ counter,
    uint64_t *__restrict__ status, uint32_t out_words, uint32_t wpl) {
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    __shared__ uint32_t s_tab[256];
    __shared__ uint32_t s_wave[THREADS / 64];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_tile, s_job;
    uint32_t *s_out = s_dyn;
    uint32_t *s_gap = s_dyn + out_words;
    uint32_t loaded = 0xFFFFFFFFu;
    while (true) {
        if (threadIdx.x == 0) {
            const uint32_t g = atomicAdd(counter, 1u);
            s_tile = g;
            if (g < total_tiles) s_job = find_job(jobs, count, g);
        }
        __syncthreads();
        const uint32_t g = s_tile;
        if (g >= total_tiles) break;
        const uint32_t j = s_job;
        const EncJobDev job = jobs[j];
        if (job.table != loaded) {  // uniform: every thread sees the same job
            for (int i = threadIdx.x; i < 256; i += THREADS) s_tab[i] = tables[256u * job.table + i];
            loaded = job.table;
            __syncthreads();
        }
        EncBuffer buf;
        buf.in = job.in;
        buf.n = job.n;
        buf.payload = job.payload;
        buf.payload_words = job.payload_words;
        buf.gap_words = job.gap_words;
        buf.gap = job.gap;
        buf.status = status + job.base;
        buf.num_tiles = job.num_tiles;
        buf.reserved = 0;
        const EncStage<K> st = enc_fetch<THREADS, K>(buf, g - job.base);
        NextTile none{nullptr, nullptr, 0u};
        EncChain unused{};
        EncStage<K> no_prefetch;
        enc_tile<THREADS, K, false, false>(buf, g - job.base, st, s_tab, s_out, s_gap, s_wave,
                                           &s_excl, wpl, none, unused, no_prefetch);
    }
}

// Chained batched encode (the chained encoder for every job).  Ranges of all jobs are numbered
// in job order (job.base = the job's first); a job's status words are [base, base + ranges).
// The prep kernel stores every range's bit count (its row of range counts dotted with the
// job's code lengths) as a look-back aggregate, a job's first range as an inclusive prefix.
// One wave per range; it also records the range's job (range_job[u]), so that the encoder
// needs no search per range.
__global__ __launch_bounds__(256) void enc_range_prep_batch_kernel(
    const EncJobDev *__restrict__ jobs, uint32_t count, uint32_t total_ranges,
    const uint32_t *__restrict__ tables, const uint32_t *__restrict__ range_hist,
    uint32_t range_bytes, uint64_t *__restrict__ status, uint32_t *__restrict__ range_job,
    uint32_t *__restrict__ counter) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * 4u;
    for (uint32_t u = blockIdx.x * 4u + (threadIdx.x >> 6); u < total_ranges; u += waves) {
        const uint32_t j = find_job(jobs, count, u);
        const EncJobDev job = jobs[j];
        const uint32_t r = u - job.base;
        if (lane == 0) range_job[u] = j;
        if ((uint64_t)r * range_bytes >= job.n) {  // only when the plan does not match
            if (lane == 0) status_store(&status[u], kFlagA);
            continue;
        }
        const uint32_t *len = tables + 256ull * job.table + 4 * lane;
        const u32x4 h = *reinterpret_cast<const u32x4 *>(
            range_hist + 256ull * (job.range_base + r) + 4 * lane);
        const uint64_t bits = wave_sum_u64(
            (uint64_t)h[0] * (len[0] & 31u) + (uint64_t)h[1] * (len[1] & 31u) +
            (uint64_t)h[2] * (len[2] & 31u) + (uint64_t)h[3] * (len[3] & 31u));
        if (lane == 0) status_store(&status[u], (r == 0 ? kFlagP : kFlagA) | (bits & kValueMask));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

// A workgroup takes a range from the counter, reloads the code table when the job changes, and
// encodes the range's tiles in order (enc_tile's chained mode; the range's first tile claims
// the workgroup's next range).
template <int THREADS, int K>
__global__ __launch_bounds__(THREADS) void enc_range_batch_kernel(
    const EncJobDev *__restrict__ jobs, uint32_t count, uint32_t total_ranges,
    const uint32_t *__restrict__ tables, uint32_t *__restrict__ counter,
    uint64_t *__restrict__ status, const uint32_t *__restrict__ range_job, uint32_t out_words,
    uint32_t wpl, uint32_t tiles_per_range) {
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    __shared__ uint32_t s_tab[256];
    __shared__ uint32_t s_wave[THREADS / 64];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_ticket;
    uint32_t *s_out = s_dyn;
    uint32_t *s_gap = s_dyn + out_words;
    if (threadIdx.x == 0) s_ticket = atomicAdd(counter, 1u);
    __syncthreads();
    EncChain ch{};
    ch.next_range = s_ticket;
    ch.tiles_per_range = tiles_per_range;
    NextTile next{counter, &s_ticket, 0u};
    uint32_t loaded = 0xFFFFFFFFu;
    while (ch.next_range < total_ranges) {  // uniform
        const uint32_t j = range_job[ch.next_range];
        const EncJobDev job = jobs[j < count ? j : count - 1];
        if (job.table != loaded) {  // the previous tile ended with a barrier
            for (int i = threadIdx.x; i < 256; i += THREADS) s_tab[i] = tables[256u * job.table + i];
            loaded = job.table;
            __syncthreads();
        }
        EncBuffer buf;
        buf.in = job.in;
        buf.n = job.n;
        buf.payload = job.payload;
        buf.payload_words = job.payload_words;
        buf.gap_words = job.gap_words;
        buf.gap = job.gap;
        buf.status = status + job.base;
        buf.num_tiles = job.num_tiles;
        buf.reserved = 0;
        ch.range = ch.next_range - job.base;
        ch.ranges = 0;  // the range's last tile prefetches nothing (the next range is another's)
        const uint32_t lo = ch.range * tiles_per_range;
        if (lo >= job.num_tiles) {  // a range without tiles (only when the plan does not match)
            __syncthreads();
            if (threadIdx.x == 0) s_ticket = atomicAdd(counter, 1u);
            __syncthreads();
            ch.next_range = s_ticket;
            continue;
        }
        ch.end = lo + tiles_per_range < job.num_tiles ? lo + tiles_per_range : job.num_tiles;
        EncStage<K> ahead = enc_fetch<THREADS, K>(buf, lo);
        for (uint32_t t = lo; t < ch.end; t++) {
            ch.first = t == lo;
            const EncStage<K> st = ahead;
            enc_tile<THREADS, K, true, true>(buf, t, st, s_tab, s_out, s_gap, s_wave, &s_excl,
                                             wpl, next, ch, ahead);
        }
    }
}

struct DecJobDev {
    const uint32_t *payload;
    const uint32_t *gap;
    uint8_t *out;
    uint64_t n;
    const uint32_t *table;
    uint32_t payload_words;
    uint32_t gap_elements;
    uint32_t prefix_bits;
    uint32_t l1_entries;
    uint32_t table_words;
    uint32_t max_symbols;
    uint32_t max_length;
    uint32_t base;     // first tile (and look-back status word) of this job
    uint32_t num_tiles;
    uint32_t reserved;
};

// Batched decode: the three single-buffer launches over the tiles of all jobs.  Workgroup b
// takes the contiguous tiles [b*C, (b+1)*C) (C = ceil(total / grid)), so consecutive tiles of a
// workgroup mostly share a job and its table is loaded once per job.  The offsets kernel scans
// all tiles; a tile's offset inside its job is its prefix minus that of the job's first tile.
// a job descriptor read into scalar registers (every lane reads the same one)
__device__ __forceinline__ DecJobDev dec_load_job(const DecJobDev *p) {
    static_assert(sizeof(DecJobDev) % 4 == 0, "DecJobDev is whole words");
    DecJobDev job;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(p);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&job);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(DecJobDev) / 4; i++)
        dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return job;
}

struct DecCursor {
    const DecJobDev *jobs;
    uint32_t count, j;
    DecJobDev job;
    // the job owning tile g (g at or after the current job's first tile)
    __device__ __forceinline__ void seek(uint32_t g) {
        while (g >= job.base + job.num_tiles && j + 1 < count) job = dec_load_job(&jobs[++j]);
    }
};

__device__ __forceinline__ DecBuffer dec_job_buffer(const DecJobDev &job) {
    DecBuffer buf;
    buf.payload = job.payload;
    buf.payload_words = job.payload_words;
    buf.gap_elements = job.gap_elements;
    buf.gap = job.gap;
    buf.out = job.out;
    buf.n = job.n;
    buf.status = nullptr;
    buf.num_tiles = job.num_tiles;
    buf.last_limit = dec_last_limit(job.payload_words, job.gap_elements);
    return buf;
}

// whether a job's decode table has one level (no second-level entries after the first level)
__device__ __forceinline__ bool dec_job_single(const DecJobDev &job) {
    return job.table_words == job.l1_entries + kDecLengthWords;
}

// loads job `j`'s table into LDS (after a barrier: the previous table may still be in use);
// `rotate`: one-level entries rotated right by 8 for the count walk (EntRotated)
template <int THREADS>
__device__ __forceinline__ void dec_load_job_table(uint32_t *s_tab, const DecJobDev &job,
                                                   bool rotate = false) {
    __syncthreads();
    const uint32_t entries = job.table_words - kDecLengthWords;
    for (uint32_t i = threadIdx.x; i < job.table_words; i += THREADS) {
        const uint32_t e = job.table[i];
        s_tab[i] = rotate && i < entries ? __builtin_amdgcn_alignbit(e, e, 8) : e;
    }
    __syncthreads();
}

// A tile's inputs in the batched kernels: loaded one tile ahead while the next tile of the
// workgroup's run belongs to the same job (the usual case), else at the tile's start.
struct DecBatchStage {
    DecStage st;
    uint32_t cnt;  // the write kernel's segment count (0 in the count kernel)
    bool ready;    // loaded ahead (uniform)
};

template <int THREADS, bool WITH_COUNT>
__device__ __forceinline__ DecBatchStage dec_batch_fetch(const DecJobDev &job, uint32_t g,
                                                         const uint8_t *seg_counts) {
    DecBatchStage s;
    s.st = dec_fetch<THREADS>(dec_job_buffer(job), g - job.base);
    s.cnt = WITH_COUNT ? seg_counts[(uint64_t)g * THREADS + threadIdx.x] : 0u;
    s.ready = true;
    return s;
}

// this tile's inputs (`ahead` if it was loaded ahead), then starts loading the next tile's when
// it is in the same job
template <int THREADS, bool WITH_COUNT>
__device__ __forceinline__ DecBatchStage dec_batch_stage(DecBatchStage &ahead,
                                                         const DecJobDev &job, uint32_t g,
                                                         uint32_t g1,
                                                         const uint8_t *seg_counts) {
    const DecBatchStage now =
        ahead.ready ? ahead : dec_batch_fetch<THREADS, WITH_COUNT>(job, g, seg_counts);
    if (g + 1 < g1 && g + 1 < job.base + job.num_tiles)
        ahead = dec_batch_fetch<THREADS, WITH_COUNT>(job, g + 1, seg_counts);
    else
        ahead.ready = false;
    return now;
}

template <int THREADS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void dec_count_batch_kernel(const DecJobDev *__restrict__ jobs, uint32_t count, uint32_t total,
                            uint8_t *__restrict__ seg_counts, uint32_t *__restrict__ wave_sums) {
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    uint32_t *s_tab = s_dyn;
    const uint32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t g0 = blockIdx.x * chunk;
    const uint32_t g1 = g0 + chunk < total ? g0 + chunk : total;
    if (g0 >= g1) return;  // uniform
    DecCursor cur{jobs, count, find_job(jobs, count, g0), {}};
    cur.job = dec_load_job(&jobs[cur.j]);
    DecBatchStage ahead{};
    uint32_t loaded = 0xFFFFFFFFu;
    bool single = false;
    for (uint32_t g = g0; g < g1; g++) {
        cur.seek(g);
        if (cur.j != loaded) {  // uniform
            single = dec_job_single(cur.job);
            dec_load_job_table<THREADS>(s_tab, cur.job, single);
            loaded = cur.j;
        }
        const DecStage st =
            dec_batch_stage<THREADS, false>(ahead, cur.job, g, g1, nullptr).st;
        const DecBuffer buf = dec_job_buffer(cur.job);
        const uint32_t tile = g - cur.job.base;
        const uint32_t seg = tile * THREADS + threadIdx.x;
        const DecSeg sg = dec_segment(buf, st, seg);
        const uint8_t *s_len =
            reinterpret_cast<const uint8_t *>(s_tab + cur.job.table_words - kDecLengthWords);
        const uint32_t P = cur.job.prefix_bits, l1 = cur.job.l1_entries;
        uint32_t cnt = 0u;
        if (seg < buf.gap_elements)
            cnt = single ? dec_count<true, EntRotated>(sg, s_tab, s_len, P, l1)
                         : dec_count<false, EntStored>(sg, s_tab, s_len, P, l1);
        seg_counts[(uint64_t)g * THREADS + threadIdx.x] = (uint8_t)cnt;
        const uint32_t wsum = wave_incl_sum(cnt);
        if (lane_id() == 63) wave_sums[(uint64_t)g * (THREADS / 64) + (threadIdx.x >> 6)] = wsum;
    }
}

template <int THREADS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void dec_write_batch_kernel(const DecJobDev *__restrict__ jobs, uint32_t count, uint32_t total,
                            uint32_t image_offset, uint32_t image_bytes,
                            const uint8_t *__restrict__ seg_counts,
                            const uint64_t *__restrict__ tile_offsets) {
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    __shared__ uint32_t s_wave[THREADS / 64];
    uint32_t *s_tab = s_dyn;
    uint8_t *s_img = reinterpret_cast<uint8_t *>(s_dyn + image_offset);
    const uint32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t g0 = blockIdx.x * chunk;
    const uint32_t g1 = g0 + chunk < total ? g0 + chunk : total;
    if (g0 >= g1) return;  // uniform
    DecCursor cur{jobs, count, find_job(jobs, count, g0), {}};
    cur.job = dec_load_job(&jobs[cur.j]);
    DecBatchStage ahead{};
    uint32_t loaded = 0xFFFFFFFFu;
    bool single = false;
    for (uint32_t g = g0; g < g1; g++) {
        cur.seek(g);
        if (cur.j != loaded) {  // uniform; its barriers also end the previous tile's copy-out
            single = dec_job_single(cur.job);
            dec_load_job_table<THREADS>(s_tab, cur.job);
            loaded = cur.j;
        }
        const DecBatchStage now = dec_batch_stage<THREADS, true>(ahead, cur.job, g, g1, seg_counts);
        const DecBuffer buf = dec_job_buffer(cur.job);
        const uint32_t tile = g - cur.job.base;
        const uint64_t O = tile_offsets[g] - tile_offsets[cur.job.base];
        const DecTable t{cur.job.prefix_bits, cur.job.l1_entries};
        const uint32_t seg = tile * THREADS + threadIdx.x;
        if (single)
            dec_write_tile<THREADS, true>(buf, t, s_tab, s_img, image_bytes, s_wave, now.st, seg,
                                          now.cnt, O);
        else
            dec_write_tile<THREADS, false>(buf, t, s_tab, s_img, image_bytes, s_wave, now.st,
                                           seg, now.cnt, O);
    }
}

// ---------------------------------------------------------------------------------------------
// Synthetic data (counter-based splitmix64 + integer inverse CDF).
struct SynthTable {
    uint64_t cdf[256];
    uint8_t perm[256];
};

__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void synth_kernel(uint8_t *__restrict__ out, uint64_t n,
                                                    uint64_t start, uint64_t seed,
                                                    SynthTable t) {
    __shared__ uint64_t s_cdf[256];
    __shared__ uint8_t s_perm[256];
    s_cdf[threadIdx.x] = t.cdf[threadIdx.x];
    s_perm[threadIdx.x] = t.perm[threadIdx.x];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 16;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16; i0 < n;
         i0 += stride) {
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint64_t r = splitmix64_at(seed, start + i0 + j) >> 11;
            uint32_t lo = 0, hi = 255;
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const uint32_t mid = (lo + hi) >> 1;
                if (r < s_cdf[mid]) hi = mid; else lo = mid + 1;
            }
            w[j >> 2] |= (uint32_t)s_perm[lo] << (8 * (j & 3));
        }
        if (i0 + 16 <= n) {
            const u32x4 v = {w[0], w[1], w[2], w[3]};
            *reinterpret_cast<u32x4 *>(out + i0) = v;
        } else {
            for (uint64_t j = 0; i0 + j < n; j++) out[i0 + j] = (uint8_t)(w[j >> 2] >> (8 * (j & 3)));
        }
    }
}

// ---------------------------------------------------------------------------------------------
// launch helpers

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

static int device_cus() {
    static int cached = -2;
    if (cached == -2) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
            cached = -1;
        else
            cached = prop.multiProcessorCount;
    }
    return cached;
}

constexpr size_t kMaxLds = 160 * 1024;

// persistent grid: as many workgroups as fit (or blocks_per_cu per CU when set), at most `work`
template <typename K>
static unsigned grid_for(K kernel, int threads, size_t dyn_lds, uint32_t blocks_per_cu,
                         uint64_t work) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, dyn_lds) !=
            hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    if (blocks_per_cu > 0 && (int)blocks_per_cu < per_cu) per_cu = (int)blocks_per_cu;
    const int cus = device_cus();
    const uint64_t g = (uint64_t)per_cu * (uint64_t)(cus > 0 ? cus : 1);
    return (unsigned)(work < g ? work : g);
}

// Raises the kernel's dynamic-LDS limit to what this launch needs.  Returns false (with HIP's
// error state cleared) when static + dynamic LDS exceed the CU's 160 KiB.
template <typename K>
static bool allow_dynamic_lds(K kernel, size_t dyn_bytes) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(kernel)) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (attr.sharedSizeBytes + dyn_bytes > kMaxLds) return false;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(kMaxLds - attr.sharedSizeBytes)) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

static hgap_status launch_zero(uint32_t *a, uint64_t na, uint32_t *b, uint64_t nb,
                               hipStream_t s) {
    if (na + nb == 0) return HGAP_OK;
    (void)hipGetLastError();  // report only this launch's error below
    uint64_t blocks = (na + nb + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, na, b, nb);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

static uint32_t window_per_lane(const hgap_launch_config &c) {
    uint32_t w = (c.lookback_window + 63) / 64;
    return w < 1 ? 1 : (w > kMaxWindowPerLane ? kMaxWindowPerLane : w);
}

// tiles_per_range = 0: one tile per ticket and a look-back per tile; otherwise the chained
// encoder over ranges of that many tiles
template <int THREADS, int K, bool CHAINED>
static hgap_status launch_enc_kernel(const EncBuffer &b, const EncTable &t, uint32_t *counter,
                                     uint32_t max_len, const hgap_launch_config &c,
                                     uint32_t tiles_per_range, hipStream_t s) {
    constexpr uint32_t tile = THREADS * K;
    const uint32_t out_words = enc_out_words(tile, max_len);
    const size_t lds = 4ull * (out_words + enc_gap_words(tile, max_len));
    const uint64_t units =
        CHAINED ? (b.num_tiles + tiles_per_range - 1) / tiles_per_range : b.num_tiles;
    auto kern = enc_kernel<THREADS, K, CHAINED>;
    if (!allow_dynamic_lds(kern, lds)) return HGAP_ERR_UNSUPPORTED;
    const unsigned grid = grid_for(kern, THREADS, lds, c.blocks_per_cu, units);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(THREADS), lds, s, b, t, counter, out_words,
                       window_per_lane(c), tiles_per_range);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

template <int THREADS, int K>
static hgap_status launch_enc_t(const EncBuffer &b, const EncTable &t, uint32_t *counter,
                                uint32_t max_len, const hgap_launch_config &c,
                                uint32_t tiles_per_range, hipStream_t s) {
    if (tiles_per_range)
        return launch_enc_kernel<THREADS, K, true>(b, t, counter, max_len, c, tiles_per_range, s);
    return launch_enc_kernel<THREADS, K, false>(b, t, counter, max_len, c, 0u, s);
}

static hgap_status launch_enc(const EncBuffer &b, const EncTable &t, uint32_t *counter,
                              uint32_t max_len, const hgap_launch_config &c,
                              uint32_t tiles_per_range, hipStream_t s) {
#define HGAP_ENC_CASE(T, K)                                                      \
    if (c.enc_threads == T && c.enc_bytes_per_thread == K)                     \
        return launch_enc_t<T, K>(b, t, counter, max_len, c, tiles_per_range, s);
    HGAP_ENC_CASE(256, 16) HGAP_ENC_CASE(256, 32) HGAP_ENC_CASE(256, 64)
    HGAP_ENC_CASE(512, 16) HGAP_ENC_CASE(512, 32) HGAP_ENC_CASE(512, 64)
    HGAP_ENC_CASE(1024, 16) HGAP_ENC_CASE(1024, 32) HGAP_ENC_CASE(1024, 64)
#undef HGAP_ENC_CASE
    return HGAP_ERR_UNSUPPORTED;
}

// Decoder workgroup LDS: the table image (rounded up to 4 words), then the output image.
static uint32_t dec_image_offset(uint32_t table_words) { return (table_words + 3u) & ~3u; }

template <typename K>
static size_t static_lds(K kernel) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(kernel)) != hipSuccess) {
        (void)hipGetLastError();
        return kMaxLds;
    }
    return attr.sharedSizeBytes;
}

// Image region of a decoder launch for segments of up to `max_symbols` codewords: the region
// of the most workgroups of `threads` lanes per CU (32 waves per CU, LDS blocks of 512 bytes)
// that holds such a tile next to the table and the kernel's static LDS (capped at that need),
// else the largest region one workgroup can have; 0 when not even min(need, 16 * threads + 48)
// bytes fit (the shape does not fit).  Tiles whose symbols exceed the region go through it in
// windows (one more walk per window for the lanes that meet it).
static uint32_t dec_image_bytes(uint32_t threads, uint32_t table_words, uint32_t max_symbols,
                                size_t static_bytes) {
    const uint32_t need = dec_image_need(threads, max_symbols);
    const uint32_t minimum = std::min(need, 16u * threads + 48u);
    const size_t fixed = static_bytes + 4ull * dec_image_offset(table_words);
    for (uint32_t wgs = 32u / (threads / 64u); wgs >= 1; wgs--) {
        const size_t per_wg = (kMaxLds / wgs) & ~(size_t)511;
        if (per_wg < fixed + need && wgs > 1) continue;
        if (per_wg < fixed + minimum) return 0;
        const size_t budget = (per_wg - fixed) & ~(size_t)15;
        return (uint32_t)std::min<size_t>(need, budget);
    }
    return 0;
}

template <int THREADS>
static uint32_t dec_shape_image(uint32_t table_words, uint32_t max_symbols) {
    size_t st = static_lds(dec_write_kernel<THREADS, true>);
    st = std::max(st, static_lds(dec_write_kernel<THREADS, false>));
    st = std::max(st, static_lds(dec_write_batch_kernel<THREADS>));
    return dec_image_bytes(THREADS, table_words, max_symbols, st);
}

// Workspace of the single-buffer decoder, laid out for the smallest tile any shape uses:
// the offsets kernel's block ticket (16 bytes), its look-back status words, one count byte per
// segment (rounded up to whole 1024-segment tiles), one 32-bit sum per 64 of those segments,
// one 64-bit output offset per tile.
struct DecWork {
    uint32_t *counter;
    uint64_t *status;
    uint8_t *seg_counts;
    uint32_t *wave_sums;
    uint64_t *tile_offsets;
};

static uint64_t dec_max_tiles(uint64_t gap_elements) {
    return (gap_elements + kMinDecThreads - 1) / kMinDecThreads;
}
static uint64_t dec_status_bytes(uint64_t gap_elements) {
    const uint64_t blocks = (dec_max_tiles(gap_elements) + kOffsetTiles - 1) / kOffsetTiles;
    return (8 * blocks + 15) & ~15ull;
}
static uint64_t dec_count_bytes(uint64_t gap_elements) { return (gap_elements + 1023) & ~1023ull; }

static DecWork dec_work(void *ws, uint64_t gap_elements) {
    uint8_t *p = static_cast<uint8_t *>(ws);
    DecWork w;
    w.counter = reinterpret_cast<uint32_t *>(p);
    w.status = reinterpret_cast<uint64_t *>(p + 16);
    w.seg_counts = p + 16 + dec_status_bytes(gap_elements);
    w.wave_sums = reinterpret_cast<uint32_t *>(w.seg_counts + dec_count_bytes(gap_elements));
    w.tile_offsets = reinterpret_cast<uint64_t *>(w.seg_counts + dec_count_bytes(gap_elements) +
                                                  dec_count_bytes(gap_elements) / 16);
    return w;
}

static uint32_t dec_shape_image(uint32_t threads, uint32_t table_words, uint32_t max_symbols) {
    switch (threads) {
    case 256: return dec_shape_image<256>(table_words, max_symbols);
    case 512: return dec_shape_image<512>(table_words, max_symbols);
    case 1024: return dec_shape_image<1024>(table_words, max_symbols);
    default: return 0;
    }
}

template <int THREADS, bool SINGLE>
static hgap_status launch_dec_t(const DecBuffer &b, const uint32_t *table,
                                const hgap_dectable_info &info, const DecWork &w,
                                const hgap_launch_config &c, hipStream_t s) {
    const uint32_t table_words = info.bytes / 4u;
    const uint32_t image = dec_shape_image<THREADS>(table_words, info.max_symbols_per_segment);
    if (image == 0) return HGAP_ERR_UNSUPPORTED;
    const uint32_t image_offset = dec_image_offset(table_words);
    auto count = dec_count_kernel<THREADS, SINGLE>;
    auto write = dec_write_kernel<THREADS, SINGLE>;
    const size_t lds_count = 4ull * table_words;
    const size_t lds_write = 4ull * image_offset + image;
    if (!allow_dynamic_lds(count, lds_count) || !allow_dynamic_lds(write, lds_write))
        return HGAP_ERR_UNSUPPORTED;
    const DecTable t{info.prefix_bits, info.l1_entries};
    hipLaunchKernelGGL(count, dim3(grid_for(count, THREADS, lds_count, c.blocks_per_cu, b.num_tiles)),
                       dim3(THREADS), lds_count, s, b, t, table, table_words, w.seg_counts,
                       w.wave_sums);
    const unsigned blocks = (b.num_tiles + kOffsetTiles - 1) / kOffsetTiles;
    hipLaunchKernelGGL(dec_offsets_kernel<THREADS>, dim3(blocks), dim3(kOffsetThreads), 0, s,
                       w.wave_sums, b.num_tiles, w.tile_offsets, w.status, w.counter,
                       window_per_lane(c));
    hipLaunchKernelGGL(write, dim3(grid_for(write, THREADS, lds_write, c.blocks_per_cu, b.num_tiles)),
                       dim3(THREADS), lds_write, s, b, t, table, table_words, image_offset, image,
                       w.seg_counts, w.tile_offsets);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

template <int THREADS>
static hgap_status launch_dec_threads(const DecBuffer &b, const uint32_t *table,
                                      const hgap_dectable_info &info, const DecWork &w,
                                      const hgap_launch_config &c, hipStream_t s) {
    return info.l2_entries == 0 ? launch_dec_t<THREADS, true>(b, table, info, w, c, s)
                                : launch_dec_t<THREADS, false>(b, table, info, w, c, s);
}

static hgap_status launch_dec(const DecBuffer &b, const uint32_t *table,
                              const hgap_dectable_info &info, const DecWork &w,
                              const hgap_launch_config &c, hipStream_t s) {
    switch (c.dec_threads) {
    case 256: return launch_dec_threads<256>(b, table, info, w, c, s);
    case 512: return launch_dec_threads<512>(b, table, info, w, c, s);
    case 1024: return launch_dec_threads<1024>(b, table, info, w, c, s);
    default: return HGAP_ERR_UNSUPPORTED;
    }
}

static bool config_valid(const hgap_launch_config &c) {
    const bool et = c.enc_threads == 256 || c.enc_threads == 512 || c.enc_threads == 1024;
    const bool ek = c.enc_bytes_per_thread == 16 || c.enc_bytes_per_thread == 32 ||
                    c.enc_bytes_per_thread == 64;
    const bool dt = c.dec_threads == 256 || c.dec_threads == 512 || c.dec_threads == 1024;
    return et && ek && dt && c.lookback_window >= 1 && c.lookback_window <= 256 &&
           c.prefix_bits <= HGAP_MAX_CODE_LEN;
}

static hgap_launch_config resolve(const hgap_launch_config *cfg) {
    hgap_launch_config c;
    hgap_default_config(&c);
    if (cfg) c = *cfg;
    return c;
}


// ---- batched launches -----------------------------------------------------------------------
static_assert(sizeof(HistJobDev) == sizeof(hgap_hist_job), "hgap_hist_job layout");
static_assert(sizeof(EncJobDev) == sizeof(hgap_enc_job), "hgap_enc_job layout");
static_assert(sizeof(DecJobDev) == sizeof(hgap_dec_job), "hgap_dec_job layout");

template <int THREADS, int K>
static hgap_status launch_enc_batch_t(const EncJobDev *jobs, uint32_t count, uint32_t total,
                                      const uint32_t *tables, uint32_t *ws, uint32_t max_len,
                                      const hgap_launch_config &c, hipStream_t s) {
    constexpr uint32_t tile = THREADS * K;
    const uint32_t out_words = enc_out_words(tile, max_len);
    const size_t lds = 4ull * (out_words + enc_gap_words(tile, max_len));
    auto kern = enc_batch_kernel<THREADS, K>;
    if (!allow_dynamic_lds(kern, lds)) return HGAP_ERR_UNSUPPORTED;
    const unsigned grid = grid_for(kern, THREADS, lds, c.blocks_per_cu, total);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(THREADS), lds, s, jobs, count, total, tables, ws,
                       reinterpret_cast<uint64_t *>(ws + 4), out_words, window_per_lane(c));
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

static hgap_status launch_enc_batch(const EncJobDev *jobs, uint32_t count, uint32_t total,
                                    const uint32_t *tables, uint32_t *ws, uint32_t max_len,
                                    const hgap_launch_config &c, hipStream_t s) {
#define HGAP_ENC_CASE(T, K)                                                                    \
    if (c.enc_threads == T && c.enc_bytes_per_thread == K)                                   \
        return launch_enc_batch_t<T, K>(jobs, count, total, tables, ws, max_len, c, s);
    HGAP_ENC_CASE(256, 16) HGAP_ENC_CASE(256, 32) HGAP_ENC_CASE(256, 64)
    HGAP_ENC_CASE(512, 16) HGAP_ENC_CASE(512, 32) HGAP_ENC_CASE(512, 64)
    HGAP_ENC_CASE(1024, 16) HGAP_ENC_CASE(1024, 32) HGAP_ENC_CASE(1024, 64)
#undef HGAP_ENC_CASE
    return HGAP_ERR_UNSUPPORTED;
}

template <int THREADS, int K>
static hgap_status launch_enc_range_batch_t(const EncJobDev *jobs, uint32_t count,
                                            uint32_t total_ranges, const uint32_t *tables,
                                            uint32_t *ws, uint32_t max_len,
                                            const hgap_launch_config &c, hipStream_t s) {
    constexpr uint32_t tile = THREADS * K;
    const uint32_t out_words = enc_out_words(tile, max_len);
    const size_t lds = 4ull * (out_words + enc_gap_words(tile, max_len));
    auto kern = enc_range_batch_kernel<THREADS, K>;
    if (!allow_dynamic_lds(kern, lds)) return HGAP_ERR_UNSUPPORTED;
    const unsigned grid = grid_for(kern, THREADS, lds, c.blocks_per_cu, total_ranges);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(THREADS), lds, s, jobs, count, total_ranges, tables,
                       ws, reinterpret_cast<uint64_t *>(ws + 4), ws + 4 + 2ull * total_ranges,
                       out_words, window_per_lane(c), (uint32_t)(HGAP_RANGE_BYTES / tile));
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

static hgap_status launch_enc_range_batch(const EncJobDev *jobs, uint32_t count,
                                          uint32_t total_ranges, const uint32_t *tables,
                                          uint32_t *ws, uint32_t max_len,
                                          const hgap_launch_config &c, hipStream_t s) {
#define HGAP_ENC_CASE(T, K)                                                                    \
    if (c.enc_threads == T && c.enc_bytes_per_thread == K)                                   \
        return launch_enc_range_batch_t<T, K>(jobs, count, total_ranges, tables, ws, max_len,  \
                                              c, s);
    HGAP_ENC_CASE(256, 16) HGAP_ENC_CASE(256, 32) HGAP_ENC_CASE(256, 64)
    HGAP_ENC_CASE(512, 16) HGAP_ENC_CASE(512, 32) HGAP_ENC_CASE(512, 64)
    HGAP_ENC_CASE(1024, 16) HGAP_ENC_CASE(1024, 32) HGAP_ENC_CASE(1024, 64)
#undef HGAP_ENC_CASE
    return HGAP_ERR_UNSUPPORTED;
}

// Workspace of a batched decode: the offsets kernel's ticket and status words, one count byte
// per segment, one sum per 64 segments, one 64-bit prefix per tile (of the whole batch).
struct DecBatchWork {
    DecWork w;
    uint64_t bytes;
};

static DecBatchWork dec_batch_work(void *ws, uint64_t total_tiles, uint32_t threads) {
    const uint64_t blocks = (total_tiles + kOffsetTiles - 1) / kOffsetTiles;
    const uint64_t status = (8 * blocks + 15) & ~15ull;
    const uint64_t counts = (total_tiles * threads + 15) & ~15ull;
    const uint64_t sums = (total_tiles * (threads / 64) * 4 + 15) & ~15ull;
    uint8_t *p = static_cast<uint8_t *>(ws);
    DecBatchWork r;
    r.w.counter = reinterpret_cast<uint32_t *>(p);
    r.w.status = reinterpret_cast<uint64_t *>(p + 16);
    r.w.seg_counts = p + 16 + status;
    r.w.wave_sums = reinterpret_cast<uint32_t *>(p + 16 + status + counts);
    r.w.tile_offsets = reinterpret_cast<uint64_t *>(p + 16 + status + counts + sums);
    r.bytes = 16 + status + counts + sums + 8 * total_tiles;
    return r;
}

template <int THREADS>
static hgap_status launch_dec_batch_threads(const DecJobDev *jobs, uint32_t count,
                                            uint32_t total, uint32_t max_table_words,
                                            uint32_t max_symbols, const DecWork &w,
                                            const hgap_launch_config &c, hipStream_t s) {
    const uint32_t image = dec_shape_image<THREADS>(max_table_words, max_symbols);
    if (image == 0) return HGAP_ERR_UNSUPPORTED;
    const uint32_t image_offset = dec_image_offset(max_table_words);
    auto count_k = dec_count_batch_kernel<THREADS>;
    auto write_k = dec_write_batch_kernel<THREADS>;
    const size_t lds_count = 4ull * max_table_words;
    const size_t lds_write = 4ull * image_offset + image;
    if (!allow_dynamic_lds(count_k, lds_count) || !allow_dynamic_lds(write_k, lds_write))
        return HGAP_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(count_k, dim3(grid_for(count_k, THREADS, lds_count, c.blocks_per_cu, total)),
                       dim3(THREADS), lds_count, s, jobs, count, total, w.seg_counts, w.wave_sums);
    const unsigned blocks = (total + kOffsetTiles - 1) / kOffsetTiles;
    hipLaunchKernelGGL(dec_offsets_kernel<THREADS>, dim3(blocks), dim3(kOffsetThreads), 0, s,
                       w.wave_sums, total, w.tile_offsets, w.status, w.counter,
                       window_per_lane(c));
    hipLaunchKernelGGL(write_k, dim3(grid_for(write_k, THREADS, lds_write, c.blocks_per_cu, total)),
                       dim3(THREADS), lds_write, s, jobs, count, total, image_offset, image,
                       w.seg_counts, w.tile_offsets);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

template <typename K>
static bool lds_fits(K kernel, size_t dyn_bytes) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(kernel)) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.sharedSizeBytes + dyn_bytes <= kMaxLds;
}

template <int THREADS, int K>
static bool enc_shape_fits(uint32_t max_len) {
    constexpr uint32_t tile = THREADS * K;
    const size_t lds = 4ull * (enc_out_words(tile, max_len) + enc_gap_words(tile, max_len));
    return lds_fits(enc_kernel<THREADS, K, false>, lds) &&
           lds_fits(enc_batch_kernel<THREADS, K>, lds);
}

static bool enc_fits(const hgap_launch_config &c, uint32_t max_len) {
#define HGAP_ENC_CASE(T, K) \
    if (c.enc_threads == T && c.enc_bytes_per_thread == K) return enc_shape_fits<T, K>(max_len);
    HGAP_ENC_CASE(256, 16) HGAP_ENC_CASE(256, 32) HGAP_ENC_CASE(256, 64)
    HGAP_ENC_CASE(512, 16) HGAP_ENC_CASE(512, 32) HGAP_ENC_CASE(512, 64)
    HGAP_ENC_CASE(1024, 16) HGAP_ENC_CASE(1024, 32) HGAP_ENC_CASE(1024, 64)
#undef HGAP_ENC_CASE
    return false;
}

// How evenly the chained encoder's ranges fill the device at one shape: a workgroup encodes
// whole ranges, so R ranges on S resident workgroups take ceil(R / S) rounds, and the fill is
// R / (S * rounds).  Without a shape from the caller, hgap_encode_ranges takes 512 x 32 instead
// of the 256 x 32 default when that fills clearly better: 512 MiB (2048 ranges) fills 1536
// resident 256 x 32 workgroups to 67% but 768 resident 512 x 32 ones to 89%, and encoded
// 0.24-0.27 ms instead of 0.34-0.38 ms on three 512 MiB buffers.  At 1 GiB both fill 89% and
// 256 x 32 was up to 4% faster.  0 when the shape does not fit.
template <int THREADS, int K>
static double chained_fill(uint64_t n, uint32_t max_len) {
    constexpr uint32_t tile = THREADS * K;
    const size_t lds = 4ull * (enc_out_words(tile, max_len) + enc_gap_words(tile, max_len));
    auto kern = enc_kernel<THREADS, K, true>;
    if (!allow_dynamic_lds(kern, lds)) return 0.0;
    const uint64_t slots = grid_for(kern, THREADS, lds, 0, ~0ull);
    const uint64_t ranges = hgap_range_count(n);
    if (slots == 0 || ranges == 0) return 0.0;
    const uint64_t rounds = (ranges + slots - 1) / slots;
    return (double)ranges / (double)(rounds * slots);
}

static bool dec_fits(const hgap_launch_config &c, const hgap_dectable_info &info) {
    return dec_shape_image(c.dec_threads, info.bytes / 4u, info.max_symbols_per_segment) > 0;
}

}  // namespace hgap

using namespace hgap;

extern "C" {

hgap_status hgap_config_supported(const hgap_launch_config *cfg, uint32_t max_len,
                                  const hgap_dectable_info *info) {
    const hgap_launch_config c = resolve(cfg);
    if (!config_valid(c) || max_len > HGAP_MAX_CODE_LEN) return HGAP_ERR_INVALID_ARG;
    if (max_len > 0 && !enc_fits(c, max_len)) return HGAP_ERR_UNSUPPORTED;
    if (info && !dec_fits(c, *info)) return HGAP_ERR_UNSUPPORTED;
    return HGAP_OK;
}

uint32_t hgap_decode_image_bytes(uint32_t dec_threads, uint32_t table_words,
                                 uint32_t max_symbols) {
    return dec_shape_image(dec_threads, table_words, max_symbols);
}


int hgap_device_cu_count(void) { return device_cus(); }

void hgap_default_config(hgap_launch_config *cfg) {
    if (!cfg) return;
    cfg->enc_threads = kDefaultEncThreads;
    cfg->enc_bytes_per_thread = kDefaultEncBytesPerThread;
    cfg->dec_threads = kDefaultDecThreads;
    cfg->lookback_window = kDefaultLookbackWindow;
    cfg->blocks_per_cu = 0;
    cfg->prefix_bits = 0;
}

void hgap_default_chained_config(hgap_launch_config *cfg) {
    if (!cfg) return;
    hgap_default_config(cfg);
    cfg->enc_threads = kDefaultChainedThreads;
    cfg->enc_bytes_per_thread = kDefaultChainedBytesPerThread;
}

hgap_status hgap_histogram(const void *d_in, uint64_t n, uint32_t *d_hist,
                           hgap_stream_t stream) {
    if (!d_hist || (n > 0 && !d_in)) return HGAP_ERR_INVALID_ARG;
    if (!aligned16(d_in) || !aligned16(d_hist)) return HGAP_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hgap_status st = launch_zero(d_hist, 256, nullptr, 0, s);
    if (st != HGAP_OK || n == 0) return st;
    uint64_t want = ((n >> 4) + kHistThreadsSingle * 4 - 1) / (kHistThreadsSingle * 4);
    if (want < 1) want = 1;
    const unsigned blocks = grid_for(hist_kernel, kHistThreadsSingle, 0, 0, want);
    hipLaunchKernelGGL(hist_kernel, dim3(blocks), dim3(kHistThreadsSingle), 0, s,
                       static_cast<const uint8_t *>(d_in), n, d_hist);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

size_t hgap_encode_workspace_bytes(uint64_t n) {
    const uint64_t tiles = (n + kMinEncTileBytes - 1) / kMinEncTileBytes;
    return 16 + 8 * tiles;
}

uint32_t hgap_range_count(uint64_t n) {
    return (uint32_t)((n + HGAP_RANGE_BYTES - 1) / HGAP_RANGE_BYTES);
}

hgap_status hgap_histogram_ranges(const void *d_in, uint64_t n, uint32_t *d_hist,
                                  uint32_t *d_range_hist, hgap_stream_t stream) {
    if (!d_hist || (n > 0 && (!d_in || !d_range_hist))) return HGAP_ERR_INVALID_ARG;
    if (!aligned16(d_in) || !aligned16(d_hist) || !aligned16(d_range_hist))
        return HGAP_ERR_INVALID_ARG;
    if (n > 0xFFFFFFFFull) return HGAP_ERR_SIZE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hgap_status st = launch_zero(d_hist, 256, nullptr, 0, s);
    if (st != HGAP_OK || n == 0) return st;
    const uint32_t ranges = hgap_range_count(n);
    const unsigned blocks = grid_for(hist_range_kernel, kHistThreadsSingle, 0, 0, ranges);
    hipLaunchKernelGGL(hist_range_kernel, dim3(blocks), dim3(kHistThreadsSingle), 0, s,
                       static_cast<const uint8_t *>(d_in), n, d_hist, d_range_hist, ranges,
                       (uint32_t)HGAP_RANGE_BYTES);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

static_assert(HGAP_RANGE_BYTES % (1024u * 64u) == 0, "ranges hold whole tiles of every shape");

// hgap_encode_cfg, and with d_range_hist the chained encoder
static hgap_status encode_common(const void *d_in, uint64_t n, const hgap_codebook *cb,
                                 const hgap_sizes *sizes, const uint32_t *d_range_hist,
                                 uint32_t *d_payload, uint32_t *d_gap, void *d_workspace,
                                 size_t workspace_bytes, const hgap_launch_config *cfg,
                                 hgap_stream_t stream) {
    if (!cb || !sizes) return HGAP_ERR_INVALID_ARG;
    const hgap_launch_config c = resolve(cfg);
    if (!config_valid(c)) return HGAP_ERR_INVALID_ARG;
    if (n == 0) return (sizes->payload_words == 0 && sizes->gap_elements == 0) ? HGAP_OK
                                                                            : HGAP_ERR_SIZE;
    if (!d_in || !d_payload || !d_gap || !d_workspace) return HGAP_ERR_INVALID_ARG;
    if (!aligned16(d_in) || !aligned16(d_payload) || !aligned16(d_gap) ||
        !aligned16(d_workspace) || !aligned16(d_range_hist))
        return HGAP_ERR_INVALID_ARG;
    if (n > 0xFFFFFFFFull) return HGAP_ERR_SIZE;
    if (workspace_bytes < hgap_encode_workspace_bytes(n)) return HGAP_ERR_WORKSPACE;
    if ((uint64_t)sizes->payload_words * 32 < sizes->bits ||
        (sizes->bits > 0 && (uint64_t)sizes->payload_words * 32 >= sizes->bits + 32))
        return HGAP_ERR_SIZE;
    EncTable t;
    uint32_t max_len = 1;
    for (int b = 0; b < 256; b++) {
        const uint32_t l = cb->length[b];
        if (l > HGAP_MAX_CODE_LEN || (l > 0 && (cb->code[b] >> l) != 0)) return HGAP_ERR_CODEBOOK;
        t.e[b] = (cb->code[b] << 5) | l;
        if (l > max_len) max_len = l;
    }
    const uint64_t tile = (uint64_t)c.enc_threads * c.enc_bytes_per_thread;
    const uint64_t tiles = (n + tile - 1) / tile;
    const uint32_t gap_words = (uint32_t)(((uint64_t)sizes->gap_elements + 7) / 8);
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t *ws = static_cast<uint32_t *>(d_workspace);
    uint64_t *status = reinterpret_cast<uint64_t *>(ws + 4);
    uint32_t tiles_per_range = 0;
    if (d_range_hist) {
        // the prep kernel zeroes the gap array and the counter and fills one status word per
        // range (the workspace holds one per tile, at least as many)
        tiles_per_range = (uint32_t)(HGAP_RANGE_BYTES / tile);
        const uint32_t ranges = hgap_range_count(n);
        (void)hipGetLastError();
        uint64_t blocks = std::max<uint64_t>((ranges + 3) / 4, (gap_words + 255) / 256);
        blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), 2048);
        hipLaunchKernelGGL(enc_range_prep_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                           d_range_hist, ranges, t, status, ws, d_gap, gap_words);
        if (hipGetLastError() != hipSuccess) return HGAP_ERR_HIP;
    } else {
        hgap_status st = launch_zero(d_gap, gap_words, ws, 4 + 2 * tiles, s);
        if (st != HGAP_OK) return st;
    }
    EncBuffer b;
    b.in = static_cast<const uint8_t *>(d_in);
    b.n = n;
    b.payload = d_payload;
    b.payload_words = sizes->payload_words;
    b.gap_words = gap_words;
    b.gap = d_gap;
    b.status = status;
    b.num_tiles = (uint32_t)tiles;
    b.reserved = 0;
    return launch_enc(b, t, ws, max_len, c, tiles_per_range, s);
}

hgap_status hgap_encode_cfg(const void *d_in, uint64_t n, const hgap_codebook *cb,
                            const hgap_sizes *sizes, uint32_t *d_payload, uint32_t *d_gap,
                            void *d_workspace, size_t workspace_bytes,
                            const hgap_launch_config *cfg, hgap_stream_t stream) {
    return encode_common(d_in, n, cb, sizes, nullptr, d_payload, d_gap, d_workspace,
                         workspace_bytes, cfg, stream);
}

// cfg NULL: the default shape first, then smaller tiles when its LDS image does not fit
hgap_status hgap_encode_ranges(const void *d_in, uint64_t n, const hgap_codebook *cb,
                               const hgap_sizes *sizes, const uint32_t *d_range_hist,
                               uint32_t *d_payload, uint32_t *d_gap, void *d_workspace,
                               size_t workspace_bytes, const hgap_launch_config *cfg,
                               hgap_stream_t stream) {
    if (n > 0 && (!d_range_hist || !aligned16(d_range_hist))) return HGAP_ERR_INVALID_ARG;
    if (cfg)
        return encode_common(d_in, n, cb, sizes, d_range_hist, d_payload, d_gap, d_workspace,
                             workspace_bytes, cfg, stream);
    if (n < HGAP_CHAINED_MIN_BYTES)
        return hgap_encode(d_in, n, cb, sizes, d_payload, d_gap, d_workspace, workspace_bytes,
                           stream);
    hgap_launch_config c;
    hgap_default_chained_config(&c);
    if (!cb) return HGAP_ERR_INVALID_ARG;
    uint32_t max_len = 1;
    for (int b = 0; b < 256; b++) max_len = std::max<uint32_t>(max_len, cb->length[b]);
    if (max_len <= HGAP_MAX_CODE_LEN &&
        chained_fill<512, 32>(n, max_len) > chained_fill<256, 32>(n, max_len) + 0.05)
        c.enc_threads = 512;
    hgap_status st = HGAP_ERR_UNSUPPORTED;
    for (uint32_t threads = c.enc_threads; threads >= 256; threads /= 2) {
        c.enc_threads = threads;
        st = encode_common(d_in, n, cb, sizes, d_range_hist, d_payload, d_gap, d_workspace,
                           workspace_bytes, &c, stream);
        if (st != HGAP_ERR_UNSUPPORTED) break;
    }
    return st;
}

// Default shape first; when its LDS image does not fit (long codes), smaller tiles.
hgap_status hgap_encode(const void *d_in, uint64_t n, const hgap_codebook *cb,
                        const hgap_sizes *sizes, uint32_t *d_payload, uint32_t *d_gap,
                        void *d_workspace, size_t workspace_bytes, hgap_stream_t stream) {
    hgap_launch_config c;
    hgap_default_config(&c);
    hgap_status st = HGAP_ERR_UNSUPPORTED;
    for (uint32_t threads = c.enc_threads; threads >= 256; threads /= 2) {
        c.enc_threads = threads;
        st = hgap_encode_cfg(d_in, n, cb, sizes, d_payload, d_gap, d_workspace, workspace_bytes,
                             &c, stream);
        if (st != HGAP_ERR_UNSUPPORTED) break;
    }
    return st;
}

size_t hgap_decode_workspace_bytes(uint32_t gap_elements) {
    return 16 + dec_status_bytes(gap_elements) + dec_count_bytes(gap_elements) +
           dec_count_bytes(gap_elements) / 16 + 8 * dec_max_tiles(gap_elements);
}

hgap_status hgap_decode_cfg(const uint32_t *d_payload, uint32_t payload_words,
                            const uint32_t *d_gap, uint32_t gap_elements, const void *d_table,
                            const hgap_dectable_info *info, void *d_out, uint64_t n,
                            void *d_workspace, size_t workspace_bytes,
                            const hgap_launch_config *cfg, hgap_stream_t stream) {
    if (!info) return HGAP_ERR_INVALID_ARG;
    const hgap_launch_config c = resolve(cfg);
    if (!config_valid(c)) return HGAP_ERR_INVALID_ARG;
    if (n == 0) return HGAP_OK;
    if (!d_payload || !d_gap || !d_table || !d_out || !d_workspace) return HGAP_ERR_INVALID_ARG;
    if (!aligned16(d_payload) || !aligned16(d_gap) || !aligned16(d_table) || !aligned16(d_out) ||
        !aligned16(d_workspace))
        return HGAP_ERR_INVALID_ARG;
    if ((uint64_t)gap_elements != ((uint64_t)payload_words + 3) / 4) return HGAP_ERR_SIZE;
    if (info->symbol_count == 0 || info->prefix_bits < 1 || info->prefix_bits > 16 ||
        info->l1_entries != (1u << info->prefix_bits) || info->min_length < 1 ||
        info->bytes != 4u * (info->l1_entries + info->l2_entries + kDecLengthWords))
        return HGAP_ERR_CODEBOOK;
    if (workspace_bytes < hgap_decode_workspace_bytes(gap_elements)) return HGAP_ERR_WORKSPACE;
    const uint64_t tiles = ((uint64_t)gap_elements + c.dec_threads - 1) / c.dec_threads;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const DecWork w = dec_work(d_workspace, gap_elements);
    // the offsets kernel's ticket and look-back status start at zero
    const uint64_t blocks = (tiles + kOffsetTiles - 1) / kOffsetTiles;
    hgap_status st = launch_zero(w.counter, 4 + 2 * blocks, nullptr, 0, s);
    if (st != HGAP_OK) return st;
    DecBuffer b;
    b.payload = d_payload;
    b.payload_words = payload_words;
    b.gap_elements = gap_elements;
    b.gap = d_gap;
    b.out = static_cast<uint8_t *>(d_out);
    b.n = n;
    b.status = nullptr;
    b.num_tiles = (uint32_t)tiles;
    b.last_limit = dec_last_limit(payload_words, gap_elements);
    // size the LDS image for this stream's density (1.25 x its mean codewords per segment, + 2)
    // rather than the table's worst case: denser tiles go through the region in windows
    hgap_dectable_info sized = *info;
    const uint64_t expect = (5ull * n + 4ull * gap_elements - 1) / (4ull * gap_elements) + 2;
    if (expect < sized.max_symbols_per_segment) sized.max_symbols_per_segment = (uint32_t)expect;
    return launch_dec(b, static_cast<const uint32_t *>(d_table), sized, w, c, s);
}

// Default shape first; when table + staging do not fit in LDS, fewer segments per tile.
hgap_status hgap_decode(const uint32_t *d_payload, uint32_t payload_words, const uint32_t *d_gap,
                        uint32_t gap_elements, const void *d_table,
                        const hgap_dectable_info *info, void *d_out, uint64_t n,
                        void *d_workspace, size_t workspace_bytes, hgap_stream_t stream) {
    hgap_launch_config c;
    hgap_default_config(&c);
    hgap_status st = HGAP_ERR_UNSUPPORTED;
    for (uint32_t threads = c.dec_threads; threads >= 256; threads /= 2) {
        c.dec_threads = threads;
        st = hgap_decode_cfg(d_payload, payload_words, d_gap, gap_elements, d_table, info, d_out,
                             n, d_workspace, workspace_bytes, &c, stream);
        if (st != HGAP_ERR_UNSUPPORTED) break;
    }
    return st;
}


uint64_t hgap_batch_plan_histogram(hgap_hist_job *jobs, uint32_t count) {
    uint64_t base = 0;
    uint32_t ranges = 0;
    for (uint32_t j = 0; j < count; j++) {
        jobs[j].base = base;
        jobs[j].range_base = ranges;
        base += jobs[j].n;
        ranges += hgap_range_count(jobs[j].n);
    }
    return base;
}

hgap_status hgap_batch_histogram_ranges(const hgap_hist_job *d_jobs, uint32_t count,
                                        uint32_t total_ranges, uint32_t *d_range_hist,
                                        hgap_stream_t stream) {
    if (count == 0) return HGAP_OK;
    if (!d_jobs || (total_ranges > 0 && (!d_range_hist || !aligned16(d_range_hist))))
        return HGAP_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();
    const HistJobDev *jobs = reinterpret_cast<const HistJobDev *>(d_jobs);
    hipLaunchKernelGGL(hist_batch_zero_kernel, dim3(count < 4096 ? count : 4096), dim3(256), 0, s,
                       jobs, count);
    if (hipGetLastError() != hipSuccess) return HGAP_ERR_HIP;
    if (total_ranges == 0) return HGAP_OK;
    const unsigned blocks =
        grid_for(hist_range_batch_kernel, kHistThreadsSingle, 0, 0, total_ranges);
    hipLaunchKernelGGL(hist_range_batch_kernel, dim3(blocks), dim3(kHistThreadsSingle), 0, s,
                       jobs, count, total_ranges, (uint32_t)HGAP_RANGE_BYTES, d_range_hist);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

hgap_status hgap_batch_histogram(const hgap_hist_job *d_jobs, uint32_t count, uint64_t total,
                                 hgap_stream_t stream) {
    if (count == 0) return HGAP_OK;
    if (!d_jobs) return HGAP_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();
    const HistJobDev *jobs = reinterpret_cast<const HistJobDev *>(d_jobs);
    hipLaunchKernelGGL(hist_batch_zero_kernel, dim3(count < 4096 ? count : 4096), dim3(256), 0, s,
                       jobs, count);
    if (hipGetLastError() != hipSuccess) return HGAP_ERR_HIP;
    if (total == 0) return HGAP_OK;
    uint64_t want = (total + (1u << 20) - 1) >> 20;  // at least 1 MiB per workgroup
    const unsigned blocks = grid_for(hist_batch_kernel, kHistThreads, 0, 0, want ? want : 1);
    hipLaunchKernelGGL(hist_batch_kernel, dim3(blocks), dim3(kHistThreads), 0, s, jobs, count,
                       total);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

hgap_status hgap_encode_table(const hgap_codebook *cb, uint32_t table[256]) {
    if (!cb || !table) return HGAP_ERR_INVALID_ARG;
    for (int b = 0; b < 256; b++) {
        const uint32_t l = cb->length[b];
        if (l > HGAP_MAX_CODE_LEN || (l > 0 && (cb->code[b] >> l) != 0)) return HGAP_ERR_CODEBOOK;
        table[b] = (cb->code[b] << 5) | l;
    }
    return HGAP_OK;
}

uint32_t hgap_batch_plan_encode(hgap_enc_job *jobs, uint32_t count,
                                const hgap_launch_config *cfg) {
    const hgap_launch_config c = resolve(cfg);
    const uint64_t tile = (uint64_t)c.enc_threads * c.enc_bytes_per_thread;
    uint64_t base = 0;
    for (uint32_t j = 0; j < count; j++) {
        jobs[j].base = (uint32_t)base;
        jobs[j].num_tiles = (uint32_t)((jobs[j].n + tile - 1) / tile);
        base += jobs[j].num_tiles;
    }
    return (uint32_t)base;
}

uint32_t hgap_batch_plan_encode_ranges(hgap_enc_job *jobs, uint32_t count,
                                       const hgap_launch_config *cfg) {
    const hgap_launch_config c = resolve(cfg);
    const uint64_t tile = (uint64_t)c.enc_threads * c.enc_bytes_per_thread;
    uint64_t base = 0;
    for (uint32_t j = 0; j < count; j++) {
        jobs[j].base = (uint32_t)base;
        jobs[j].num_tiles = (uint32_t)((jobs[j].n + tile - 1) / tile);
        base += hgap_range_count(jobs[j].n);
    }
    return (uint32_t)base;
}

uint32_t hgap_batch_plan_decode(hgap_dec_job *jobs, uint32_t count,
                                const hgap_launch_config *cfg) {
    const hgap_launch_config c = resolve(cfg);
    uint64_t base = 0;
    for (uint32_t j = 0; j < count; j++) {
        jobs[j].base = (uint32_t)base;
        jobs[j].num_tiles = jobs[j].n ? (jobs[j].gap_elements + c.dec_threads - 1) / c.dec_threads
                                      : 0;
        base += jobs[j].num_tiles;
    }
    return (uint32_t)base;
}

size_t hgap_batch_workspace_bytes(uint32_t total_tiles) { return 16 + 8ull * total_tiles; }

size_t hgap_batch_encode_ranges_workspace_bytes(uint32_t total_ranges) {
    return 16 + 12ull * total_ranges;
}

size_t hgap_batch_decode_workspace_bytes(uint32_t total_tiles, const hgap_launch_config *cfg) {
    const hgap_launch_config c = resolve(cfg);
    if (!config_valid(c)) return 0;
    return dec_batch_work(nullptr, total_tiles, c.dec_threads).bytes;
}

hgap_status hgap_batch_encode(const hgap_enc_job *d_jobs, uint32_t count, uint32_t total_tiles,
                              uint32_t max_len, const uint32_t *d_tables, void *d_workspace,
                              size_t workspace_bytes, const hgap_launch_config *cfg,
                              hgap_stream_t stream) {
    const hgap_launch_config c = resolve(cfg);
    if (!config_valid(c)) return HGAP_ERR_INVALID_ARG;
    if (count == 0 || total_tiles == 0) return HGAP_OK;
    if (!d_jobs || !d_tables || !d_workspace || !aligned16(d_workspace))
        return HGAP_ERR_INVALID_ARG;
    if (max_len < 1 || max_len > HGAP_MAX_CODE_LEN) return HGAP_ERR_CODEBOOK;
    if (workspace_bytes < hgap_batch_workspace_bytes(total_tiles)) return HGAP_ERR_WORKSPACE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t *ws = static_cast<uint32_t *>(d_workspace);
    const EncJobDev *jobs = reinterpret_cast<const EncJobDev *>(d_jobs);
    hgap_status st = launch_zero(ws, 4 + 2ull * total_tiles, nullptr, 0, s);
    if (st != HGAP_OK) return st;
    hipLaunchKernelGGL(gap_batch_zero_kernel<EncJobDev>, dim3(count < 4096 ? count : 4096),
                       dim3(256), 0, s, jobs, count);
    if (hipGetLastError() != hipSuccess) return HGAP_ERR_HIP;
    return launch_enc_batch(jobs, count, total_tiles, d_tables, ws, max_len, c, s);
}

hgap_status hgap_batch_encode_ranges(const hgap_enc_job *d_jobs, uint32_t count,
                                     uint32_t total_ranges, uint32_t max_len,
                                     const uint32_t *d_tables, const uint32_t *d_range_hist,
                                     void *d_workspace, size_t workspace_bytes,
                                     const hgap_launch_config *cfg, hgap_stream_t stream) {
    const hgap_launch_config c = resolve(cfg);
    if (!config_valid(c)) return HGAP_ERR_INVALID_ARG;
    if (count == 0 || total_ranges == 0) return HGAP_OK;
    if (!d_jobs || !d_tables || !d_range_hist || !d_workspace || !aligned16(d_workspace) ||
        !aligned16(d_range_hist))
        return HGAP_ERR_INVALID_ARG;
    if (max_len < 1 || max_len > HGAP_MAX_CODE_LEN) return HGAP_ERR_CODEBOOK;
    if (workspace_bytes < hgap_batch_encode_ranges_workspace_bytes(total_ranges))
        return HGAP_ERR_WORKSPACE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t *ws = static_cast<uint32_t *>(d_workspace);
    const EncJobDev *jobs = reinterpret_cast<const EncJobDev *>(d_jobs);
    (void)hipGetLastError();
    hipLaunchKernelGGL(gap_batch_zero_kernel<EncJobDev>, dim3(count < 4096 ? count : 4096),
                       dim3(256), 0, s, jobs, count);
    // workspace: counter (16 bytes), a status word per range, the job of each range
    const uint32_t blocks = std::min<uint32_t>((total_ranges + 3) / 4, 4096);
    hipLaunchKernelGGL(enc_range_prep_batch_kernel, dim3(blocks), dim3(256), 0, s, jobs, count,
                       total_ranges, d_tables, d_range_hist, (uint32_t)HGAP_RANGE_BYTES,
                       reinterpret_cast<uint64_t *>(ws + 4), ws + 4 + 2ull * total_ranges, ws);
    if (hipGetLastError() != hipSuccess) return HGAP_ERR_HIP;
    return launch_enc_range_batch(jobs, count, total_ranges, d_tables, ws, max_len, c, s);
}

hgap_status hgap_batch_decode(const hgap_dec_job *d_jobs, uint32_t count, uint32_t total_tiles,
                              uint32_t max_table_words, uint32_t max_symbols, void *d_workspace,
                              size_t workspace_bytes, const hgap_launch_config *cfg,
                              hgap_stream_t stream) {
    const hgap_launch_config c = resolve(cfg);
    if (!config_valid(c)) return HGAP_ERR_INVALID_ARG;
    if (count == 0 || total_tiles == 0) return HGAP_OK;
    if (!d_jobs || !d_workspace || !aligned16(d_workspace)) return HGAP_ERR_INVALID_ARG;
    if (max_symbols < 1 || max_symbols > HGAP_SEGMENT_BITS ||
        max_table_words < 2 + kDecLengthWords)
        return HGAP_ERR_CODEBOOK;
    const DecBatchWork bw = dec_batch_work(d_workspace, total_tiles, c.dec_threads);
    if (workspace_bytes < bw.bytes) return HGAP_ERR_WORKSPACE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const DecJobDev *jobs = reinterpret_cast<const DecJobDev *>(d_jobs);
    // the offsets kernel's ticket and look-back status start at zero
    const uint64_t blocks = ((uint64_t)total_tiles + kOffsetTiles - 1) / kOffsetTiles;
    hgap_status st = launch_zero(bw.w.counter, 4 + 2 * blocks, nullptr, 0, s);
    if (st != HGAP_OK) return st;
    switch (c.dec_threads) {
    case 256:
        return launch_dec_batch_threads<256>(jobs, count, total_tiles, max_table_words,
                                             max_symbols, bw.w, c, s);
    case 512:
        return launch_dec_batch_threads<512>(jobs, count, total_tiles, max_table_words,
                                             max_symbols, bw.w, c, s);
    case 1024:
        return launch_dec_batch_threads<1024>(jobs, count, total_tiles, max_table_words,
                                              max_symbols, bw.w, c, s);
    default: return HGAP_ERR_UNSUPPORTED;
    }
}

hgap_status hgap_synth_fill(void *d_out, uint64_t n, uint64_t start, uint64_t seed,
                            const uint64_t cdf[256], const uint8_t perm[256],
                            hgap_stream_t stream) {
    if (n == 0) return HGAP_OK;
    if (!d_out || !cdf || !perm || !aligned16(d_out)) return HGAP_ERR_INVALID_ARG;
    SynthTable t;
    std::memcpy(t.cdf, cdf, sizeof(t.cdf));
    std::memcpy(t.perm, perm, sizeof(t.perm));
    uint64_t blocks = (n + 256 * 16 - 1) / (256 * 16);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(synth_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<uint8_t *>(d_out), n, start,
                       seed, t);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

}  // extern "C"
