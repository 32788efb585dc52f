// hgap_host.cpp -- host-side steps of the Huffman-with-gap-arrays path: codebook (boundary
// package-merge), output sizes, decode-table image and the stream header.
//
// Reference citations are file:line relative to the reference repository root.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "hgap.h"
#include "hgap_internal.h"

namespace hgap {

// ---------------------------------------------------------------------------------------------
// Boundary package-merge with the reference's exact tie rules
// (encoder/src/package_merge.cpp:12-80 add_node, :107-166 boundary_PM).
//
// Levels 0..15; level j emits packages, each the pair of the two lightest remaining items of
// its input (the leaves, ascending, plus level j-1's packages).  Packages are produced lazily:
// a level makes its next package only when the level above consumes the current one.  A
// package wins ties against a leaf.  Level 0 packages leaves only and packages an odd last
// leaf by itself.  The top level (15) counts leaves and consumed level-14 packages until it
// holds 2k-2 items; each symbol's length is the number of levels whose selection includes it,
// read back through the chain of "newest consumed package of the level below".
class PackageMerge {
public:
    PackageMerge() = default;
    void reset(const uint64_t *weights, int k) {
        w_ = weights;
        k_ = k;
    }

    void run(uint8_t *len) {
        for (int i = 0; i < k_; i++) len[i] = 0;
        if (k_ == 1) {  // package_merge.cpp:119-129,159 with no iterations
            len[0] = 1;
            return;
        }
        for (int j = 0; j < kLevels; j++) {
            Package &p = nodes_[j][0];
            p.weight = w_[0] + w_[1];
            p.leaves = 2;
            p.below = -1;
            head_[j] = 0;
        }
        constexpr int feed = kLevels - 2;
        int top_leaves = 2;
        int last_taken = -1;
        for (int step = 2; step < 2 * (k_ - 1); step++) {
            const int cand = head_[feed];
            const bool take_leaf = top_leaves < k_ && nodes_[feed][cand].weight > w_[top_leaves];
            if (take_leaf) {
                top_leaves++;
            } else {
                emit(feed);
                last_taken = cand;
            }
        }
        for (int i = 0; i < top_leaves && i < k_; i++) len[i]++;
        for (int level = feed, node = last_taken; node >= 0 && level >= 0; level--) {
            const Package &p = nodes_[level][node];
            for (int i = 0; i < p.leaves && i < k_; i++) len[i]++;
            node = p.below;
        }
    }

private:
    static constexpr int kLevels = HGAP_MAX_CODE_LEN;
    static constexpr int kCap = 4 * 256 + 8;
    struct Package {
        uint64_t weight;
        int leaves;  // leaves of this level used up to and including this package
        int below;   // newest level j-1 package consumed so far, or -1
    };

    // Appends level j's next package (package_merge.cpp:12-80).
    void emit(int j) {
        const Package cur = nodes_[j][head_[j]];
        Package nx{0, cur.leaves, cur.below};
        int leaf = cur.leaves;
        uint64_t sum = 0;
        if (j == 0) {
            for (int side = 0; side < 2 && leaf < k_; side++) {
                sum += w_[leaf++];
                nx.leaves++;
            }
            nx.below = -1;
        } else {
            for (int side = 0; side < 2; side++) {
                const Package &lower = nodes_[j - 1][head_[j - 1]];
                if (leaf >= k_ || lower.weight <= w_[leaf]) {
                    sum += lower.weight;
                    nx.below = head_[j - 1];
                    emit(j - 1);
                } else {
                    sum += w_[leaf++];
                    nx.leaves++;
                    if (side == 0) nx.below = cur.below;
                }
            }
        }
        nx.weight = sum;
        const int slot = head_[j] + 1;
        if (slot < kCap) {
            nodes_[j][slot] = nx;
            head_[j] = slot;
        }
    }

    const uint64_t *w_ = nullptr;
    int k_ = 0;
    int head_[kLevels];
    Package nodes_[kLevels][kCap];
};

void package_merge_lengths(const uint64_t *w, int k, uint8_t *len) {
    if (k <= 0) return;
    // one node arena per host thread, reused across calls (it is ~260 KiB: allocating it per
    // call costs more than the merge itself)
    static thread_local PackageMerge *pm = nullptr;
    if (!pm) pm = new PackageMerge();
    pm->reset(w, k);
    pm->run(len);
}

// Direct decode-table entry: sym1 | bits << 8 | sym2 << 16 | symbols << 24 (one codeword:
// bits = its length, symbols = 1; two codewords: bits = len1 + len2, symbols = 2).  The two
// output bytes sit where byte stores of the low and high halves take them.
static inline uint32_t single_entry(uint32_t sym, uint32_t len) {
    return sym | (len << 8) | (1u << 24);
}
static inline uint32_t pair_entry(uint32_t sym1, uint32_t sym2, uint32_t bits) {
    return sym1 | (bits << 8) | (sym2 << 16) | (2u << 24);
}
static inline uint32_t entry_bits(uint32_t e) { return (e >> 8) & 0xFFu; }

static bool assign_codes(hgap_codebook *cb) {
    // package_merge.cpp:168-181 / decoder/src/get_table.cpp:17-20: code_0 = 0,
    // code_{t+1} = (code_t + 1) << (len_{t+1} - len_t), in code-assignment order.
    std::memset(cb->code, 0, sizeof(cb->code));
    std::memset(cb->length, 0, sizeof(cb->length));
    uint32_t code = 0;
    bool seen[256] = {};
    for (uint32_t t = 0; t < cb->symbol_count; t++) {
        const uint32_t l = cb->order_length[t];
        const uint8_t s = cb->order_symbol[t];
        if (l == 0 || l > HGAP_MAX_CODE_LEN || seen[s]) return false;
        if ((code >> l) != 0) return false;  // ran out of codes of this length
        seen[s] = true;
        cb->code[s] = code;
        cb->length[s] = l;
        const uint32_t nl = (t + 1 < cb->symbol_count) ? cb->order_length[t + 1] : l;
        if (nl < l) return false;
        code = (code + 1u) << (nl - l);
    }
    return true;
}

}  // namespace hgap

using namespace hgap;

extern "C" {

const char *hgap_status_string(int status) {
    switch (status) {
    case HGAP_OK: return "ok";
    case HGAP_ERR_INVALID_ARG: return "invalid argument";
    case HGAP_ERR_SIZE: return "inconsistent or out-of-range size";
    case HGAP_ERR_CODEBOOK: return "invalid codebook";
    case HGAP_ERR_WORKSPACE: return "workspace too small";
    case HGAP_ERR_HIP: return "HIP runtime error";
    case HGAP_ERR_FORMAT: return "malformed stream";
    case HGAP_ERR_CAPACITY: return "buffer too small";
    case HGAP_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
    }
}

hgap_status hgap_codebook_build(const uint32_t hist[256], hgap_codebook *cb) {
    if (!hist || !cb) return HGAP_ERR_INVALID_ARG;
    std::memset(cb, 0, sizeof(*cb));
    // store_symbols (symbols.cpp:29-43) then a stable sort by count (huff.cpp:17-21,79).
    // A stable sort by count of the byte-ordered list is the sort by (count, byte): sort keys
    // count << 8 | byte.
    uint64_t key[256];
    int k = 0;
    for (int b = 0; b < 256; b++)
        if (hist[b] != 0) key[k++] = ((uint64_t)hist[b] << 8) | (uint64_t)b;
    std::sort(key, key + k);
    uint8_t sym[256];
    uint64_t cnt[256];
    for (int i = 0; i < k; i++) {
        cnt[i] = key[i] >> 8;
        sym[i] = (uint8_t)(key[i] & 0xFFu);
    }
    cb->symbol_count = (uint32_t)k;
    if (k == 0) return HGAP_OK;
    uint8_t len[256];
    package_merge_lengths(cnt, k, len);
    // Code-assignment order: the sorted list walked from its end (package_merge.cpp:172,
    // huff.cpp:113).
    for (int t = 0; t < k; t++) {
        cb->order_symbol[t] = sym[k - 1 - t];
        cb->order_length[t] = len[k - 1 - t];
    }
    return assign_codes(cb) ? HGAP_OK : HGAP_ERR_CODEBOOK;
}

// Codebooks for many histograms on a pool of host threads; each is independent and built
// exactly like hgap_codebook_build.
}  // extern "C"

namespace {

// Runs fn(i) for i in [0, count) on up to `threads` host threads (0 = all cores), at least
// kItemsPerThread items per thread: starting a thread costs more than several codebooks (a
// codebook takes about 33 us; 4 of them on 4 new threads took 540 us, on the caller's 205 us).
// Returns the first non-OK status any call returned.
constexpr uint32_t kItemsPerThread = 8;

template <typename F>
hgap_status parallel_for(uint32_t count, uint32_t threads, F fn) {
    if (count == 0) return HGAP_OK;
    const uint32_t hw = std::thread::hardware_concurrency();
    if (threads == 0) threads = hw ? hw : 1;
    const uint32_t most = (count + kItemsPerThread - 1) / kItemsPerThread;
    if (threads > most) threads = most;
    std::atomic<uint32_t> next{0};
    std::atomic<int> err{HGAP_OK};
    auto work = [&]() {
        for (uint32_t i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
            const hgap_status st = fn(i);
            if (st != HGAP_OK) {
                int expected = HGAP_OK;
                err.compare_exchange_strong(expected, st);
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    return (hgap_status)err.load();
}

}  // namespace

extern "C" {

hgap_status hgap_codebook_build_batch(const uint32_t *hists, uint32_t count, hgap_codebook *cbs,
                                      uint32_t threads) {
    if (count == 0) return HGAP_OK;
    if (!hists || !cbs) return HGAP_ERR_INVALID_ARG;
    return parallel_for(count, threads, [&](uint32_t i) {
        return hgap_codebook_build(hists + 256ull * i, &cbs[i]);
    });
}

hgap_status hgap_batch_prepare(const uint32_t *hists, uint32_t count, hgap_codebook *cbs,
                               hgap_sizes *sizes, uint32_t *enc_tables, uint32_t threads) {
    if (count == 0) return HGAP_OK;
    if (!hists || !cbs || !sizes) return HGAP_ERR_INVALID_ARG;
    return parallel_for(count, threads, [&](uint32_t i) {
        const uint32_t *h = hists + 256ull * i;
        hgap_status st = hgap_codebook_build(h, &cbs[i]);
        if (st == HGAP_OK) st = hgap_encode_sizes(&cbs[i], h, &sizes[i]);
        if (st == HGAP_OK && enc_tables) st = hgap_encode_table(&cbs[i], enc_tables + 256ull * i);
        return st;
    });
}

hgap_status hgap_dectable_build_batch(const hgap_codebook *cbs, uint32_t count,
                                      const uint32_t *prefix_bits, hgap_dectable_info *infos,
                                      void *arena, const uint64_t *offsets, size_t arena_bytes,
                                      uint32_t threads) {
    if (count == 0) return HGAP_OK;
    if (!cbs || !infos || (arena && !offsets)) return HGAP_ERR_INVALID_ARG;
    return parallel_for(count, threads, [&](uint32_t i) {
        const uint32_t p = prefix_bits ? prefix_bits[i] : 0u;
        if (!arena) return hgap_dectable_build(&cbs[i], p, &infos[i], nullptr, 0);
        if (offsets[i] > arena_bytes || (offsets[i] & 3u)) return HGAP_ERR_INVALID_ARG;
        return hgap_dectable_build(&cbs[i], p, &infos[i],
                                   static_cast<uint8_t *>(arena) + offsets[i],
                                   arena_bytes - offsets[i]);
    });
}

hgap_status hgap_codebook_from_order(const uint8_t *order_symbol, const uint8_t *order_length,
                                     uint32_t k, hgap_codebook *cb) {
    if (!cb || (k > 0 && (!order_symbol || !order_length)) || k > 256)
        return HGAP_ERR_INVALID_ARG;
    std::memset(cb, 0, sizeof(*cb));
    cb->symbol_count = k;
    for (uint32_t t = 0; t < k; t++) {
        cb->order_symbol[t] = order_symbol[t];
        cb->order_length[t] = order_length[t];
    }
    return assign_codes(cb) ? HGAP_OK : HGAP_ERR_CODEBOOK;
}

hgap_status hgap_encode_sizes(const hgap_codebook *cb, const uint32_t hist[256],
                              hgap_sizes *sizes) {
    if (!cb || !hist || !sizes) return HGAP_ERR_INVALID_ARG;
    uint64_t bits = 0;
    for (int b = 0; b < 256; b++) {
        if (hist[b] == 0) continue;
        if (cb->length[b] == 0) return HGAP_ERR_CODEBOOK;
        bits += (uint64_t)hist[b] * cb->length[b];
    }
    const uint64_t words = (bits + 31) / 32;
    const uint64_t gaps = (bits + HGAP_SEGMENT_BITS - 1) / HGAP_SEGMENT_BITS;
    if (words > 0xFFFFFFFFull) return HGAP_ERR_SIZE;
    sizes->bits = bits;
    sizes->payload_words = (uint32_t)words;
    sizes->gap_elements = (uint32_t)gaps;
    sizes->gap_words = (uint32_t)((gaps + HGAP_GAP_PER_WORD - 1) / HGAP_GAP_PER_WORD);
    sizes->reserved = 0;
    return HGAP_OK;
}

// Decode table image.  The first level is indexed by the next `prefix_bits` stream bits.  An
// entry is either a pointer (bit 31 set: width << 24 | offset) to a second-level block indexed
// by the following `width` bits, or a direct entry
//     sym1 | bits << 8 | sym2 << 16 | symbols << 24
// holding one codeword (symbols = 1, bits = its length) or, when a second complete codeword
// (sym2) also fits inside the indexed bits, two (symbols = 2, bits = len1 + len2).  The
// decoder adds `bits` to its position and `symbols` to its output count without decoding the
// fields further.  Second-level entries are single codewords.  Unused entries (incomplete
// codes) decode as the first symbol with the longest length, so decoding always advances by
// at least min_length bits.  After the two levels come kDecLengthWords words of code lengths,
// one byte per symbol value: the decoder reads the length of a pair's second codeword there
// when a segment's last step has to drop a second codeword that starts past the segment.
hgap_status hgap_dectable_build(const hgap_codebook *cb, uint32_t prefix_bits,
                                hgap_dectable_info *info, void *image, size_t image_cap) {
    if (!cb || !info) return HGAP_ERR_INVALID_ARG;
    const uint32_t k = cb->symbol_count;
    if (k == 0) {
        std::memset(info, 0, sizeof(*info));
        return HGAP_OK;
    }
    // Validate and recompute codes in assignment order.
    hgap_codebook tmp;
    hgap_status st = hgap_codebook_from_order(cb->order_symbol, cb->order_length, k, &tmp);
    if (st != HGAP_OK) return st;
    uint32_t minl = 32, maxl = 0;
    for (uint32_t t = 0; t < k; t++) {
        const uint32_t l = tmp.order_length[t];
        if (l < minl) minl = l;
        if (l > maxl) maxl = l;
    }
    uint32_t P = prefix_bits ? prefix_bits : kDefaultPrefixBits;
    if (P > HGAP_MAX_CODE_LEN) return HGAP_ERR_INVALID_ARG;
    if (P > maxl) P = maxl;
    if (P < 1) P = 1;
    const uint32_t l1 = 1u << P;
    // second-level sizing: one block per distinct P-bit prefix of long codes, wide enough for
    // the longest code sharing that prefix
    uint32_t l2 = 0;
    {
        int64_t prev_prefix = -1;
        uint32_t block_width = 0;
        for (uint32_t t = 0; t < k; t++) {
            const uint32_t l = tmp.order_length[t];
            if (l <= P) continue;
            const uint32_t c = tmp.code[tmp.order_symbol[t]];
            const int64_t prefix = c >> (l - P);
            if (prefix != prev_prefix) {
                l2 += block_width ? (1u << block_width) : 0;
                block_width = 0;
                prev_prefix = prefix;
            }
            if (l - P > block_width) block_width = l - P;
        }
        l2 += block_width ? (1u << block_width) : 0;
    }
    info->prefix_bits = P;
    info->l1_entries = l1;
    info->l2_entries = l2;
    info->bytes = 4u * (l1 + l2 + kDecLengthWords);
    info->min_length = minl;
    info->max_length = maxl;
    info->max_symbols_per_segment = (HGAP_SEGMENT_BITS + minl - 1) / minl;
    info->symbol_count = k;
    if (!image) return HGAP_OK;
    if (image_cap < info->bytes) return HGAP_ERR_CAPACITY;
    uint32_t *tab = static_cast<uint32_t *>(image);
    const uint32_t filler = single_entry(tmp.order_symbol[0], maxl);
    for (uint32_t i = 0; i < l1 + l2; i++) tab[i] = filler;
    uint32_t *sub = tab + l1;
    uint32_t next_off = 0;
    int64_t open_prefix = -1;
    uint32_t open_off = 0, open_width = 0;
    for (uint32_t t = 0; t < k; t++) {
        const uint32_t l = tmp.order_length[t];
        const uint8_t s = tmp.order_symbol[t];
        const uint32_t c = tmp.code[s];
        const uint32_t entry = single_entry(s, l);
        if (l <= P) {
            const uint32_t first = c << (P - l), count = 1u << (P - l);
            for (uint32_t e = 0; e < count; e++) tab[first + e] = entry;
            continue;
        }
        const uint32_t prefix = c >> (l - P);
        if ((int64_t)prefix != open_prefix) {
            // width of this block = longest code with this prefix (codes are length-sorted)
            uint32_t width = l - P;
            for (uint32_t u = t + 1; u < k; u++) {
                const uint32_t lu = tmp.order_length[u];
                if ((tmp.code[tmp.order_symbol[u]] >> (lu - P)) != prefix) break;
                if (lu - P > width) width = lu - P;
            }
            open_prefix = prefix;
            open_off = next_off;
            open_width = width;
            next_off += 1u << width;
            tab[prefix] = 0x80000000u | (width << 24) | open_off;
        }
        const uint32_t sl = l - P;
        const uint32_t suffix = c & ((1u << sl) - 1u);
        const uint32_t first = suffix << (open_width - sl), count = 1u << (open_width - sl);
        for (uint32_t e = 0; e < count; e++) sub[open_off + first + e] = entry;
    }
    // pair up: a direct entry whose remaining P - len1 bits hold another complete codeword
    // (a direct entry of length <= P - len1 at the left-aligned remainder) decodes both
    if (P > minl) {
        std::vector<uint32_t> one(tab, tab + l1);
        for (uint32_t idx = 0; idx < l1; idx++) {
            const uint32_t e1 = one[idx];
            if (e1 >> 31) continue;
            const uint32_t len1 = entry_bits(e1);
            if (len1 + minl > P) continue;
            const uint32_t rest = (idx << len1) & (l1 - 1u);
            const uint32_t e2 = one[rest];
            if (e2 >> 31) continue;
            const uint32_t len2 = entry_bits(e2);
            if (len1 + len2 > P) continue;
            tab[idx] = pair_entry(e1 & 0xFFu, e2 & 0xFFu, len1 + len2);
        }
    }
    uint8_t *lengths = reinterpret_cast<uint8_t *>(tab + l1 + l2);
    std::memset(lengths, 0, 4u * kDecLengthWords);
    for (uint32_t t = 0; t < k; t++) lengths[tmp.order_symbol[t]] = tmp.order_length[t];
    return HGAP_OK;
}

uint64_t hgap_stream_header_bytes(uint32_t k) { return 8ull + 2ull * k + 12ull; }

uint64_t hgap_stream_bytes(uint32_t k, uint32_t payload_words, uint32_t gap_elements) {
    return hgap_stream_header_bytes(k) +
           4ull * (((uint64_t)gap_elements + HGAP_GAP_PER_WORD - 1) / HGAP_GAP_PER_WORD) +
           4ull * payload_words;
}

static void put_le(uint8_t *p, uint64_t v, int bytes) {
    for (int i = 0; i < bytes; i++) p[i] = (uint8_t)(v >> (8 * i));
}
static uint64_t get_le(const uint8_t *p, int bytes) {
    uint64_t v = 0;
    for (int i = 0; i < bytes; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

hgap_status hgap_stream_write_header(const hgap_codebook *cb, uint32_t n,
                                     const hgap_sizes *sizes, uint8_t *dst, size_t cap,
                                     size_t *written) {
    if (!cb || !sizes || !dst) return HGAP_ERR_INVALID_ARG;
    const uint64_t need = hgap_stream_header_bytes(cb->symbol_count);
    if (cap < need) return HGAP_ERR_CAPACITY;
    uint8_t *p = dst;
    put_le(p, cb->symbol_count, 8);  // size_t symbol_count (huff.cpp:111-113)
    p += 8;
    for (uint32_t t = 0; t < cb->symbol_count; t++) {  // (symbol, length) pairs (:114-119)
        *p++ = cb->order_symbol[t];
        *p++ = cb->order_length[t];
    }
    put_le(p, n, 4);                       // inputfilesize (:121)
    put_le(p + 4, sizes->payload_words, 4);  // outputfilesize (:122)
    put_le(p + 8, sizes->gap_elements, 4);   // gap_array_elements_num (:123)
    if (written) *written = (size_t)need;
    return HGAP_OK;
}

hgap_status hgap_stream_parse_header(const uint8_t *src, size_t size, hgap_codebook *cb,
                                     uint32_t *n, hgap_sizes *sizes, uint64_t *gap_offset,
                                     uint64_t *payload_offset) {
    if (!src || !cb || !n || !sizes) return HGAP_ERR_INVALID_ARG;
    if (size < 8) return HGAP_ERR_FORMAT;
    const uint64_t k = get_le(src, 8);
    if (k > 256) return HGAP_ERR_FORMAT;
    const uint64_t hdr = hgap_stream_header_bytes((uint32_t)k);
    if (size < hdr) return HGAP_ERR_FORMAT;
    uint8_t syms[256], lens[256];
    for (uint64_t t = 0; t < k; t++) {
        syms[t] = src[8 + 2 * t];
        lens[t] = src[8 + 2 * t + 1];
    }
    if (hgap_codebook_from_order(syms, lens, (uint32_t)k, cb) != HGAP_OK) return HGAP_ERR_FORMAT;
    const uint8_t *p = src + 8 + 2 * k;
    *n = (uint32_t)get_le(p, 4);
    sizes->payload_words = (uint32_t)get_le(p + 4, 4);
    sizes->gap_elements = (uint32_t)get_le(p + 8, 4);
    sizes->gap_words = (uint32_t)(((uint64_t)sizes->gap_elements + 7) / 8);
    sizes->bits = 0;  // not stored in the stream
    sizes->reserved = 0;
    // W and G are both ceil(bits / ...) of the same bit count
    const uint64_t W = sizes->payload_words, G = sizes->gap_elements;
    if ((W + 3) / 4 != G) return HGAP_ERR_FORMAT;
    const uint64_t goff = hdr, poff = hdr + 4ull * sizes->gap_words;
    if (size < poff + 4ull * W) return HGAP_ERR_FORMAT;
    if (gap_offset) *gap_offset = goff;
    if (payload_offset) *payload_offset = poff;
    return HGAP_OK;
}

}  // extern "C"
