/*
 * hgap_oracle.c -- CPU restatement of the reference's histogram -> codebook -> encode ->
 * decode path, written as plain serial C.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP product, never the product.
 * Reference citations are file:line relative to the reference repository root.
 */
#include "hgap_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------ */
/* histogram: restates cu_histgram + mergeHistogram (encoder/src/encoder.cu:33-116): a count
 * of every byte value over the first n bytes. */
void orc_histogram(const uint8_t *in, uint64_t n, uint64_t hist[256]) {
    memset(hist, 0, 256 * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) hist[in[i]]++;
}

/* ------------------------------------------------------------------------------------------ */
/* symbol list: store_symbols (encoder/src/symbols.cpp:29-43) lists non-zero bins in byte
 * order, then qsort(compare) (encoder/src/huff.cpp:17-21,79) sorts by count ascending.  glibc's
 * qsort is a stable merge sort, so equal counts stay in ascending byte order.  Restated as a
 * stable insertion sort. */
int orc_sort_symbols(const uint64_t hist[256], uint8_t sym[256], uint64_t cnt[256]) {
    int k = 0;
    for (int b = 0; b < 256; b++) {
        if (hist[b] == 0) continue;
        int pos = k;
        while (pos > 0 && cnt[pos - 1] > hist[b]) {
            cnt[pos] = cnt[pos - 1];
            sym[pos] = sym[pos - 1];
            pos--;
        }
        cnt[pos] = hist[b];
        sym[pos] = (uint8_t)b;
        k++;
    }
    return k;
}

/* ------------------------------------------------------------------------------------------ */
/* Boundary package-merge, restating boundary_PM + add_node
 * (encoder/src/package_merge.cpp:12-80 and :107-166).
 *
 * Level j (0..15) keeps the list of packages it has produced so far; the newest one is the
 * "lookahead" that level j+1 may consume next.  Every package is the merge of two items taken
 * in weight order from level j's input: the leaves (sorted ascending) and level j-1's packages.
 * Ties go to the package (`package <= leaf`, package_merge.cpp:37,52,143).  Each package
 * remembers how many leaves its level has used (`used`) and the newest level j-1 package it
 * consumed (`below`).  Level 0 has no level below; an odd last leaf is packaged alone
 * (package_merge.cpp:64-75).  The top level (15) only counts: it takes 2k-2 items and the
 * lengths are read back along the `below` chain of the last package it took. */
typedef struct {
    uint64_t weight;
    int used;
    int below; /* index into level j-1's list, or -1 */
} orc_pack;

typedef struct {
    const uint64_t *w;
    int k;
    int cap;       /* nodes per level the reference allocates: 2(k-1) */
    int store;     /* nodes per level allocated here */
    orc_pack *lists[ORC_MAX_LEN];
    int newest[ORC_MAX_LEN];
    int flags;
} orc_pm_state;

static void orc_pm_produce(orc_pm_state *s, int j) {
    const orc_pack prev = s->lists[j][s->newest[j]];
    orc_pack nx;
    int leaf = prev.used;
    uint64_t a = 0, b = 0;
    nx.used = prev.used;
    if (j > 0) {
        /* first item: the lookahead package of level j-1, or the next leaf */
        const orc_pack *p = &s->lists[j - 1][s->newest[j - 1]];
        if (leaf >= s->k || p->weight <= s->w[leaf]) {
            a = p->weight;
            nx.below = s->newest[j - 1];
            orc_pm_produce(s, j - 1);
        } else {
            a = s->w[leaf++];
            nx.used++;
            nx.below = prev.below;
        }
        /* second item: compared against the (possibly new) lookahead of level j-1 */
        p = &s->lists[j - 1][s->newest[j - 1]];
        if (leaf >= s->k || p->weight <= s->w[leaf]) {
            b = p->weight;
            nx.below = s->newest[j - 1];
            orc_pm_produce(s, j - 1);
        } else {
            b = s->w[leaf++];
            nx.used++;
        }
    } else {
        if (leaf < s->k) {
            a = s->w[leaf++];
            nx.used++;
        } else {
            s->flags |= ORC_PM_UNINIT_LEFT;
        }
        if (leaf < s->k) {
            b = s->w[leaf++];
            nx.used++;
        }
        nx.below = -1;
    }
    nx.weight = a + b;
    int slot = s->newest[j] + 1;
    if (slot >= s->cap) s->flags |= ORC_PM_LIST_OVERFLOW;
    if (slot >= s->store) { /* cannot happen for k <= 256 with the allocation below */
        s->flags |= ORC_PM_LIST_OVERFLOW;
        return;
    }
    s->lists[j][slot] = nx;
    s->newest[j] = slot;
}

/* cnt[0..k-1] ascending (as orc_sort_symbols returns).  Writes len[0..k-1].  Returns 0, or -1
 * for k == 0 (the reference cannot run on an empty input: package_merge.cpp:113 allocates a
 * negative size). */
int orc_boundary_pm(const uint64_t *cnt, int k, uint8_t *len, int *flags) {
    if (flags) *flags = 0;
    if (k <= 0) return -1;
    for (int i = 0; i < k; i++) len[i] = 0;
    if (k == 1) {
        /* package_merge.cpp:119-129,159: no loop iterations, the top level's two initial
         * leaves give the only symbol length 1 */
        len[0] = 1;
        return 0;
    }
    orc_pm_state s;
    s.w = cnt;
    s.k = k;
    s.cap = 2 * (k - 1);
    s.store = 4 * k + 8;
    s.flags = 0;
    orc_pack *pool = (orc_pack *)malloc(sizeof(orc_pack) * (size_t)s.store * ORC_MAX_LEN);
    if (!pool) return -2;
    for (int j = 0; j < ORC_MAX_LEN; j++) {
        s.lists[j] = pool + (size_t)j * s.store;
        s.lists[j][0].weight = cnt[0] + cnt[1];
        s.lists[j][0].used = 2;
        s.lists[j][0].below = -1;
        s.newest[j] = 0;
    }
    const int src = ORC_MAX_LEN - 2; /* level feeding the top level */
    int leaves = 2, taken = -1;
    for (int step = 2; step < 2 * (k - 1); step++) {
        int lookahead = s.newest[src];
        if (leaves < k && !(s.lists[src][lookahead].weight <= cnt[leaves])) {
            leaves++;
        } else {
            orc_pm_produce(&s, src);
            taken = lookahead;
        }
    }
    for (int i = 0; i < leaves && i < k; i++) len[i]++;
    int level = src, node = taken;
    while (node >= 0 && level >= 0) {
        const orc_pack *p = &s.lists[level][node];
        for (int i = 0; i < p->used && i < k; i++) len[i]++;
        node = p->below;
        level--;
    }
    free(pool);
    if (flags) *flags = s.flags;
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Code assignment, restating package_merge.cpp:168-181: walk the count-ascending list from
 * its end (most frequent first); code_0 = 0 and code_{i+1} = (code_i + 1) << (len_{i+1} -
 * len_i).  The file lists symbols in that same walk order (encoder/src/huff.cpp:111-119). */
void orc_assign_codes(const uint8_t *sym, const uint8_t *len, int k,
                      uint32_t code_by_byte[256], uint32_t len_by_byte[256],
                      uint8_t *order_sym, uint8_t *order_len) {
    memset(code_by_byte, 0, 256 * sizeof(uint32_t));
    memset(len_by_byte, 0, 256 * sizeof(uint32_t));
    if (k <= 0) return;
    uint32_t code = 0;
    for (int t = 0; t < k; t++) {
        int i = k - 1 - t;
        uint32_t l = len[i];
        code_by_byte[sym[i]] = code;
        len_by_byte[sym[i]] = l;
        order_sym[t] = sym[i];
        order_len[t] = (uint8_t)l;
        uint32_t nl = (i == 0) ? l : len[i - 1];
        code = (code + 1u) << (nl - l);
    }
}

int orc_build_codebook(const uint64_t hist[256], uint32_t code_by_byte[256],
                       uint32_t len_by_byte[256], uint8_t order_sym[256],
                       uint8_t order_len[256], int *flags) {
    uint8_t sym[256], len[256];
    uint64_t cnt[256];
    int k = orc_sort_symbols(hist, sym, cnt);
    if (flags) *flags = 0;
    if (k == 0) {
        memset(code_by_byte, 0, 256 * sizeof(uint32_t));
        memset(len_by_byte, 0, 256 * sizeof(uint32_t));
        return 0;
    }
    if (orc_boundary_pm(cnt, k, len, flags) != 0) return -1;
    orc_assign_codes(sym, len, k, code_by_byte, len_by_byte, order_sym, order_len);
    return k;
}

/* ------------------------------------------------------------------------------------------ */
/* Sizes, restating get_outputfilesize (encoder/src/symbols.cpp:45-55) and the size arithmetic
 * of encoder/src/huff.cpp:84-88. */
void orc_sizes(const uint64_t hist[256], const uint32_t len_by_byte[256], uint64_t *bits,
               uint64_t *payload_words, uint64_t *gap_elements, uint64_t *gap_words) {
    uint64_t b = 0;
    for (int i = 0; i < 256; i++) b += hist[i] * (uint64_t)len_by_byte[i];
    *bits = b;
    *payload_words = (b + 31) / 32;
    *gap_elements = (b + ORC_SEGMENT - 1) / ORC_SEGMENT;
    *gap_words = (*gap_elements + ORC_GAP_PER_WORD - 1) / ORC_GAP_PER_WORD;
}

/* ------------------------------------------------------------------------------------------ */
/* Encoder, restating what cuencoder + cu_get_gaparray produce (encoder/src/encoder.cu:142-379):
 *  - payload: the codewords concatenated MSB-first into 32-bit words (bit b of the stream is
 *    bit 31-(b%32) of word b/32), tail bits zero;
 *  - gap element s (s = segment where a codeword starts, encoder.cu:307-312): when that codeword
 *    reaches or crosses the next 128-bit boundary, its end position & 15, i.e. the number of its
 *    bits that lie in segment s+1; otherwise 0.  Eight 4-bit elements per u32, low nibble first
 *    (encoder.cu:371-377). */
int orc_encode(const uint8_t *in, uint64_t n, const uint32_t code_by_byte[256],
               const uint32_t len_by_byte[256], uint32_t *payload, uint64_t payload_words,
               uint32_t *gap, uint64_t gap_elements) {
    uint64_t gap_words = (gap_elements + 7) / 8;
    memset(payload, 0, payload_words * sizeof(uint32_t));
    memset(gap, 0, gap_words * sizeof(uint32_t));
    uint64_t pos = 0;     /* bits written */
    uint64_t acc = 0;     /* pending bits, right-aligned */
    int pending = 0;
    uint64_t word = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t l = len_by_byte[in[i]];
        uint32_t c = code_by_byte[in[i]];
        if (l == 0) return -1; /* byte not in the codebook */
        uint64_t end = pos + l;
        if ((pos >> 7) != (end >> 7)) {
            uint64_t s = pos >> 7;
            if (s >= gap_elements) return -2;
            gap[s >> 3] |= (uint32_t)(end & 15u) << ((s & 7u) * 4u);
        }
        acc = (acc << l) | c;
        pending += (int)l;
        if (pending >= 32) {
            pending -= 32;
            if (word >= payload_words) return -2;
            payload[word++] = (uint32_t)(acc >> pending);
            acc &= (pending ? ((1ull << pending) - 1ull) : 0ull);
        }
        pos = end;
    }
    if (pending > 0) {
        if (word >= payload_words) return -2;
        payload[word++] = (uint32_t)(acc << (32 - pending));
    }
    return (word == payload_words) ? 0 : -3;
}

/* ------------------------------------------------------------------------------------------ */
/* Decoder, restating the segment semantics of gpu_dec_l1_l2 (decoder/src/decoder.cu:501-569):
 * segment s starts at bit 128*s + (s ? gap element s-1 : 0) and decodes every codeword whose
 * first bit lies inside the segment and inside the payload (word index < payload_words);
 * segment outputs are concatenated in order and the result truncated to n bytes
 * (decoder.cu:794 copies exactly original_size bytes).  The table is a flat 2^16-entry map
 * from the next 16 stream bits to (symbol, length), built from the file-order symbol list with
 * the canonical rule of decoder/src/get_table.cpp:17-20.
 * Returns the number of bytes produced before truncation, or a negative error. */
static inline uint32_t orc_bits16(const uint32_t *payload, uint64_t words, uint64_t pos) {
    uint64_t w = pos >> 5;
    uint32_t off = (uint32_t)(pos & 31u);
    uint64_t hi = (w < words) ? payload[w] : 0u;
    uint64_t lo = (w + 1 < words) ? payload[w + 1] : 0u;
    uint64_t v = (hi << 32) | lo;
    return (uint32_t)((v << off) >> 48);
}

int64_t orc_decode_segments(const uint32_t *payload, uint64_t payload_words, const uint32_t *gap,
                            uint64_t gap_elements, const uint8_t *order_sym,
                            const uint8_t *order_len, int k, uint8_t *out, uint64_t n,
                            uint32_t *seg_counts) {
    if (k <= 0) return (n == 0 && gap_elements == 0) ? 0 : -1;
    uint16_t *tab = (uint16_t *)calloc(65536, sizeof(uint16_t)); /* (len << 8) | sym, 0 = none */
    if (!tab) return -2;
    uint32_t code = 0;
    for (int t = 0; t < k; t++) {
        uint32_t l = order_len[t];
        if (l == 0 || l > ORC_MAX_LEN) { free(tab); return -3; }
        uint32_t first = code << (16 - l), count = 1u << (16 - l);
        if ((uint64_t)first + count > 65536u) { free(tab); return -3; }
        for (uint32_t e = 0; e < count; e++) tab[first + e] = (uint16_t)((l << 8) | order_sym[t]);
        uint32_t nl = (t + 1 < k) ? order_len[t + 1] : l;
        if (nl < l) { free(tab); return -3; }
        code = (code + 1u) << (nl - l);
    }
    uint64_t produced = 0;
    const uint64_t limit_bits = payload_words * 32ull;
    for (uint64_t s = 0; s < gap_elements; s++) {
        uint64_t start = 0;
        if (s > 0) start = (gap[(s - 1) >> 3] >> (((s - 1) & 7u) * 4u)) & 15u;
        uint64_t pos = s * ORC_SEGMENT + start;
        uint64_t end = (s + 1) * ORC_SEGMENT;
        uint32_t c = 0;
        while (pos < end && pos < limit_bits) {
            uint16_t e = tab[orc_bits16(payload, payload_words, pos)];
            if (e == 0) { free(tab); return -4; }
            if (produced < n) out[produced] = (uint8_t)(e & 0xFF);
            produced++;
            c++;
            pos += e >> 8;
        }
        if (seg_counts) seg_counts[s] = c;
    }
    free(tab);
    return (int64_t)produced;
}

/* ------------------------------------------------------------------------------------------ */
/* Stream layout, restating the writer encoder/src/huff.cpp:111-125 and the reader
 * decoder/src/huff.cpp:27-39,71-81: u64 k; k x {u8 symbol, u8 length}; u32 n; u32 W; u32 G;
 * u32 gap[ceil(G/8)]; u32 payload[W] -- all little-endian. */
int64_t orc_stream_size(int k, uint64_t payload_words, uint64_t gap_elements) {
    return 8 + 2 * (int64_t)k + 12 + 4 * (int64_t)((gap_elements + 7) / 8) +
           4 * (int64_t)payload_words;
}

static void orc_put32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}
static uint32_t orc_get32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

int64_t orc_stream_write(uint8_t *dst, uint64_t cap, int k, const uint8_t *order_sym,
                         const uint8_t *order_len, uint32_t n, uint32_t payload_words,
                         uint32_t gap_elements, const uint32_t *gap, const uint32_t *payload) {
    int64_t size = orc_stream_size(k, payload_words, gap_elements);
    if ((uint64_t)size > cap) return -1;
    uint8_t *p = dst;
    uint64_t kk = (uint64_t)k;
    for (int i = 0; i < 8; i++) *p++ = (uint8_t)(kk >> (8 * i));
    for (int i = 0; i < k; i++) { *p++ = order_sym[i]; *p++ = order_len[i]; }
    orc_put32(p, n); p += 4;
    orc_put32(p, payload_words); p += 4;
    orc_put32(p, gap_elements); p += 4;
    uint32_t gw = (gap_elements + 7) / 8;
    for (uint32_t i = 0; i < gw; i++) { orc_put32(p, gap[i]); p += 4; }
    for (uint32_t i = 0; i < payload_words; i++) { orc_put32(p, payload[i]); p += 4; }
    return size;
}

int orc_stream_parse(const uint8_t *src, uint64_t size, int *k, uint8_t *order_sym,
                     uint8_t *order_len, uint32_t *n, uint32_t *payload_words,
                     uint32_t *gap_elements, uint64_t *gap_offset, uint64_t *payload_offset) {
    if (size < 8) return -1;
    uint64_t kk = 0;
    for (int i = 0; i < 8; i++) kk |= (uint64_t)src[i] << (8 * i);
    if (kk > 256) return -2;
    uint64_t off = 8;
    if (size < off + 2 * kk + 12) return -1;
    for (uint64_t i = 0; i < kk; i++) { order_sym[i] = src[off++]; order_len[i] = src[off++]; }
    *k = (int)kk;
    *n = orc_get32(src + off); off += 4;
    *payload_words = orc_get32(src + off); off += 4;
    *gap_elements = orc_get32(src + off); off += 4;
    uint64_t gw = ((uint64_t)*gap_elements + 7) / 8;
    *gap_offset = off;
    *payload_offset = off + 4 * gw;
    if (size < *payload_offset + 4ull * *payload_words) return -1;
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Counter-based synthetic data: element i of stream `seed` is the splitmix64 output
 * finalize(seed + (i+1) * golden); its top 53 bits pick a rank by inverse CDF over cdf[256]
 * (first rank whose threshold exceeds the value; thresholds are cumulative probabilities in
 * units of 2^-53, last = 2^53) and the byte is perm[rank]. */
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_synth_fill(uint8_t *out, uint64_t n, uint64_t start, uint64_t seed,
                    const uint64_t cdf[256], const uint8_t perm[256]) {
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r = orc_splitmix64_at(seed, start + i) >> 11;
        int lo = 0, hi = 255;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (r < cdf[mid]) hi = mid; else lo = mid + 1;
        }
        out[i] = perm[lo];
    }
}
