/*
 * hgap_oracle_mt.c -- all-cores (OpenMP) variants of the oracle's histogram, encoder and
 * segment decoder.  TEST INFRASTRUCTURE ONLY (the multi-threaded CPU baseline of bench.py and
 * a checker in tests/); the product never links it.
 *
 * They restate the same reference semantics as the serial functions in hgap_oracle.c and
 * produce identical outputs (tests/test_oracle_cpu.py compares them word for word):
 *   orc_histogram_mt        encoder/src/encoder.cu:33-116 (byte counts)
 *   orc_encode_mt           encoder/src/encoder.cu:142-379 (payload words + gap nibbles)
 *   orc_decode_segments_mt  decoder/src/decoder.cu:501-569 (segment decode, concatenated)
 * Work split: the encoder cuts the input into chunks, sums each chunk's code lengths, scans
 * the sums into bit offsets and packs every chunk at its offset (words shared by two chunks are
 * zeroed first and OR-ed; gap words are OR-ed atomically).  The decoder cuts the segments into
 * ranges, decodes each range once into a private buffer and concatenates the buffers.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "hgap_oracle.h"

static int orc_threads(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

void orc_histogram_mt(const uint8_t *in, uint64_t n, uint64_t hist[256], int threads) {
    const int T = orc_threads(threads);
    uint64_t *part = (uint64_t *)calloc((size_t)T * 256, sizeof(uint64_t));
#pragma omp parallel num_threads(T)
    {
        uint64_t *h = part + 256 * (size_t)omp_get_thread_num();
#pragma omp for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; i++) h[in[i]]++;
    }
    for (int b = 0; b < 256; b++) {
        uint64_t s = 0;
        for (int t = 0; t < T; t++) s += part[256 * (size_t)t + b];
        hist[b] = s;
    }
    free(part);
}

static void orc_or_word(uint32_t *w, uint32_t v) {
    if (v) __atomic_fetch_or(w, v, __ATOMIC_RELAXED);
}

int orc_encode_mt(const uint8_t *in, uint64_t n, const uint32_t code_by_byte[256],
                  const uint32_t len_by_byte[256], uint32_t *payload, uint64_t payload_words,
                  uint32_t *gap, uint64_t gap_elements, int threads) {
    const int T = orc_threads(threads);
    const int64_t C = (int64_t)T * 8;
    const uint64_t chunk = (n + (uint64_t)C - 1) / (uint64_t)C;
    uint64_t *off = (uint64_t *)calloc((size_t)C + 1, sizeof(uint64_t));
    if (!off) return -1;
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t c = 0; c < C; c++) {
        const uint64_t a = (uint64_t)c * chunk, b = a + chunk < n ? a + chunk : n;
        uint64_t s = 0;
        for (uint64_t i = a; i < b; i++) s += len_by_byte[in[i]];
        off[c + 1] = s;
    }
    for (int64_t c = 0; c < C; c++) off[c + 1] += off[c];
    const uint64_t bits = off[C];
    if ((bits + 31) / 32 != payload_words || (bits + 127) / 128 != gap_elements) {
        free(off);
        return -3;
    }
    memset(gap, 0, ((gap_elements + 7) / 8) * sizeof(uint32_t));
    /* words a chunk shares with its neighbours: zeroed before any chunk ORs into them */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t c = 0; c < C; c++) {
        if (off[c + 1] == off[c]) continue;
        payload[off[c] >> 5] = 0;
        payload[(off[c + 1] - 1) >> 5] = 0;
    }
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t c = 0; c < C; c++) {
        const uint64_t a = (uint64_t)c * chunk, b = a + chunk < n ? a + chunk : n;
        if (off[c + 1] == off[c]) continue;
        const uint64_t first = off[c] >> 5, last = (off[c + 1] - 1) >> 5;
        uint64_t pos = off[c];
        uint64_t word = first;
        int pending = (int)(pos & 31u); /* bits of `word` before this chunk: left as zero */
        uint64_t acc = 0;
        for (uint64_t i = a; i < b; i++) {
            const uint32_t l = len_by_byte[in[i]], cw = code_by_byte[in[i]];
            const uint64_t end = pos + l;
            if ((pos >> 7) != (end >> 7) && (end & 15u)) {
                const uint64_t s = pos >> 7;
                orc_or_word(&gap[s >> 3], (uint32_t)(end & 15u) << ((s & 7u) * 4u));
            }
            acc = (acc << l) | cw;
            pending += (int)l;
            if (pending >= 32) {
                pending -= 32;
                const uint32_t v = (uint32_t)(acc >> pending);
                if (word == first || word == last)
                    orc_or_word(&payload[word], v);
                else
                    payload[word] = v;
                word++;
                acc &= pending ? ((1ull << pending) - 1ull) : 0ull;
            }
            pos = end;
        }
        if (pending > 0) orc_or_word(&payload[word], (uint32_t)(acc << (32 - pending)));
    }
    free(off);
    return 0;
}

static inline uint32_t orc_mt_bits16(const uint32_t *payload, uint64_t words, uint64_t pos) {
    const uint64_t w = pos >> 5;
    const uint32_t o = (uint32_t)(pos & 31u);
    const uint64_t hi = (w < words) ? payload[w] : 0u;
    const uint64_t lo = (w + 1 < words) ? payload[w + 1] : 0u;
    return (uint32_t)((((hi << 32) | lo) << o) >> 48);
}

int64_t orc_decode_segments_mt(const uint32_t *payload, uint64_t payload_words,
                               const uint32_t *gap, uint64_t gap_elements,
                               const uint8_t *order_sym, const uint8_t *order_len, int k,
                               uint8_t *out, uint64_t n, int threads) {
    if (k <= 0) return (n == 0 && gap_elements == 0) ? 0 : -1;
    const int T = orc_threads(threads);
    uint16_t *tab = (uint16_t *)calloc(65536, sizeof(uint16_t)); /* (len << 8) | sym */
    if (!tab) return -2;
    uint32_t code = 0;
    for (int t = 0; t < k; t++) {
        const uint32_t l = order_len[t];
        if (l == 0 || l > ORC_MAX_LEN) { free(tab); return -3; }
        const uint32_t first = code << (16 - l), count = 1u << (16 - l);
        if ((uint64_t)first + count > 65536u) { free(tab); return -3; }
        for (uint32_t e = 0; e < count; e++) tab[first + e] = (uint16_t)((l << 8) | order_sym[t]);
        const uint32_t nl = (t + 1 < k) ? order_len[t + 1] : l;
        if (nl < l) { free(tab); return -3; }
        code = (code + 1u) << (nl - l);
    }
    const int64_t R = (int64_t)T * 8;
    const uint64_t per = (gap_elements + (uint64_t)R - 1) / (uint64_t)R;
    uint8_t **buf = (uint8_t **)calloc((size_t)R, sizeof(uint8_t *));
    uint64_t *cnt = (uint64_t *)calloc((size_t)R + 1, sizeof(uint64_t));
    int bad = 0;
    const uint64_t limit_bits = payload_words * 32ull;
#pragma omp parallel for num_threads(T) schedule(dynamic, 1) reduction(| : bad)
    for (int64_t r = 0; r < R; r++) {
        const uint64_t s0 = (uint64_t)r * per;
        const uint64_t s1 = s0 + per < gap_elements ? s0 + per : gap_elements;
        if (s0 >= s1) continue;
        uint64_t cap = (s1 - s0) * 24 + 64, m = 0;
        uint8_t *o = (uint8_t *)malloc(cap);
        for (uint64_t s = s0; s < s1 && o; s++) {
            uint64_t pos = s * ORC_SEGMENT;
            if (s > 0) pos += (gap[(s - 1) >> 3] >> (((s - 1) & 7u) * 4u)) & 15u;
            const uint64_t end = (s + 1) * ORC_SEGMENT;
            if (m + ORC_SEGMENT + 1 > cap) {
                cap = 2 * cap + ORC_SEGMENT;
                o = (uint8_t *)realloc(o, cap);
                if (!o) break;
            }
            while (pos < end && pos < limit_bits) {
                const uint16_t e = tab[orc_mt_bits16(payload, payload_words, pos)];
                if (e == 0) { bad = 1; break; }
                o[m++] = (uint8_t)(e & 0xFF);
                pos += e >> 8;
            }
        }
        if (!o) bad = 1;
        buf[r] = o;
        cnt[r + 1] = m;
    }
    for (int64_t r = 0; r < R; r++) cnt[r + 1] += cnt[r];
    if (!bad) {
#pragma omp parallel for num_threads(T) schedule(static)
        for (int64_t r = 0; r < R; r++) {
            if (!buf[r] || cnt[r] >= n) continue;
            const uint64_t len = cnt[r + 1] - cnt[r];
            memcpy(out + cnt[r], buf[r], cnt[r] + len <= n ? len : n - cnt[r]);
        }
    }
    for (int64_t r = 0; r < R; r++) free(buf[r]);
    const int64_t produced = (int64_t)cnt[R];
    free(buf);
    free(cnt);
    free(tab);
    return bad ? -4 : produced;
}
