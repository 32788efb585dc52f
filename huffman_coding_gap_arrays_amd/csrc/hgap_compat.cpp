// hgap_compat.cpp -- the reference's entry points under their own names, implemented on top of
// the asynchronous C ABI.  Blocking, default stream, exit(EXIT_FAILURE) on device errors like
// the reference's CUERROR macro (encoder/include/encoder.cuh:4-10).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hgap.h"
#include "hgap_compat.h"
#include "hgap_internal.h"

namespace {

void die(const char *what, const char *detail, int line) {
    printf("Error: %s : %s at %s line %d\n", what, detail, __FILE__, line);
    fflush(stdout);
    exit(EXIT_FAILURE);
}

#define COMPAT_HIP(call)                                                     \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) die(#call, hipGetErrorString(e_), __LINE__);  \
    } while (0)

#define COMPAT_OK(call)                                                      \
    do {                                                                     \
        hgap_status s_ = (call);                                             \
        if (s_ != HGAP_OK) die(#call, hgap_status_string(s_), __LINE__);     \
    } while (0)

// canonical codes in stream order (decoder/src/get_table.cpp:17-20)
bool stream_codes(const Symbol *symbols, int k, uint32_t *codes) {
    uint32_t code = 0;
    for (int i = 0; i < k; i++) {
        const uint32_t l = symbols[i].length;
        const uint32_t nl = (i + 1 < k) ? symbols[i + 1].length : l;
        if (l == 0 || l > HGAP_MAX_CODE_LEN || nl < l) return false;
        codes[i] = code;
        code = (code + 1u) << (nl - l);
    }
    return true;
}

// Reads a stage table in the layout get_twolevel_table writes back into a codebook: first-level
// entry idx holds a symbol whose code is idx >> (prefix_bit - length); pointer j covers prefix
// l1table_size + j.
hgap_status stage_table_codebook(const uint8_t *stage, int tablesize, uint32_t prefix_bit,
                                 uint32_t symbol_count, const TableInfo &info,
                                 hgap_codebook *cb) {
    const uint64_t need = 4ull * info.ptrtable_size + 256ull + info.l1table_size +
                          info.l2table_size;
    if (need > (uint64_t)tablesize || prefix_bit < 1 || prefix_bit > HGAP_MAX_CODE_LEN)
        return HGAP_ERR_FORMAT;
    const uint8_t *lens = stage + 4ull * info.ptrtable_size;
    const uint8_t *l1 = lens + 256;
    const uint8_t *l2 = l1 + info.l1table_size;
    struct Rec {
        uint32_t len, code;
        uint8_t sym;
    };
    std::vector<Rec> recs;
    bool seen[256] = {};
    const uint32_t P = prefix_bit;
    const uint32_t l1n = std::min<uint32_t>(info.l1table_size, 1u << P);
    for (uint32_t idx = 0; idx < l1n; idx++) {
        const uint8_t s = l1[idx];
        const uint32_t L = lens[s];
        if (L == 0 || L > P) return HGAP_ERR_FORMAT;
        if ((idx & ((1u << (P - L)) - 1u)) == 0 && !seen[s]) {
            seen[s] = true;
            recs.push_back({L, idx >> (P - L), s});
        }
    }
    for (uint32_t j = 0; j < info.ptrtable_size; j++) {
        uint32_t v;
        std::memcpy(&v, stage + 4ull * j, 4);
        const uint32_t width = v >> 16, off = v & 0xFFFFu;
        if (width < 1 || width > HGAP_MAX_CODE_LEN - P ||
            (uint64_t)off + (1u << width) > info.l2table_size)
            return HGAP_ERR_FORMAT;
        const uint32_t prefix = info.l1table_size + j;
        for (uint32_t x = 0; x < (1u << width); x++) {
            const uint8_t s = l2[off + x];
            const uint32_t L = lens[s];
            if (L <= P || L - P > width) return HGAP_ERR_FORMAT;
            const uint32_t sl = L - P;
            if ((x & ((1u << (width - sl)) - 1u)) == 0 && !seen[s]) {
                seen[s] = true;
                recs.push_back({L, (prefix << sl) | (x >> (width - sl)), s});
            }
        }
    }
    if (recs.size() != symbol_count) return HGAP_ERR_FORMAT;
    std::sort(recs.begin(), recs.end(), [](const Rec &a, const Rec &b) {
        return a.len != b.len ? a.len < b.len : a.code < b.code;
    });
    uint8_t osym[256], olen[256];
    for (size_t t = 0; t < recs.size(); t++) {
        osym[t] = recs[t].sym;
        olen[t] = (uint8_t)recs[t].len;
    }
    hgap_status st = hgap_codebook_from_order(osym, olen, symbol_count, cb);
    if (st != HGAP_OK) return st;
    for (const Rec &r : recs)
        if (cb->code[r.sym] != r.code) return HGAP_ERR_FORMAT;
    return HGAP_OK;
}

}  // namespace

extern "C" {

void histgram(unsigned int *d_Histogram, unsigned int *d_Data, unsigned int byteCount) {
    COMPAT_OK(hgap_histogram(d_Data, byteCount, d_Histogram, nullptr));
    COMPAT_HIP(hipDeviceSynchronize());
}

void encode(unsigned int *outputfile, unsigned int outputfilesize, unsigned int *d_inputfile,
            unsigned int inputfilesize, unsigned int gap_array_elements_num,
            struct Codetable *codetable) {
    hgap_codebook cb;
    std::memset(&cb, 0, sizeof(cb));
    for (int b = 0; b < 256; b++) {
        cb.code[b] = codetable[b].code;
        cb.length[b] = codetable[b].length;
    }
    hgap_sizes sz;
    sz.payload_words = outputfilesize;
    sz.gap_elements = gap_array_elements_num;
    sz.gap_words = (gap_array_elements_num + 7) / 8;
    sz.bits = 32ull * outputfilesize;  // exact bit count is not passed by the reference caller
    sz.reserved = 0;
    const size_t ws_bytes = hgap_encode_workspace_bytes(inputfilesize);
    uint32_t *d_payload = nullptr, *d_gap = nullptr;
    void *d_ws = nullptr;
    COMPAT_HIP(hipMalloc(&d_payload, 4ull * std::max(outputfilesize, 1u)));
    COMPAT_HIP(hipMalloc(&d_gap, 4ull * std::max(sz.gap_words, 1u)));
    COMPAT_HIP(hipMalloc(&d_ws, ws_bytes));
    COMPAT_OK(hgap_encode(d_inputfile, inputfilesize, &cb, &sz, d_payload, d_gap, d_ws, ws_bytes,
                          nullptr));
    COMPAT_HIP(hipMemcpy(outputfile, d_payload, 4ull * outputfilesize, hipMemcpyDeviceToHost));
    COMPAT_HIP(hipMemcpy(outputfile + outputfilesize, d_gap, 4ull * sz.gap_words,
                         hipMemcpyDeviceToHost));
    COMPAT_HIP(hipFree(d_inputfile));  // ownership passes to encode (encoder.cu:451)
    COMPAT_HIP(hipFree(d_payload));
    COMPAT_HIP(hipFree(d_gap));
    COMPAT_HIP(hipFree(d_ws));
}

int store_symbols(unsigned int *num_of_symbols, struct Symbol *symbols) {
    int k = 0;
    for (int b = 0; b < 256; b++) {
        if (num_of_symbols[b] == 0) continue;
        symbols[k].symbol = (unsigned char)b;
        symbols[k].num = num_of_symbols[b];
        k++;
    }
    return k;
}

unsigned long long get_outputfilesize(struct Symbol *symbols, unsigned int symbol_count) {
    unsigned long long bits = 0;
    for (unsigned int i = 0; i < symbol_count; i++)
        bits += (unsigned long long)symbols[i].length * symbols[i].num;
    return bits;
}

void boundary_PM(struct Symbol *symbols, unsigned int symbol_count,
                 struct Codetable *codetable) {
    if (symbol_count == 0 || symbol_count > 256) return;
    uint64_t w[256];
    uint8_t len[256];
    for (unsigned int i = 0; i < symbol_count; i++) w[i] = symbols[i].num;
    hgap::package_merge_lengths(w, (int)symbol_count, len);
    for (unsigned int i = 0; i < symbol_count; i++) symbols[i].length += len[i];
    // package_merge.cpp:168-181: codes from the most frequent (last) symbol down
    uint32_t code = 0;
    for (int i = (int)symbol_count - 1; i >= 0; i--) {
        const uint32_t l = symbols[i].length;
        codetable[symbols[i].symbol].code = code;
        codetable[symbols[i].symbol].length = l;
        const uint32_t nl = (i == 0) ? l : symbols[i - 1].length;
        code = (code + 1u) << (nl - l);
    }
}

unsigned int get_table_info(struct Symbol *symbols, int symbol_count, int prefix_bit,
                            struct TableInfo *tableinfo) {
    tableinfo->l1table_size = tableinfo->l2table_size = tableinfo->ptrtable_size = 0;
    if (symbol_count <= 0 || symbol_count > 256) return (unsigned int)prefix_bit;
    uint32_t codes[256];
    if (!stream_codes(symbols, symbol_count, codes)) return (unsigned int)prefix_bit;
    const int P = prefix_bit;
    int64_t group = -1;
    uint32_t group_width = 0;
    for (int i = 0; i < symbol_count; i++) {
        const int L = symbols[i].length;
        if (L <= P) continue;
        const int64_t prefix = codes[i] >> (L - P);
        if (tableinfo->ptrtable_size == 0) tableinfo->l1table_size = (uint32_t)prefix;
        if (prefix != group) {
            if (group >= 0) tableinfo->l2table_size += 1u << group_width;
            tableinfo->ptrtable_size++;
            group = prefix;
            group_width = 0;
        }
        group_width = std::max<uint32_t>(group_width, (uint32_t)(L - P));
    }
    if (group >= 0) {
        tableinfo->l2table_size += 1u << group_width;
        return (unsigned int)P;
    }
    // no code longer than prefix_bit: one level indexed by the longest length
    // (get_table.cpp:40-44)
    const int maxl = symbols[symbol_count - 1].length;
    tableinfo->l1table_size = 1u << maxl;
    return (unsigned int)maxl;
}

unsigned int get_twolevel_table(void *stage_table, int prefix_bit, struct Symbol *symbols,
                                int symbol_count, struct TableInfo tableinfo) {
    uint8_t *base = static_cast<uint8_t *>(stage_table);
    uint8_t *lens = base + 4ull * tableinfo.ptrtable_size;
    uint8_t *l1 = lens + 256;
    uint8_t *l2 = l1 + tableinfo.l1table_size;
    std::memset(lens, 0, 256);
    if (symbol_count <= 0 || symbol_count > 256) return 0;
    uint32_t codes[256];
    if (!stream_codes(symbols, symbol_count, codes)) return 0;
    const int P = prefix_bit;
    unsigned int boundary = 0;
    uint32_t ptr = 0, off = 0;
    for (int i = 0; i < symbol_count;) {
        const int L = symbols[i].length;
        const uint8_t s = symbols[i].symbol;
        lens[s] = (uint8_t)L;
        if (L <= P) {
            const uint32_t first = codes[i] << (P - L), count = 1u << (P - L);
            std::memset(l1 + first, s, count);
            i++;
            continue;
        }
        const uint32_t prefix = codes[i] >> (L - P);
        if (boundary == 0) boundary = prefix;
        int end = i;
        uint32_t width = 0;
        while (end < symbol_count && (codes[end] >> (symbols[end].length - P)) == prefix) {
            width = std::max<uint32_t>(width, (uint32_t)(symbols[end].length - P));
            end++;
        }
        for (int t = i; t < end; t++) {
            const uint32_t sl = (uint32_t)(symbols[t].length - P);
            lens[symbols[t].symbol] = symbols[t].length;
            const uint32_t suffix = codes[t] & ((1u << sl) - 1u);
            std::memset(l2 + off + (suffix << (width - sl)), symbols[t].symbol,
                        1u << (width - sl));
        }
        const uint32_t v = (width << 16) | off;
        std::memcpy(base + 4ull * ptr, &v, 4);
        ptr++;
        off += 1u << width;
        i = end;
    }
    return boundary;
}

void decoder_l1_l2(unsigned int *input, int inputfilesize, unsigned int *output,
                   int outputfilesize, int gap_element_num, void *dectable, int tablesize,
                   unsigned int prefix_bit, unsigned int symbol_count,
                   struct TableInfo tableinfo) {
    hgap_codebook cb;
    COMPAT_OK(stage_table_codebook(static_cast<const uint8_t *>(dectable), tablesize, prefix_bit,
                                   symbol_count, tableinfo, &cb));
    hgap_dectable_info info;
    COMPAT_OK(hgap_dectable_build(&cb, 0, &info, nullptr, 0));
    std::vector<uint8_t> image(info.bytes);
    COMPAT_OK(hgap_dectable_build(&cb, 0, &info, image.data(), image.size()));
    const uint32_t W = (uint32_t)inputfilesize, G = (uint32_t)gap_element_num;
    const uint32_t gap_words = (G + 7) / 8;
    const uint64_t n = (uint64_t)(uint32_t)outputfilesize;
    const size_t ws_bytes = hgap_decode_workspace_bytes(G);
    uint32_t *d_payload = nullptr, *d_gap = nullptr;
    void *d_table = nullptr, *d_out = nullptr, *d_ws = nullptr;
    COMPAT_HIP(hipMalloc(&d_payload, 4ull * std::max(W, 1u)));
    COMPAT_HIP(hipMalloc(&d_gap, 4ull * std::max(gap_words, 1u)));
    COMPAT_HIP(hipMalloc(&d_table, std::max<size_t>(image.size(), 16)));
    COMPAT_HIP(hipMalloc(&d_out, std::max<uint64_t>(n, 16)));
    COMPAT_HIP(hipMalloc(&d_ws, ws_bytes));
    hipEvent_t ev[4];
    for (auto &e : ev) COMPAT_HIP(hipEventCreate(&e));
    COMPAT_HIP(hipEventRecord(ev[0], nullptr));
    COMPAT_HIP(hipMemcpy(d_gap, input, 4ull * gap_words, hipMemcpyHostToDevice));
    COMPAT_HIP(hipMemcpy(d_payload, input + gap_words, 4ull * W, hipMemcpyHostToDevice));
    COMPAT_HIP(hipMemcpy(d_table, image.data(), image.size(), hipMemcpyHostToDevice));
    COMPAT_HIP(hipEventRecord(ev[1], nullptr));
    COMPAT_OK(hgap_decode(d_payload, W, d_gap, G, d_table, &info, d_out, n, d_ws, ws_bytes,
                          nullptr));
    COMPAT_HIP(hipEventRecord(ev[2], nullptr));
    COMPAT_HIP(hipMemcpy(output, d_out, n, hipMemcpyDeviceToHost));
    COMPAT_HIP(hipEventRecord(ev[3], nullptr));
    COMPAT_HIP(hipEventSynchronize(ev[3]));
    float t[3];
    for (int i = 0; i < 3; i++) COMPAT_HIP(hipEventElapsedTime(&t[i], ev[i], ev[i + 1]));
    printf("HtoD,%f, dec,%f, DtoH,%f", t[0], t[1], t[2]);
    printf(",SEGMENTSIZE,%d,THREAD_NUM,%u\n", HGAP_SEGMENT_BITS, hgap::kDefaultDecThreads);
    for (auto &e : ev) COMPAT_HIP(hipEventDestroy(e));
    COMPAT_HIP(hipFree(d_payload));
    COMPAT_HIP(hipFree(d_gap));
    COMPAT_HIP(hipFree(d_table));
    COMPAT_HIP(hipFree(d_out));
    COMPAT_HIP(hipFree(d_ws));
}

}  // extern "C"
