// hgap_cli.cpp -- command-line encoder/decoder over the native C ABI, with the reference's
// usage and file format (encoder/src/huff.cpp:23-144, decoder/src/huff.cpp:15-105).
//
//   hgap_encoder <input> <output>      compress a file
//   hgap_decoder <input> [output]      decompress; without [output] writes ./decodedfile like
//                                      the reference (decoder/src/huff.cpp:25)
//
// Built twice from this file (HGAP_CLI_DECODER selects the decoder).  Errors print a message
// and exit with status 1; timings are printed as name,ms pairs on one line.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hgap.h"

namespace {

[[noreturn]] void fail(const char *what, const char *detail) {
    fprintf(stderr, "%s: %s\n", what, detail);
    exit(EXIT_FAILURE);
}

void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(what, hipGetErrorString(e));
}

void ok(hgap_status s, const char *what) {
    if (s != HGAP_OK) fail(what, hgap_status_string(s));
}

std::vector<uint8_t> read_file(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) fail("cannot open", path);
    std::vector<uint8_t> buf;
    if (fseek(f, 0, SEEK_END) != 0) fail("cannot seek", path);
    const long size = ftell(f);
    if (size < 0) fail("cannot size", path);
    rewind(f);
    buf.resize((size_t)size);
    if (size > 0 && fread(buf.data(), 1, buf.size(), f) != buf.size()) fail("short read", path);
    fclose(f);
    return buf;
}

void write_file(const char *path, const void *data, size_t n) {
    FILE *f = fopen(path, "wb");
    if (!f) fail("cannot create", path);
    if (n > 0 && fwrite(data, 1, n, f) != n) fail("short write", path);
    if (fclose(f) != 0) fail("cannot close", path);
}

void *dmalloc(size_t n) {
    void *p = nullptr;
    hip_ok(hipMalloc(&p, n < 16 ? 16 : n), "hipMalloc");
    return p;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

[[maybe_unused]] int run_encoder(const char *in_path, const char *out_path) {
    std::vector<uint8_t> in = read_file(in_path);
    const uint64_t n = in.size();
    if (n > 0x7FFFFFFFull) fail(in_path, "larger than the format's 2 GiB limit");
    double t0 = now_ms();
    void *d_in = dmalloc(n);
    uint32_t *d_hist = static_cast<uint32_t *>(dmalloc(1024));
    hip_ok(hipMemcpy(d_in, in.data(), n, hipMemcpyHostToDevice), "hipMemcpy H2D");
    double t1 = now_ms();
    ok(hgap_histogram(d_in, n, d_hist, nullptr), "hgap_histogram");
    uint32_t hist[256];
    hip_ok(hipMemcpy(hist, d_hist, sizeof(hist), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    double t2 = now_ms();
    hgap_codebook cb;
    hgap_sizes sz;
    ok(hgap_codebook_build(hist, &cb), "hgap_codebook_build");
    ok(hgap_encode_sizes(&cb, hist, &sz), "hgap_encode_sizes");
    double t3 = now_ms();
    const size_t ws_bytes = hgap_encode_workspace_bytes(n);
    uint32_t *d_payload = static_cast<uint32_t *>(dmalloc(4ull * sz.payload_words));
    uint32_t *d_gap = static_cast<uint32_t *>(dmalloc(4ull * sz.gap_words));
    void *d_ws = dmalloc(ws_bytes);
    ok(hgap_encode(d_in, n, &cb, &sz, d_payload, d_gap, d_ws, ws_bytes, nullptr), "hgap_encode");
    hip_ok(hipDeviceSynchronize(), "encode");
    double t4 = now_ms();
    const uint64_t hdr = hgap_stream_header_bytes(cb.symbol_count);
    std::vector<uint8_t> out(hgap_stream_bytes(cb.symbol_count, sz.payload_words, sz.gap_elements));
    size_t wrote = 0;
    ok(hgap_stream_write_header(&cb, (uint32_t)n, &sz, out.data(), out.size(), &wrote),
       "hgap_stream_write_header");
    hip_ok(hipMemcpy(out.data() + hdr, d_gap, 4ull * sz.gap_words, hipMemcpyDeviceToHost),
           "hipMemcpy gap");
    hip_ok(hipMemcpy(out.data() + hdr + 4ull * sz.gap_words, d_payload, 4ull * sz.payload_words,
                     hipMemcpyDeviceToHost),
           "hipMemcpy payload");
    double t5 = now_ms();
    write_file(out_path, out.data(), out.size());
    double t6 = now_ms();
    printf("file,%s,SEGMENTSIZE,%d\n", in_path, HGAP_SEGMENT_BITS);
    printf("HtD,%f,hist,%f,dic(cpu),%f,enc,%f,DtH,%f,write(cpu),%f\n", t1 - t0, t2 - t1, t3 - t2,
           t4 - t3, t5 - t4, t6 - t5);
    printf("ratio=outputfile/inputfile,%lf,outputfilesize,%zu,inputfilesize,%llu\n",
           n ? (double)out.size() / (double)n : 0.0, out.size(), (unsigned long long)n);
    (void)hipFree(d_in);
    (void)hipFree(d_hist);
    (void)hipFree(d_payload);
    (void)hipFree(d_gap);
    (void)hipFree(d_ws);
    return 0;
}

[[maybe_unused]] int run_decoder(const char *in_path, const char *out_path) {
    std::vector<uint8_t> in = read_file(in_path);
    hgap_codebook cb;
    hgap_sizes sz;
    uint32_t n = 0;
    uint64_t goff = 0, poff = 0;
    ok(hgap_stream_parse_header(in.data(), in.size(), &cb, &n, &sz, &goff, &poff),
       "hgap_stream_parse_header");
    std::vector<uint8_t> out(n);
    if (n == 0) {
        write_file(out_path, out.data(), 0);
        return 0;
    }
    hgap_dectable_info info;
    ok(hgap_dectable_build(&cb, 0, &info, nullptr, 0), "hgap_dectable_build");
    std::vector<uint8_t> image(info.bytes);
    ok(hgap_dectable_build(&cb, 0, &info, image.data(), image.size()), "hgap_dectable_build");
    double t0 = now_ms();
    uint32_t *d_gap = static_cast<uint32_t *>(dmalloc(4ull * sz.gap_words));
    uint32_t *d_payload = static_cast<uint32_t *>(dmalloc(4ull * sz.payload_words));
    void *d_table = dmalloc(image.size());
    void *d_out = dmalloc(n);
    const size_t ws_bytes = hgap_decode_workspace_bytes(sz.gap_elements);
    void *d_ws = dmalloc(ws_bytes);
    hip_ok(hipMemcpy(d_gap, in.data() + goff, 4ull * sz.gap_words, hipMemcpyHostToDevice), "H2D");
    hip_ok(hipMemcpy(d_payload, in.data() + poff, 4ull * sz.payload_words, hipMemcpyHostToDevice),
           "H2D");
    hip_ok(hipMemcpy(d_table, image.data(), image.size(), hipMemcpyHostToDevice), "H2D");
    double t1 = now_ms();
    ok(hgap_decode(d_payload, sz.payload_words, d_gap, sz.gap_elements, d_table, &info, d_out, n,
                   d_ws, ws_bytes, nullptr),
       "hgap_decode");
    hip_ok(hipDeviceSynchronize(), "decode");
    double t2 = now_ms();
    hip_ok(hipMemcpy(out.data(), d_out, n, hipMemcpyDeviceToHost), "D2H");
    double t3 = now_ms();
    write_file(out_path, out.data(), out.size());
    printf("file,%s,HtoD,%f,dec,%f,DtoH,%f,SEGMENTSIZE,%d\n", in_path, t1 - t0, t2 - t1, t3 - t2,
           HGAP_SEGMENT_BITS);
    (void)hipFree(d_gap);
    (void)hipFree(d_payload);
    (void)hipFree(d_table);
    (void)hipFree(d_out);
    (void)hipFree(d_ws);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
#ifdef HGAP_CLI_DECODER
    if (argc < 2 || argc > 3) {
        fprintf(stderr, "usage: %s input [output]   (default output: decodedfile)\n", argv[0]);
        return EXIT_FAILURE;
    }
    return run_decoder(argv[1], argc == 3 ? argv[2] : "decodedfile");
#else
    if (argc != 3) {
        fprintf(stderr, "Please enter : %s input output\n", argv[0]);
        return EXIT_FAILURE;
    }
    return run_encoder(argv[1], argv[2]);
#endif
}
