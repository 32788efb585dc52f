/*
 * hgap_compat.h -- the reference's own entry points, exported under their reference names by
 * libhgap.so, so a program written against the reference (encoder/src/huff.cpp,
 * decoder/src/huff.cpp) links against this library instead of the CUDA objects.
 *
 * Semantics follow the reference: host buffers where it takes host buffers, blocking calls on
 * the default stream, and on a device error a message on stdout and exit(EXIT_FAILURE)
 * (encoder/include/encoder.cuh:4-10).  New code should use the asynchronous API in hgap.h.
 */
#ifndef HGAP_COMPAT_H
#define HGAP_COMPAT_H

/* The reference defines these in encoder/include/constants.hpp:34-43 and
 * decoder/include/constants.hpp:42-57 under the include guard CONSTANTS_; the layouts below
 * are identical, and are skipped when a reference constants.hpp was included first. */
#ifndef CONSTANTS_
struct Symbol {
    unsigned char symbol;
    unsigned char length;
    unsigned int num;
};
struct Codetable {
    unsigned int code;
    unsigned int length;
};
struct TableInfo {
    unsigned int l1table_size;
    unsigned int l2table_size;
    unsigned int ptrtable_size;
};
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* encoder/src/encoder.cu:118-121.  d_Data: device bytes (as u32 words), byteCount bytes;
 * d_Histogram: device u32[256].  Blocking. */
void histgram(unsigned int *d_Histogram, unsigned int *d_Data, unsigned int byteCount);

/* encoder/src/encoder.cu:382-388.  outputfile: host buffer of outputfilesize payload words
 * followed by ceil(gap_array_elements_num / 8) gap words (encoder.cu:399).  d_inputfile:
 * device input of inputfilesize bytes; as in the reference (encoder.cu:451) this call frees
 * it.  codetable: host Codetable[256] by byte value. */
void encode(unsigned int *outputfile, unsigned int outputfilesize, unsigned int *d_inputfile,
            unsigned int inputfilesize, unsigned int gap_array_elements_num,
            struct Codetable *codetable);

/* encoder/src/symbols.cpp:29-43: non-zero bins, in byte order; returns their number. */
int store_symbols(unsigned int *num_of_symbols, struct Symbol *symbols);

/* encoder/src/symbols.cpp:45-55: sum of length x count over the first symbol_count entries. */
unsigned long long get_outputfilesize(struct Symbol *symbols, unsigned int symbol_count);

/* encoder/src/package_merge.cpp:107-182: symbols sorted by count ascending (the caller's
 * qsort, encoder/src/huff.cpp:79); adds each symbol's code length to symbols[i].length and
 * fills codetable[byte]. */
void boundary_PM(struct Symbol *symbols, unsigned int symbol_count, struct Codetable *codetable);

/* decoder/src/get_table.cpp:3-46 (the reference passes TableInfo by reference).  symbols in
 * stream order.  Returns the first-level width; fills the three table sizes. */
unsigned int get_table_info(struct Symbol *symbols, int symbol_count, int prefix_bit,
                            struct TableInfo *tableinfo);

/* decoder/src/get_table.cpp:48-139: writes the stage table
 * [u32 ptr[ptrtable_size]][u8 length[256]][u8 l1[l1table_size]][u8 l2[l2table_size]]
 * and returns the first second-level prefix (0 when every code fits the first level). */
unsigned int get_twolevel_table(void *stage_table, int prefix_bit, struct Symbol *symbols,
                                int symbol_count, struct TableInfo tableinfo);

/* decoder/src/decoder.cu:732-743.  input: host [gap words][payload words], inputfilesize =
 * payload words; output: host buffer of >= outputfilesize bytes; dectable: host stage table of
 * tablesize bytes from get_twolevel_table.  Prints "HtoD,<ms>, dec,<ms>, DtoH,<ms>,..." like
 * the reference. */
void decoder_l1_l2(unsigned int *input, int inputfilesize, unsigned int *output,
                   int outputfilesize, int gap_element_num, void *dectable, int tablesize,
                   unsigned int prefix_bit, unsigned int symbol_count,
                   struct TableInfo tableinfo);

#ifdef __cplusplus
}
/* decoder/src/get_table.cpp:3-7 signature, for C++ callers written against the reference */
inline unsigned int get_table_info(struct Symbol *symbols, int symbol_count, int prefix_bit,
                                   struct TableInfo &tableinfo) {
    return get_table_info(symbols, symbol_count, prefix_bit, &tableinfo);
}
#endif

#endif /* HGAP_COMPAT_H */
