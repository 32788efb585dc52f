// ref_driver.cpp -- a thin extern "C" driver around the reference's own host C++ files
// (encoder/src/package_merge.cpp, encoder/src/symbols.cpp, decoder/src/get_table.cpp),
// compiled from where they lie by oracle/Makefile.ref into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: used to write golden vectors under tests/golden/ and to check the
// oracle's restatement.  The reference's own main() (encoder/src/huff.cpp) cannot be compiled
// here (it needs cuda.h), so the two lines of it this driver mirrors are restated:
//   huff.cpp:17-21 compare() = P->num - Q->num, and huff.cpp:78-79 store_symbols + qsort.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>

#include "constants.hpp"      // encoder/include (struct Symbol, struct Codetable)
#include "package_merge.hpp"  // encoder/include
#include "symbols.hpp"        // encoder/include

extern "C" {

// The link wraps malloc (ld --wrap=malloc), so package_merge.cpp:113, the reference's only
// allocation, lands here.  The reference never frees it.  The fill byte lets a test show that
// results do not depend on the contents of memory boundary_PM reads before writing.  The
// allocation gets 4 KiB of slack: for one distinct symbol the reference asks for 0 bytes and
// then writes one 48-byte node (package_merge.cpp:113-132).
static unsigned char g_fill = 0;
void ref_set_fill(int byte) { g_fill = (unsigned char)byte; }

void *__real_malloc(size_t bytes);
void *__wrap_malloc(size_t bytes) {
    void *p = __real_malloc(bytes + 4096);
    if (p) std::memset(p, g_fill, bytes + 4096);
    return p;
}

static int ref_compare(const void *p, const void *q) {
    const Symbol *P = (const Symbol *)p;
    const Symbol *Q = (const Symbol *)q;
    return P->num - Q->num;
}

// hist[256] -> k; symbols in sorted (count-ascending) order with their lengths; codetable by
// byte value; and the file order (huff.cpp:113-118: sorted list walked from its end).
int ref_build_codebook(const uint32_t *hist, uint8_t *sorted_sym, uint8_t *sorted_len,
                       uint32_t *code_by_byte, uint32_t *len_by_byte, uint8_t *file_sym,
                       uint8_t *file_len, uint64_t *bits) {
    unsigned int num[MAX_CODE_NUM];
    for (int i = 0; i < MAX_CODE_NUM; i++) num[i] = hist[i];
    Symbol symbols[MAX_CODE_NUM] = {};
    Codetable table[MAX_CODE_NUM] = {};
    int k = store_symbols(num, symbols);
    if (k <= 0) return k;
    qsort(symbols, k, sizeof(Symbol), ref_compare);
    boundary_PM(symbols, k, table);
    for (int i = 0; i < k; i++) {
        sorted_sym[i] = symbols[i].symbol;
        sorted_len[i] = symbols[i].length;
        file_sym[i] = symbols[k - 1 - i].symbol;
        file_len[i] = symbols[k - 1 - i].length;
    }
    for (int i = 0; i < MAX_CODE_NUM; i++) {
        code_by_byte[i] = table[i].code;
        len_by_byte[i] = table[i].length;
    }
    *bits = get_outputfilesize(symbols, k);
    return k;
}

}  // extern "C"
