// ref_table_driver.cpp -- extern "C" driver around the reference decoder's host table builder
// (decoder/src/get_table.cpp), compiled from where it lies by oracle/Makefile.ref.
//
// TEST INFRASTRUCTURE ONLY: writes golden vectors for get_table_info / get_twolevel_table.
#include <cstdint>
#include <cstring>

#include "constants.hpp"  // decoder/include (struct Symbol, struct TableInfo)
#include "get_table.h"    // decoder/include

extern "C" {

// file-order (symbol, length) list -> prefix_bit returned by get_table_info, the three table
// sizes, and the stage table image written by get_twolevel_table into `stage` (cap bytes).
// Returns the stage table size in bytes, or -1 when it does not fit.
int ref_two_level_table(const uint8_t *file_sym, const uint8_t *file_len, int k, int prefix_in,
                        uint32_t *prefix_out, uint32_t *sizes, uint32_t *boundary,
                        uint8_t *stage, int cap, int fill) {
    Symbol symbols[256] = {};
    for (int i = 0; i < k; i++) {
        symbols[i].symbol = file_sym[i];
        symbols[i].length = file_len[i];
    }
    TableInfo info = {0, 0, 0};
    unsigned int prefix = get_table_info(symbols, k, prefix_in, info);
    int size = (int)(sizeof(int) * info.ptrtable_size + MAX_CODE_NUM + info.l1table_size +
                     info.l2table_size);
    *prefix_out = prefix;
    sizes[0] = info.l1table_size;
    sizes[1] = info.l2table_size;
    sizes[2] = info.ptrtable_size;
    if (size > cap) return -1;
    std::memset(stage, fill, (size_t)size);
    *boundary = get_twolevel_table(stage, (int)prefix, symbols, k, info);
    return size;
}

}  // extern "C"
