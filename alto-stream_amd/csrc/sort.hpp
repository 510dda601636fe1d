// Device radix sort / scan utilities.
#pragma once
#include "common.hpp"

namespace alto {

struct SortScratch {
    DevBuf t0, t1, tp, counts, scan_tmp;
};

int exclusive_scan_u32(uint32_t *a, uint64_t n, DevBuf &tmp, hipStream_t s);
int radix_sort_pairs(uint64_t *k0, uint64_t *k1, uint32_t *pay, uint64_t n, int bits,
                     SortScratch &ws, hipStream_t s);
int iota_u32(uint32_t *p, uint64_t n, hipStream_t s);

}  // namespace alto
