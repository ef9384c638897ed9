// Device exclusive prefix sum over int arrays (three passes: block scans,
// scan of block totals, add-back).  out must hold n + 1 entries; out[n] is
// the total.  tmp must hold scan_tmp_ints(n) ints.
#pragma once
#include <hip/hip_runtime.h>

namespace mel {
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
inline int scan_tmp_ints(int n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1 + SCAN_TILE; }
void exclusive_scan_int(const int* in, int* out, int n, int* tmp, hipStream_t s);
}  // namespace mel
