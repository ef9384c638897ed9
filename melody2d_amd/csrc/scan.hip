#include "scan.hpp"

namespace mel {

// exclusive scan of one tile of SCAN_TILE ints held SCAN_ITEMS per thread
__device__ int block_scan_tile(int* vals, int* sh) {
    int t = threadIdx.x;
    int local = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) { int v = vals[i]; vals[i] = local; local += v; }
    sh[t] = local;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        int v = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int incl = sh[t];
    int total = sh[SCAN_BLOCK - 1];
    int excl = incl - local;
    for (int i = 0; i < SCAN_ITEMS; i++) vals[i] += excl;
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tiles(const int* in, int* out, int n, int* tile_tot) {
    __shared__ int sh[SCAN_BLOCK];
    int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    for (int i = 0; i < SCAN_ITEMS; i++) v[i] = (base + i < n) ? in[base + i] : 0;
    int tot = block_scan_tile(v, sh);
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n) out[base + i] = v[i];
    if (threadIdx.x == 0) tile_tot[blockIdx.x] = tot;
}

// single block: exclusive scan of the tile totals (any count), in place
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_totals(int* tile_tot, int ntiles, int* out, int n) {
    __shared__ int sh[SCAN_BLOCK];
    int carry = 0;
    for (int start = 0; start < ntiles; start += SCAN_TILE) {
        int base = start + threadIdx.x * SCAN_ITEMS;
        int v[SCAN_ITEMS];
        for (int i = 0; i < SCAN_ITEMS; i++) v[i] = (base + i < ntiles) ? tile_tot[base + i] : 0;
        int tot = block_scan_tile(v, sh);
        for (int i = 0; i < SCAN_ITEMS; i++)
            if (base + i < ntiles) tile_tot[base + i] = v[i] + carry;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = carry;
}

__global__ void k_scan_add(int* out, int n, const int* tile_tot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_tot[i / SCAN_TILE];
}

void exclusive_scan_int(const int* in, int* out, int n, int* tmp, hipStream_t s) {
    int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles == 0) ntiles = 1;
    k_scan_tiles<<<ntiles, SCAN_BLOCK, 0, s>>>(in, out, n, tmp);
    k_scan_totals<<<1, SCAN_BLOCK, 0, s>>>(tmp, ntiles, out, n);
    if (n > 0) k_scan_add<<<(n + 255) / 256, 256, 0, s>>>(out, n, tmp);
}

}  // namespace mel
