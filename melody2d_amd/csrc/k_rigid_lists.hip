// Contact elements of every rigid body, for the rigid resultants of the
// step (Body::Update_contact_forces rigid branches, Code/Body.h:1866-1875
// and 1929-1939): the elements a rigid body is the slave of are contiguous
// (elements are grouped by slave body, Code/Proximity.h:175-330), so a
// binary search over the slave ids finds them; the elements it is the
// master of are scattered, so after every contact update they are collected
// per rigid body (an atomic cursor per body, then each body's list sorted
// back into element order by one workgroup).  The rigid workgroups of the
// step then read only their own elements instead of scanning all of them.
#include <hip/hip_runtime.h>
#include <climits>

#include "mel_device.hpp"
#include "k_rigid_lists.hpp"

namespace mel {

namespace {

constexpr int RL_BLOCK = 256;

__global__ void k_rm_clear(DevModel m) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m.n_rigid) m.rm_cnt[r] = 0;
}

// every element whose master is rigid goes to that body's list
__global__ void k_rm_append(DevModel m) {
    const int nc = *m.c_count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nc; e += gridDim.x * blockDim.x) {
        const int r = m.body_rindex[m.C.bodyM[e]];
        if (r < 0) continue;
        const int pos = atomicAdd(&m.rm_cnt[r], 1);
        if (pos < RIGID_LIST_CAP) m.rm_list[(size_t)r * RIGID_LIST_CAP + pos] = e;
        else atomicOr(&m.status_flags[0], 32);
    }
}

__device__ inline int lower_bound_int(const int* a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// lists up to this length are sorted by rank (each entry counts the
// smaller ones: one pass, no barriers); longer ones by a bitonic network
constexpr int RANK_SORT_MAX = 1024;

// one workgroup per rigid body: its list sorted into element order (the
// entries are distinct element ids), and its slave range
__global__ void __launch_bounds__(RL_BLOCK) k_rm_sort(DevModel m) {
    __shared__ int s[RIGID_LIST_CAP];
    const int r = blockIdx.x;
    const int t = threadIdx.x;
    const int n = min(m.rm_cnt[r], RIGID_LIST_CAP);
    int* lst = m.rm_list + (size_t)r * RIGID_LIST_CAP;
    int p = 1;
    while (p < n) p <<= 1;
    if (n <= RANK_SORT_MAX) {
        for (int i = t; i < n; i += RL_BLOCK) s[i] = lst[i];
        __syncthreads();
        for (int i = t; i < n; i += RL_BLOCK) {
            const int v = s[i];
            int rank = 0;
            for (int j = 0; j < n; j++) rank += s[j] < v;
            lst[rank] = v;
        }
        p = 0;      // nothing left for the network
    }
    for (int i = t; i < p; i += RL_BLOCK) s[i] = i < n ? lst[i] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= p; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < p; i += RL_BLOCK) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const int a = s[i], b = s[l];
                    if ((a > b) == up) { s[i] = b; s[l] = a; }
                }
            }
            __syncthreads();
        }
    }
    if (p > 0)
        for (int i = t; i < n; i += RL_BLOCK) lst[i] = s[i];
    if (t == 0) {
        const int nc = *m.c_count;
        const int b = m.rigid_ids[r];
        m.rs_range[2 * r] = lower_bound_int(m.C.bodyS, nc, b);
        m.rs_range[2 * r + 1] = lower_bound_int(m.C.bodyS, nc, b + 1);
    }
}

}  // namespace

hipError_t build_rigid_lists(const DevModel& d, hipStream_t s, bool collect) {
    if (d.n_rigid == 0) return hipSuccess;
    if (collect) {
        k_rm_clear<<<(d.n_rigid + RL_BLOCK - 1) / RL_BLOCK, RL_BLOCK, 0, s>>>(d);
        k_rm_append<<<std::min((d.c_cap + RL_BLOCK - 1) / RL_BLOCK, 1024), RL_BLOCK, 0, s>>>(d);
    }
    k_rm_sort<<<d.n_rigid, RL_BLOCK, 0, s>>>(d);
    return hipGetLastError();
}

}  // namespace mel
