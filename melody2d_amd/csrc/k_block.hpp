// Node blocks for the fused internal-force + update kernel (k_block.hip).
//
// A block owns up to BLOCK_THREADS * NODES_PER_THREAD nodes of one body (a
// whole grain when it fits) and lists every Gauss point of that body whose
// support touches an owned node, in ascending Gauss-point order, cut into
// chunks of at most ch_cap points / e_cap support entries.  Per chunk:
//   gptr   entry offsets of its Gauss points (uint16, chunk local)
//   be_*   support entries: touched-node index (uint16) and dphi/dx, dphi/dy
//   tptr/tpos  per owned node, the chunk-local positions of its entries in
//          Gauss-point order (the transposed index of the gather)
// A body larger than one block is split; its blocks read displacement
// parameters owned by sibling blocks, so they write their updated ones to
// stage_u (per own-node position) and k_commit_split publishes them after
// the launch.
#pragma once
#include <cstdint>
#include "mel_device.hpp"

namespace mel {

constexpr int BLOCK_THREADS = 256;
constexpr int NODES_PER_THREAD = 2;
constexpr int BLOCK_MAX_OWN = BLOCK_THREADS * NODES_PER_THREAD;

struct BlockArrays {
    int n_blocks;
    int max_touch, max_own, e_cap, ch_cap;
    int *blk_body, *blk_own_off, *blk_nown, *blk_touch_off, *blk_ntouch, *blk_chunk_off, *blk_nchunk;
    int *own_nodes, *touch_nodes;
    int *ch_gp_start, *ch_ngp, *ch_ne, *ch_tptr_off, *ch_tpos_off, *ch_gptr_off;
    long long* ch_e_start;
    double* bgp_c0;
    int* bgp_region;
    uint16_t *gptr, *be_idx, *tptr, *tpos;
    double *be_fx, *be_fy;
    double* blk_work;          // 14 per block: (sum, compensation) for the 7 works
    int* blk_split;            // 1 when the block's body spans several blocks
    double* stage_u;           // 2 per own-node position (split blocks only)
    int* split_pos;            // own-node positions of split blocks
    int n_split;
    int* kin_nodes;            // deformable border nodes: kinematics needed every step
    int n_kin;
};

size_t block_lds_bytes(const BlockArrays& B);

__global__ void k_block_force_update(DevModel m, BlockArrays B);
__global__ void k_commit_split(DevModel m, BlockArrays B);
__global__ void k_kinematics_list(DevModel m, const int* nodes, int n);
__global__ void k_node_outputs(DevModel m);

}  // namespace mel
