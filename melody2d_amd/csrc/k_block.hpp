// Node blocks for the fused internal-force + update kernel (k_block.hip).
//
// A block owns up to BLOCK_THREADS * NODES_PER_THREAD nodes of one body (a
// whole grain when it fits) and lists every Gauss point of that body whose
// support touches an owned node, in ascending Gauss-point order, cut into
// chunks of at most CHUNK_GPS points / CHUNK_ENTRIES support entries.  Every
// chunk's entry arrays start at a multiple of 8 entries (16-byte vectors of
// uint16) and are padded to a multiple of 8:
//   be_fxy  dphi/dx, dphi/dy per support entry (double2)
//   be_idx  block-local touched-node index per entry
//   tpos    per owned node (in tptr order), the chunk-local positions of its
//           entries in Gauss-point order: the transposed index of the gather
// Per Gauss point: gptr (entry offsets, chunk local), c0 = w * jac, region;
// per owned node: tptr (offsets into tpos, chunk local).
// A body larger than one block is split; its blocks read displacement
// parameters owned by sibling blocks, so they write their updated ones to
// stage_u (per own-node position) and k_commit_split publishes them after
// the launch.
#pragma once
#include <cstdint>
#include "mel_device.hpp"

namespace mel {

constexpr int BLOCK_THREADS = 128;
constexpr int NODES_PER_THREAD = 2;
constexpr int BLOCK_MAX_OWN = BLOCK_THREADS * NODES_PER_THREAD;
constexpr int ENTRIES_PER_THREAD = 12;
constexpr int CHUNK_ENTRIES = BLOCK_THREADS * ENTRIES_PER_THREAD;   // entries per chunk
constexpr int CHUNK_GPS = BLOCK_THREADS;                            // Gauss points per chunk
constexpr int MAX_BLOCK_CHUNKS = 64;                                // chunks per node block

struct BlockArrays {
    int n_blocks;
    int max_touch, max_own;
    int *blk_body, *blk_own_off, *blk_nown, *blk_touch_off, *blk_ntouch, *blk_chunk_off, *blk_nchunk;
    int* blk_single_region;    // 1 when every Gauss point of the block is in one region
    int *own_nodes, *touch_nodes;
    int4* ch_meta;             // per chunk: (gp start, n gp, n entries, entry start / 8)
    int2* ch_ptr_off;          // per chunk: (gptr offset, tptr offset)
    double* bgp_c0;
    int* bgp_region;
    int *gptr, *tptr;          // chunk-local entry offsets per Gauss point / per owned node
    uint16_t *be_idx, *tpos;
    double2* be_fxy;
    double* blk_work;          // 14 per block: (sum, compensation) for the 7 works
    int* blk_split;            // 1 when the block's body spans several blocks
    double* stage_u;           // 2 per own-node position (split blocks only)
    int* split_pos;            // own-node positions of split blocks
    int n_split;
    int* kin_nodes;            // deformable border nodes: kinematics needed every step
    int n_kin;
};

__global__ void k_block_force_update(DevModel m, BlockArrays B);
__global__ void k_commit_split(DevModel m, BlockArrays B);
__global__ void k_kinematics_list(DevModel m, const int* nodes, int n);
__global__ void k_node_outputs(DevModel m);

}  // namespace mel
