// Node blocks for the fused internal-force + update kernel (k_block.hip).
//
// A block owns up to BLOCK_THREADS * NODES_PER_THREAD nodes of one body (a
// whole grain when it fits) and lists every Gauss point of that body whose
// support touches an owned node, cut into chunks of at most BLOCK_THREADS
// Gauss points and CHUNK_ENTRIES owned-node entries (contiguous Gauss-point
// ranges in the reference order, or dealt round-robin for balance).  In a
// chunk every Gauss point is a lane; lanes are sorted by support size and
// grouped in waves of 64.  A wave's support entries are stored entry-major
// (entry i of all its lanes contiguous, 16-byte aligned loads coalesce),
// every lane padded to the wave's largest support with zero derivatives:
//   be_fxy  dphi/dx, dphi/dy (double2; float2 in be_fxy32 for the FP32-mixed form)
//   be_pk   touched-node index | destination slot << 16, where owned node o
//           owns slots tptr[o] .. tptr[o+1] - 1 (its entries in Gauss-point
//           order) and CHUNK_ENTRIES is a sink for entries of other blocks'
//           nodes and for padding
// Per lane: support size gl_cnt, c0 = w * jac, region.
// A body larger than one block is split; its blocks read displacement
// parameters owned by sibling blocks, which is safe because the kernel only
// computes internal forces: the parameters are updated by k_node_update, a
// separate launch.
#pragma once
#include <cstdint>
#include "mel_device.hpp"

namespace mel {

constexpr int BLOCK_THREADS = 128;
constexpr int NODES_PER_THREAD = 2;
constexpr int BLOCK_MAX_OWN = BLOCK_THREADS * NODES_PER_THREAD;
constexpr int CHUNK_ENTRIES = 2048;                                 // owned-node entry slots per chunk
constexpr int KREG_ENTRIES = 20;                                    // support entries held in registers per lane (FP64)
constexpr int KREG_ENTRIES_F32 = 24;                                // the same for the FP32-mixed form
constexpr int KREG_ENTRIES_WIDE = 30;                               // the wide window (both forms)
constexpr int MAX_BLOCK_CHUNKS = 64;                                // chunks per node block
// LDS scratch (doubles) of a workgroup that steps one rigid body
// (rigid_body_wg): reductions, the body's record, its driven DOFs, BC flags
constexpr int RIGID_WG_LDS = 6 * BLOCK_THREADS + 8 + RIGID_FIELDS + 8 + 16;

// per node block, a header of four int4 read at once when the block starts
// (everything the internal-force pass needs before its first loads):
//   [0] first owned node, owned nodes, touched nodes, touched-node list offset
//       (-1: the touched nodes are the owned ones, first owned node on)
//   [1] chunk offset, chunks, flags (BLK_*), owned-node list offset
//   [2] descriptor of the first chunk (ch_meta)
//   [3] slot counts of the first chunk (ch_ptr_off), border-node list offset,
//       border nodes finished by the block
constexpr int BLK_SINGLE = 1;          // every Gauss point of the block in one region
constexpr int BLK_INACTIVE = 2;        // body status 1: no update
constexpr int BLK_NEO_HOOKEAN = 4;     // material law (else linear)
constexpr int BLK_OWN_CONTIGUOUS = 8;  // owned nodes are first owned node + 0, 1, ...
constexpr int BLK_NO_GRAVITY = 16;     // body status 2: no body force

struct BlockArrays {
    int n_blocks;
    int4* blk_hdr;             // 4 per block, see above
    double2* blk_par;          // per block: (Mu, K2) of its material
    int max_touch, max_own;
    int *blk_body, *blk_own_off, *blk_nown, *blk_touch_off, *blk_ntouch, *blk_chunk_off, *blk_nchunk;
    int* blk_single_region;    // 1 when every Gauss point of the block is in one region
    int *own_nodes, *touch_nodes;
    int any_multi_region;
    int4* ch_meta;             // per chunk: (lane start, n lanes, entry start, K0 | K1 << 8 | n0 << 16 | n1 << 24)
    int2* ch_ptr_off;          // per chunk: (slots used, tptr offset)
    double* bgp_c0;
    int* bgp_region;
    int* gl_cnt;               // support size per lane
    int* tptr;                 // chunk-local slot offsets per owned node
    uint32_t* be_pk;
    double2* be_fxy;           // FP64 form
    float2* be_fxy32;          // FP32-mixed form
    double* blk_work;          // 14 per block: (sum, compensation) for the 7 works
    int *blk_bn_off, *blk_bn_n, *bn_list;   // border nodes finished by each (unsplit) block
    int *split_bn, n_split_bn;              // border nodes of split bodies
    int *split_kin, n_split_kin;            // their nodes (kinematics after the update)
    int *bc_bn, n_bc_bn;                    // deformable border nodes on borders with a BC
    uint16_t* bt_lbn;          // per point-force spreading entry: its border node's index in the
                               // block's list (one-launch step only)
    // per entry of bn_list, what the finish of its block reads first:
    int4* fin_rec0;            // (border node, node, DOF_TO_DISP row start, row end)
    int4* fin_rec1;            // (slot list start, slot list end, first slot, second slot)
    double2* fin_recX;         // initial position of the node
    // per slot, the inputs of its tributary length besides positions:
    int4* sl_rec0;             // (stencil nodes 0, 1, 2, index j on the border)
    int4* sl_rec1;             // (periodic shifts 0, 1, 2, slots of the border | 1 << 30 when simple)
    int lds_finish_ok;         // every block meets the one-launch step's finish preconditions
};

__global__ void k_block_force_update(DevModel m, BlockArrays B);
__global__ void k_block_force_update32(DevModel m, BlockArrays B);
__global__ void k_block_force_update_wide(DevModel m, BlockArrays B);
__global__ void k_block_force_update32_wide(DevModel m, BlockArrays B);
__global__ void k_update_finish(DevModel m, BlockArrays B, double t_next);
__global__ void k_kinematics_list(DevModel m, const int* nodes, int n);
__global__ void k_node_outputs(DevModel m, int with_body_forces);

}  // namespace mel
