// The step in two launches (k_fused.hip).
#pragma once
#include "mel_device.hpp"
#include "k_block.hpp"

namespace mel {

struct StepLaunch {
    int n_cwg;       // workgroups of contact elements and BC point forces (they loop over the
                     // element count on the device)
};

// workgroups of launch A: contact elements and BC point forces, node blocks
inline int step_forces_grid(const DevModel& d, const BlockArrays& B, const StepLaunch& S) {
    (void)d;
    return S.n_cwg + B.n_blocks;
}

// rigid-body nodes of launch B: each thread loads the data of up to
// RIGID_NODES_PREFETCH of its nodes before it waits for their bodies
constexpr int RIGID_NODES_PREFETCH = 4;
// workgroups for n_rigid_nodes nodes when `free_slots` workgroup slots are
// left beside the rigid bodies, node blocks and the BC update (so that every
// workgroup of the launch is resident from its start): as few as fit, at
// most one node per thread
inline int step_rigid_node_wgs(int n_rigid_nodes, int free_slots) {
    const int most = (n_rigid_nodes + BLOCK_THREADS - 1) / BLOCK_THREADS;
    const int least = (n_rigid_nodes + BLOCK_THREADS * RIGID_NODES_PREFETCH - 1) / (BLOCK_THREADS * RIGID_NODES_PREFETCH);
    if (most == 0) return 0;
    return free_slots >= most ? most : (free_slots >= least ? free_slots : least);
}

// workgroups of launch B: rigid bodies, node blocks, rigid-body nodes, the
// BC update
inline int step_finish_grid(const DevModel& d, const BlockArrays& B, int n_rwg) {
    return d.n_rigid + B.n_blocks + n_rwg + 1;
}

// dynamic LDS of k_step_finish: a node block's staged point forces and
// parameters, or a rigid body's reduction scratch
inline size_t step_finish_lds(int max_bn, int max_own) {
    const size_t rigid = (size_t)RIGID_WG_LDS * sizeof(double);
    const size_t blk = 6 * (size_t)(max_bn + max_own) * sizeof(double);
    return blk > rigid ? blk : rigid;
}

__global__ void k_step_forces(DevModel m, BlockArrays B, StepLaunch S);
__global__ void k_step_forces32(DevModel m, BlockArrays B, StepLaunch S);
__global__ void k_step_finish(DevModel m, BlockArrays B, int n_rwg);

}  // namespace mel
