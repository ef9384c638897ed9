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

// rigid-body nodes per thread of launch B, and the workgroups they take
constexpr int RIGID_NODES_PER_THREAD = 1;
__host__ __device__ inline int step_rigid_node_wgs(int n_rigid_nodes) {
    const int per = BLOCK_THREADS * RIGID_NODES_PER_THREAD;
    return (n_rigid_nodes + per - 1) / per;
}

// workgroups of launch B: rigid bodies, node blocks, rigid-body nodes, the
// BC update
inline int step_finish_grid(const DevModel& d, const BlockArrays& B) {
    return d.n_rigid + B.n_blocks + step_rigid_node_wgs(d.n_rigid_nodes) + 1;
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
__global__ void k_step_finish(DevModel m, BlockArrays B);

}  // namespace mel
