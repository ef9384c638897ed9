// One-launch explicit step (k_fused.hip).
#pragma once
#include "mel_device.hpp"
#include "k_block.hpp"

namespace mel {

struct FusedStep {
    int nc;          // contact elements
    int n_cwg;       // workgroups of contact elements and BC point forces
    double t_next;   // time at the end of the step
};

// workgroups of one fused step launch
inline int fused_step_grid(const DevModel& d, const BlockArrays& B, const FusedStep& S) {
    return S.n_cwg + d.n_rigid + B.n_blocks + 1;
}

__global__ void k_step_fused(DevModel m, BlockArrays B, FusedStep S);
__global__ void k_step_fused32(DevModel m, BlockArrays B, FusedStep S);
__global__ void k_step_forces(DevModel m, BlockArrays B, FusedStep S);
__global__ void k_step_forces32(DevModel m, BlockArrays B, FusedStep S);
__global__ void k_step_finish(DevModel m, BlockArrays B, FusedStep S);

// dynamic LDS of k_step_finish: rigid-body reduction scratch or a node
// block's staged point forces and parameters
inline size_t step_finish_lds(int max_bn, int max_own) {
    size_t rigid = (6 * (size_t)BLOCK_THREADS + 8) * sizeof(double);
    size_t block = 6 * (size_t)(max_bn + max_own) * sizeof(double);
    return rigid > block ? rigid : block;
}

}  // namespace mel
