// The step in two launches (k_fused.hip).
#pragma once
#include "mel_device.hpp"
#include "k_block.hpp"

namespace mel {

struct StepLaunch {
    int n_cwg;       // workgroups of contact elements and BC point forces (they loop over the
                     // element count on the device)
};

// workgroups of launch A
inline int step_forces_grid(const DevModel& d, const BlockArrays& B, const StepLaunch& S) {
    return S.n_cwg + d.n_rigid + B.n_blocks + 1;
}

// workgroups of launch B: node blocks, then one thread per rigid-body node
inline int step_finish_grid(const DevModel& d, const BlockArrays& B) {
    return B.n_blocks + (d.n_rigid_nodes + BLOCK_THREADS - 1) / BLOCK_THREADS;
}

// dynamic LDS of k_step_finish: a node block's staged point forces and
// parameters
inline size_t step_finish_lds(int max_bn, int max_own) {
    return 6 * (size_t)(max_bn + max_own) * sizeof(double);
}

__global__ void k_step_forces(DevModel m, BlockArrays B, StepLaunch S);
__global__ void k_step_forces32(DevModel m, BlockArrays B, StepLaunch S);
__global__ void k_step_finish(DevModel m, BlockArrays B);

}  // namespace mel
