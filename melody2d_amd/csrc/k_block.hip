// Fused internal force + nodal update, one workgroup per node block.
//
// Replaces, per block of owned nodes (a whole grain when it fits):
//   Body::Update_internal_forces  Code/Body.h:937-1049 (+ Code/Model.h:19-73)
//   Node::Sum_up_forces           Code/Node.h:212-225
//   Node::Apply_Newton            Code/Node.h:227-231
//   Body::Apply_Euler (works)     Code/Body.h:2051-2068, Node::Apply_Euler Code/Node.h:239-245
//
// Data path per block: the displacement parameters of the touched nodes are
// staged in LDS once.  Chunk by chunk, every Gauss point is a lane that holds
// its own support entries (dphi/dx, dphi/dy, node index, destination slot)
// in registers, loaded with coalesced 16-byte loads from the wave's
// entry-major stream while the previous chunk is summed.  The lane gathers
// all its displacements from LDS at once, forms F and the stress terms
// (Code/Model.h:19-73) summing in the reference's entry order, and writes
// each entry's contribution c0 * (dphi . P) to the entry's slot in
// owned-node order.  Then each thread sums, for its owned nodes, their
// contiguous slots - a node's entries in Gauss-point order - per region when
// the reference order is requested.  A block is 128 threads and needs < 40 KB
// of LDS, so four blocks share a CU and a 1000-grain model is resident at
// once.  HBM traffic per entry: 16 (dphi/dx, dphi/dy) + 4 (index and slot),
// times the padding of each wave to its largest support.
#include "mel_device.hpp"
#include "k_block.hpp"
#include "k_block_dev.hpp"

namespace mel {

__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_block_force_update(DevModel m, BlockArrays B) {
    __shared__ __align__(16) blockdev::Fp64::slot s_cxy[CHUNK_ENTRIES + 1];
    double fi[2][NODES_PER_THREAD];
    block_force_update<blockdev::Fp64>(m, B, blockIdx.x, s_cxy, fi);
}

__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_block_force_update32(DevModel m, BlockArrays B) {
    __shared__ __align__(16) blockdev::Fp32::slot s_cxy[2 * (CHUNK_ENTRIES + 1)];
    double fi[2][NODES_PER_THREAD];
    block_force_update<blockdev::Fp32>(m, B, blockIdx.x, s_cxy, fi);
}

// the same with the wide register window (supports of more than
// KREG_ENTRIES / KREG_ENTRIES_F32 entries)
__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_block_force_update_wide(DevModel m, BlockArrays B) {
    __shared__ __align__(16) blockdev::Fp64W::slot s_cxy[CHUNK_ENTRIES + 1];
    double fi[2][NODES_PER_THREAD];
    block_force_update<blockdev::Fp64W>(m, B, blockIdx.x, s_cxy, fi);
}

__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_block_force_update32_wide(DevModel m, BlockArrays B) {
    __shared__ __align__(16) blockdev::Fp32W::slot s_cxy[2 * (CHUNK_ENTRIES + 1)];
    double fi[2][NODES_PER_THREAD];
    block_force_update<blockdev::Fp32W>(m, B, blockIdx.x, s_cxy, fi);
}

// DOF_TO_DISP gather (Body::Update_kinematics, Code/Body.h:2176-2212) and
// current positions for the listed nodes
__global__ void k_kinematics_list(DevModel m, const int* nodes, int n) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    kinematics_node(m, nodes[q]);
}

// force fields of deformable nodes rebuilt from the last step's components
// (same expression as the update), for output; body forces m g
// (Body::Update_body_forces) unless with_body_forces is 0
__global__ void k_node_outputs(DevModel m, int with_body_forces) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.n_nodes) return;
    if (m.body_rigid_flag[m.node_body[j]]) return;
    node_outputs(m, j, with_body_forces);
}

}  // namespace mel
