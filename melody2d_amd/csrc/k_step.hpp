// Kernel declarations shared by the host orchestration (mel_context.cpp).
#pragma once
#include "mel_device.hpp"

namespace mel {

constexpr int BUCKET_CAP = 32;     // point-force records per border node per step
constexpr int BLOCK = 256;
constexpr int SEG_CHUNK = 8;      // border slots per proximity culling chunk
constexpr int SC_GROUP = 8;       // chunks per second-level culling box (a wavefront of the box kernel)
static_assert(SEG_CHUNK * SC_GROUP == 64, "a chunk group is one wavefront of k_prox_chunk_boxes");
constexpr int ROW_CHUNK = 256;    // narrow-phase rows per counting workgroup
constexpr int NB_THREADS = 256;   // neighbour search: threads per body
constexpr int NB_PER = 4;         // neighbour search: candidate bodies per thread and pass

__global__ void k_border_length(DevModel m);
__global__ void k_bc_eval(DevModel m, int plus_dt);
__global__ void k_damping(DevModel m);
__global__ void k_alid(DevModel m, double t);
__global__ void k_rigid_nodes(DevModel m, int with_forces);
__global__ void k_step_contacts(DevModel m, const int* bc_bn, int n_bc);
__global__ void k_step_gather(DevModel m, const int* split_bn, int n_split);
__global__ void k_flush_contact_work(DevModel m);
__global__ void k_clear_contact_work(DevModel m);
__global__ void k_body_energy(DevModel m, double* out);
__global__ void k_eval_items(DevModel m, const int* bc_bn, int n_bc, int do_contacts, int do_bc);
__global__ void k_bucket_sums(DevModel m);
__global__ void k_clear_point_forces(DevModel m, int contacts, int bcs);
__global__ void k_clear_node_forces(DevModel m, int internal, int damping);

// proximity (k_proximity.hip)
__global__ void k_prox_chunk_boxes(DevModel m);
__global__ void k_prox_aabb(DevModel m);
__global__ void k_cells_clear(DevModel m);
__global__ void k_cells_assign(DevModel m);
__global__ void k_cells_scan_blocks(DevModel m);
__global__ void k_cells_scan_top(DevModel m, int nblocks);
__global__ void k_cells_scatter(DevModel m);
__global__ void k_prox_neighbours(DevModel m);
__global__ void k_prox_chunk_offsets(DevModel m);
__global__ void k_prox_chunk_bodies(DevModel m, int cap);
__global__ void k_prox_count(DevModel m, int n_count);
__global__ void k_prox_elem_offsets(DevModel m);
__global__ void k_prox_write(DevModel m);
__global__ void k_prox_history(DevModel m);
__global__ void k_prox_chain(DevModel m);
__global__ void k_init_czm(DevModel m);
__global__ void k_project_all(DevModel m);
__global__ void k_nodal_stresses(DevModel m, const double* dphix, const double* dphiy, const double* bpar,
                                 double* out);

}  // namespace mel
