// Kernels of one explicit step (FP64 drop-in).  Each kernel names the
// reference function it replaces.  Compiled with -ffp-contract=off: the
// expression and summation orders below follow the reference wherever the
// result can be made bit-identical (internal forces, Neumann/Dirichlet
// spreading, damping, kinematics); contact point forces are summed in a
// fixed element order instead of the reference's body/send order.
#include "mel_device.hpp"
#include "k_step.hpp"
#include "k_step_dev.hpp"

namespace mel {

__global__ void k_border_length(DevModel m) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m.n_slots) return;
    int b = m.br_body[m.sl_border[s]];
    if (m.body_rigid_flag[b] || m.body_status[b] == 1) return;
    m.sl_len[s] = slot_length(m, s);
}

__global__ void k_bc_eval(DevModel m, int plus_dt) {
    int br = blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= m.n_borders) return;
    double t = *m.time;
    if (plus_dt) t = t + m.dt;
    bc_eval_at(m, br, t);
}

// Body::Update_damping_forces (Code/Body.h:1057-1154): node gather over the
// COO entries in entry order, reading velocity parameters.
__global__ void k_damping(DevModel m) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.n_nodes) return;
    double fx = 0., fy = 0.;
    for (int q = m.mt_ptr[j]; q < m.mt_ptr[j + 1]; q++) {
        int code = m.mt_entry[q];
        int i = code >> 1, side = code & 1;
        int dof0 = m.m_dof[2 * i], dof1 = m.m_dof[2 * i + 1];
        double d = m.m_damp[i];
        int n0 = dof0 / 2, r0 = dof0 - 2 * n0, n1 = dof1 / 2, r1 = dof1 - 2 * n1;
        if (side == 0) {          // contribution to node0's direction r0 from node1's r1 velocity
            double v = r1 == 0 ? m.vpx[n1] : m.vpy[n1];
            if (r0 == 0) fx -= v * d; else fy -= v * d;
        } else {                  // contribution to node1's direction r1 from node0's r0 velocity
            double v = r0 == 0 ? m.vpx[n0] : m.vpy[n0];
            if (r1 == 0) fx -= v * d; else fy -= v * d;
        }
    }
    m.fkx[j] = fx;
    m.fky[j] = fy;
}

__global__ void k_rigid_nodes(DevModel m, int with_forces) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.n_rigid_nodes) return;
    rigid_node(m, m.rigid_node_ids[q], with_forces);
}

// ---------------------------------------------------------------------------
// The step's launches.  Work of one dependency level shares a launch, each
// kind of work on its own index range:
//   k_step_contacts   Update_contacts + Update_contact_forces (per element,
//                     point-force records appended to per-border-node
//                     buckets) | Update_bc_forces deformable (per border node
//                     on a border with a BC)
//   k_step_gather     rigid resultants and rigid Newton + Euler (one
//                     workgroup per rigid body) | bucket sums and lengths of
//                     the border nodes of split bodies
//   k_update_finish   per node block: bucket sums, Newton + Euler + works,
//                     lengths, border-node kinematics | rigid nodes |
//                     Update_bc(t + dt) | time = t + dt
// ---------------------------------------------------------------------------
__global__ void k_step_contacts(DevModel m, const int* bc_bn, int n_bc) {
    const int nc = *m.c_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc + n_bc; i += gridDim.x * blockDim.x)
        contact_or_bc_item(m, i, nc, bc_bn, n_bc);
}

__global__ void __launch_bounds__(BLOCK_THREADS) k_step_gather(DevModel m, const int* split_bn, int n_split) {
    __shared__ double sc[RIGID_WG_LDS];
    if ((int)blockIdx.x < m.n_rigid) {
        // resultants, then Newton + Euler of the same body by thread 0
        rigid_body_wg(m, blockIdx.x, sc, nullptr);
        return;
    }
    // border nodes of bodies split over several node blocks: bucket sums and
    // lengths here (the positions are those of the start of the step)
    int q = (blockIdx.x - m.n_rigid) * blockDim.x + threadIdx.x;
    if (q < n_split) {
        int bn = split_bn[q];
        bucket_sum_node(m, bn);
        int b = m.node_body[m.bn_node[bn]];
        if (!m.body_rigid_flag[b] && m.body_status[b] != 1) bn_lengths(m, bn);
    }
}

__global__ void __launch_bounds__(BLOCK_THREADS) k_update_finish(DevModel m, BlockArrays B, double t_next) {
    if ((int)blockIdx.x < B.n_blocks) {
        node_block_finish(m, B, blockIdx.x);
        return;
    }
    int i = (blockIdx.x - B.n_blocks) * blockDim.x + threadIdx.x;
    if (i == 0) *m.time = t_next;
    if (i < m.n_rigid_nodes) {
        rigid_node(m, m.rigid_node_ids[i], 1);
        return;
    }
    i -= m.n_rigid_nodes;
    if (i < m.n_borders) bc_eval_at(m, i, t_next);
}

}  // namespace mel
