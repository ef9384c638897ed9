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

// Update_borders of the initialisation block (Code/main.cpp:106-107), which
// runs for every body, inactive ones included
__global__ void k_border_length(DevModel m) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m.n_slots) return;
    int b = m.br_body[m.sl_border[s]];
    if (m.body_rigid_flag[b]) return;
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
    if (m.body_status[m.node_body[j]] == 1) return;   // inactive body: its damping forces stay
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

// Body::Update_alid_forces (Code/Body.h:1163-1296) at time t for one node
// of a body with absorbing layers: the mass-proportional term from the
// node's physical velocity, then the stiffness-proportional terms of the
// matrix entries naming the node, in entry order, from velocity parameters
// relative to the layers' imposed velocity (pvx, pvy)
__device__ inline double alid_velocity(const DevModel& m, int h0, int h1, double t, bool& on) {
    on = h1 > h0;
    double pv = 0.;
    for (int n = h0; n + 1 < h1; n++) {
        double t0 = m.hist_t[n], p0 = m.hist_v[n], t1 = m.hist_t[n + 1], p1 = m.hist_v[n + 1];
        if ((t0 <= t) & (t <= t1)) { pv = p0 + (p1 - p0) / (t1 - t0) * (t - t0); break; }
    }
    return pv;
}

__global__ void k_alid(DevModel m, double t) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.n_alid_nodes) return;
    const int j = m.alid_nodes[q];
    const int b = m.node_body[j];
    if (m.body_status[b] == 1) return;
    const int* h = m.body_alid + 4 * b;
    bool onx, ony;
    const double pvx = alid_velocity(m, h[0], h[1], t, onx);
    const double pvy = alid_velocity(m, h[2], h[3], t, ony);
    const double beta = m.alid_b[j];
    double fx = onx ? -beta * m.mx[j] * (m.vx[j] - pvx) : 0.;
    double fy = ony ? -beta * m.my[j] * (m.vy[j] - pvy) : 0.;
    for (int k = m.mt_ptr[j]; k < m.mt_ptr[j + 1]; k++) {
        const int code = m.mt_entry[k];
        const int e = code >> 1, side = code & 1;
        const int dof0 = m.m_dof[2 * e], dof1 = m.m_dof[2 * e + 1];
        const int n0 = dof0 / 2, r0 = dof0 - 2 * n0, n1 = dof1 / 2, r1 = dof1 - 2 * n1;
        const double d = m.m_stiff[e] / (1. / m.alid_a[n0] + 1. / m.alid_a[n1]);
        // side 0: node0's direction r0 from node1's r1 velocity parameter;
        // side 1: node1's direction r1 from node0's r0 one (off-diagonal entries)
        const int rt = side == 0 ? r0 : r1, rs = side == 0 ? r1 : r0, ns = side == 0 ? n1 : n0;
        if (!(rt == 0 ? onx : ony)) continue;
        const double v = rs == 0 ? m.vpx[ns] - pvx : m.vpy[ns] - pvy;
        const double f = -v * d;
        if (rt == 0) fx += f;
        else fy += f;
    }
    m.fax[j] = fx;
    m.fay[j] = fy;
}

// rigid-body nodes from their bodies' records; with_forces < 0: the
// positions Load_dynamic forms for every rigid body, inactive ones included
// (Code/IO.h:2050)
__global__ void k_rigid_nodes(DevModel m, int with_forces) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.n_rigid_nodes) return;
    const int i = m.rigid_node_ids[q];
    if (with_forces >= 0) {
        rigid_node(m, i, with_forces);
        return;
    }
    const double* R = m.body_rigid + m.node_body[i] * RIGID_FIELDS;
    const double dr = R[R_DISP + 2] - 0.;
    rigid_node_at(m, i, R, cos(dr), sin(dr), with_forces, m.X[i], m.Y[i]);
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
// (the element path needs ~230 registers: two waves per SIMD keep it out of scratch)
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_step_contacts(DevModel m, const int* bc_bn, int n_bc) {
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
        int b = m.node_body[m.bn_node[bn]];
        bucket_gather(m, bn, b);
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

// ---------------------------------------------------------------------------
// Contact work (Code/Body.h:1847-1848): every element holds the half of its
// work accumulated since the last flush; a flush adds, per body, the halves
// of the elements naming it as slave and as master (a self-contact element
// twice), in a fixed order, then clears them.  One workgroup per body.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_flush_contact_work(DevModel m) {
    __shared__ double red[BLOCK];
    const int b = blockIdx.x;
    const int nc = *m.c_count;
    double v = 0.;
    for (int e = threadIdx.x; e < nc; e += BLOCK) {
        const double w = m.C.wacc[e];
        v += m.C.bodyS[e] == b ? w : 0.;
        v += m.C.bodyM[e] == b ? w : 0.;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] != 0.) m.body_rigid[b * RIGID_FIELDS + R_WORK + W_CONTACT] += red[0];
}

__global__ void k_clear_contact_work(DevModel m) {
    const int nc = *m.c_count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nc; e += gridDim.x * blockDim.x) m.C.wacc[e] = 0.;
}

// ---------------------------------------------------------------------------
// Per-body energy scalars, out[3 * b + k]: k = 0 kinetic energy from the
// physical velocities and nodal masses, k = 1 kinetic energy from the
// velocity parameters (rigid bodies: from the body's velocity, mass and
// inertia), k = 2 boundary work of the last step, sum over the border slots
// of (px vx + py vy) * length * dt: the MONITOR_ENERGY block
// (Code/Monitoring.h:508-537) and monitor type 9 (Code/Monitoring.h:395-412).
// One workgroup per body; each sum is per-thread partials in index order,
// then a fixed tree.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_body_energy(DevModel m, double* out) {
    __shared__ double red[3][BLOCK];
    const int b = blockIdx.x;
    const int t = threadIdx.x;
    const int n0 = m.node_off[b], nn = m.body_nnodes[b];
    const bool rigid = m.body_rigid_flag[b] != 0;
    double ke = 0., kp = 0., dw = 0.;
    for (int i = t; i < nn; i += BLOCK) {
        const int j = n0 + i;
        ke += 0.5 * m.vx[j] * m.vx[j] * m.mx[j];
        ke += 0.5 * m.vy[j] * m.vy[j] * m.my[j];
        if (!rigid) {
            kp += 0.5 * m.vpx[j] * m.vpx[j] * m.mx[j];
            kp += 0.5 * m.vpy[j] * m.vpy[j] * m.my[j];
        }
    }
    const int br0 = m.border_off[b], nbr = m.body_nborders[b];
    for (int k = 0; k < nbr; k++) {
        const int s0 = m.br_slot0[br0 + k], ns = m.br_nslots[br0 + k];
        for (int q = t; q < ns; q += BLOCK) {
            const int s = s0 + q, node = m.sl_node[s];
            dw += (m.sl_xbp[s] * m.vx[node] + m.sl_ybp[s] * m.vy[node]) * m.sl_len[s] * m.dt;
        }
    }
    red[0][t] = ke; red[1][t] = kp; red[2][t] = dw;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < 3; k++) red[k][t] += red[k][t + h];
        __syncthreads();
    }
    if (t == 0) {
        double p = red[1][0];
        if (rigid) {
            const double* R = m.body_rigid + b * RIGID_FIELDS;
            const double mass = m.body_mass[b], inertia = m.body_inertia[b];
            p = 0.;
            p += 0.5 * R[R_VEL] * R[R_VEL] * mass;
            p += 0.5 * R[R_VEL + 1] * R[R_VEL + 1] * mass;
            p += 0.5 * R[R_VEL + 2] * R[R_VEL + 2] * inertia;
        }
        out[3 * b] = red[0][0]; out[3 * b + 1] = p; out[3 * b + 2] = red[2][0];
    }
}

// ---------------------------------------------------------------------------
// One force evaluation without integration: items i < nc are contact
// elements (Update_contacts + Update_contact_forces, when do_contacts),
// the next n_bc are deformable border nodes with a BC (Update_bc_forces,
// when do_bc); then k_bucket_sums gathers the contact point forces of every
// border node.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_eval_items(DevModel m, const int* bc_bn, int n_bc, int do_contacts, int do_bc) {
    const int nc = do_contacts ? *m.c_count : 0;
    const int nb = do_bc ? n_bc : 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc + nb; i += gridDim.x * blockDim.x)
        contact_or_bc_item(m, i, nc, bc_bn, nb);
}

__global__ void k_bucket_sums(DevModel m) {
    int bn = blockIdx.x * blockDim.x + threadIdx.x;
    if (bn >= m.n_bnodes) return;
    // an inactive body's nodes keep their contact forces (emptied bucket)
    if (m.body_status[m.node_body[m.bn_node[bn]]] == 1) bucket_total(m, bn);
    else bucket_sum_node(m, bn);
}

__global__ void k_clear_point_forces(DevModel m, int contacts, int bcs) {
    int bn = blockIdx.x * blockDim.x + threadIdx.x;
    if (bn >= m.n_bnodes) return;
    if (contacts) { m.bn_pcx[bn] = 0.; m.bn_pcy[bn] = 0.; }
    if (bcs) { m.bn_pdx[bn] = 0.; m.bn_pdy[bn] = 0.; m.bn_pnx[bn] = 0.; m.bn_pny[bn] = 0.; }
}

__global__ void k_clear_node_forces(DevModel m, int internal, int damping) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.n_nodes) return;
    if (internal) { m.fix[j] = 0.; m.fiy[j] = 0.; }
    if (damping) { m.fkx[j] = 0.; m.fky[j] = 0.; }
}

// Body::Compute_nodal_stresses (Code/Body.h:2310-2519) for every node, from
// the derivatives of its DOF_TO_DISP row (dphix, dphiy, aligned with d_idx)
// and its body's material (bpar: type 0 linear, 1 neo-Hookean, -1 rigid;
// Mu; Lambda or Kappa), written field-major: MEL_STRESS_FIELD_COUNT fields
// of n_nodes values (SigmaI, SigmaII, SigmaIII, jacobian, Sxx, Syy, Sxy,
// Szz, Sigmaxx, Sigmayy, Sigmaxy, Sigmazz, SigmaTresca, SigmaVM, SigmaSph,
// Exx, Eyy, Exy, NormE)
__global__ void k_nodal_stresses(DevModel m, const double* dphix, const double* dphiy, const double* bpar,
                                 double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_nodes) return;
    const int N = m.n_nodes;
    const int b = m.node_body[i];
    const double type = bpar[3 * b];
    double v[19];
    if (type < 0.) {
        for (int f = 0; f < 19; f++) v[f] = 0.;
        v[3] = 1.;
    } else {
        double J = 0., S11 = 0., S12 = 0., S22 = 0., S33 = 0.;
        double Sxx = 0., Syy = 0., Sxy = 0., Szz = 0., Exx = 0., Eyy = 0., Exy = 0.;
        if (type == 0.) {
            double E11 = 0., E12 = 0., E22 = 0.;
            for (int k = m.d_ptr[i]; k < m.d_ptr[i + 1]; k++) {
                const int j = m.d_idx[k];
                const double ux = m.upx[j], uy = m.upy[j];
                E11 += dphix[k] * ux;
                E12 += dphix[k] * uy + dphiy[k] * ux;
                E22 += dphiy[k] * uy;
            }
            E12 = 0.5 * E12;
            const double Mu2 = 2 * bpar[3 * b + 1], Lambda = bpar[3 * b + 2];
            J = 1. + E11 + E22;
            S33 = Lambda * (J - 1.);
            S11 = S33 + Mu2 * E11;
            S22 = S33 + Mu2 * E22;
            S12 = Mu2 * E12;
            if (J < 0.) { S11 = 0.; S12 = 0.; S22 = 0.; S33 = 0.; }
            Sxx = S11; Syy = S22; Sxy = S12; Szz = S33;
            Exx = E11; Eyy = E22; Exy = E12;
        } else {
            double F11 = 1., F12 = 0., F21 = 0., F22 = 1.;
            for (int k = m.d_ptr[i]; k < m.d_ptr[i + 1]; k++) {
                const int j = m.d_idx[k];
                const double ux = m.upx[j], uy = m.upy[j];
                F11 += dphix[k] * ux;
                F12 += dphiy[k] * ux;
                F21 += dphix[k] * uy;
                F22 += dphiy[k] * uy;
            }
            const double Mu = bpar[3 * b + 1], Kappa = bpar[3 * b + 2];
            const double F11F12 = F11 * F11 + F12 * F12, F21F22 = F21 * F21 + F22 * F22;
            J = F11 * F22 - F12 * F21;
            const double invJ = 1. / J, K = Kappa * (J - 1.);
            const double MuJ53 = Mu * pow(J, -1.666666666667);
            const double trB3 = 0.333333333333 * (F11F12 + F21F22 + 1.);
            const double s11 = K + MuJ53 * (F11F12 - trB3), s22 = K + MuJ53 * (F21F22 - trB3);
            const double s12 = MuJ53 * (F11 * F21 + F12 * F22);
            S11 = (F12 * (F12 * s22 - F22 * s12) - F22 * (F12 * s12 - F22 * s11)) * invJ;
            S22 = (F11 * (F11 * s22 - F21 * s12) - F21 * (F11 * s12 - F21 * s11)) * invJ;
            S12 = (F21 * (F12 * s12 - F22 * s11) - F11 * (F12 * s22 - F22 * s12)) * invJ;
            S33 = (K + MuJ53 * (1. - trB3)) * J;
            if (J < 0.) { F11 = 0.; F12 = 0.; F21 = 0.; F22 = 0.; S11 = 0.; S12 = 0.; S22 = 0.; S33 = 0.; }
            Sxx = F11 * (S11 * F11 + S12 * F12) + F12 * (S12 * F11 + S22 * F12);
            Syy = F21 * (S11 * F21 + S12 * F22) + F22 * (S12 * F21 + S22 * F22);
            Sxy = F21 * (S11 * F11 + S12 * F12) + F22 * (S12 * F11 + S22 * F12);
            Szz = S33;
            Exx = 0.5 * (F11 * F11 + F21 * F21 - 1.);
            Eyy = 0.5 * (F12 * F12 + F22 * F22 - 1.);
            Exy = 0.5 * (F12 * F11 + F22 * F21);
        }
        const double NormE = pow(Exx * Exx + Eyy * Eyy + 2. * Exy * Exy, 0.5);
        const double root = __dsqrt_rn((Sxx + Syy) * (Sxx + Syy) - 4. * (Sxx * Syy - Sxy * Sxy));
        const double A = 0.5 * (Sxx + Syy + root), Bv = 0.5 * (Sxx + Syy - root);
        double s1, s2, s3;
        if (A > Bv && A > Szz) { s1 = A; if (Bv > Szz) { s2 = Bv; s3 = Szz; } else { s2 = Szz; s3 = Bv; } }
        else if (Bv > A && Bv > Szz) { s1 = Bv; if (A > Szz) { s2 = A; s3 = Szz; } else { s2 = Szz; s3 = A; } }
        else { s1 = Szz; if (Bv > A) { s2 = Bv; s3 = A; } else { s2 = A; s3 = Bv; } }
        const double vm = __dsqrt_rn((Sxx - Syy) * (Sxx - Syy) + (Sxx - Szz) * (Sxx - Szz) + (Szz - Syy) * (Szz - Syy) +
                                     6. * Sxy * Sxy) * 0.70710678118655;
        v[0] = s1; v[1] = s2; v[2] = s3; v[3] = J; v[4] = S11; v[5] = S22; v[6] = S12; v[7] = S33;
        v[8] = Sxx; v[9] = Syy; v[10] = Sxy; v[11] = Szz; v[12] = s1 - s3; v[13] = vm;
        v[14] = 0.333333333 * (Sxx + Syy + Szz); v[15] = Exx; v[16] = Eyy; v[17] = Exy; v[18] = NormE;
    }
    for (int f = 0; f < 19; f++) out[(size_t)f * N + i] = v[f];
}

}  // namespace mel
