// Kernels of the adaptive schemes of Solver_step (Code/Solver.h:98-486).
// An attempt of step dt is two halves of the launch sequence (contacts and
// BC forces with the law's time step dt/2, internal forces, then the half's
// end below); after the second half the error estimate is reduced per body
// and over the model on the device, and the host reads the few scalars that
// decide acceptance.  Compiled with -ffp-contract=off; expression orders
// follow the reference.
#include "k_adaptive.hpp"
#include "k_step_dev.hpp"

namespace mel {

// Body::Update_bc (Code/Body.h:335-395) at time t for every active body:
// border values and the rigid bodies' driven DOFs (the start of the second
// half, Code/Solver.h:227)
__global__ void k_ad_bc(DevModel m, double t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m.n_borders) {
        if (m.body_status[m.br_body[i]] != 1) bc_eval_at(m, i, t);
        return;
    }
    i -= m.n_borders;
    if (i >= m.n_rigid) return;
    const int b = m.rigid_ids[i];
    if (m.body_status[b] == 1) return;
    double* R = m.body_rigid + b * RIGID_FIELDS;
    double dv[8];
    for (int k = 0; k < 8; k++) dv[k] = R[R_DRIVEN + k];
    rigid_driven_dofs(m, b, t, dv);
    for (int k = 0; k < 8; k++) R[R_DRIVEN + k] = dv[k];
}

// The end of a half for one rigid body, by thread 0 of its workgroup once
// the resultants F are known: Sum_up_forces + Apply_Newton with the body's
// mass-scaling factor (Code/Body.h:1984-2037); first half: the temporary
// predictor of Apply_Euler_temporary (Code/Body.h:2126-2166) with the
// current driven DOFs; second half: Update_bc(t + dt) (Code/Solver.h:268);
// then Apply_Euler(dt/2) with works (Code/Body.h:2071-2117) and current
// positions; second half: Compute_error and Compute_mass_scaling of the
// rigid branch (Code/Body.h:2538-2551, 2569-2585).  The record is worked on
// in registers and stored back once.
__device__ void rigid_half(const DevModel& m, const AdArrays& A, int b, const RigidForces& F, const double* rec,
                           const RigidScalars& rs, int half, double dt, double t_end) {
    if (rs.status == 1) return;
    double* Rg = m.body_rigid + b * RIGID_FIELDS;
    double* ab = A.body + (size_t)b * AD_BODY_FIELDS;
    double R[RIGID_FIELDS];
#pragma unroll
    for (int k = 0; k < RIGID_FIELDS; k++) R[k] = rec[k];
    for (int q = 0; q < 3; q++) { R[R_FCON + q] = F.fcon[q]; R[R_FNEU + q] = F.fneu[q]; R[R_FDIR + q] = F.fdir[q]; }
    R[R_FBODY] = F.fbody[0];
    R[R_FBODY + 1] = F.fbody[1];
    for (int q = 0; q < 3; q++)
        R[R_FTOT + q] = R[R_FCON + q] + R[R_FBODY + q] + R[R_FDIR + q] + R[R_FNEU + q] + R[R_FDAMP + q];
    const double fac = ab[AB_FACTOR];
    R[R_ACC] = R[R_FTOT] * rs.inv_mass / fac;
    R[R_ACC + 1] = R[R_FTOT + 1] * rs.inv_mass / fac;
    R[R_ACC + 2] = R[R_FTOT + 2] * rs.inv_inertia / fac;
    R[R_NODE_ACC] = R[R_ACC]; R[R_NODE_ACC + 1] = R[R_ACC + 1];
    R[R_FDIR_NODE] = R[R_FDIR]; R[R_FDIR_NODE + 1] = R[R_FDIR + 1];
    double* dv = R + R_DRIVEN;
    if (half == 1) {
        double* dt_ = ab + AB_DISP_T;
        double* vt = ab + AB_VEL_T;
        if (dv[0] == 0.) {
            vt[0] = R[R_VEL] + dt * R[R_ACC];
            dt_[0] = R[R_DISP] + dt * vt[0];
        } else if (dv[0] == 1.) {
            vt[0] = dv[2];
            dt_[0] = dv[1];
        }
        if (dv[4] == 0.) {
            vt[1] = R[R_VEL + 1] + dt * R[R_ACC + 1];
            dt_[1] = R[R_DISP + 1] + dt * vt[1];
        } else if (dv[4] == 1.) {
            vt[1] = dv[6];
            dt_[1] = dv[5];
        }
        if (dv[0] == 0. && dv[4] == 0. && rs.periodicity != 2) {
            vt[2] = R[R_VEL + 2] + dt * R[R_ACC + 2];
            dt_[2] = R[R_DISP + 2] + dt * vt[2];
        } else {
            vt[2] = 0.;
            dt_[2] = 0.;
        }
    } else {
        rigid_driven_dofs(m, b, t_end, dv);
    }
    const double h = dt * 0.5;
    double hx = R[R_VEL] + 0.5 * h * R[R_ACC];
    double hy = R[R_VEL + 1] + 0.5 * h * R[R_ACC + 1];
    double hr = R[R_VEL + 2] + 0.5 * h * R[R_ACC + 2];
    R[R_WORK + W_BODY] += h * (hx * R[R_FBODY] + hy * R[R_FBODY + 1] + hr * R[R_FBODY + 2]);
    R[R_WORK + W_DIRICHLET] += h * (hx * R[R_FDIR] + hy * R[R_FDIR + 1] + hr * R[R_FDIR + 2]);
    R[R_WORK + W_NEUMANN] += h * (hx * R[R_FNEU] + hy * R[R_FNEU + 1] + hr * R[R_FNEU + 2]);
    R[R_WORK + W_DAMPING] += h * (hx * R[R_FDAMP] + hy * R[R_FDAMP + 1] + hr * R[R_FDAMP + 2]);
    const double mass = rs.mass;
    if (dv[0] == 0.) {
        R[R_VEL] += h * R[R_ACC];
        R[R_DISP] = R[R_DISP] + h * R[R_VEL];
    } else if (dv[0] == 1.) {
        R[R_ACC] = dv[3]; R[R_VEL] = dv[2]; R[R_DISP] = dv[1];
        R[R_FDIR] = mass * R[R_ACC] - R[R_FTOT];
    }
    if (dv[4] == 0.) {
        R[R_VEL + 1] += h * R[R_ACC + 1];
        R[R_DISP + 1] = R[R_DISP + 1] + h * R[R_VEL + 1];
    } else if (dv[4] == 1.) {
        R[R_ACC + 1] = dv[7]; R[R_VEL + 1] = dv[6]; R[R_DISP + 1] = dv[5];
        R[R_FDIR + 1] = mass * R[R_ACC + 1] - R[R_FTOT + 1];
    }
    if (dv[0] == 0. && dv[4] == 0. && rs.periodicity != 2) {
        R[R_VEL + 2] += h * R[R_ACC + 2];
        R[R_DISP + 2] = R[R_DISP + 2] + h * R[R_VEL + 2];
    } else {
        R[R_VEL + 2] = 0.;
        R[R_DISP + 2] = 0.;
    }
    R[R_CUR] = R[R_XC] + R[R_DISP];
    R[R_CUR + 1] = R[R_YC] + R[R_DISP + 1];
    R[R_CUR + 2] = 0. + R[R_DISP + 2];
    if (half == 2) {
        const double* dt_ = ab + AB_DISP_T;
        double te = 0.;
        if (dv[0] != 1.) te += (R[R_DISP] - dt_[0]) * (R[R_DISP] - dt_[0]);
        if (dv[4] != 1.) te += (R[R_DISP + 1] - dt_[1]) * (R[R_DISP + 1] - dt_[1]);
        te = 0.1 * sqrt(te) / A.body_nd[b];
        if (te == 0.) te = 1.e-16;
        ab[AB_TOTAL] = te;
        ab[AB_MAX] = te;
        ab[AB_NODE] = -1.;
        if (A.mass_scaling) {
            const double thr = A.target_error * A.err_factor_ms;
            double f = ab[AB_FACTOR];
            if (te >= thr) {
                const double d = pow((te - thr), A.ctrl_ms);
                ab[AB_DFACTOR] = d;
                f += d;
                if (f >= A.max_ms) f = A.max_ms;
            } else {
                if (f != 1) f *= A.decrease_ms;
                if (f < 1) f = 1;
            }
            ab[AB_FACTOR] = f;
            ab[AB_MASS_MS] = mass * f;
        }
    }
#pragma unroll
    for (int k = 0; k < RIGID_FIELDS; k++)
        if (k != R_WORK + W_CONTACT) Rg[k] = R[k];
}

// rigid bodies (one workgroup each) | border nodes of split bodies: bucket
// sums and lengths (positions of the start of the half)
__global__ void __launch_bounds__(BLOCK_THREADS) k_ad_gather(DevModel m, AdArrays A, const int* split_bn,
                                                             int n_split, int half, double dt, double t_end) {
    __shared__ double sc[RIGID_WG_LDS];
    if ((int)blockIdx.x < m.n_rigid) {
        const int b = m.rigid_ids[blockIdx.x];
        double* s_rec = sc + 6 * BLOCK_THREADS + 8;
        const double* Rg = m.body_rigid + b * RIGID_FIELDS;
        const RigidScalars rs = rigid_scalars(m, b);
        for (int k = threadIdx.x; k < RIGID_FIELDS; k += blockDim.x) s_rec[k] = Rg[k];
        const RigidForces F = rigid_forces_block(m, blockIdx.x, sc, nullptr);   // its barriers publish s_rec
        __syncthreads();
        if (threadIdx.x == 0) rigid_half(m, A, b, F, s_rec, rs, half, dt, t_end);
        return;
    }
    int q = (blockIdx.x - m.n_rigid) * blockDim.x + threadIdx.x;
    if (q < n_split) {
        int bn = split_bn[q];
        int b = m.node_body[m.bn_node[bn]];
        bucket_gather(m, bn, b);
        if (!m.body_rigid_flag[b] && m.body_status[b] != 1) bn_lengths(m, bn);
    }
}

// The end of a half for one node block (a whole body or a piece of one):
// the send of the contact point forces to the block's border nodes,
// Sum_up_forces + Apply_Newton with the nodes' mass-scaling factors
// (Code/Node.h:212-231); first half: Apply_Euler_temporary(dt)
// (Code/Node.h:247-251) from the start-of-step parameters; Apply_Euler(dt/2)
// with works (Code/Body.h:2051-2068); Update_borders of the block's slots
// and the kinematics of its border nodes.
__device__ void node_block_half(const DevModel& m, const BlockArrays& B, const AdArrays& A, int blk, int half,
                                double dt) {
    __shared__ double red[BLOCK_THREADS / 64][6];
    const int t = threadIdx.x;
    const int body = B.blk_body[blk];
    const int nown = B.blk_nown[blk];
    const int* own = B.own_nodes + B.blk_own_off[blk];
    const int* bns = B.bn_list + B.blk_bn_off[blk];
    const int nbn = B.blk_bn_n[blk];
    if (m.body_status[body] == 1) {
        for (int i = t; i < nbn; i += BLOCK_THREADS) bucket_drain_node(m, bns[i]);
        return;
    }
    for (int i = t; i < nbn; i += BLOCK_THREADS) bucket_sum_node(m, bns[i]);
    __syncthreads();
    double w[6] = {0., 0., 0., 0., 0., 0.};   // internal body dirichlet neumann damping alid
    double gx = m.gravity_x, gy = m.gravity_y;
    if (m.body_status[body] == 2) { gx = 0.; gy = 0.; }
    const double h = dt * 0.5;
    const bool alid = m.alid_on != 0;
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) {
        int o = t + q * BLOCK_THREADS;
        if (o >= nown) break;
        int j = own[o];
        double ix = m.fix[j], iy = m.fiy[j];
        double cx = 0., cy = 0., dx = 0., dy = 0., nx = 0., ny = 0.;
        for (int e = m.bt_ptr[j]; e < m.bt_ptr[j + 1]; e++) {
            int bn = m.bt_bn[e];
            double f = m.bt_phi[e];
            cx += m.bn_pcx[bn] * f; cy += m.bn_pcy[bn] * f;
            dx += m.bn_pdx[bn] * f; dy += m.bn_pdy[bn] * f;
            nx += m.bn_pnx[bn] * f; ny += m.bn_pny[bn] * f;
        }
        double bx = m.mx[j] * gx, by = m.my[j] * gy;
        double kx = m.fkx[j], ky = m.fky[j];
        const double lx = alid ? m.fax[j] : 0., ly = alid ? m.fay[j] : 0.;
        double fx = ix + cx + 0. + bx + dx + nx + kx + lx;
        double fy = iy + cy + 0. + by + dy + ny + ky + ly;
        double apx = fx * m.imx[j] / A.fsx[j], apy = fy * m.imy[j] / A.fsy[j];
        double vpx = m.vpx[j], vpy = m.vpy[j];
        const double upx = m.upx[j], upy = m.upy[j];
        if (half == 1) {
            A.uptx[j] = upx + dt * (vpx + dt * apx);
            A.upty[j] = upy + dt * (vpy + dt * apy);
        }
        double hx = vpx + 0.5 * h * apx, hy = vpy + 0.5 * h * apy;
        w[0] += h * (hx * ix + hy * iy);
        w[1] += h * (hx * bx + hy * by);
        w[2] += h * (hx * dx + hy * dy);
        w[3] += h * (hx * nx + hy * ny);
        w[4] += h * (hx * kx + hy * ky);
        w[5] += h * (hx * lx + hy * ly);
        vpx = vpx + h * apx;
        vpy = vpy + h * apy;
        const double nux = upx + h * vpx, nuy = upy + h * vpy;
        if (!isfinite(nux) || !isfinite(nuy)) atomicOr(&m.status_flags[1], 1);
        m.upx[j] = nux;
        m.upy[j] = nuy;
        m.vpx[j] = vpx; m.vpy[j] = vpy;
        m.apx[j] = apx; m.apy[j] = apy;
    }
    const int nw = alid ? 6 : 5;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        if (q >= nw) break;
        double v = w[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((t & 63) == 0) red[t >> 6][q] = v;
    }
    __syncthreads();
    if (t == 0) {
        const int slot[6] = {W_INTERNAL, W_BODY, W_DIRICHLET, W_NEUMANN, W_DAMPING, W_ALID};
        double* acc = B.blk_work + 14 * (size_t)blk;
        for (int q = 0; q < nw; q++) {
            double v = 0.;
            for (int wv = 0; wv < BLOCK_THREADS / 64; wv++) v += red[wv][q];
            int s = slot[q];
            double sum = acc[2 * s], x = v, nt = sum + x;
            acc[2 * s + 1] += fabs(sum) >= fabs(x) ? (sum - nt) + x : (x - nt) + sum;
            acc[2 * s] = nt;
        }
    }
    for (int i = t; i < nbn; i += BLOCK_THREADS) bn_lengths(m, bns[i]);
    __syncthreads();
    for (int i = t; i < nbn; i += BLOCK_THREADS) kinematics_node(m, m.bn_node[bns[i]]);
}

// node blocks | rigid nodes | second half: Update_bc(t + dt) of the borders
__global__ void __launch_bounds__(BLOCK_THREADS) k_ad_finish(DevModel m, BlockArrays B, AdArrays A, int half,
                                                             double dt, double t_end) {
    if ((int)blockIdx.x < B.n_blocks) {
        node_block_half(m, B, A, blockIdx.x, half, dt);
        return;
    }
    int i = (blockIdx.x - B.n_blocks) * blockDim.x + threadIdx.x;
    if (i < m.n_rigid_nodes) {
        rigid_node(m, m.rigid_node_ids[i], 1);
        return;
    }
    i -= m.n_rigid_nodes;
    if (half == 2 && i < m.n_borders && m.body_status[m.br_body[i]] != 1) bc_eval_at(m, i, t_end);
}

// Update_kinematics of every deformable node (Code/Body.h:2176-2212), the
// temporary displacements of Update_kinematics_temporary
// (Code/Body.h:2235-2251), Node::Compute_error (Code/Node.h:253-272) and,
// for the mass-scaling schemes, Node::Compute_mass_scaling
// (Code/Node.h:274-320)
__global__ void k_ad_error(DevModel m, AdArrays A) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.n_kin) return;
    const int i = A.kin_nodes[q];
    const int b = m.node_body[i];
    if (m.body_status[b] == 1) return;
    kinematics_node(m, i);
    double tx = 0., ty = 0.;
    for (int k = m.d_ptr[i]; k < m.d_ptr[i + 1]; k++) {
        const double f = m.d_phi[k];
        const int j = m.d_idx[k];
        tx += f * A.uptx[j];
        ty += f * A.upty[j];
    }
    const int drv = A.body_drv[b];
    const double ex = (drv & 1) ? 0. : m.ux[i] - tx;
    const double ey = (drv & 2) ? 0. : m.uy[i] - ty;
    const double ide = A.ide[i];
    A.xerr[i] = ex;
    A.yerr[i] = ey;
    A.enorm[i] = sqrt(ex * ex + ey * ey) * ide;
    if (!A.mass_scaling) return;
    const double thr = A.target_error * A.err_factor_ms;
    double fx = A.fsx[i], fy = A.fsy[i];
    if (fabs(ex * ide) >= thr) {
        fx += pow((fabs(ex * ide) - thr), A.ctrl_ms);
        if (fx > A.max_ms) fx = A.max_ms;
    } else {
        if (fx != 1) fx *= A.decrease_ms;
        if (fx < 1) fx = 1;
    }
    if (fabs(ey * ide) >= thr) {
        fy += pow((fabs(ey * ide) - thr), A.ctrl_ms);
        if (fy > A.max_ms) fy = A.max_ms;
    } else {
        if (fy != 1) fy *= A.decrease_ms;
        if (fy < 1) fy = 1;
    }
    A.fsx[i] = fx;
    A.fsy[i] = fy;
    A.fsn[i] = (fx + fy) * 0.5;
    const double mx = m.mx[i], my = m.my[i];
    A.msm[i] = sqrt(mx * mx * fx * fx + my * my * fy * fy);
}

// Body::Compute_error / Compute_mass_scaling reductions of a deformable
// body (one workgroup): the total error and the scaled mass summed serially
// in node order (the reference's order, so the scalars match it bit for
// bit), the largest error (floored at 1e-16) at its first node and the
// largest node factor by the workgroup
__global__ void __launch_bounds__(BLOCK) k_ad_body_reduce(DevModel m, AdArrays A) {
    __shared__ double smax[BLOCK];
    __shared__ int sidx[BLOCK];
    __shared__ double sfac[BLOCK];
    const int b = blockIdx.x;
    if (m.body_rigid_flag[b] || m.body_status[b] == 1) return;
    const int n0 = m.node_off[b], n = m.body_nnodes[b];
    double* ab = A.body + (size_t)b * AD_BODY_FIELDS;
    double mx = 0.;
    int mi = 0;
    double mf = 0.;
    for (int k = threadIdx.x; k < n; k += BLOCK) {
        double e = A.enorm[n0 + k];
        if (e == 0.) e = 1.e-16;
        if (e > mx) { mx = e; mi = k; }
        if (A.mass_scaling && A.fsn[n0 + k] > mf) mf = A.fsn[n0 + k];
    }
    smax[threadIdx.x] = mx;
    sidx[threadIdx.x] = mi;
    sfac[threadIdx.x] = mf;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = smax[threadIdx.x + w];
            const int oi = sidx[threadIdx.x + w];
            if (o > smax[threadIdx.x] || (o == smax[threadIdx.x] && o > 0. && oi < sidx[threadIdx.x])) {
                smax[threadIdx.x] = o;
                sidx[threadIdx.x] = oi;
            }
            if (sfac[threadIdx.x + w] > sfac[threadIdx.x]) sfac[threadIdx.x] = sfac[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double tot = 0.;
        for (int k = 0; k < n; k++) tot += A.enorm[n0 + k];
        ab[AB_TOTAL] = tot;
        ab[AB_MAX] = smax[0];
        ab[AB_NODE] = smax[0] > 0. ? (double)sidx[0] : 0.;
        if (A.mass_scaling) {
            double ms = 0.;
            for (int k = 0; k < n; k++) ms += A.msm[n0 + k];
            ab[AB_MASS_MS] = ms;
            ab[AB_MAXFACTOR] = sfac[0];
        }
    }
}

// the scheme's scalars over the active bodies in body order
// (Code/Solver.h:285-326): largest error, mean error, total scaled mass,
// largest scaled-to-initial mass ratio; AO_NAN set when some body's total
// error is NaN (the run stops there, Code/main.cpp:227-237)
__global__ void k_ad_global(DevModel m, AdArrays A) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double mean = 0., mxe = 0., tm = 0., mm = 0., nan = 0.;
    int nn = 0;
    for (int b = 0; b < m.n_bodies; b++) {
        if (m.body_status[b] == 1) continue;
        const double* ab = A.body + (size_t)b * AD_BODY_FIELDS;
        mean += ab[AB_TOTAL];
        nn += m.body_nnodes[b];
        if (ab[AB_MAX] > mxe) mxe = ab[AB_MAX];
        if (isnan(ab[AB_TOTAL])) nan = 1.;
        if (A.mass_scaling) {
            tm += ab[AB_MASS_MS];
            if (ab[AB_MASS_MS] * m.body_inv_mass[b] > mm) mm = ab[AB_MASS_MS] * m.body_inv_mass[b];
        }
    }
    mean /= nn;
    A.out[AO_MAX_ERROR] = mxe;
    A.out[AO_MEAN_ERROR] = mean;
    A.out[AO_TOTAL_MASS] = tm;
    A.out[AO_MAX_MASS] = mm;
    A.out[AO_NAN] = nan;
}

// Store / Restore (Code/Body.h:2593-2652): descriptor y, its bytes strided
// over the x workgroups; 8-byte words when the size allows, else 4-byte
__global__ void k_ad_copy(const CopyDesc* d, int n_desc, const int* c_count) {
    const int k = blockIdx.y;
    if (k >= n_desc) return;
    const CopyDesc D = d[k];
    const unsigned long long n = D.per_contact ? (unsigned long long)*c_count : D.n;
    const unsigned long long bytes = n * D.elem;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    const unsigned long long first = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if ((bytes & 7ull) == 0) {
        const unsigned long long* s = reinterpret_cast<const unsigned long long*>(D.src);
        unsigned long long* o = reinterpret_cast<unsigned long long*>(D.dst);
        for (unsigned long long w = first; w < bytes / 8; w += stride) o[w] = s[w];
    } else {
        const unsigned* s = reinterpret_cast<const unsigned*>(D.src);
        unsigned* o = reinterpret_cast<unsigned*>(D.dst);
        for (unsigned long long w = first; w < bytes / 4; w += stride) o[w] = s[w];
    }
}

// the rigid records' restored fields: current position, displacement and
// velocity (Body::Restore) and the works
__global__ void k_ad_rigid_copy(const double* src, double* dst, int n_bodies) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = i / 16, k = i % 16;
    if (b >= n_bodies) return;
    const int f = k < 9 ? R_CUR + k : (k < 16 ? R_WORK + (k - 9) : -1);
    if (f < 0) return;
    dst[(size_t)b * RIGID_FIELDS + f] = src[(size_t)b * RIGID_FIELDS + f];
}

__global__ void k_set_time(double* time, double t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *time = t;
}

}  // namespace mel
