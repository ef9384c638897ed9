// Fused internal force + nodal update, one workgroup per node block.
//
// Replaces, per block of owned nodes (a whole grain when it fits):
//   Body::Update_internal_forces  Code/Body.h:937-1049 (+ Code/Model.h:19-73)
//   Node::Sum_up_forces           Code/Node.h:212-225
//   Node::Apply_Newton            Code/Node.h:227-231
//   Body::Apply_Euler (works)     Code/Body.h:2051-2068, Node::Apply_Euler Code/Node.h:239-245
//
// Data path per block: the displacement parameters of the touched nodes are
// staged in LDS once; the block's Gauss-point support entries (local node
// index + dphi/dx, dphi/dy) stream from HBM in chunks with coalesced loads
// into LDS; one thread per Gauss point forms F and the stress terms from
// LDS; then one thread per owned node walks the chunk-local transposed index
// (positions of its entries, in Gauss-point order) and accumulates
// -c0 (dphi . P) per region, exactly in the order of the reference's
// per-region accumulators.  The per-entry HBM traffic is 2 + 16 + 2 bytes.
#include "mel_device.hpp"
#include "k_block.hpp"

namespace mel {

namespace {

struct Lds {
    double *ux, *uy, *P, *c0;
    double *ex, *ey;
    int* greg;
    uint16_t *eidx, *egp, *tpos, *tptr, *gptr;
};

__device__ inline Lds carve(unsigned char* smem, const BlockArrays& B) {
    Lds L;
    unsigned char* p = smem;
    auto take = [&](size_t bytes) { unsigned char* q = p; p += (bytes + 15) & ~size_t(15); return q; };
    L.ux = (double*)take(sizeof(double) * B.max_touch);
    L.uy = (double*)take(sizeof(double) * B.max_touch);
    L.ex = (double*)take(sizeof(double) * B.e_cap);
    L.ey = (double*)take(sizeof(double) * B.e_cap);
    L.P = (double*)take(sizeof(double) * 4 * B.ch_cap);
    L.c0 = (double*)take(sizeof(double) * B.ch_cap);
    L.greg = (int*)take(sizeof(int) * B.ch_cap);
    L.eidx = (uint16_t*)take(2 * B.e_cap);
    L.egp = (uint16_t*)take(2 * B.e_cap);
    L.tpos = (uint16_t*)take(2 * B.e_cap);
    L.tptr = (uint16_t*)take(2 * (B.max_own + 1));
    L.gptr = (uint16_t*)take(2 * (B.ch_cap + 1));
    return L;
}

// stress terms of Code/Model.h:19-73 for one Gauss point, written so that
// the gather computes c0 * (phix * A + phiy * B) and c0 * (phix * C + phiy * D)
__device__ inline void gp_terms(int mat_type, double Mu, double K2, const Lds& L, int k0, int k1, double4& out) {
    if (mat_type == 0) {
        double E11 = 0., E12 = 0., E22 = 0.;
        for (int k = k0; k < k1; k++) {
            int j = L.eidx[k];
            double fx = L.ex[k], fy = L.ey[k], ux = L.ux[j], uy = L.uy[j];
            E11 += fx * ux;
            E12 += fx * uy + fy * ux;
            E22 += fy * uy;
        }
        E12 = 0.5 * E12;
        double J = 1. + E11 + E22;
        double S33 = K2 * (J - 1.);
        double M2 = 2 * Mu;
        double S11 = S33 + M2 * E11, S22 = S33 + M2 * E22, S12 = M2 * E12;
        out = make_double4(S11, S12, S12, S22);
    } else {
        double F11 = 1., F12 = 0., F21 = 0., F22 = 1.;
        for (int k = k0; k < k1; k++) {
            int j = L.eidx[k];
            double fx = L.ex[k], fy = L.ey[k], ux = L.ux[j], uy = L.uy[j];
            F11 += fx * ux;
            F12 += fy * ux;
            F21 += fx * uy;
            F22 += fy * uy;
        }
        double F11F12 = F11 * F11 + F12 * F12;
        double F21F22 = F21 * F21 + F22 * F22;
        double J = F11 * F22 - F12 * F21;
        double invJ = 1. / J;
        double K = K2 * (J - 1.);
        double MuJ53 = Mu * pow(J, -1.666666666667);
        double trB3 = 0.333333333333 * (F11F12 + F21F22 + 1.);
        double s11 = K + MuJ53 * (F11F12 - trB3);
        double s22 = K + MuJ53 * (F21F22 - trB3);
        double s12 = MuJ53 * (F11 * F21 + F12 * F22);
        double S11 = (F12 * (F12 * s22 - F22 * s12) - F22 * (F12 * s12 - F22 * s11)) * invJ;
        double S22 = (F11 * (F11 * s22 - F21 * s12) - F21 * (F11 * s12 - F21 * s11)) * invJ;
        double S12 = (F21 * (F12 * s12 - F22 * s11) - F11 * (F12 * s22 - F22 * s12)) * invJ;
        out = make_double4(F11 * S11 + F12 * S12, F12 * S22 + F11 * S12, F21 * S11 + F22 * S12,
                           F22 * S22 + F21 * S12);
    }
}

}  // namespace

__global__ void __launch_bounds__(BLOCK_THREADS) k_block_force_update(DevModel m, BlockArrays B) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[BLOCK_THREADS / 64][5];
    const int blk = blockIdx.x;
    const int t = threadIdx.x;
    const int body = B.blk_body[blk];
    if (m.body_status[body] == 1) return;   // inactive body: no update (uniform over the block)
    const bool split = B.blk_split[blk] != 0;
    const int nown = B.blk_nown[blk];
    const int ntouch = B.blk_ntouch[blk];
    const int* own = B.own_nodes + B.blk_own_off[blk];
    const int* touch = B.touch_nodes + B.blk_touch_off[blk];
    Lds L = carve(smem, B);
    const int mat = m.body_material[body];
    const int mtype = m.mat_type[mat];
    const double Mu = m.mat_p[6 * mat + 3], K2 = m.mat_p[6 * mat + 4];

    for (int i = t; i < ntouch; i += BLOCK_THREADS) {
        int g = touch[i];
        L.ux[i] = m.upx[g];
        L.uy[i] = m.upy[g];
    }
    double tx[NODES_PER_THREAD], ty[NODES_PER_THREAD], rx[NODES_PER_THREAD], ry[NODES_PER_THREAD];
    int cur[NODES_PER_THREAD];
    for (int q = 0; q < NODES_PER_THREAD; q++) { tx[q] = ty[q] = rx[q] = ry[q] = 0.; cur[q] = -1; }

    const int c0i = B.blk_chunk_off[blk], c1i = c0i + B.blk_nchunk[blk];
    for (int c = c0i; c < c1i; c++) {
        const int ngp = B.ch_ngp[c], ne = B.ch_ne[c];
        const int gs = B.ch_gp_start[c];
        const size_t es = B.ch_e_start[c];
        __syncthreads();
        for (int i = t; i < ne; i += BLOCK_THREADS) {
            L.eidx[i] = B.be_idx[es + i];
            L.ex[i] = B.be_fx[es + i];
            L.ey[i] = B.be_fy[es + i];
            L.tpos[i] = B.tpos[B.ch_tpos_off[c] + i];
        }
        for (int i = t; i <= nown; i += BLOCK_THREADS) L.tptr[i] = B.tptr[B.ch_tptr_off[c] + i];
        for (int i = t; i <= ngp; i += BLOCK_THREADS) L.gptr[i] = B.gptr[B.ch_gptr_off[c] + i];
        for (int i = t; i < ngp; i += BLOCK_THREADS) {
            L.c0[i] = B.bgp_c0[gs + i];
            L.greg[i] = B.bgp_region[gs + i];
        }
        __syncthreads();
        for (int gl = t; gl < ngp; gl += BLOCK_THREADS) {
            int k0 = L.gptr[gl], k1 = L.gptr[gl + 1];
            double4 P;
            gp_terms(mtype, Mu, K2, L, k0, k1, P);
            L.P[4 * gl] = P.x; L.P[4 * gl + 1] = P.y; L.P[4 * gl + 2] = P.z; L.P[4 * gl + 3] = P.w;
            for (int k = k0; k < k1; k++) L.egp[k] = (uint16_t)gl;
        }
        __syncthreads();
        for (int q = 0; q < NODES_PER_THREAD; q++) {
            int o = t + q * BLOCK_THREADS;
            if (o >= nown) break;
            for (int p = L.tptr[o]; p < L.tptr[o + 1]; p++) {
                int k = L.tpos[p];
                int gl = L.egp[k];
                int r = L.greg[gl];
                if (r != cur[q]) {
                    if (cur[q] >= 0) { tx[q] += rx[q]; ty[q] += ry[q]; }
                    rx[q] = 0.; ry[q] = 0.; cur[q] = r;
                }
                double c0 = L.c0[gl], fx = L.ex[k], fy = L.ey[k];
                rx[q] -= c0 * (fx * L.P[4 * gl] + fy * L.P[4 * gl + 1]);
                ry[q] -= c0 * (fx * L.P[4 * gl + 2] + fy * L.P[4 * gl + 3]);
            }
        }
    }

    // Sum_up_forces + Apply_Newton + Apply_Euler for the owned nodes
    double w[5] = {0., 0., 0., 0., 0.};   // internal body dirichlet neumann damping
    double gx = m.gravity_x, gy = m.gravity_y;
    if (m.body_status[body] == 2) { gx = 0.; gy = 0.; }
    const double dt = m.dt;
    for (int q = 0; q < NODES_PER_THREAD; q++) {
        int o = t + q * BLOCK_THREADS;
        if (o >= nown) break;
        int j = own[o];
        if (cur[q] >= 0) { tx[q] += rx[q]; ty[q] += ry[q]; }
        double ix = tx[q], iy = ty[q];
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
        double fx = ix + cx + 0. + bx + dx + nx + kx + 0.;
        double fy = iy + cy + 0. + by + dy + ny + ky + 0.;
        double apx = fx * m.imx[j], apy = fy * m.imy[j];
        double vpx = m.vpx[j], vpy = m.vpy[j];
        double hx = vpx + 0.5 * dt * apx, hy = vpy + 0.5 * dt * apy;
        w[0] += dt * (hx * ix + hy * iy);
        w[1] += dt * (hx * bx + hy * by);
        w[2] += dt * (hx * dx + hy * dy);
        w[3] += dt * (hx * nx + hy * ny);
        w[4] += dt * (hx * kx + hy * ky);
        vpx = vpx + dt * apx;
        vpy = vpy + dt * apy;
        double nux = m.upx[j] + dt * vpx, nuy = m.upy[j] + dt * vpy;
        if (split) {
            size_t pos = (size_t)B.blk_own_off[blk] + o;
            B.stage_u[2 * pos] = nux;
            B.stage_u[2 * pos + 1] = nuy;
        } else {
            m.upx[j] = nux;
            m.upy[j] = nuy;
        }
        m.vpx[j] = vpx; m.vpy[j] = vpy;
        m.apx[j] = apx; m.apy[j] = apy;
        m.fix[j] = ix; m.fiy[j] = iy;
    }
    // block work increments: fixed-order reduction (wavefront butterflies,
    // then the waves in order), accumulated with compensation across steps
    for (int q = 0; q < 5; q++) {
        double v = w[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((t & 63) == 0) red[t >> 6][q] = v;
    }
    __syncthreads();
    if (t == 0) {
        static const int slot[5] = {W_INTERNAL, W_BODY, W_DIRICHLET, W_NEUMANN, W_DAMPING};
        double* acc = B.blk_work + 14 * (size_t)blk;
        for (int q = 0; q < 5; q++) {
            double v = 0.;
            for (int wv = 0; wv < BLOCK_THREADS / 64; wv++) v += red[wv][q];
            // Neumaier accumulation: acc[2s] sum, acc[2s+1] compensation
            int s = slot[q];
            double sum = acc[2 * s], x = v, nt = sum + x;
            acc[2 * s + 1] += fabs(sum) >= fabs(x) ? (sum - nt) + x : (x - nt) + sum;
            acc[2 * s] = nt;
        }
    }
}

// publish the displacement parameters of split bodies once every block of
// the launch has read the old ones
__global__ void k_commit_split(DevModel m, BlockArrays B) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= B.n_split) return;
    int pos = B.split_pos[q];
    int j = B.own_nodes[pos];
    if (m.body_status[m.node_body[j]] == 1) return;
    m.upx[j] = B.stage_u[2 * (size_t)pos];
    m.upy[j] = B.stage_u[2 * (size_t)pos + 1];
}

// DOF_TO_DISP gather (Body::Update_kinematics, Code/Body.h:2176-2212) and
// current positions for the listed nodes
__global__ void k_kinematics_list(DevModel m, const int* nodes, int n) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int i = nodes[q];
    double dx = 0., dy = 0., vx = 0., vy = 0., ax = 0., ay = 0.;
    for (int k = m.d_ptr[i]; k < m.d_ptr[i + 1]; k++) {
        int j = m.d_idx[k];
        double f = m.d_phi[k];
        dx += f * m.upx[j]; dy += f * m.upy[j];
        vx += f * m.vpx[j]; vy += f * m.vpy[j];
        ax += f * m.apx[j]; ay += f * m.apy[j];
    }
    m.ux[i] = dx; m.uy[i] = dy; m.vx[i] = vx; m.vy[i] = vy; m.ax[i] = ax; m.ay[i] = ay;
    m.xc[i] = m.X[i] + dx;
    m.yc[i] = m.Y[i] + dy;
}

// force fields of deformable nodes rebuilt from the last step's components
// (same expression as the update), for output
__global__ void k_node_outputs(DevModel m) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.n_nodes) return;
    int b = m.node_body[j];
    if (m.body_rigid_flag[b]) return;
    double cx = 0., cy = 0., dx = 0., dy = 0., nx = 0., ny = 0.;
    for (int e = m.bt_ptr[j]; e < m.bt_ptr[j + 1]; e++) {
        int bn = m.bt_bn[e];
        double f = m.bt_phi[e];
        cx += m.bn_pcx[bn] * f; cy += m.bn_pcy[bn] * f;
        dx += m.bn_pdx[bn] * f; dy += m.bn_pdy[bn] * f;
        nx += m.bn_pnx[bn] * f; ny += m.bn_pny[bn] * f;
    }
    double gx = m.gravity_x, gy = m.gravity_y;
    if (m.body_status[b] == 2) { gx = 0.; gy = 0.; }
    double bx = m.mx[j] * gx, by = m.my[j] * gy;
    m.fcx[j] = cx; m.fcy[j] = cy; m.fdx[j] = dx; m.fdy[j] = dy; m.fnx[j] = nx; m.fny[j] = ny;
    m.fbx[j] = bx; m.fby[j] = by;
    m.fx[j] = m.fix[j] + cx + 0. + bx + dx + nx + m.fkx[j] + 0.;
    m.fy[j] = m.fiy[j] + cy + 0. + by + dy + ny + m.fky[j] + 0.;
}

}  // namespace mel
