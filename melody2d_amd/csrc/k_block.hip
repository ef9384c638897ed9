// Fused internal force + nodal update, one workgroup per node block.
//
// Replaces, per block of owned nodes (a whole grain when it fits):
//   Body::Update_internal_forces  Code/Body.h:937-1049 (+ Code/Model.h:19-73)
//   Node::Sum_up_forces           Code/Node.h:212-225
//   Node::Apply_Newton            Code/Node.h:227-231
//   Body::Apply_Euler (works)     Code/Body.h:2051-2068, Node::Apply_Euler Code/Node.h:239-245
//
// Data path per block: the displacement parameters of the touched nodes are
// staged in LDS once.  The block's support entries stream from HBM chunk by
// chunk with 16-byte loads into registers; the loads of chunk c + 1 are in
// flight while chunk c is computed.  Per chunk, one thread per Gauss point
// forms F and the stress terms from LDS and overwrites its entries with their
// contributions c0 * (dphi . P); then each thread sums, for its owned nodes,
// the contributions listed by the chunk's transposed index (a node's entries
// in Gauss-point order), per region when the reference order is requested.
// A block is 128 threads and needs < 40 KB of LDS, so four blocks share a CU
// and a 1000-grain model is resident at once.  HBM traffic per entry: 16
// (dphi/dx, dphi/dy) + 2 (node index) + 2 (transposed position).
#include "mel_device.hpp"
#include "k_block.hpp"

namespace mel {

namespace {

constexpr int VEC16 = CHUNK_ENTRIES / 8;    // uint16x8 vectors per chunk array
constexpr int VEC_PER_THREAD = (VEC16 + BLOCK_THREADS - 1) / BLOCK_THREADS;
// native vector types (LLVM vectors) keep the in-flight chunk in registers
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
static_assert(CHUNK_GPS <= BLOCK_THREADS, "one Gauss point per thread");

// stress terms of Code/Model.h:19-73 for one Gauss point, written so that a
// support entry contributes c0 * (phix * P.x + phiy * P.y), c0 * (phix * P.z + phiy * P.w)
__device__ inline double4 gp_terms(int mat_type, double Mu, double K2, const uint16_t* eidx, const double2* exy,
                                   const double2* uv, int k0, int k1) {
    if (mat_type == 0) {
        double E11 = 0., E12 = 0., E22 = 0.;
#pragma unroll 4
        for (int k = k0; k < k1; k++) {
            int j = eidx[k];
            double2 f = exy[k];
            double2 w = uv[j];
            double u = w.x, v = w.y;
            E11 += f.x * u;
            E12 += f.x * v + f.y * u;
            E22 += f.y * v;
        }
        E12 = 0.5 * E12;
        double J = 1. + E11 + E22;
        double S33 = K2 * (J - 1.);
        double M2 = 2 * Mu;
        double S11 = S33 + M2 * E11, S22 = S33 + M2 * E22, S12 = M2 * E12;
        return make_double4(S11, S12, S12, S22);
    }
    double F11 = 1., F12 = 0., F21 = 0., F22 = 1.;
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        int j = eidx[k];
        double2 f = exy[k];
        double2 w = uv[j];
        double u = w.x, v = w.y;
        F11 += f.x * u;
        F12 += f.y * u;
        F21 += f.x * v;
        F22 += f.y * v;
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
    return make_double4(F11 * S11 + F12 * S12, F12 * S22 + F11 * S12, F21 * S11 + F22 * S12, F22 * S22 + F21 * S12);
}

// one chunk's data in flight, held in registers until it is written to LDS
struct ChunkRegs {
    f64x2 f[ENTRIES_PER_THREAD];
    u32x4 iv[VEC_PER_THREAD], tv[VEC_PER_THREAD];
    int gk0, gk1, greg;
    double gc0;
    int tp0[NODES_PER_THREAD], tp1[NODES_PER_THREAD];
};

__device__ inline void load_chunk(const BlockArrays& B, int4 meta, int2 po, int nown, int t, ChunkRegs& R) {
    const int ngp = meta.y, ne8 = (meta.z + 7) & ~7;
    const size_t E = 8 * (size_t)meta.w;
    // small per-thread items first (the counter of outstanding loads is in
    // order); int32 offsets need no unpacking, so nothing waits here
    if (t < ngp) {
        R.gk0 = B.gptr[po.x + t];
        R.gk1 = B.gptr[po.x + t + 1];
        R.gc0 = B.bgp_c0[meta.x + t];
        R.greg = B.bgp_region[meta.x + t];
    }
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) {
        int o = t + q * BLOCK_THREADS;
        if (o < nown) {
            R.tp0[q] = B.tptr[po.y + o];
            R.tp1[q] = B.tptr[po.y + o + 1];
        }
    }
#pragma unroll
    for (int q = 0; q < VEC_PER_THREAD; q++) {
        int v = t + q * BLOCK_THREADS;
        if (v < ne8 / 8) {
            R.iv[q] = reinterpret_cast<const u32x4*>(B.be_idx + E)[v];
            R.tv[q] = reinterpret_cast<const u32x4*>(B.tpos + E)[v];
        }
    }
#pragma unroll
    for (int q = 0; q < ENTRIES_PER_THREAD; q++) {
        int i = t + q * BLOCK_THREADS;
        if (i < ne8) R.f[q] = reinterpret_cast<const f64x2*>(B.be_fxy)[E + i];
    }
}

}  // namespace

__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_block_force_update(DevModel m, BlockArrays B) {
    extern __shared__ __align__(16) double2 s_uv[];        // (ux, uy) of the touched nodes
    __shared__ __align__(16) double2 s_exy[CHUNK_ENTRIES];  // dphi/dx, dphi/dy, then contributions
    __shared__ __align__(16) uint16_t s_eidx[CHUNK_ENTRIES];
    __shared__ __align__(16) uint16_t s_tpos[CHUNK_ENTRIES];
    __shared__ __align__(16) uint16_t s_ereg[CHUNK_ENTRIES];
    __shared__ int4 s_meta[MAX_BLOCK_CHUNKS];
    __shared__ int2 s_po[MAX_BLOCK_CHUNKS];
    __shared__ double red[BLOCK_THREADS / 64][5];
    const int blk = blockIdx.x;
    const int t = threadIdx.x;
    const int body = B.blk_body[blk];
    if (m.body_status[body] == 1) return;   // inactive body: no update (uniform over the block)
    const bool split = B.blk_split[blk] != 0;
    const bool single = B.blk_single_region[blk] != 0;
    const int nown = B.blk_nown[blk];
    const int ntouch = B.blk_ntouch[blk];
    const int* own = B.own_nodes + B.blk_own_off[blk];
    const int* touch = B.touch_nodes + B.blk_touch_off[blk];
    const int mat = m.body_material[body];
    const int mtype = m.mat_type[mat];
    const double Mu = m.mat_p[6 * mat + 3], K2 = m.mat_p[6 * mat + 4];

    const int nch = B.blk_nchunk[blk];
    const int4* gmeta = B.ch_meta + B.blk_chunk_off[blk];
    const int2* gpo = B.ch_ptr_off + B.blk_chunk_off[blk];
    ChunkRegs R;
    if (nch > 0) load_chunk(B, gmeta[0], gpo[0], nown, t, R);
    // chunk descriptors in LDS: read without waiting on the entry stream
    for (int i = t; i < nch; i += BLOCK_THREADS) {
        s_meta[i] = gmeta[i];
        s_po[i] = gpo[i];
    }
    for (int i = t; i < ntouch; i += BLOCK_THREADS) {
        int g = touch[i];
        s_uv[i] = make_double2(m.upx[g], m.upy[g]);
    }
    double tx[NODES_PER_THREAD], ty[NODES_PER_THREAD], rx[NODES_PER_THREAD], ry[NODES_PER_THREAD];
    int cur[NODES_PER_THREAD];
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) { tx[q] = ty[q] = rx[q] = ry[q] = 0.; cur[q] = -1; }

    for (int c = 0; c < nch; c++) {
        __syncthreads();                      // previous chunk's LDS reads are done
        const int4 meta = s_meta[c];
        const int ngp = meta.y, ne8 = (meta.z + 7) & ~7;
#pragma unroll
        for (int q = 0; q < ENTRIES_PER_THREAD; q++) {
            int i = t + q * BLOCK_THREADS;
            if (i < ne8) reinterpret_cast<f64x2*>(s_exy)[i] = R.f[q];
        }
#pragma unroll
        for (int q = 0; q < VEC_PER_THREAD; q++) {
            int v = t + q * BLOCK_THREADS;
            if (v < ne8 / 8) {
                reinterpret_cast<u32x4*>(s_eidx)[v] = R.iv[q];
                reinterpret_cast<u32x4*>(s_tpos)[v] = R.tv[q];
            }
        }
        const int gk0 = R.gk0, gk1 = R.gk1, greg = R.greg;
        const double gc0 = R.gc0;
        int tp0[NODES_PER_THREAD], tp1[NODES_PER_THREAD];
#pragma unroll
        for (int q = 0; q < NODES_PER_THREAD; q++) { tp0[q] = R.tp0[q]; tp1[q] = R.tp1[q]; }
        __syncthreads();
        // next chunk's loads stay in flight during this chunk's work
        if (c + 1 < nch) load_chunk(B, s_meta[c + 1], s_po[c + 1], nown, t, R);
        if (t < ngp) {
            double4 P = gp_terms(mtype, Mu, K2, s_eidx, s_exy, s_uv, gk0, gk1);
            // contributions overwrite the entries in place (each thread owns
            // its Gauss point's entries)
#pragma unroll 4
            for (int k = gk0; k < gk1; k++) {
                double2 f = s_exy[k];
                s_exy[k] = make_double2(gc0 * (f.x * P.x + f.y * P.y), gc0 * (f.x * P.z + f.y * P.w));
            }
            if (!single)
                for (int k = gk0; k < gk1; k++) s_ereg[k] = (uint16_t)greg;
        }
        __syncthreads();
        if (single) {
            // both owned nodes of the thread advance together
            const int n0 = t < nown ? tp1[0] - tp0[0] : 0;
            const int n1 = NODES_PER_THREAD > 1 && t + BLOCK_THREADS < nown ? tp1[NODES_PER_THREAD - 1] - tp0[NODES_PER_THREAD - 1] : 0;
            if (n0 > 0) cur[0] = 0;
            if (n1 > 0) cur[NODES_PER_THREAD - 1] = 0;
            const int nmax = max(n0, n1);
#pragma unroll 2
            for (int i = 0; i < nmax; i++) {
                if (i < n0) {
                    double2 v = s_exy[s_tpos[tp0[0] + i]];
                    rx[0] -= v.x;
                    ry[0] -= v.y;
                }
                if (i < n1) {
                    double2 v = s_exy[s_tpos[tp0[NODES_PER_THREAD - 1] + i]];
                    rx[NODES_PER_THREAD - 1] -= v.x;
                    ry[NODES_PER_THREAD - 1] -= v.y;
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < NODES_PER_THREAD; q++) {
                int o = t + q * BLOCK_THREADS;
                if (o >= nown) break;
                for (int p = tp0[q]; p < tp1[q]; p++) {
                    int k = s_tpos[p];
                    double2 v = s_exy[k];
                    int r = s_ereg[k];
                    if (r != cur[q]) {
                        if (cur[q] >= 0) { tx[q] += rx[q]; ty[q] += ry[q]; }
                        rx[q] = 0.; ry[q] = 0.; cur[q] = r;
                    }
                    rx[q] -= v.x;
                    ry[q] -= v.y;
                }
            }
        }
    }

    // Sum_up_forces + Apply_Newton + Apply_Euler for the owned nodes
    double w[5] = {0., 0., 0., 0., 0.};   // internal body dirichlet neumann damping
    double gx = m.gravity_x, gy = m.gravity_y;
    if (m.body_status[body] == 2) { gx = 0.; gy = 0.; }
    const double dt = m.dt;
#pragma unroll
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
#pragma unroll
    for (int q = 0; q < 5; q++) {
        double v = w[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((t & 63) == 0) red[t >> 6][q] = v;
    }
    __syncthreads();
    if (t == 0) {
        const int slot[5] = {W_INTERNAL, W_BODY, W_DIRICHLET, W_NEUMANN, W_DAMPING};
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
