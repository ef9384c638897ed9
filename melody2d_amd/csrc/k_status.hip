// Update_status (Code/Proximity.h:550-1007) on the device, and the contact
// sends of inactive bodies.
//
// The rules run in order in one workgroup: the parallel part forms what a
// rule reads (node kinematics of the body, node force fields, per border
// node damage values), thread 0 then sums in the reference's node order and
// compares.  x, y and r stay in thread 0's registers from one rule to the
// next, as the reference's locals do: a deformable body's r and a force kind
// without a branch read what the previous rule left.
//
// A body the rules deactivate stops being updated, detected and reset
// (Code/Solver.h:36-82, Code/Proximity.h:72-131) but Send_contact_forces
// still runs for it (Code/Solver.h:57-58): the master point forces of its
// last step are added again every step.  They are the master records of
// its elements of that step, the previous proximity result once the
// contact update has run, and become ghost records: border-node records
// appended to the buckets by k_ghost_sends, rigid-master records read by
// the rigid-body resultants.  Its elements of the new proximity result are
// marked inactive (no projection, no law), its node blocks skipped, and its
// rigid nodes get the force copies of its last step (Body::Sum_up_forces
// sets them, Code/Body.h:1999-2012, and no longer runs).
#include "k_status.hpp"
#include "k_step_dev.hpp"

namespace mel {

namespace {

// Update_status's comparison: 0 below (<=), 1 equal, 2 above (>=)
__device__ inline bool holds(int cmp, double v, double thr) {
    return (cmp == 0 && v <= thr) || (cmp == 1 && v == thr) || (cmp == 2 && v >= thr);
}

// the component a rule compares: x, y, the magnitude (the reference's
// pow(x*x+y*y, 0.5), which is the correctly rounded square root up to
// glibc's pow error) or r; Position has no magnitude (component 2 is r)
__device__ inline double component(int type, int comp, double x, double y, double r) {
    if (comp == 0) return x;
    if (comp == 1) return y;
    if (type == 0 || comp == 3) return r;
    return __dsqrt_rn(x * x + y * y);
}

// a deformable node's force of one kind (Read_spy_line's order: total,
// internal, contact, body, Dirichlet, Neumann, damping), rebuilt with the
// output expression; false for the ALID kind, which Update_status skips
__device__ inline bool node_force(const DevModel& m, int j, int kind, double& x, double& y) {
    if (kind > 6) return false;
    node_outputs(m, j, 1);
    const double* px[7] = {m.fx, m.fix, m.fcx, m.fbx, m.fdx, m.fnx, m.fkx};
    const double* py[7] = {m.fy, m.fiy, m.fcy, m.fby, m.fdy, m.fny, m.fky};
    x = px[kind][j];
    y = py[kind][j];
    return true;
}

__device__ inline void deactivate(const DevModel& m, int b, int* s_any) {
    if (m.body_status[b] != 1) {
        m.body_status[b] = 1;
        m.body_flip[b] = 1;
        *s_any = 1;
    }
}

// first element of slave body b in the current proximity result (elements
// are grouped by slave body, in body order)
__device__ inline int first_element_of(const ContactArrays& C, int nc, int b) {
    int lo = 0, hi = nc;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (C.bodyS[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// exclusive prefix of v over the workgroup; s_w: STATUS_THREADS / 64 ints
__device__ inline int excl_scan(int v, int* s_w, int& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wv] = x;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int k = 0; k < STATUS_THREADS / 64; k++) {
        const int t = s_w[k];
        if (k < wv) before += t;
        total += t;
    }
    __syncthreads();
    return before + x - v;
}

}  // namespace

__global__ void __launch_bounds__(STATUS_THREADS) k_update_status(DevModel m, BlockArrays B) {
    __shared__ int s_any, s_lo, s_hi;
    __shared__ int s_w[STATUS_THREADS / 64];
    const int t = threadIdx.x;
    if (t == 0) s_any = 0;
    __syncthreads();
    double x = 0., y = 0., r = 0.;     // thread 0's, carried over the rules
    for (int i = 0; i < m.n_rules; i++) {
        const double* D = m.rules + RULE_STRIDE * i;
        const int type = (int)D[0];
        if (type >= 0 && type <= 3) {
            // [1] component, [2] body, [3] node (< 0: the body), [4] comparison, [5] threshold
            const int b = (int)D[2];
            const bool rigid = m.body_rigid_flag[b] != 0;
            const int n0 = m.node_off[b], nn = m.body_nnodes[b];
            const bool body_level = D[3] < 0.;
            // deformable nodes: kinematics from the parameters; a rigid
            // node: from its body's record (an inactive body's stay)
            if (!rigid && body_level)
                for (int k = t; k < nn; k += STATUS_THREADS) kinematics_node(m, n0 + k);
            __syncthreads();
            if (t == 0) {
                const double* px = type == 0 ? m.xc : (type == 1 ? m.ux : (type == 2 ? m.vx : m.ax));
                const double* py = type == 0 ? m.yc : (type == 1 ? m.uy : (type == 2 ? m.vy : m.ay));
                if (!body_level) {
                    const int j = n0 + (int)D[3];
                    if (rigid) rigid_node(m, j, 1);
                    else kinematics_node(m, j);
                    x = px[j];
                    y = py[j];
                } else if (!rigid) {
                    double mass = 0.;
                    x = 0.;
                    y = 0.;
                    for (int k = 0; k < nn; k++) {
                        x += px[n0 + k] * m.mx[n0 + k];
                        y += py[n0 + k] * m.my[n0 + k];
                        mass += m.mx[n0 + k];
                    }
                    x /= mass;
                    y /= mass;
                } else {
                    const double* R = m.body_rigid + (size_t)b * RIGID_FIELDS;
                    const int off = type == 0 ? R_CUR : (type == 1 ? R_DISP : (type == 2 ? R_VEL : R_ACC));
                    x = R[off];
                    y = R[off + 1];
                    r = R[off + 2];
                }
                if (holds((int)D[4], component(type, (int)D[1], x, y, r), D[5])) deactivate(m, b, &s_any);
            }
        } else if (type == 4) {
            // [1] component, [2] force kind, [3] body, [4] node, [5] comparison, [6] threshold
            const int kind = (int)D[2], b = (int)D[3];
            const int n0 = m.node_off[b];
            const bool body_level = D[4] < 0.;
            // the body-level branch tests body [2] (the force kind) and loops
            // over its node count, reading the nodes of body [3]
            const bool kind_rigid = body_level && m.body_rigid_flag[kind] != 0;
            const int nk = body_level && !kind_rigid ? m.body_nnodes[kind] : 0;
            if (nk > 0 && kind <= 6)
                for (int k = t; k < nk; k += STATUS_THREADS) node_outputs(m, n0 + k, 1);
            __syncthreads();
            if (t == 0) {
                if (!body_level) {
                    node_force(m, n0 + (int)D[4], kind, x, y);
                } else if (!kind_rigid) {
                    x = 0.;
                    y = 0.;
                    if (kind <= 6) {
                        const double* px[7] = {m.fx, m.fix, m.fcx, m.fbx, m.fdx, m.fnx, m.fkx};
                        const double* py[7] = {m.fy, m.fiy, m.fcy, m.fby, m.fdy, m.fny, m.fky};
                        for (int k = 0; k < nk; k++) {
                            x += px[kind][n0 + k];
                            y += py[kind][n0 + k];
                        }
                    }
                } else {
                    // no rigid branch for the internal (1) and ALID (7) kinds
                    const int off[8] = {R_FTOT, -1, R_FCON, R_FBODY, R_FDIR, R_FNEU, R_FDAMP, -1};
                    if (off[kind] >= 0) {
                        const double* R = m.body_rigid + (size_t)b * RIGID_FIELDS;
                        x = R[off[kind]];
                        y = R[off[kind] + 1];
                        r = R[off[kind] + 2];
                    }
                }
                if (holds((int)D[5], component(4, (int)D[1], x, y, r), D[6])) deactivate(m, b, &s_any);
            }
        } else if (type == 8) {
            // Body::Update_damage (Code/Body.h:2667-2700), then [2] comparison, [3] threshold
            const int b = (int)D[1];
            if (t == 0) {
                const int nc = *m.c_count;
                s_lo = first_element_of(m.C, nc, b);
                s_hi = first_element_of(m.C, nc, b + 1);
            }
            __syncthreads();
            const int lo = s_lo, hi = s_hi;
            const int g0 = m.border_off[b], nbr = m.body_nborders[b];
            // per border node: the smallest first internal variable of its
            // elements that have one, 1 when none
            for (int k = 0; k < nbr; k++) {
                const int s0 = m.br_slot0[g0 + k], ns = m.br_nslots[g0 + k];
                for (int j = t; j < ns; j += STATUS_THREADS) {
                    double d = 1.;
                    for (int e = lo; e < hi; e++)
                        if (m.C.borderS[e] == k && m.C.bnodeS[e] == j && m.C.nbint[e] > 0 && m.C.int0[e] < d)
                            d = m.C.int0[e];
                    m.slot_damage[s0 + j] = d;
                }
            }
            __syncthreads();
            if (t == 0) {
                double total_nodes = 0., damage = 0.;
                const double previous = m.body_damage[b];
                for (int k = 0; k < nbr; k++) {
                    const int s0 = m.br_slot0[g0 + k], ns = m.br_nslots[g0 + k];
                    total_nodes += ns;
                    for (int j = 0; j < ns; j++) damage += m.slot_damage[s0 + j];
                }
                damage = damage / total_nodes;
                if (damage < previous) damage = previous;
                m.body_damage[b] = damage;
                if (holds((int)D[2], damage, D[3])) deactivate(m, b, &s_any);
            }
        }
        __syncthreads();
    }
    if (!s_any) return;

    // the deactivated bodies' last master records (their elements of the
    // previous proximity result), in element order
    const int nco = *m.c_count_old;
    int base_d = m.gh_count[0], base_r = m.gh_count[1];
    for (int e0 = 0; e0 < nco; e0 += STATUS_THREADS) {
        const int e = e0 + t;
        int nd = 0, nr = 0;
        bool m1 = false, m2 = false;
        if (e < nco && m.body_flip[m.Cold.bodyS[e]]) {
            m1 = m.rec_bn[3 * e + 1] >= 0;
            m2 = m.rec_bn[3 * e + 2] >= 0;
            nd = (int)m1 + (int)m2;
            const double* q = m.rig_rec + 6 * (size_t)e;
            if (m.body_rigid_flag[m.Cold.bodyM[e]] && (q[3] != 0. || q[4] != 0. || q[5] != 0.)) nr = 1;
        }
        int total;
        const int ex = excl_scan(nd | nr << 16, s_w, total);
        int od = base_d + (ex & 0xffff), orr = base_r + (ex >> 16);
        if (m1) {
            if (od < m.ghost_cap) {
                m.gh_bn[od] = m.rec_bn[3 * e + 1]; m.gh_fx[od] = m.rec_fx[3 * e + 1]; m.gh_fy[od] = m.rec_fy[3 * e + 1];
            }
            od++;
        }
        if (m2 && od < m.ghost_cap) {
            m.gh_bn[od] = m.rec_bn[3 * e + 2]; m.gh_fx[od] = m.rec_fx[3 * e + 2]; m.gh_fy[od] = m.rec_fy[3 * e + 2];
        }
        if (nr && orr < m.ghost_cap) {
            const double* q = m.rig_rec + 6 * (size_t)e;
            m.gh_rb[orr] = m.Cold.bodyM[e];
            m.gh_rf[3 * orr] = q[3]; m.gh_rf[3 * orr + 1] = q[4]; m.gh_rf[3 * orr + 2] = q[5];
        }
        base_d += total & 0xffff;
        base_r += total >> 16;
    }
    if (t == 0) {
        if (base_d > m.ghost_cap || base_r > m.ghost_cap) atomicOr(&m.status_flags[0], 64);
        m.gh_count[0] = min(base_d, m.ghost_cap);
        m.gh_count[1] = min(base_r, m.ghost_cap);
    }
    // their elements of the current proximity result
    const int nc = *m.c_count;
    for (int e = t; e < nc; e += STATUS_THREADS)
        if (m.body_flip[m.C.bodyS[e]]) {
            int4 inf = m.C.info[e];
            if (inf.z & CI_VALID) {
                inf.z |= CI_INACTIVE_S;
                m.C.info[e] = inf;
            }
        }
    // their node blocks
    for (int k = t; k < B.n_blocks; k += STATUS_THREADS)
        if (m.body_flip[B.blk_body[k]]) B.blk_hdr[4 * k + 1].z |= BLK_INACTIVE;
    // their rigid nodes: the force copies of the last step
    for (int q = t; q < m.n_rigid_nodes; q += STATUS_THREADS) {
        const int j = m.rigid_node_ids[q];
        const int b = m.node_body[j];
        if (!m.body_flip[b]) continue;
        const double* R = m.body_rigid + (size_t)b * RIGID_FIELDS;
        const double dr = R[R_DISP + 2] - 0.;
        rigid_node_at(m, j, R, cos(dr), sin(dr), 1, m.X[j], m.Y[j]);
    }
    __syncthreads();
    for (int b = t; b < m.n_bodies; b += STATUS_THREADS) m.body_flip[b] = 0;
}

__global__ void k_ghost_sends(DevModel m) {
    const int ng = min(m.gh_count[0], m.ghost_cap);
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
        const int id = 3 * m.c_cap + g;
        const double fx = m.gh_fx[g], fy = m.gh_fy[g];
        m.rec_fx[id] = fx;
        m.rec_fy[id] = fy;
        bucket_append(m, m.gh_bn[g], id, fx, fy);
    }
}

// ---------------------------------------------------------------------------
// Monitoring (Code/Monitoring.h:8-559)
// ---------------------------------------------------------------------------
namespace {

// a body's work of kind q as Body::*_work holds it after the step: the
// node blocks' compensated accumulators added in block order (deformable
// bodies), the body record (rigid bodies, and every body's contact work)
// plus the contact work its elements hold since the last flush, reduced
// in the order of the flush.  Every thread calls it; thread 0 has the value.
__device__ double body_work(const DevModel& m, const BlockArrays& B, int b, int q, double* red) {
    const int t = threadIdx.x;
    double v = 0.;
    if (q == W_CONTACT) {
        const int nc = *m.c_count;
        for (int e = t; e < nc; e += STATUS_THREADS) {
            const double w = m.C.wacc[e];
            v += m.C.bodyS[e] == b ? w : 0.;
            v += m.C.bodyM[e] == b ? w : 0.;
        }
    }
    red[t] = v;
    __syncthreads();
    for (int h = STATUS_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    double out = 0.;
    if (t == 0) {
        const double* R = m.body_rigid + (size_t)b * RIGID_FIELDS;
        out = R[R_WORK + q];
        if (q == W_CONTACT) {
            if (red[0] != 0.) out += red[0];
        } else if (!m.body_rigid_flag[b]) {
            double S = 0., c = 0.;
            bool has = false;
            for (int k = 0; k < B.n_blocks; k++) {
                if (B.blk_body[k] != b) continue;
                has = true;
                for (int h = 0; h < 2; h++) {
                    const double x = B.blk_work[14 * (size_t)k + 2 * q + h], tt = S + x;
                    c += fabs(S) >= fabs(x) ? (S - tt) + x : (x - tt) + S;
                    S = tt;
                }
            }
            if (has) out = S + c;
        }
    }
    __syncthreads();
    return out;
}

}  // namespace

__global__ void __launch_bounds__(STATUS_THREADS) k_monitor(DevModel m, BlockArrays B, MonitorArgs a) {
    __shared__ double red[STATUS_THREADS];
    __shared__ int s_lo, s_hi, s_dummy;
    const int t = threadIdx.x;
    const int row = *a.n_rows;
    if (row >= a.cap) return;       // the host drains before the rows fill up
    double* out = a.rows + (size_t)row * a.stride;
    int nv = 0;                      // thread 0's value count
    auto push = [&](double v) {
        if (nv + 1 < a.stride) out[1 + nv] = v;
        nv++;
    };
    double x = 0., y = 0., r = 0.;  // thread 0's, carried over the entries
    if (t == 0) {
        const double it = *a.iter + 1.;
        *a.iter = it;
        push(it);
        push(*m.time);
        push(m.dt);
        push(a.neval);
        push(a.max_error);
        push(a.mean_error);
    }
    for (int i = 0; i < a.n_items; i++) {
        const double* D = a.items + RULE_STRIDE * i;
        const int type = (int)D[0];
        if (type >= 0 && type <= 3) {
            const int b = (int)D[2];
            const bool rigid = m.body_rigid_flag[b] != 0;
            const int n0 = m.node_off[b], nn = m.body_nnodes[b];
            const bool body_level = D[3] < 0.;
            if (!rigid && body_level)
                for (int k = t; k < nn; k += STATUS_THREADS) kinematics_node(m, n0 + k);
            __syncthreads();
            if (t == 0) {
                const double* px = type == 0 ? m.xc : (type == 1 ? m.ux : (type == 2 ? m.vx : m.ax));
                const double* py = type == 0 ? m.yc : (type == 1 ? m.uy : (type == 2 ? m.vy : m.ay));
                if (!body_level) {
                    const int j = n0 + (int)D[3];
                    if (rigid) rigid_node(m, j, 1);
                    else kinematics_node(m, j);
                    x = px[j];
                    y = py[j];
                } else if (!rigid) {
                    double mass = 0.;
                    x = 0.;
                    y = 0.;
                    for (int k = 0; k < nn; k++) {
                        x += px[n0 + k] * m.mx[n0 + k];
                        y += py[n0 + k] * m.my[n0 + k];
                        mass += m.mx[n0 + k];
                    }
                    x /= mass;
                    y /= mass;
                } else {
                    const double* R = m.body_rigid + (size_t)b * RIGID_FIELDS;
                    const int off = type == 0 ? R_CUR : (type == 1 ? R_DISP : (type == 2 ? R_VEL : R_ACC));
                    x = R[off];
                    y = R[off + 1];
                    r = R[off + 2];
                }
                const int c = (int)D[1];
                if (c >= 0 && c <= (type == 0 ? 2 : 3)) push(component(type, c, x, y, r));
            }
        } else if (type == 4) {
            // [1] component, [2] force kind, [3] body, [4] node; the body-level
            // branch tests body [3] and loops over the node count of body [2]
            const int kind = (int)D[2], b = (int)D[3];
            const int n0 = m.node_off[b];
            const bool rigid = m.body_rigid_flag[b] != 0;
            const bool body_level = D[4] < 0.;
            const int nk = body_level && !rigid ? m.body_nnodes[kind] : 0;
            for (int k = t; k < nk; k += STATUS_THREADS) node_outputs(m, n0 + k, 1);
            __syncthreads();
            if (t == 0) {
                const double* px[8] = {m.fx, m.fix, m.fcx, m.fbx, m.fdx, m.fnx, m.fkx, m.fax};
                const double* py[8] = {m.fy, m.fiy, m.fcy, m.fby, m.fdy, m.fny, m.fky, m.fay};
                if (!body_level) {
                    const int j = n0 + (int)D[4];
                    if (rigid) rigid_node(m, j, 1);
                    else node_outputs(m, j, 1);
                    if (kind <= 7) { x = px[kind][j]; y = py[kind][j]; }
                } else if (!rigid) {
                    x = 0.;
                    y = 0.;
                    if (kind <= 7)
                        for (int k = 0; k < nk; k++) { x += px[kind][n0 + k]; y += py[kind][n0 + k]; }
                } else {
                    const int off[8] = {R_FTOT, -1, R_FCON, R_FBODY, R_FDIR, R_FNEU, R_FDAMP, -1};
                    if (kind <= 7 && off[kind] >= 0) {
                        const double* R = m.body_rigid + (size_t)b * RIGID_FIELDS;
                        x = R[off[kind]];
                        y = R[off[kind] + 1];
                        r = R[off[kind] + 2];
                    }
                }
                const int c = (int)D[1];
                if (c >= 0 && c <= 3) push(component(4, c, x, y, r));
            }
        } else if (type == 5 || type == 6) {
            // Node::jacobian and stresses of the last Compute_nodal_stresses
            // (Code/Body.h:2310-2519; 1 and 0 before the first, Code/Node.h:172-185)
            if (t == 0) {
                const int b = (int)D[type == 5 ? 1 : 2], j = m.node_off[b] + (int)D[type == 5 ? 2 : 3];
                // Sigmaxx Sigmayy Sigmaxy Sigmazz Tresca VM I II III in the field-major layout
                const int field6[9] = {8, 9, 10, 11, 12, 13, 0, 1, 2};
                const int c = type == 6 ? (int)D[1] : 0;
                if (type == 5) push(a.stress ? a.stress[3 * (size_t)m.n_nodes + j] : 1.);
                else if (c >= 0 && c <= 8) push(a.stress ? a.stress[(size_t)field6[c] * m.n_nodes + j] : 0.);
            }
        } else if (type == 7) {
            // the first element of slave body [2] on border [3], border node [4]
            if (t == 0) {
                const int b = (int)D[2], c = (int)D[1];
                const int nc = *m.c_count;
                const int lo = first_element_of(m.C, nc, b), hi = first_element_of(m.C, nc, b + 1);
                for (int e = lo; e < hi; e++) {
                    if (m.C.borderS[e] == (int)D[3] && m.C.bnodeS[e] == (int)D[4]) {
                        const double v[6] = {m.C.gapn[e], m.C.gapt[e], m.C.xsi[e], m.C.xn[e], m.C.yn[e], m.C.int0[e]};
                        if (c >= 0 && c <= 5) push(v[c]);
                        break;
                    }
                }
            }
        } else if (type == 8) {
            // Body::Update_damage, then the damage
            const int b = (int)D[1];
            if (t == 0) {
                const int nc = *m.c_count;
                s_lo = first_element_of(m.C, nc, b);
                s_hi = first_element_of(m.C, nc, b + 1);
            }
            __syncthreads();
            const int lo = s_lo, hi = s_hi;
            const int g0 = m.border_off[b], nbr = m.body_nborders[b];
            for (int k = 0; k < nbr; k++) {
                const int s0 = m.br_slot0[g0 + k], ns = m.br_nslots[g0 + k];
                for (int j = t; j < ns; j += STATUS_THREADS) {
                    double d = 1.;
                    for (int e = lo; e < hi; e++)
                        if (m.C.borderS[e] == k && m.C.bnodeS[e] == j && m.C.nbint[e] > 0 && m.C.int0[e] < d)
                            d = m.C.int0[e];
                    m.slot_damage[s0 + j] = d;
                }
            }
            __syncthreads();
            if (t == 0) {
                double total_nodes = 0., damage = 0.;
                const double previous = m.body_damage[b];
                for (int k = 0; k < nbr; k++) {
                    const int s0 = m.br_slot0[g0 + k], ns = m.br_nslots[g0 + k];
                    total_nodes += ns;
                    for (int j = 0; j < ns; j++) damage += m.slot_damage[s0 + j];
                }
                damage = damage / total_nodes;
                if (damage < previous) damage = previous;
                m.body_damage[b] = damage;
                push(damage);
            }
        } else if (type == 9) {
            // kinetic energy from the velocity parameters (rigid bodies: from
            // the body's velocity, mass, inertia); the elastic energies are
            // zero (the laws set their energy to 0, Code/Model.h:28, 73)
            if (t == 0) {
                const int c = (int)D[1], bsel = (int)D[2];
                x = 0.;
                if (bsel >= 0 && c == 0) x += a.body_energy[3 * bsel + 1];
                else if (bsel == -1 && c == 0)
                    for (int n = 0; n < m.n_bodies; n++) x += a.body_energy[3 * n + 1];
                push(x);
            }
        } else if (type == 10) {
            const int c = (int)D[1], bsel = (int)D[2];
            if (bsel >= 0 && c >= 0 && c <= 6) {
                const double w = body_work(m, B, bsel, c, red);
                if (t == 0) push(w);
            } else if (bsel == -1) {
                double acc = 0.;
                if (c >= 0 && c <= 6)
                    for (int n = 0; n < m.n_bodies; n++) {
                        const double w = body_work(m, B, n, c, red);
                        if (t == 0) acc += w;
                    }
                if (t == 0) { x = acc; push(x); }
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        if (a.energy) {
            double dwork = 0., kinetic = 0.;
            for (int n = 0; n < m.n_bodies; n++) {
                dwork += a.body_energy[3 * n + 2];
                kinetic += a.body_energy[3 * n];
            }
            push(dwork);
            push(0.);
            push(kinetic);
        }
        out[0] = nv;
        *a.n_rows = row + 1;
    }
    (void)s_dummy;
}

}  // namespace mel
