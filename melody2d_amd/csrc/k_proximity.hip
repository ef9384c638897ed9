// Contact detection, Update_proximity (Code/Proximity.h:48-473), on the
// device.  The element order of the reference is kept exactly:
//   per slave body: [non-self]  borderS -> neighbour -> border nodeS -> segmentM
//                   [self]      borderS -> border nodeS -> borderM -> segmentM
// Each (body, borderS, neighbour, nodeS) "row" (or (body, borderS, nodeS,
// borderM) self row) scans its master segments in order; rows are counted,
// scanned and written, so the element index of a candidate equals its rank in
// the reference's loop nest.  History transfer (Code/Proximity.h:271-319)
// looks up, per slave slot, the previous elements in their old order.
#include "mel_device.hpp"
#include "k_step.hpp"
#include "scan.hpp"

namespace mel {

constexpr int OLD_CAP = 16;    // previous elements per slave slot

// per chunk of SEG_CHUNK border slots: the box of the slot nodes inflated by
// the detection distance (its part of the border AABB of
// Code/Proximity.h:63-109) and the box of the chunk's segments, used to skip
// whole chunks of master segments in the narrow phase
__global__ void k_prox_chunk_boxes(DevModel m) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_seg_chunks) return;
    int br = m.sc_border[c];
    int b = m.br_body[br];
    if (m.body_status[b] == 1) return;
    int s0 = m.br_slot0[br], ns = m.br_nslots[br];
    int first = (c - m.br_chunk0[br]) * SEG_CHUNK;
    double d = m.body_detect[b];
    double xmin = 1.e100, xmax = -1.e100, ymin = 1.e100, ymax = -1.e100;
    for (int j = first; j < min(first + SEG_CHUNK, ns); j++) {
        int nd = m.sl_node[s0 + j];
        double x = m.xc[nd], y = m.yc[nd];
        if (x - d < xmin) xmin = x - d;
        if (x + d > xmax) xmax = x + d;
        if (y - d < ymin) ymin = y - d;
        if (y + d > ymax) ymax = y + d;
    }
    m.sc_nxmin[c] = xmin; m.sc_nxmax[c] = xmax; m.sc_nymin[c] = ymin; m.sc_nymax[c] = ymax;
    double P = m.period;
    xmin = 1.e100; xmax = -1.e100; ymin = 1.e100; ymax = -1.e100;
    for (int seg = first; seg < min(first + SEG_CHUNK, ns - 1); seg++) {
        int s = s0 + seg;
        int n1 = m.sl_n1[s], n2 = m.sl_n2[s];
        double x1 = m.xc[n1] + P * m.sl_s1[s], x2 = m.xc[n2] + P * m.sl_s2[s];
        double y1 = m.yc[n1], y2 = m.yc[n2];
        xmin = fmin(xmin, fmin(x1, x2)); xmax = fmax(xmax, fmax(x1, x2));
        ymin = fmin(ymin, fmin(y1, y2)); ymax = fmax(ymax, fmax(y1, y2));
    }
    m.sc_xmin[c] = xmin; m.sc_xmax[c] = xmax; m.sc_ymin[c] = ymin; m.sc_ymax[c] = ymax;
}

// border AABBs inflated by the detection distance + periodic shift range
// (Code/Proximity.h:63-109): min / max over the chunk boxes, which equals
// the reference's running min / max over the border nodes
__global__ void k_prox_aabb(DevModel m) {
    int br = blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= m.n_borders) return;
    int b = m.br_body[br];
    if (m.body_status[b] == 1) return;
    double xmin = 1.e100, xmax = -1.e100, ymin = 1.e100, ymax = -1.e100;
    for (int c = m.br_chunk0[br]; c < m.br_chunk0[br + 1]; c++) {
        if (m.sc_nxmin[c] < xmin) xmin = m.sc_nxmin[c];
        if (m.sc_nxmax[c] > xmax) xmax = m.sc_nxmax[c];
        if (m.sc_nymin[c] < ymin) ymin = m.sc_nymin[c];
        if (m.sc_nymax[c] > ymax) ymax = m.sc_nymax[c];
    }
    double inv = 1. / m.period;
    m.pb_xmin[br] = xmin; m.pb_xmax[br] = xmax; m.pb_ymin[br] = ymin; m.pb_ymax[br] = ymax;
    m.pb_shmin[br] = (int)floor((xmin - m.xmin_period) * inv);
    m.pb_shmax[br] = (int)floor((xmax - m.xmin_period) * inv);
}

// neighbour lists (Code/Proximity.h:112-169).  One 64-wide block per body1;
// lanes take consecutive body2 so the order of insertion is the reference's
// (border1, body2, border2, shift) order with first-occurrence de-duplication.
__global__ void __launch_bounds__(64) k_prox_neighbours(DevModel m) {
    int b1 = blockIdx.x;
    int lane = threadIdx.x;
    __shared__ int count_sh;
    __shared__ int pref[64];
    if (lane == 0) count_sh = 0;
    __syncthreads();
    int* list = m.nb_list + (size_t)b1 * m.nb_cap * 3;
    if (m.body_status[b1] != 1) {
        double P = m.period;
        for (int r1 = 0; r1 < m.body_nborders[b1]; r1++) {
            int g1 = m.border_off[b1] + r1;
            double xmin1 = m.pb_xmin[g1], xmax1 = m.pb_xmax[g1], ymin1 = m.pb_ymin[g1], ymax1 = m.pb_ymax[g1];
            int shmin1 = m.pb_shmin[g1], shmax1 = m.pb_shmax[g1];
            int before = count_sh;     // entries from earlier border1 iterations
            for (int base = 0; base < m.n_bodies; base += 64) {
                int b2 = base + lane;
                int mine = 0;
                // pass 1: count this lane's new neighbours
                for (int pass = 0; pass < 2; pass++) {
                    int w = 0;
                    if (b2 < m.n_bodies && b2 != b1 && m.body_status[b2] != 1) {
                        for (int r2 = 0; r2 < m.body_nborders[b2]; r2++) {
                            int g2 = m.border_off[b2] + r2;
                            if ((m.pb_ymax[g2] < ymin1) || (ymax1 < m.pb_ymin[g2])) continue;
                            for (int sh = shmin1 - m.pb_shmax[g2]; sh <= shmax1 - m.pb_shmin[g2]; sh++) {
                                if ((m.pb_xmax[g2] + sh * P > xmin1) && (m.pb_xmin[g2] + sh * P < xmax1)) {
                                    bool exists = false;
                                    for (int k = 0; k < before && k < m.nb_cap; k++)
                                        if (list[3 * k] == b2 && list[3 * k + 1] == r2 && list[3 * k + 2] == sh) { exists = true; break; }
                                    if (exists) continue;
                                    if (pass == 1) {
                                        int pos = count_sh + pref[lane] + w;
                                        if (pos < m.nb_cap) {
                                            list[3 * pos] = b2; list[3 * pos + 1] = r2; list[3 * pos + 2] = sh;
                                        }
                                    }
                                    w++;
                                }
                            }
                        }
                    }
                    if (pass == 0) {
                        mine = w;
                        pref[lane] = mine;
                        __syncthreads();
                        // exclusive scan over the 64 lanes
                        for (int off = 1; off < 64; off <<= 1) {
                            int v = lane >= off ? pref[lane - off] : 0;
                            __syncthreads();
                            pref[lane] += v;
                            __syncthreads();
                        }
                        pref[lane] -= mine;
                        __syncthreads();
                    }
                }
                __syncthreads();
                if (lane == 63) count_sh += pref[63] + mine;
                __syncthreads();
            }
        }
    }
    if (lane == 0) {
        if (count_sh > m.nb_cap) atomicOr(&m.status_flags[0], 2);
        m.nb_count[b1] = min(count_sh, m.nb_cap);
    }
}

__device__ inline int body_slots(const DevModel& m, int b) {
    int g0 = m.border_off[b], g1 = g0 + m.body_nborders[b];
    return m.br_slot0[g1 - 1] + m.br_nslots[g1 - 1] - m.br_slot0[g0];
}

__device__ inline bool self_enabled(const DevModel& m, int b) {
    return m.self_contact && !m.body_rigid_flag[b] && m.body_periodicity[b] != 2;
}

// rows per body (non-self + self)
__global__ void k_prox_rows(DevModel m) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.n_bodies) return;
    int r = 0;
    if (m.body_status[b] != 1 && m.body_nborders[b] > 0) {
        int ns = body_slots(m, b);
        r = m.nb_count[b] * ns;
        if (self_enabled(m, b)) r += ns * m.body_nborders[b];
    }
    m.body_rows[b] = r;
}

struct RowInfo {
    int body, self, borderS, k, jS, borderM;
};

__device__ inline bool decode_row(const DevModel& m, int row, RowInfo& ri) {
    int total = m.body_row_off[m.n_bodies];
    if (row >= total) return false;
    // binary search for the body
    int lo = 0, hi = m.n_bodies;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (m.body_row_off[mid] <= row) lo = mid;
        else hi = mid;
    }
    int b = lo;
    int local = row - m.body_row_off[b];
    int nb = m.nb_count[b];
    int g0 = m.border_off[b];
    int s0 = m.br_slot0[g0];
    int ns = body_slots(m, b);
    ri.body = b;
    if (local < nb * ns) {
        ri.self = 0;
        for (int r = 0; r < m.body_nborders[b]; r++) {
            int n = m.br_nslots[g0 + r];
            int pre = m.br_slot0[g0 + r] - s0;
            if (local < nb * (pre + n)) {
                int q = local - nb * pre;
                ri.borderS = r;
                ri.k = q / n;
                ri.jS = q - ri.k * n;
                ri.borderM = -1;
                return true;
            }
        }
        return false;
    }
    local -= nb * ns;
    ri.self = 1;
    int nbr = m.body_nborders[b];
    int q = local / nbr;          // slot within body
    ri.borderM = local - q * nbr;
    for (int r = 0; r < nbr; r++) {
        int n = m.br_nslots[g0 + r];
        int pre = m.br_slot0[g0 + r] - s0;
        if (q < pre + n) {
            ri.borderS = r;
            ri.jS = q - pre;
            ri.k = -1;
            return true;
        }
    }
    return false;
}

// scans one row; write == false counts, write == true emits elements at base
template <bool WRITE>
__device__ int scan_row(const DevModel& m, const RowInfo& ri, int base) {
    int bS = ri.body;
    int gS = m.border_off[bS] + ri.borderS;
    int slotS = m.br_slot0[gS] + ri.jS;
    int nodeS = m.sl_node[slotS];
    double Xs = m.xc[nodeS], Ys = m.yc[nodeS];
    double detect = m.body_detect[bS], cdist = m.body_contact[bS];
    double P = m.period;
    int bM, rM, shM;
    if (!ri.self) {
        const int* nb = m.nb_list + ((size_t)bS * m.nb_cap + ri.k) * 3;
        bM = nb[0]; rM = nb[1]; shM = nb[2];
    } else {
        bM = bS; rM = ri.borderM; shM = 0;
    }
    int gM = m.border_off[bM] + rM;
    int nM = m.br_nslots[gM];
    int s0 = m.br_slot0[gM];
    double emass;
    if (!ri.self) {
        double mM = m.body_mass[bM], mS = m.body_mass[bS];
        if (mM == 0.) emass = 0.;
        else if (mS == 0.) emass = 0.;
        else emass = 1. / (m.body_inv_mass[bM] + m.body_inv_mass[bS]);
    } else {
        emass = m.body_mass[bS];
    }
    int cnt = 0;
    // chunks whose segment box lies farther than the detection distance are
    // skipped: detect_segment only accepts points within that distance of a
    // segment; the margin covers rounding of the shifted coordinates
    const double Xq = Xs - P * shM;
    const double dd = 1.000001 * detect + 1e-10 * (1. + fabs(Xs) + fabs(Ys) + fabs(P * shM));
    const int cM0 = m.br_chunk0[gM], cM1 = m.br_chunk0[gM + 1];
    for (int ch = cM0; ch < cM1; ch++) {
    if (!(Xq > m.sc_xmin[ch] - dd && Xq < m.sc_xmax[ch] + dd && Ys > m.sc_ymin[ch] - dd && Ys < m.sc_ymax[ch] + dd))
        continue;
    const int sg0 = (ch - cM0) * SEG_CHUNK, sg1 = min(sg0 + SEG_CHUNK, nM - 1);
    for (int seg = sg0; seg < sg1; seg++) {
        int s = s0 + seg;
        int n1 = m.sl_n1[s], n2 = m.sl_n2[s];
        if (ri.self && (n1 == nodeS || n2 == nodeS)) continue;
        double X1 = m.xc[n1] + P * (shM + m.sl_s1[s]), Y1 = m.yc[n1];
        double X2 = m.xc[n2] + P * (shM + m.sl_s2[s]), Y2 = m.yc[n2];
        if (!detect_segment(Xs, Ys, X1, Y1, X2, Y2, detect)) continue;
        int n0 = m.sl_n0[s], n3 = m.sl_n3[s];
        double X0 = m.xc[n0] + P * (shM + m.sl_s0[s]), Y0 = m.yc[n0];
        double X3 = m.xc[n3] + P * (shM + m.sl_s3[s]), Y3 = m.yc[n3];
        Proj p;
        p.gap = 0.; p.xsi = 0.; p.xn = 0.; p.yn = 0.; p.xt = 0.; p.yt = 0.;
        p.N0 = 0.; p.N1 = 0.; p.N2 = 0.; p.N3 = 0.; p.flag = 0;
        bool accept;
        if (ri.self && (n0 == nodeS || n3 == nodeS)) {
            closest_self(Xs, Ys, X1, Y1, X2, Y2, p);
            accept = false;       // the self projection always ends with flag -3
        } else {
            closest_linear(Xs, Ys, X0, Y0, X1, Y1, X2, Y2, X3, Y3, p);
            accept = ri.self ? (p.flag == 0) : (p.flag == -2 || p.flag == -1 || p.flag == 0);
            accept = accept && (p.gap > -cdist);
        }
        if (!accept) continue;
        if (WRITE) {
            int e = base + cnt;
            if (e < m.c_cap) {
                const ContactArrays& C = m.C;
                C.bodyS[e] = bS; C.borderS[e] = ri.borderS; C.bnodeS[e] = ri.jS; C.nodeS[e] = nodeS;
                C.bodyM[e] = bM; C.borderM[e] = rM; C.shiftM[e] = shM; C.segM[e] = seg;
                C.nM0[e] = n0; C.nM1[e] = n1; C.nM2[e] = n2; C.nM3[e] = n3;
                if (ri.self) { C.sM0[e] = 0; C.sM1[e] = 0; C.sM2[e] = 0; C.sM3[e] = 0; }
                else { C.sM0[e] = m.sl_s0[s]; C.sM1[e] = m.sl_s1[s]; C.sM2[e] = m.sl_s2[s]; C.sM3[e] = m.sl_s3[s]; }
                C.N0[e] = p.N0; C.N1[e] = p.N1; C.N2[e] = p.N2; C.N3[e] = p.N3;
                C.gapn[e] = p.gap; C.gapt[e] = 0.; C.xsi[e] = p.xsi;
                C.xn[e] = p.xn; C.yn[e] = p.yn; C.xt[e] = p.xt; C.yt[e] = p.yt;
                C.emass[e] = emass; C.len[e] = m.sl_len[slotS]; C.fx[e] = 0.; C.fy[e] = 0.;
                C.nbint[e] = 0; C.int0[e] = 0.;
                C.vgapn[e] = 0.; C.vgapt[e] = 0.;
                C.fresh[e] = ri.self ? 2 : 1;
            }
        }
        cnt++;
    }
    }
    return cnt;
}

__global__ void k_prox_count(DevModel m, int n_rows) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n_rows) return;
    RowInfo ri;
    if (!decode_row(m, row, ri)) { m.row_count[row] = 0; return; }
    m.row_count[row] = scan_row<false>(m, ri, 0);
}

__global__ void k_prox_write(DevModel m, int n_rows) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n_rows) return;
    RowInfo ri;
    if (!decode_row(m, row, ri)) return;
    if (m.row_count[row] == 0) return;
    scan_row<true>(m, ri, m.row_off[row]);
}

// index of the previous elements by slave slot, for history transfer
__global__ void k_prox_old_index(DevModel m) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *m.c_count_old) return;
    const ContactArrays& O = m.Cold;
    int slot = m.br_slot0[m.border_off[O.bodyS[e]] + O.borderS[e]] + O.bnodeS[e];
    int pos = atomicAdd(&m.slot_old_cursor[slot], 1);
    if (pos < OLD_CAP) m.slot_old_list[slot * OLD_CAP + pos] = e;
    else atomicOr(&m.status_flags[0], 4);
}

__device__ inline void copy_element(const ContactArrays& D, int e, const ContactArrays& S, int o) {
    D.bodyS[e] = S.bodyS[o]; D.borderS[e] = S.borderS[o]; D.bnodeS[e] = S.bnodeS[o]; D.nodeS[e] = S.nodeS[o];
    D.bodyM[e] = S.bodyM[o]; D.borderM[e] = S.borderM[o]; D.shiftM[e] = S.shiftM[o]; D.segM[e] = S.segM[o];
    D.nbint[e] = S.nbint[o];
    D.nM0[e] = S.nM0[o]; D.nM1[e] = S.nM1[o]; D.nM2[e] = S.nM2[o]; D.nM3[e] = S.nM3[o];
    D.sM0[e] = S.sM0[o]; D.sM1[e] = S.sM1[o]; D.sM2[e] = S.sM2[o]; D.sM3[e] = S.sM3[o];
    D.N0[e] = S.N0[o]; D.N1[e] = S.N1[o]; D.N2[e] = S.N2[o]; D.N3[e] = S.N3[o];
    D.gapn[e] = S.gapn[o]; D.vgapn[e] = S.vgapn[o]; D.gapt[e] = S.gapt[o]; D.vgapt[e] = S.vgapt[o];
    D.xsi[e] = S.xsi[o]; D.xn[e] = S.xn[o]; D.yn[e] = S.yn[o]; D.xt[e] = S.xt[o]; D.yt[e] = S.yt[o];
    D.emass[e] = S.emass[o]; D.len[e] = S.len[o]; D.fx[e] = S.fx[o]; D.fy[e] = S.fy[o]; D.int0[e] = S.int0[o];
}

// history transfer: first previous element of the same slave slot with the
// same master body and border and |segment difference| <= 1
__global__ void k_prox_history(DevModel m) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *m.c_count) return;
    const ContactArrays& C = m.C;
    if (C.fresh[e] != 1) return;          // self elements never inherit (Code/Proximity.h:410-425)
    int slot = m.br_slot0[m.border_off[C.bodyS[e]] + C.borderS[e]] + C.bnodeS[e];
    int n = min(m.slot_old_cursor[slot], OLD_CAP);
    const int* lst = m.slot_old_list + slot * OLD_CAP;
    int best = -1;
    for (int q = 0; q < n; q++) {
        int o = lst[q];
        if (best >= 0 && o > best) continue;
        if (m.Cold.bodyM[o] == C.bodyM[e] && m.Cold.borderM[o] == C.borderM[e] &&
            abs(m.Cold.segM[o] - C.segM[e]) <= 1)
            best = o;
    }
    if (best >= 0) {
        copy_element(C, e, m.Cold, best);
        C.fresh[e] = 0;
    }
}

// fresh elements carry the relative velocities of the element created just
// before them in the same pass (the reused new_contact record of
// Code/Proximity.h:211 and :350), zero at the start of a pass
__global__ void k_prox_chain(DevModel m) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    int n = *m.c_count;
    if (e >= n) return;
    const ContactArrays& C = m.C;
    int f = C.fresh[e];
    if (f == 0) return;
    int b = C.bodyS[e];
    int phase = f == 2;
    double vn = 0., vt = 0.;
    for (int q = e - 1; q >= 0; q--) {
        if (C.bodyS[q] != b) break;
        int fq = C.fresh[q];
        int ph = (fq == 2) || (fq == 0 && C.bodyM[q] == b);
        if (ph != phase) break;
        if (fq == 0) { vn = C.vgapn[q]; vt = C.vgapt[q]; break; }
    }
    C.vgapn[e] = vn;
    C.vgapt[e] = vt;
}

__global__ void k_prox_finish(DevModel m, int n_rows) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int total = n_rows > 0 ? m.row_off[n_rows - 1] + m.row_count[n_rows - 1] : 0;
    if (total > m.c_cap) { atomicOr(&m.status_flags[0], 8); total = m.c_cap; }
    *m.c_count = total;
}

// runs after the host swapped the element buffers: C is empty, Cold holds
// the previous elements whose count is still in c_count
__global__ void k_prox_clear_old(DevModel m) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m.n_slots) m.slot_old_cursor[s] = 0;
    if (s == 0) *m.c_count_old = *m.c_count;
}

// Initialize_CZM (Code/Proximity.h:481-541) after the elements have been
// re-projected: damage 0 below gapinit, 1 otherwise, for CZMlinear pairs.
__global__ void k_init_czm(DevModel m) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *m.c_count) return;
    const ContactArrays& C = m.C;
    int ms = m.body_material[C.bodyS[e]], mm = m.body_material[C.bodyM[e]];
    int li = m.law_first[ms * m.n_materials + mm];
    if (li < 0 || m.law_type[li] != 6) return;
    C.nbint[e] = 1;
    C.int0[e] = C.gapn[e] < m.law_p[12 * li + 5] ? 0. : 1.;
}

__global__ void k_project_all(DevModel m) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *m.c_count) return;
    project_element(m, m.C, e);
}

}  // namespace mel
