// Contact detection, Update_proximity (Code/Proximity.h:48-473), on the
// device.  The element order of the reference is kept exactly:
//   per slave body: [non-self]  borderS -> neighbour -> border nodeS -> segmentM
//                   [self]      borderS -> border nodeS -> borderM -> segmentM
// Each (body, borderS, neighbour, nodeS) "row" (or (body, borderS, nodeS,
// borderM) self row) scans its master segments in order; rows are counted,
// scanned and written, so the element index of a candidate equals its rank in
// the reference's loop nest.  History transfer (Code/Proximity.h:271-319)
// looks up, per slave slot, the previous elements in their old order.
#include <climits>
#include "mel_device.hpp"
#include "k_rigid_lists.hpp"
#include "k_step.hpp"

namespace mel {

constexpr int OLD_CAP = 16;    // previous elements per slave slot

// exclusive prefix of v over the workgroup (blockDim.x a multiple of 64, at
// most 1024): wavefront scans, then the wavefront totals in order.
// s_w: at least blockDim.x / 64 ints of LDS.  Every thread must call it.
__device__ inline int wg_excl_scan(int v, int* s_w, int& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wv] = x;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int k = 0; k < nw; k++) {
        int t = s_w[k];
        if (k < wv) before += t;
        total += t;
    }
    __syncthreads();
    return before + x - v;
}

// out[0..n) = exclusive prefix of in[0..n), out[n] = total, by one workgroup
__device__ inline void wg_scan_array(const int* in, int* out, int n, int* s_w) {
    int base = 0;
    for (int t0 = 0; t0 < n; t0 += blockDim.x) {
        const int i = t0 + threadIdx.x;
        const int v = i < n ? in[i] : 0;
        int tot;
        const int ex = wg_excl_scan(v, s_w, tot);
        if (i < n) out[i] = base + ex;
        base += tot;
    }
    if (threadIdx.x == 0) out[n] = base;
}

// per chunk of SEG_CHUNK border slots, SEG_CHUNK lanes (lane j takes slot
// and segment j of the chunk): the box of the slot nodes inflated by the
// detection distance (its part of the border AABB of
// Code/Proximity.h:63-109), the box of the chunk's segments, used to skip
// whole chunks of master segments in the narrow phase, and the segments'
// end-node records read by the narrow phase.  Minima and maxima do not
// depend on the order they are taken in.
__global__ void k_prox_chunk_boxes(DevModel m) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    // the cell list's counters start again (nothing reads them before
    // k_cells_assign)
    for (int q = g; q < m.cell_n; q += gridDim.x * blockDim.x) {
        m.cell_count[q] = 0;
        m.cell_fill[q] = 0;
    }
    if (g == 0) *m.big_count = 0;
    const int c = g / SEG_CHUNK, j = g % SEG_CHUNK;
    const bool live = c < m.n_seg_chunks;
    const int br = live ? m.sc_border[c] : 0;
    const int b = m.br_body[br];
    const bool on = live && m.body_status[b] != 1;
    const int s0 = m.br_slot0[br], ns = m.br_nslots[br];
    const int k = (c - m.br_chunk0[br]) * SEG_CHUNK + j;
    const double d = m.body_detect[b];
    double nxmin = 1.e100, nxmax = -1.e100, nymin = 1.e100, nymax = -1.e100;
    double xmin = 1.e100, xmax = -1.e100, ymin = 1.e100, ymax = -1.e100;
    // the slots' old-element cursors for the history transfer start again
    // (the count of the elements about to become the previous ones is kept)
    if (live && k < ns) m.slot_old_cursor[s0 + k] = 0;
    if (g == 0) *m.c_count_old = *m.c_count;
    if (on && k < ns) {
        const int nd = m.sl_node[s0 + k];
        const double x = m.xc[nd], y = m.yc[nd];
        nxmin = x - d; nxmax = x + d; nymin = y - d; nymax = y + d;
        m.slot_xy[s0 + k] = make_double2(x, y);
    }
    if (on && k < ns - 1) {
        const int s = s0 + k;
        const int n1 = m.sl_n1[s], n2 = m.sl_n2[s], s1 = m.sl_s1[s], s2 = m.sl_s2[s];
        const double xa = m.xc[n1], ya = m.yc[n1], xb = m.xc[n2], yb = m.yc[n2];
        const double P = m.period;
        const double x1 = xa + P * s1, x2 = xb + P * s2;
        xmin = fmin(x1, x2); xmax = fmax(x1, x2); ymin = fmin(ya, yb); ymax = fmax(ya, yb);
        m.seg_ns[s] = make_int4(n1, n2, s1, s2);
        m.seg_xy[s] = make_double4(xa, ya, xb, yb);
    }
    for (int off = SEG_CHUNK / 2; off > 0; off >>= 1) {
        nxmin = fmin(nxmin, __shfl_xor(nxmin, off, 64)); nxmax = fmax(nxmax, __shfl_xor(nxmax, off, 64));
        nymin = fmin(nymin, __shfl_xor(nymin, off, 64)); nymax = fmax(nymax, __shfl_xor(nymax, off, 64));
        xmin = fmin(xmin, __shfl_xor(xmin, off, 64)); xmax = fmax(xmax, __shfl_xor(xmax, off, 64));
        ymin = fmin(ymin, __shfl_xor(ymin, off, 64)); ymax = fmax(ymax, __shfl_xor(ymax, off, 64));
    }
    if (on && j == 0) {
        m.sc_nxmin[c] = nxmin; m.sc_nxmax[c] = nxmax; m.sc_nymin[c] = nymin; m.sc_nymax[c] = nymax;
        m.sc_box[c] = make_double4(xmin, xmax, ymin, ymax);
    }
    // the wavefront's SC_GROUP chunks: their union (chunks of neighbouring
    // borders may share it, which only makes it larger)
    for (int off = SEG_CHUNK; off < 64; off <<= 1) {
        xmin = fmin(xmin, __shfl_xor(xmin, off, 64)); xmax = fmax(xmax, __shfl_xor(xmax, off, 64));
        ymin = fmin(ymin, __shfl_xor(ymin, off, 64)); ymax = fmax(ymax, __shfl_xor(ymax, off, 64));
    }
    if ((threadIdx.x & 63) == 0 && c < m.n_seg_chunks) m.sc_box2[g / 64] = make_double4(xmin, xmax, ymin, ymax);
}

// border AABBs inflated by the detection distance + periodic shift range
// (Code/Proximity.h:63-109): min / max over the chunk boxes, which equals
// the reference's running min / max over the border nodes; one wavefront
// per border
__global__ void k_prox_aabb(DevModel m) {
    const int br = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x & 63;
    const bool on = br < m.n_borders && m.body_status[m.br_body[br < m.n_borders ? br : 0]] != 1;
    double xmin = 1.e100, xmax = -1.e100, ymin = 1.e100, ymax = -1.e100;
    double sxmin = 1.e100, sxmax = -1.e100, symin = 1.e100, symax = -1.e100;
    if (on)
        for (int c = m.br_chunk0[br] + lane; c < m.br_chunk0[br + 1]; c += 64) {
            xmin = fmin(xmin, m.sc_nxmin[c]); xmax = fmax(xmax, m.sc_nxmax[c]);
            ymin = fmin(ymin, m.sc_nymin[c]); ymax = fmax(ymax, m.sc_nymax[c]);
            const double4 sb = m.sc_box[c];
            sxmin = fmin(sxmin, sb.x); sxmax = fmax(sxmax, sb.y); symin = fmin(symin, sb.z); symax = fmax(symax, sb.w);
        }
    for (int off = 32; off > 0; off >>= 1) {
        xmin = fmin(xmin, __shfl_xor(xmin, off, 64)); xmax = fmax(xmax, __shfl_xor(xmax, off, 64));
        ymin = fmin(ymin, __shfl_xor(ymin, off, 64)); ymax = fmax(ymax, __shfl_xor(ymax, off, 64));
        sxmin = fmin(sxmin, __shfl_xor(sxmin, off, 64)); sxmax = fmax(sxmax, __shfl_xor(sxmax, off, 64));
        symin = fmin(symin, __shfl_xor(symin, off, 64)); symax = fmax(symax, __shfl_xor(symax, off, 64));
    }
    if (!on || lane != 0) return;
    m.pb_segbox[br] = make_double4(sxmin, sxmax, symin, symax);
    double inv = 1. / m.period;
    m.pb_xmin[br] = xmin; m.pb_xmax[br] = xmax; m.pb_ymin[br] = ymin; m.pb_ymax[br] = ymax;
    m.pb_shmin[br] = (int)floor((xmin - m.xmin_period) * inv);
    m.pb_shmax[br] = (int)floor((xmax - m.xmin_period) * inv);
}

// ---------------------------------------------------------------------------
// Broad phase: a cell list over the border AABBs (Code/Proximity.h:63-109
// computes them; the reference then tests every border pair,
// Code/Proximity.h:112-169).  Borders are binned by the centre of their
// AABB on a uniform grid (x taken modulo the period, so cells wrap across
// the periodic boundary); a border larger than a cell is "big" and kept in
// a list of its own.  The candidates of a border are the borders binned in
// the cells its AABB covers grown by one cell, and every big border: two
// AABBs of at most a cell each that overlap have centres at most a cell
// apart, so nothing the reference finds is missed.  Every candidate gets
// the reference's exact tests, and the body's neighbour list is put in the
// reference's order: sorted by (border1, body2, border2, shift), keeping the
// first of equal (body2, border2, shift) (Code/Proximity.h:149-159).
// ---------------------------------------------------------------------------
__global__ void k_cells_clear(DevModel m) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m.cell_n; c += gridDim.x * blockDim.x) {
        m.cell_count[c] = 0;
        m.cell_fill[c] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *m.big_count = 0;
}

// column of a (wrapped) cell coordinate q, which may be any finite double
__device__ inline int wrap_col(const DevModel& m, double q) {
    double r = fmod(q, (double)m.cell_nx);
    if (r < 0.) r += m.cell_nx;
    const int c = (int)r;
    return min(max(c, 0), m.cell_nx - 1);
}

__device__ inline int cell_col(const DevModel& m, double x) {
    return wrap_col(m, floor((x - m.xmin_period) / m.cell_dx));
}

__device__ inline int cell_row(const DevModel& m, double y) {
    const double r = fmin(fmax(floor((y - m.cell_y0) / m.cell_dy), 0.), (double)(m.cell_ny - 1));
    return (int)r;
}

// bin one border: its cell, or the big list (-2); inactive bodies: -1
__global__ void k_cells_assign(DevModel m) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m.n_borders) return;
    if (m.body_status[m.br_body[g]] == 1) { m.border_cell[g] = -1; return; }
    const double xmin = m.pb_xmin[g], xmax = m.pb_xmax[g], ymin = m.pb_ymin[g], ymax = m.pb_ymax[g];
    // a border whose box is not finite (no slot, or a state gone bad, which
    // the update reports by itself) is checked against everything
    const bool finite = isfinite(xmin) && isfinite(xmax) && isfinite(ymin) && isfinite(ymax) &&
                        fabs(xmin) < 1e30 && fabs(xmax) < 1e30 && fabs(ymin) < 1e30 && fabs(ymax) < 1e30;
    if (!finite || xmax - xmin > m.cell_dx || ymax - ymin > m.cell_dy) {
        const int k = atomicAdd(m.big_count, 1);
        if (k < m.big_cap) m.big_list[k] = g;
        else atomicOr(&m.status_flags[0], 2);
        m.border_cell[g] = -2;
        return;
    }
    const int c = cell_row(m, 0.5 * (ymin + ymax)) * m.cell_nx + cell_col(m, 0.5 * (xmin + xmax));
    m.border_cell[g] = c;
    atomicAdd(&m.cell_count[c], 1);
}

// exclusive scan of the cell counts: per 1024-cell block, then the block
// totals, then the block offsets added
__global__ void __launch_bounds__(1024) k_cells_scan_blocks(DevModel m) {
    __shared__ int s_w[16];
    const int c = blockIdx.x * 1024 + threadIdx.x;
    const int v = c < m.cell_n ? m.cell_count[c] : 0;
    int tot;
    const int ex = wg_excl_scan(v, s_w, tot);
    if (c < m.cell_n) m.cell_start[c] = ex;
    if (threadIdx.x == 0) m.cell_bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_cells_scan_top(DevModel m, int nblocks) {
    __shared__ int s_w[16];
    wg_scan_array(m.cell_bsum, m.cell_boff, nblocks, s_w);
}

__global__ void k_cells_scatter(DevModel m) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m.n_borders) return;
    const int c = m.border_cell[g];
    if (c < 0) return;
    const int pos = m.cell_boff[c >> 10] + m.cell_start[c] + atomicAdd(&m.cell_fill[c], 1);
    m.cell_items[pos] = g;
}

__device__ inline int cell_first(const DevModel& m, int c) { return m.cell_boff[c >> 10] + m.cell_start[c]; }

constexpr int NB_CAND = 2048;   // candidate neighbours per body before ordering

// neighbour lists (Code/Proximity.h:112-169), one workgroup of NB_THREADS
// per body1 over the candidates of the cell list
__device__ inline int body_slots(const DevModel& m, int b) {
    int g0 = m.border_off[b], g1 = g0 + m.body_nborders[b];
    return m.br_slot0[g1 - 1] + m.br_nslots[g1 - 1] - m.br_slot0[g0];
}

__device__ inline bool self_enabled(const DevModel& m, int b) {
    return m.self_contact && !m.body_rigid_flag[b] && m.body_periodicity[b] != 2;
}

// rows of body b: non-self rows (borderS, neighbour k, nodeS), then self
// rows (borderS, nodeS, borderM)
__device__ inline int body_rows(const DevModel& m, int b) {
    if (m.body_status[b] == 1 || m.body_nborders[b] == 0) return 0;
    const int ns = body_slots(m, b);
    int r = m.nb_count[b] * ns;
    if (self_enabled(m, b)) r += ns * m.body_nborders[b];
    return r;
}

// row chunks per body, and the body's row layout for the counting pass:
// (rows, slots, neighbours, self | borders << 1) and (first border, first slot)
__device__ inline void body_row_layout(const DevModel& m, int b) {
    const int rows = body_rows(m, b);
    m.body_chunks[b] = (rows + ROW_CHUNK - 1) / ROW_CHUNK;
    const int g0 = m.border_off[b];
    const int nbr = m.body_nborders[b];
    m.body_rowrec[b] = make_int4(rows, nbr > 0 ? body_slots(m, b) : 0, m.nb_count[b],
                                 (self_enabled(m, b) ? 1 : 0) | (nbr << 1));
    m.body_gs0[b] = make_int2(g0, nbr > 0 ? m.br_slot0[g0] : 0);
}

__global__ void __launch_bounds__(NB_THREADS) k_prox_neighbours(DevModel m) {
    const int b1 = blockIdx.x;
    const int t = threadIdx.x;
    __shared__ unsigned long long cand[NB_CAND];
    __shared__ int n_cand;
    __shared__ int s_w[NB_THREADS / 64];
    if (t == 0) n_cand = 0;
    __syncthreads();
    const double P = m.period;
    const int nbig = min(*m.big_count, m.big_cap);
    auto test = [&](int r1, int g1, int g2) {
        const int b2 = m.br_body[g2];
        if (b2 == b1 || m.body_status[b2] == 1) return;
        if ((m.pb_ymax[g2] < m.pb_ymin[g1]) || (m.pb_ymax[g1] < m.pb_ymin[g2])) return;
        const int r2 = g2 - m.border_off[b2];
        for (int sh = m.pb_shmin[g1] - m.pb_shmax[g2]; sh <= m.pb_shmax[g1] - m.pb_shmin[g2]; sh++) {
            if ((m.pb_xmax[g2] + sh * P > m.pb_xmin[g1]) && (m.pb_xmin[g2] + sh * P < m.pb_xmax[g1])) {
                const int k = atomicAdd(&n_cand, 1);
                if (k < NB_CAND)
                    cand[k] = ((unsigned long long)r1 << 56) | ((unsigned long long)b2 << 32) |
                              ((unsigned long long)(r2 & 0xFFFF) << 16) | (unsigned long long)((sh + 32768) & 0xFFFF);
            }
        }
    };
    if (m.body_status[b1] != 1) {
        for (int r1 = 0; r1 < m.body_nborders[b1]; r1++) {
            const int g1 = m.border_off[b1] + r1;
            // the cells the AABB covers, grown by one cell; all columns when
            // they wrap onto themselves
            const double xa = m.pb_xmin[g1], xb = m.pb_xmax[g1], ya = m.pb_ymin[g1], yb = m.pb_ymax[g1];
            const bool finite = isfinite(xa) && isfinite(xb) && isfinite(ya) && isfinite(yb) &&
                                fabs(xa) < 1e30 && fabs(xb) < 1e30 && fabs(ya) < 1e30 && fabs(yb) < 1e30;
            const double dc0 = floor((xa - m.xmin_period) / m.cell_dx) - 1.;
            const double dc1 = floor((xb - m.xmin_period) / m.cell_dx) + 1.;
            const bool all_cols = !finite || dc1 - dc0 + 1. >= (double)m.cell_nx;
            const int ncol = all_cols ? m.cell_nx : (int)(dc1 - dc0) + 1;
            const int c0 = all_cols ? 0 : wrap_col(m, dc0);
            const int ry0 = finite ? max(cell_row(m, ya) - 1, 0) : 0;
            const int ry1 = finite ? min(cell_row(m, yb) + 1, m.cell_ny - 1) : m.cell_ny - 1;
            const int nrow = ry1 - ry0 + 1;
            // one thread per cell of the range, then its borders
            for (int q = t; q < ncol * nrow; q += NB_THREADS) {
                const int row = ry0 + q / ncol;
                const int col = (c0 + q % ncol) % m.cell_nx;
                const int c = row * m.cell_nx + col;
                const int e0 = cell_first(m, c), e1 = e0 + m.cell_count[c];
                for (int e = e0; e < e1; e++) test(r1, g1, m.cell_items[e]);
            }
            for (int q = t; q < nbig; q += NB_THREADS) test(r1, g1, m.big_list[q]);
        }
    }
    __syncthreads();
    int n = n_cand;
    if (n > NB_CAND) {
        if (t == 0) atomicOr(&m.status_flags[0], 2);
        n = NB_CAND;
    }
    // bitonic sort of the candidates (padded to a power of two)
    int p = 1;
    while (p < n) p <<= 1;
    for (int i = n + t; i < p; i += NB_THREADS) cand[i] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= p; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < p; i += NB_THREADS) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const unsigned long long a = cand[i], b = cand[l];
                    if ((a > b) == up) { cand[i] = b; cand[l] = a; }
                }
            }
            __syncthreads();
        }
    }
    // the first of equal (body2, border2, shift) in (border1, ...) order
    const unsigned long long LOW = (1ull << 56) - 1;
    int* list = m.nb_list + (size_t)b1 * m.nb_cap * 3;
    int base = 0;
    for (int i0 = 0; i0 < n; i0 += NB_THREADS) {
        const int i = i0 + t;
        int keep = 0;
        if (i < n) {
            keep = 1;
            const unsigned long long x = cand[i] & LOW;
            for (int j = 0; j < i; j++)
                if ((cand[j] & LOW) == x) { keep = 0; break; }
        }
        int tot;
        const int ex = wg_excl_scan(keep, s_w, tot);
        if (keep) {
            const int pos = base + ex;
            if (pos < m.nb_cap) {
                const unsigned long long x = cand[i];
                const int b2 = (int)((x >> 32) & 0xFFFFFF), r2 = (int)((x >> 16) & 0xFFFF);
                const int sh = (int)(x & 0xFFFF) - 32768;
                list[3 * pos] = b2; list[3 * pos + 1] = r2; list[3 * pos + 2] = sh;
                m.nb_gm[(size_t)b1 * m.nb_cap + pos] = m.border_off[b2] + r2;
            }
        }
        base += tot;
    }
    if (t == 0) {
        if (base > m.nb_cap) atomicOr(&m.status_flags[0], 2);
        m.nb_count[b1] = min(base, m.nb_cap);
        body_row_layout(m, b1);     // reads the count just written
    }
}

// One workgroup: the first row chunk of every body (the bodies' row layouts
// come from k_prox_neighbours); clears the hit list and the rigid bodies'
// master-element counters
__global__ void __launch_bounds__(1024) k_prox_chunk_offsets(DevModel m) {
    __shared__ int s_w[16];
    wg_scan_array(m.body_chunks, m.chunk_off, m.n_bodies, s_w);
    if (threadIdx.x == 0) *m.hit_n = 0;
    // the rigid bodies' master-element counters (build_rigid_lists)
    for (int r = threadIdx.x; r < m.n_rigid; r += blockDim.x) m.rm_cnt[r] = 0;
}

// One workgroup per body: the body of each of its row chunks (a wall's
// thousands of chunks spread over the workgroup's lanes)
__global__ void k_prox_chunk_bodies(DevModel m, int cap) {
    const int b = blockIdx.x;
    const int c1 = min(m.chunk_off[b + 1], cap);
    for (int c = m.chunk_off[b] + (int)threadIdx.x; c < c1; c += blockDim.x) m.chunk_body[c] = b;
}

struct RowInfo {
    int body, self, borderS, k, jS, borderM;
    int gS, slotS;     // slave border (global) and slot
};

// row `local` of body b
__device__ inline bool decode_local_row(const DevModel& m, int b, int local, RowInfo& ri) {
    const int nb = m.nb_count[b];
    const int g0 = m.border_off[b];
    const int s0 = m.br_slot0[g0];
    const int ns = body_slots(m, b);
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
                ri.gS = g0 + r;
                ri.slotS = s0 + pre + ri.jS;
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
            ri.gS = g0 + r;
            ri.slotS = s0 + q;
            return true;
        }
    }
    return false;
}

// row `local` of body b from its row records; bodies with one border
// decode without further loads
__device__ inline bool decode_row_rec(const DevModel& m, int b, int4 rr, int2 gs, int local, RowInfo& ri) {
    if ((rr.w >> 1) != 1) return decode_local_row(m, b, local, ri);
    const int ns = rr.y, nb = rr.z;
    ri.body = b;
    ri.borderS = 0;
    ri.gS = gs.x;
    if (local < nb * ns) {
        ri.self = 0;
        ri.k = local / ns;
        ri.jS = local - ri.k * ns;
        ri.borderM = -1;
    } else {
        ri.self = 1;
        ri.k = -1;
        ri.jS = local - nb * ns;
        ri.borderM = 0;
    }
    ri.slotS = gs.y + ri.jS;
    return true;
}

// Scans one row from master segment seg_start on.  WRITE == false counts the
// candidates (and reports the first one's segment in *first), WRITE == true
// emits them as elements base, base + 1, ...
constexpr int SEG_BATCH = 4;     // segment records loaded together
constexpr int BOX_BATCH = 4;     // chunk boxes loaded together
template <bool WRITE>
__device__ int scan_row(const DevModel& m, const RowInfo& ri, int base, int seg_start, int* first, int limit) {
    int bS = ri.body;
    int slotS = ri.slotS;
    int nodeS = m.sl_node[slotS];
    const double2 xs = m.slot_xy[slotS];          // position of the slot's node (this update)
    double Xs = xs.x, Ys = xs.y;
    double detect = m.body_detect[bS], cdist = m.body_contact[bS];
    double P = m.period;
    int bM, rM, shM, gM;
    if (!ri.self) {
        const int* nb = m.nb_list + ((size_t)bS * m.nb_cap + ri.k) * 3;
        bM = nb[0]; rM = nb[1]; shM = nb[2];
        gM = m.nb_gm[(size_t)bS * m.nb_cap + ri.k];
    } else {
        bM = bS; rM = ri.borderM; shM = 0;
        gM = ri.gS - ri.borderS + rM;
    }
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
    if (!WRITE) {
        // outside the whole border's segment box: no chunk box can pass
        const double4 sb = m.pb_segbox[gM];
        if (!(Xq > sb.x - dd && Xq < sb.y + dd && Ys > sb.z - dd && Ys < sb.w + dd)) return 0;
    }
    // groups of SC_GROUP chunks first, then the chunks of the groups in reach
    const int cS = cM0 + seg_start / SEG_CHUNK;
    for (int gq = cS / SC_GROUP; gq * SC_GROUP < cM1 && cnt < limit; gq += BOX_BATCH) {
    unsigned gpass = 0;       // groups of this batch whose box is within reach
#pragma unroll
    for (int u = 0; u < BOX_BATCH; u++) {
        const double4 bx = m.sc_box2[min(gq + u, (cM1 - 1) / SC_GROUP)];
        if ((gq + u) * SC_GROUP < cM1 && Xq > bx.x - dd && Xq < bx.y + dd && Ys > bx.z - dd && Ys < bx.w + dd)
            gpass |= 1u << u;
    }
    for (; gpass && cnt < limit; gpass &= gpass - 1) {
    const int grp = gq + __builtin_ctz(gpass);
    const int c1 = min(grp * SC_GROUP + SC_GROUP, cM1);
    for (int cb = max(grp * SC_GROUP, cS); cb < c1 && cnt < limit; cb += BOX_BATCH) {
    unsigned pass = 0;        // chunks of this batch whose box is within reach
#pragma unroll
    for (int u = 0; u < BOX_BATCH; u++) {
        const double4 bx = m.sc_box[min(cb + u, c1 - 1)];
        if (cb + u < c1 && Xq > bx.x - dd && Xq < bx.y + dd && Ys > bx.z - dd && Ys < bx.w + dd) pass |= 1u << u;
    }
    for (; pass && cnt < limit; pass &= pass - 1) {
    const int ch = cb + __builtin_ctz(pass);
    const int sg0 = max((ch - cM0) * SEG_CHUNK, seg_start), sg1 = min((ch - cM0) * SEG_CHUNK + SEG_CHUNK, nM - 1);
    for (int g = sg0; g < sg1 && cnt < limit; g += SEG_BATCH) {
    int4 sn[SEG_BATCH];
    double4 sx[SEG_BATCH];
#pragma unroll
    for (int u = 0; u < SEG_BATCH; u++) {
        const int s = s0 + min(g + u, sg1 - 1);
        sn[u] = m.seg_ns[s];
        sx[u] = m.seg_xy[s];
    }
#pragma unroll
    for (int u = 0; u < SEG_BATCH; u++) {
        const int seg = g + u;
        if (seg >= sg1) break;
        const int s = s0 + seg;
        int n1 = sn[u].x, n2 = sn[u].y;
        if (ri.self && (n1 == nodeS || n2 == nodeS)) continue;
        double X1 = sx[u].x + P * (shM + sn[u].z), Y1 = sx[u].y;
        double X2 = sx[u].z + P * (shM + sn[u].w), Y2 = sx[u].w;
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
        if (!WRITE && cnt == 0) *first = seg;
        if (WRITE) {
            int e = base + cnt;
            if (e < m.c_cap) {
                const ContactArrays& C = m.C;
                C.bodyS[e] = bS; C.borderS[e] = ri.borderS; C.bnodeS[e] = ri.jS; C.nodeS[e] = nodeS;
                C.bodyM[e] = bM; C.borderM[e] = rM; C.shiftM[e] = shM; C.segM[e] = seg;
                C.nM0[e] = n0; C.nM1[e] = n1; C.nM2[e] = n2; C.nM3[e] = n3;
                if (ri.self) { C.sM0[e] = 0; C.sM1[e] = 0; C.sM2[e] = 0; C.sM3[e] = 0; }
                else { C.sM0[e] = m.sl_s0[s]; C.sM1[e] = sn[u].z; C.sM2[e] = sn[u].w; C.sM3[e] = m.sl_s3[s]; }
                C.N0[e] = p.N0; C.N1[e] = p.N1; C.N2[e] = p.N2; C.N3[e] = p.N3;
                C.gapn[e] = p.gap; C.gapt[e] = 0.; C.xsi[e] = p.xsi;
                C.xn[e] = p.xn; C.yn[e] = p.yn; C.xt[e] = p.xt; C.yt[e] = p.yt;
                C.emass[e] = emass; C.len[e] = m.sl_len[slotS]; C.fx[e] = 0.; C.fy[e] = 0.;
                C.nbint[e] = 0; C.int0[e] = 0.; C.int1[e] = 0.; C.int2[e] = 0.;
                C.vgapn[e] = 0.; C.vgapt[e] = 0.; C.wacc[e] = 0.;
                C.fresh[e] = ri.self ? 2 : 1;
                int4 inf, sinf;
                element_info_from(m, bS, slotS, bM, s0, inf, sinf);
                // the stencil written above is the slot's (load_stencil)
                // except the shifts of a self element
                if (!ri.self) inf.z |= CI_STENCIL;
                C.info[e] = inf;
                C.sinfo[e] = sinf;
            }
        }
        cnt++;
    }
    }
    }
    }
    }
    }
    return cnt;
}

// index of the previous elements by slave slot, for history transfer
__device__ inline void old_index(const DevModel& m, int wg, int n_wg) {
    const int n = *m.c_count_old;
    for (int e = wg * blockDim.x + threadIdx.x; e < n; e += n_wg * blockDim.x) {
    const ContactArrays& O = m.Cold;
    int slot = m.br_slot0[m.border_off[O.bodyS[e]] + O.borderS[e]] + O.bnodeS[e];
    int pos = atomicAdd(&m.slot_old_cursor[slot], 1);
    if (pos < OLD_CAP) m.slot_old_list[slot * OLD_CAP + pos] = e;
    else atomicOr(&m.status_flags[0], 4);
    }
}

// Counting pass over row chunks of ROW_CHUNK rows (one row per thread, the
// grid loops over the chunks): candidates per chunk, and one hit record per
// row with candidates: (chunk, elements of the chunk's earlier rows, first
// candidate's segment, row of the body)
// Workgroups from n_count on build the index of the previous elements
// (old_index), which nothing before the history transfer reads.
__global__ void __launch_bounds__(ROW_CHUNK) k_prox_count(DevModel m, int n_count) {
    if ((int)blockIdx.x >= n_count) {
        old_index(m, blockIdx.x - n_count, gridDim.x - n_count);
        return;
    }
    __shared__ int s_w[ROW_CHUNK / 64];
    __shared__ int s_hbase;
    const int n_chunks = m.chunk_off[m.n_bodies];
    for (int c = blockIdx.x; c < n_chunks; c += n_count) {
        const int b = m.chunk_body[c];
        const int4 rr = m.body_rowrec[b];
        const int2 gs = m.body_gs0[b];
        const int local = (c - m.chunk_off[b]) * ROW_CHUNK + threadIdx.x;
        int cnt = 0, first = 0;
        RowInfo ri;
        if (local < rr.x && decode_row_rec(m, b, rr, gs, local, ri)) cnt = scan_row<false>(m, ri, 0, 0, &first, INT_MAX);
        int tot;
        const int ex = wg_excl_scan(cnt, s_w, tot);
        if (threadIdx.x == 0) m.chunk_cnt[c] = tot;
        // the workgroup's rows with candidates take consecutive hit records:
        // one atomic per chunk (the records' order is free, k_prox_write
        // places each row's elements by its chunk and offset)
        int nhit;
        const int hex = wg_excl_scan(cnt > 0 ? 1 : 0, s_w, nhit);
        if (threadIdx.x == 0 && nhit > 0) s_hbase = atomicAdd(m.hit_n, nhit);
        __syncthreads();
        if (cnt > 0) {
            const int h = s_hbase + hex;
            if (h < m.c_cap) {
                m.hit_rec[h] = make_int4(c, ex, first, local);
                m.hit_cnt[h] = cnt;
            } else {
                atomicOr(&m.status_flags[0], 8);
            }
        }
        __syncthreads();       // s_hbase is read before the next chunk's atomic replaces it
    }
}

// first element of every row chunk (one workgroup) and the element count
__global__ void __launch_bounds__(1024) k_prox_elem_offsets(DevModel m) {
    __shared__ int s_w[16];
    const int n_chunks = m.chunk_off[m.n_bodies];
    wg_scan_array(m.chunk_cnt, m.chunk_elem_off, n_chunks, s_w);
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = m.chunk_elem_off[n_chunks];
        if (total > m.c_cap) { atomicOr(&m.status_flags[0], 8); total = m.c_cap; }
        *m.c_count = total;
    }
}

// Writing pass: the rows with candidates, each from its first candidate on
// (up to 256 registers per lane: the emitting scan does not spill)
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(1, 2))) k_prox_write(DevModel m) {
    const int n = min(*m.hit_n, m.c_cap);
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < n; h += gridDim.x * blockDim.x) {
        const int4 r = m.hit_rec[h];
        const int b = m.chunk_body[r.x];
        RowInfo ri;
        if (!decode_row_rec(m, b, m.body_rowrec[b], m.body_gs0[b], r.w, ri)) continue;
        scan_row<true>(m, ri, m.chunk_elem_off[r.x] + r.y, r.z, nullptr, m.hit_cnt[h]);
    }
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
    D.int1[e] = S.int1[o]; D.int2[e] = S.int2[o];
    D.info[e] = S.info[o]; D.sinfo[e] = S.sinfo[o];
}

// history transfer: first previous element of the same slave slot with the
// same master body and border and |segment difference| <= 1
__global__ void k_prox_history(DevModel m) {
    const int ne = *m.c_count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
    const ContactArrays& C = m.C;
    if (C.fresh[e] != 1) continue;        // self elements never inherit (Code/Proximity.h:410-425)
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
}

// fresh elements carry the relative velocities of the element created just
// before them in the same pass (the reused new_contact record of
// Code/Proximity.h:211 and :350), zero at the start of a pass
__global__ void k_prox_chain(DevModel m) {
    const int n = *m.c_count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const ContactArrays& C = m.C;
    // an element whose master is rigid goes to that body's list (sorted
    // into element order by build_rigid_lists)
    const int rr = m.body_rindex[C.bodyM[e]];
    if (rr >= 0) {
        const int pos = atomicAdd(&m.rm_cnt[rr], 1);
        if (pos < RIGID_LIST_CAP) m.rm_list[(size_t)rr * RIGID_LIST_CAP + pos] = e;
        else atomicOr(&m.status_flags[0], 32);
    }
    int f = C.fresh[e];
    if (f == 0) continue;
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
}


// Initialize_CZM (Code/Proximity.h:481-541) after the elements have been
// re-projected: damage 0 below gapinit, 1 otherwise, for CZMlinear pairs.
__global__ void k_init_czm(DevModel m) {
    const int n = *m.c_count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const ContactArrays& C = m.C;
        int ms = m.body_material[C.bodyS[e]], mm = m.body_material[C.bodyM[e]];
        int li = m.law_first[ms * m.n_materials + mm];
        if (li < 0) continue;
        // gapinit: parameter 5 of CZMlinear, 8 of CZMfatigue, 10 of BondedMohrCoulomb
        const int lt = m.law_type[li];
        const int gp = lt == 6 ? 5 : (lt == 7 ? 8 : (lt == 5 ? 10 : -1));
        if (gp < 0) continue;
        C.nbint[e] = lt == 6 ? 1 : 3;
        C.int0[e] = C.gapn[e] < m.law_p[12 * li + gp] ? 0. : 1.;
        C.int1[e] = 0.;
        C.int2[e] = 0.;
    }
}

__global__ void k_project_all(DevModel m) {
    const int n = *m.c_count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        project_element(m, m.C, e);
}

}  // namespace mel
