// Device-side model state and the arithmetic shared by the kernels.
//
// Layout (all arrays in HBM, structure of arrays, one allocation each):
//   nodes      global index g = node_off[body] + local id; FP64 fields
//   Gauss pts  global index in body order; CSR gp -> support (global node ids)
//   gtrans     transposed CSR node -> (gp, dphi/dx, dphi/dy) for the
//              deterministic node-centric internal-force gather
//   d2d        DOF_TO_DISP CSR node -> (support node, phi)
//   borders    global border id = border_off[body] + local border
//   slots      one per border-node position (closed borders repeat their first
//              node), with the four-node segment stencil of Initialize_segments
//   bnodes     distinct border nodes; point forces (contact, Dirichlet,
//              Neumann) are summed per border node and spread through the
//              transposed DOF_TO_DISP rows restricted to border nodes (btrans)
//   contacts   structure of arrays with a fixed capacity, double-buffered
//              across proximity updates
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mel {

constexpr int RIGID_FIELDS = 64;
// per-body state slots in body_rigid[b * RIGID_FIELDS + k]; the work slots are
// used by every body, the rest by rigid bodies
enum RigidSlot {
    R_CUR = 0, R_DISP = 3, R_VEL = 6, R_ACC = 9, R_FCON = 12, R_FBODY = 15, R_FDIR = 18, R_FNEU = 21,
    R_FDAMP = 24, R_FTOT = 27, R_DRIVEN = 30, R_XC = 38, R_YC = 39, R_WORK = 40, R_NODE_ACC = 48,
    R_FDIR_NODE = 50
};
enum WorkSlot { W_INTERNAL = 0, W_CONTACT, W_BODY, W_DIRICHLET, W_NEUMANN, W_DAMPING, W_ALID };

struct ContactArrays {
    int *bodyS, *borderS, *bnodeS, *nodeS, *bodyM, *borderM, *shiftM, *segM, *nbint, *fresh;
    int *nM0, *nM1, *nM2, *nM3, *sM0, *sM1, *sM2, *sM3;
    double *N0, *N1, *N2, *N3, *gapn, *vgapn, *gapt, *vgapt, *xsi, *xn, *yn, *xt, *yt, *emass, *len, *fx, *fy,
        *int0, *int1, *int2;     // internal variables (damage; the three-variable laws' tractions)
    // half of the element's contact work accumulated since the last flush
    // into its bodies' work (one thread per element adds, so the sums do not
    // depend on scheduling)
    double* wacc;
    // what the element's step reads first (element_info), so that its
    // dependent loads are one level deep:
    //   info   slave slot, slave border node (-1: rigid slave), law index + 1
    //          (0: none) | CI_* flags, first slot of the master border
    //   sinfo  the slave slot's length stencil: nodes 0, 1, 2; shifts + 1
    //          (2 bits each) | end class << 6 (1: first slot of a Simple
    //          border, 2: last one)
    int4 *info, *sinfo;
};

constexpr int CI_LAW_MASK = (1 << 24) - 1;
constexpr int CI_VALID = 1 << 24;       // info and sinfo hold the element's values
constexpr int CI_STENCIL = 1 << 25;     // nM0-3 / sM0-3 are the stencil of (borderM, segM)
constexpr int CI_RIGID_S = 1 << 26;
constexpr int CI_RIGID_M = 1 << 27;
constexpr int CI_INACTIVE_S = 1 << 28;  // slave body of status 1

struct DevModel {
    // sizes
    int n_bodies, n_nodes, n_gauss, n_borders, n_slots, n_bnodes, n_rigid;
    int n_materials, n_laws;
    double period, xmin_period, gravity_x, gravity_y, dt;
    int self_contact;          // 1 unless NO_MONITORING (Code/Proximity.h:332)
    int law_f32;               // contact-law tractions in FP32 (the FP32-mixed form)

    // bodies
    int *body_rigid_flag, *body_material, *body_periodicity, *body_status;
    int *node_off, *body_nnodes, *gp_off, *body_ngauss, *body_nregions, *body_per_region;
    int *border_off, *body_nborders;
    double *body_detect, *body_contact, *body_mass, *body_inv_mass, *body_rigid;   // body_rigid: RIGID_FIELDS each
    double *body_inertia, *body_inv_inertia;
    int *rigid_ids;            // body ids of the rigid bodies
    int n_rigid_nodes;
    int *rigid_node_ids;       // global ids of the nodes of rigid bodies

    // materials / laws
    int *mat_type;             // per material
    double *mat_p;             // 6 per material (rho alpha beta mu lambda|kappa c)
    int *law_table;            // n_materials^2
    int *law_type, *law_fixed;
    double *law_p;             // 12 per law

    // nodes
    int *node_body;
    double *X, *Y, *xc, *yc, *ux, *uy, *vx, *vy, *ax, *ay;
    double *upx, *upy, *vpx, *vpy, *apx, *apy;
    double *mx, *my, *imx, *imy;
    double *fix, *fiy, *fcx, *fcy, *fbx, *fby, *fdx, *fdy, *fnx, *fny, *fkx, *fky, *fx, *fy;
    const int* node_body_host; // host copy used while building (null on the device)

    // Gauss points
    int *g_ptr, *g_idx;
    double *g_phix, *g_phiy, *g_c0, *g_P;   // g_P: 4 per GP (c0 * first Piola terms)
    int *g_body;
    int *t_ptr, *t_gp;
    double *t_phix, *t_phiy;

    // DOF_TO_DISP
    int *d_ptr, *d_idx;
    double *d_phi;

    // damping matrices (global COO, upper triangle, dofs global: 2*g + dir)
    int n_matrix;
    int *m_dof;                // 2 per entry
    double *m_damp;
    int *mt_ptr, *mt_entry;    // node -> contributing (entry*2 + side) list, for a deterministic gather

    // borders
    int *br_body, *br_slot0, *br_nslots, *br_before, *br_after, *br_periodicity, *br_xtype, *br_ytype;
    int n_bc_timed, *bc_timed;               // borders whose BC value follows the time (type 2, 3 or 4)
    double *br_xparam, *br_yparam;           // stiffness, damping
    int *br_hist_ptr;                        // 2 per border + 1 (x then y)
    double *hist_t, *hist_v;
    double *br_xval, *br_yval, *br_xvel, *br_yvel;   // current BC value / velocity

    // slots
    int *sl_border, *sl_j, *sl_node, *sl_n0, *sl_n1, *sl_n2, *sl_n3, *sl_s0, *sl_s1, *sl_s2, *sl_s3, *sl_bn;
    double *sl_len, *sl_len0, *sl_xcp, *sl_ycp;
    double *sl_xbp, *sl_ybp;                 // BC pressure of the last step (Border::x_bc_pressure)

    // distinct border nodes and point forces
    int *bn_node;
    int *bn_slot_ptr, *bn_slots;             // bn -> slots (ascending)
    double *bn_pcx, *bn_pcy, *bn_pdx, *bn_pdy, *bn_pnx, *bn_pny;
    int *bt_ptr, *bt_bn;                     // node -> (border node, phi) entries
    double *bt_phi;

    // contact point-force bucketing (per border node)
    int *bk_count, *bk_off, *bk_cursor, *bk_rec;
    double2 *bk_val;                         // force of each bucket record
    int *rec_bn;                             // 3 per element: slave, master 1, master 2 (-1 = none)
    double *rec_fx, *rec_fy;
    double *rig_rec;                         // 6 per element: slave rigid (fx,fy,m), master rigid (fx,fy,m)

    // contacts
    int c_cap;
    int *c_count;                            // device-side count (current buffer)
    ContactArrays C;                         // current
    ContactArrays Cold;                      // previous proximity result (history source)

    // proximity scratch
    double *pb_xmin, *pb_xmax, *pb_ymin, *pb_ymax;   // per border AABB
    int *pb_shmin, *pb_shmax;
    int n_seg_chunks;                        // chunks of SEG_CHUNK border slots
    int *br_chunk0, *sc_border;              // first chunk per border (n_borders + 1), border per chunk
    double *sc_nxmin, *sc_nxmax, *sc_nymin, *sc_nymax;   // slot-node box inflated by detection
    double4 *sc_box;                         // segment box (slot shifts applied): xmin, xmax, ymin, ymax
    double4 *sc_box2;                        // per group of SC_GROUP consecutive chunks: the union of
                                             // their segment boxes
    int4 *seg_ns;                            // per slot s, segment s: end nodes n1, n2 and their shifts
    double4 *seg_xy;                         // per slot s, segment s: x, y of n1, then of n2
    // cell list of the broad phase (k_proximity.hip)
    int cell_nx, cell_ny, cell_n;            // grid columns (wrapping over the period), rows, cells
    double cell_dx, cell_dy, cell_y0;        // cell size; first row's lower edge
    int *cell_count, *cell_fill, *cell_start, *cell_bsum, *cell_boff, *cell_items;
    int *border_cell;                        // per border: its cell, -2 big, -1 inactive
    int big_cap, *big_count, *big_list;      // borders larger than a cell
    int nb_cap;                              // neighbour capacity per body
    int *nb_count, *nb_list;                 // per body: count, 3 ints per neighbour
    int *body_chunks, *chunk_off;            // narrow-phase row chunks per body, first chunk per body
    int *chunk_body;                         // body of each row chunk
    int4 *body_rowrec;                       // per body: rows, slots, neighbours, self | borders << 1
    int2 *body_gs0;                          // per body: first border, first slot
    double2 *slot_xy;                        // per slot: its node's position at the last proximity update
    int *nb_gm;                              // per neighbour entry: the neighbour border's global index
    int *chunk_cnt, *chunk_elem_off;         // candidates per row chunk, first element per row chunk
    int4 *hit_rec;                           // rows with candidates: chunk, offset, first segment, row
    int *hit_cnt;                            // their candidate counts
    int *hit_n;
    double4 *pb_segbox;                      // per border: box of its segments (union of its chunks' boxes)
    int *slot_old_list, *slot_old_cursor;
    int *c_count_old;                        // element count of Cold
    int *law_first;                          // first law matching a material pair by name (Initialize_CZM)
    int *status_flags;                       // [0] capacity overflow bits, [1] non-finite state seen
    double *time;                            // device copy of the simulation time
    double *rigid_dv;                        // per rigid body: driven DOFs at the end of the current
                                             // step (8 each, as R_DRIVEN), formed by the force launch
    int n_rbc;                               // borders of rigid bodies with force BCs (types 2, 3, 4)
    int *rbc_border;                         // their global border ids
    int *br_rbc;                             // per global border: its index in rbc_border, or -1
    double *rbc_part;                        // per such border and run: the run's BC terms (6 each),
                                             // formed by the force launch
    unsigned *step_sync;                     // step launch B: [1] rigid-body workgroups done
    // contact elements of each rigid body (k_rigid_lists.hip)
    int *body_rindex;                        // per body: its rigid index, or -1
    int *rm_cnt;                             // per rigid body: elements it is the master of
    int *rm_list;                            // RIGID_LIST_CAP per rigid body: those elements, ascending
    int *rs_range;                           // 2 per rigid body: its slave elements [begin, end)
    // absorbing layers (Body::Update_alid_forces, Code/Body.h:1163-1296)
    int alid_on;                             // some deformable body has them
    double *fax, *fay;                       // Node::x_alid_force, y_alid_force (zero elsewhere)
    double *alid_a, *alid_b;                 // per node: alid_alpha, alid_beta (ALID models)
    int *body_alid;                          // per body: x history [h0, h1), y history [h2, h3) in hist_t/hist_v
    int *alid_nodes, n_alid_nodes;           // the nodes of the bodies with absorbing layers
    double *m_stiff;                         // per matrix entry: its stiffness (ALID models)
    // deactivation (Update_status, Code/Proximity.h:550-1007)
    int n_rules;                             // DEACTIVATION rules, RULE_STRIDE entries each as read
    const double* rules;
    double* body_damage;                     // Body::damage (Code/Body.h:2667-2700)
    int* body_flip;                          // per body: deactivated by the last evaluation
    double* slot_damage;                     // per slot: Update_damage's per-border-node value
    // An inactive body's last contact sends: Update_contact_forces and
    // Initialize_contact_forces skip it while Send_contact_forces does not
    // (Code/Solver.h:36-55), so its last master point forces are added again
    // every step.  Border-node records take bucket ids past the elements'.
    int ghost_cap;
    int* gh_count;                           // [0] border-node records, [1] rigid-master records
    int* gh_bn;                              // border node
    double *gh_fx, *gh_fy;
    int* gh_rb;                              // rigid master body
    double* gh_rf;                           // its (fx, fy, moment), 3 each
    // contact point forces an inactive body's border nodes received since it
    // stopped (its node contact forces are no longer reset, Code/Body.h:1305)
    double *bn_acx, *bn_acy;
};

constexpr int RULE_STRIDE = 8;               // entries per DEACTIVATION rule on the device

// --------------------------------------------------------------------------
// geometry: Closest_2_segment (Linear branch, Code/Tools.h:106-219) and
// Closest_2_segment_self (Code/Tools.h:319-369).
//
// pow(x, -0.5) (Code/Tools.h:108-110) decides contact classification at
// geometric ties, so x^(-1/2) is computed correctly rounded: y = 1/sqrt(x)
// (two roundings, error <= ~1 ulp), then one Newton correction with the exact
// residual r = 1 - x*y*y evaluated through FMA error-free products,
// y' = y + y*r/2 rounded once.  glibc's pow agrees with the correctly rounded
// value on all but ~0.1% of inputs.
// --------------------------------------------------------------------------
__device__ __forceinline__ double inv_sqrt(double x) {
    double y = 1.0 / __dsqrt_rn(x);
    if (!(x > 0.0) || !isfinite(y) || y == 0.0) return y;
    double p = y * y;
    double plo = __fma_rn(y, y, -p);
    double q = x * p;
    double qlo = __fma_rn(x, p, -q);
    double r = ((1.0 - q) - qlo) - x * plo;
    return __fma_rn(y, r * 0.5, y);
}

struct Proj {
    double gap, xsi, xn, yn, xt, yt, N0, N1, N2, N3;
    int flag;
};

__device__ __forceinline__ void closest_linear(double Xs, double Ys, double X0, double Y0, double X1, double Y1,
                                               double X2, double Y2, double X3, double Y3, Proj& p) {
    double il = inv_sqrt((X2 - X1) * (X2 - X1) + (Y2 - Y1) * (Y2 - Y1));
    double ilm = inv_sqrt((X1 - X0) * (X1 - X0) + (Y1 - Y0) * (Y1 - Y0));
    double ilp = inv_sqrt((X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2));
    double tx = (X2 - X1) * il, ty = (Y2 - Y1) * il;
    p.xt = tx; p.yt = ty; p.xn = ty; p.yn = -tx;
    double tmx = (X1 - X0) * ilm, tmy = (Y1 - Y0) * ilm;
    double tpx = (X3 - X2) * ilp, tpy = (Y3 - Y2) * ilp;
    if (!(((Xs - X1) * (tmx + tx) + (Ys - Y1) * (tmy + ty)) * (tx * (tmx + tx) + ty * (tmy + ty)) >= 0.)) {
        p.flag = -3; p.xsi = -1.; return;
    }
    if (!(((Xs - X2) * (tx + tpx) + (Ys - Y2) * (ty + tpy)) * (tx * (tx + tpx) + ty * (ty + tpy)) <= 0.)) {
        p.flag = 3; p.xsi = 1.; return;
    }
    double xsi = -1. + 2. * ((X2 - X1) * (Xs - X1) + (Y2 - Y1) * (Ys - Y1)) * il * il;
    p.xsi = xsi;
    if ((xsi < 1.) && (xsi > -1.)) {
        double xcl = 0.5 * (X1 + X2) + xsi * 0.5 * (X2 - X1);
        double ycl = 0.5 * (Y1 + Y2) + xsi * 0.5 * (Y2 - Y1);
        p.gap = p.xn * (Xs - xcl) + p.yn * (Ys - ycl);
        p.N0 = 0.; p.N1 = 0.5 * (1. - xsi); p.N2 = 0.5 * (1. + xsi); p.N3 = 0.;
        p.flag = 0;
    } else if (xsi == -1.) {
        p.gap = p.xn * (Xs - X1) + p.yn * (Ys - Y1);
        p.N0 = 0.; p.N1 = 1.; p.N2 = 0.; p.N3 = 0.;
        p.flag = -1;
    } else if (xsi < -1.) {
        double idm = inv_sqrt((Xs - X1) * (Xs - X1) + (Ys - Y1) * (Ys - Y1));
        double nx = (Xs - X1) * idm, ny = (Ys - Y1) * idm;
        if (-ny * (tx + tmx) + nx * (ty + tmy) < 0.) { nx = -nx; ny = -ny; }
        p.xn = nx; p.yn = ny; p.xt = -ny; p.yt = nx;
        p.gap = nx * (Xs - X1) + ny * (Ys - Y1);
        p.N0 = 0.; p.N1 = 1.; p.N2 = 0.; p.N3 = 0.;
        p.xsi = -1.;
        p.flag = -2;
    } else if (xsi == 1.) {
        p.gap = p.xn * (Xs - X2) + p.yn * (Ys - Y2);
        p.N0 = 0.; p.N1 = 0.; p.N2 = 1.; p.N3 = 0.;
        p.flag = 1;
    } else {
        double idp = inv_sqrt((Xs - X2) * (Xs - X2) + (Ys - Y2) * (Ys - Y2));
        double nx = (Xs - X2) * idp, ny = (Ys - Y2) * idp;
        if (-ny * (tx + tpx) + nx * (ty + tpy) < 0.) { nx = -nx; ny = -ny; }
        p.xn = nx; p.yn = ny; p.xt = -ny; p.yt = nx;
        p.gap = nx * (Xs - X2) + ny * (Ys - Y2);
        p.N0 = 0.; p.N1 = 0.; p.N2 = 1.; p.N3 = 0.;
        p.xsi = 1.;
        p.flag = 2;
    }
}

__device__ __forceinline__ void closest_self(double Xs, double Ys, double X1, double Y1, double X2, double Y2,
                                             Proj& p) {
    double il = inv_sqrt((X2 - X1) * (X2 - X1) + (Y2 - Y1) * (Y2 - Y1));
    double tx = (X2 - X1) * il, ty = (Y2 - Y1) * il;
    p.xt = tx; p.yt = ty;
    if (((Xs - X1) * tx + (Ys - Y1) * ty) * (tx * tx + ty * ty) >= 0.) {
        if (((Xs - X2) * tx + (Ys - Y2) * ty) * (tx * tx + ty * ty) <= 0.) {
            double xsi = -1. + 2. * ((X2 - X1) * (Xs - X1) + (Y2 - Y1) * (Ys - Y1)) * il * il;
            p.xsi = xsi;
            if ((xsi < 1.) && (xsi > -1.)) {
                p.xn = ty; p.yn = -tx;
                double xcl = 0.5 * (X1 + X2) + xsi * 0.5 * (X2 - X1);
                double ycl = 0.5 * (Y1 + Y2) + xsi * 0.5 * (Y2 - Y1);
                p.gap = p.xn * (Xs - xcl) + p.yn * (Ys - ycl);
                p.N0 = 0.; p.N1 = 0.5 * (1. - xsi); p.N2 = 0.5 * (1. + xsi); p.N3 = 0.;
            }
        } else {
            p.xsi = 1.;
        }
    } else {
        p.xsi = -1.;
    }
    p.flag = -3;
}

// Detect_Segment_Proximity, Code/Proximity.h:8-27
__device__ __forceinline__ bool detect_segment(double Xs, double Ys, double X1, double Y1, double X2, double Y2,
                                               double d) {
    d = d * d;
    if ((Xs - X1) * (Xs - X1) + (Ys - Y1) * (Ys - Y1) < d) return true;
    if ((Xs - X2) * (Xs - X2) + (Ys - Y2) * (Ys - Y2) < d) return true;
    double dx = X2 - X1, dy = Y2 - Y1;
    double a = dx * dx + dy * dy;
    double xsi = -1. + 2. * (dx * (Xs - X1) + dy * (Ys - Y1)) / a;
    if ((xsi < 1.) && (xsi > -1.)) {
        double b = dy * (Xs - 0.5 * (X1 + X2) - xsi * 0.5 * dx) - dx * (Ys - 0.5 * (Y1 + Y2) - xsi * 0.5 * dy);
        if (b * b < a * d) return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// Body::Update_contacts (Code/Body.h:1341-1634) for one element: re-project
// the slave node, walk to the previous / next master segment on flags +-2/+-3,
// accumulate the tangential gap, relative velocities.  When both projections
// return +-3 the element keeps its own stored shape values.
// ---------------------------------------------------------------------------
__device__ inline int slot_of(const DevModel& m, int body, int border, int j) {
    return m.br_slot0[m.border_off[body] + border] + j;
}

__device__ inline void load_stencil(const DevModel& m, int body, int border, int seg, int* nd, int* sh) {
    int s = slot_of(m, body, border, seg);
    nd[0] = m.sl_n0[s]; nd[1] = m.sl_n1[s]; nd[2] = m.sl_n2[s]; nd[3] = m.sl_n3[s];
    sh[0] = m.sl_s0[s]; sh[1] = m.sl_s1[s]; sh[2] = m.sl_s2[s]; sh[3] = m.sl_s3[s];
}

// info / sinfo of an element from its slave body and slot, master body and
// the first slot of its master border
__device__ inline void element_info_from(const DevModel& m, int bS, int slot, int bM, int slot0M, int4& inf,
                                         int4& sinf) {
    const int li = m.law_table[m.body_material[bS] * m.n_materials + m.body_material[bM]];
    const int rS = m.body_rigid_flag[bS];
    int fl = CI_VALID | (li + 1);
    if (rS) fl |= CI_RIGID_S;
    if (m.body_rigid_flag[bM]) fl |= CI_RIGID_M;
    if (m.body_status[bS] == 1) fl |= CI_INACTIVE_S;
    inf = make_int4(slot, rS ? -1 : m.sl_bn[slot], fl, slot0M);
    const int br = m.sl_border[slot], j = m.sl_j[slot], n = m.br_nslots[br];
    const bool simple = m.br_periodicity[br] == 1;
    const int end = (j == 0 && simple) ? 1 : ((j == n - 1 && simple) ? 2 : 0);
    sinf = make_int4(m.sl_n0[slot], m.sl_n1[slot], m.sl_n2[slot],
                     (m.sl_s0[slot] + 1) | (m.sl_s1[slot] + 1) << 2 | (m.sl_s2[slot] + 1) << 4 | end << 6);
}

// the same from the element's indices (the loads they save)
__device__ inline void element_info(const DevModel& m, const ContactArrays& C, int e, int4& inf, int4& sinf) {
    const int bS = C.bodyS[e], bM = C.bodyM[e];
    element_info_from(m, bS, slot_of(m, bS, C.borderS[e], C.bnodeS[e]), bM, slot_of(m, bM, C.borderM[e], 0), inf,
                      sinf);
}

// an element's info, computed and stored when it is not there yet
__device__ __forceinline__ void element_info_ld(const DevModel& m, const ContactArrays& C, int e, int4& inf, int4& sinf) {
    inf = C.info[e];
    sinf = C.sinfo[e];
    if (!(inf.z & CI_VALID)) {
        element_info(m, C, e, inf, sinf);
        C.info[e] = inf;
        C.sinfo[e] = sinf;
    }
}

// the re-projected element (project_core), held in registers for its law
struct ElemProj {
    int borderM, shiftM, segM, nodeS;
    int nd0, nd1, nd2, nd3, sh0, sh1, sh2, sh3;   // stencil of the final segment
    double xs, ys;            // slave node position
    double gapt, vgapn, vgapt;
    Proj p;
};

// Body::Update_contacts for one element, from its info (whose master-border
// slot and stencil flag it updates)
__device__ __forceinline__ void project_core(const DevModel& m, const ContactArrays& C, int e, int4& inf, ElemProj& r) {
    const int bS = C.bodyS[e];
    const int bodyM = C.bodyM[e];
    int borderM = C.borderM[e], shiftM = C.shiftM[e], segM = C.segM[e];
    const int nodeS = C.nodeS[e];
    Proj p;
    p.gap = C.gapn[e]; p.xsi = C.xsi[e]; p.xn = C.xn[e]; p.yn = C.yn[e]; p.xt = C.xt[e]; p.yt = C.yt[e];
    p.N0 = C.N0[e]; p.N1 = C.N1[e]; p.N2 = C.N2[e]; p.N3 = C.N3[e];
    p.flag = 0;
    double gapt = C.gapt[e];
    const double xsi_prev = p.xsi;
    int nd[4], sh[4];
    nd[0] = C.nM0[e]; nd[1] = C.nM1[e]; nd[2] = C.nM2[e]; nd[3] = C.nM3[e];
    sh[0] = C.sM0[e]; sh[1] = C.sM1[e]; sh[2] = C.sM2[e]; sh[3] = C.sM3[e];
    if (!(inf.z & CI_STENCIL)) load_stencil(m, bodyM, borderM, segM, nd, sh);
    const double xs = m.xc[nodeS], ys = m.yc[nodeS];
    const double P = m.period;
    double X[4], Y[4];
    for (int k = 0; k < 4; k++) { X[k] = m.xc[nd[k]] + P * (shiftM + sh[k]); Y[k] = m.yc[nd[k]]; }
    auto project = [&]() {
        if (bS == bodyM && (nd[0] == nodeS || nd[3] == nodeS))
            closest_self(xs, ys, X[1], Y[1], X[2], Y[2], p);
        else
            closest_linear(xs, ys, X[0], Y[0], X[1], Y[1], X[2], Y[2], X[3], Y[3], p);
    };
    project();
    auto seglen = [&]() { return __dsqrt_rn((X[2] - X[1]) * (X[2] - X[1]) + (Y[2] - Y[1]) * (Y[2] - Y[1])); };
    // the next stencil after a walk (slot = first slot of the master border + segment)
    auto restencil = [&]() {
        const int s = inf.w + segM;
        nd[0] = m.sl_n0[s]; nd[1] = m.sl_n1[s]; nd[2] = m.sl_n2[s]; nd[3] = m.sl_n3[s];
        sh[0] = m.sl_s0[s]; sh[1] = m.sl_s1[s]; sh[2] = m.sl_s2[s]; sh[3] = m.sl_s3[s];
        for (int k = 0; k < 4; k++) { X[k] = m.xc[nd[k]] + P * (shiftM + sh[k]); Y[k] = m.yc[nd[k]]; }
    };
    if (p.flag == -2 || p.flag == -3) {
        const int gbM = m.border_off[bodyM] + borderM;
        const int per = m.br_periodicity[gbM];
        gapt = gapt + 0.5 * (-1. - xsi_prev) * seglen();
        if (per == 1) {
            if (segM == 0) {
                borderM = m.br_before[gbM];
                segM = m.br_nslots[m.border_off[bodyM] + borderM] - 2;
                inf.w = slot_of(m, bodyM, borderM, 0);
            } else segM--;
        } else if (per == 2) {
            if (segM == 0) {
                if (borderM == 0) shiftM--;
                else if (borderM == 1) shiftM++;
                segM = m.br_nslots[gbM] - 2;
            } else segM--;
        } else {
            if (segM == 0) segM = m.br_nslots[gbM] - 2;
            else segM--;
        }
        restencil();
        p.xsi = 1.;
        project();
        gapt = gapt + 0.5 * (-1. + p.xsi) * seglen();
    } else if (p.flag == 2 || p.flag == 3) {
        const int gbM = m.border_off[bodyM] + borderM;
        const int per = m.br_periodicity[gbM];
        gapt = gapt + 0.5 * (-1. + xsi_prev) * seglen();
        const int last = m.br_nslots[gbM] - 2;
        if (per == 1) {
            if (segM == last) {
                borderM = m.br_after[gbM];
                segM = 0;
                inf.w = slot_of(m, bodyM, borderM, 0);
            } else segM++;
        } else if (per == 2) {
            if (segM == last) {
                if (borderM == 0) shiftM++;
                else if (borderM == 1) shiftM--;
                segM = 0;
            } else segM++;
        } else {
            if (segM == last) segM = 0;
            else segM++;
        }
        restencil();
        p.xsi = -1.;
        project();
        gapt = gapt + 0.5 * (1. + p.xsi) * seglen();
    } else {
        gapt = gapt + 0.5 * (p.xsi - xsi_prev) * seglen();
    }
    inf.z |= CI_STENCIL;
    const double vgx = m.vx[nodeS] - (p.N0 * m.vx[nd[0]] + p.N1 * m.vx[nd[1]] + p.N2 * m.vx[nd[2]] + p.N3 * m.vx[nd[3]]);
    const double vgy = m.vy[nodeS] - (p.N0 * m.vy[nd[0]] + p.N1 * m.vy[nd[1]] + p.N2 * m.vy[nd[2]] + p.N3 * m.vy[nd[3]]);
    r.nd0 = nd[0]; r.nd1 = nd[1]; r.nd2 = nd[2]; r.nd3 = nd[3];
    r.sh0 = sh[0]; r.sh1 = sh[1]; r.sh2 = sh[2]; r.sh3 = sh[3];
    r.p = p;
    r.borderM = borderM; r.shiftM = shiftM; r.segM = segM; r.nodeS = nodeS;
    r.xs = xs; r.ys = ys;
    r.gapt = gapt;
    r.vgapn = vgx * p.xn + vgy * p.yn;
    r.vgapt = vgx * p.xt + vgy * p.yt;
}

__device__ __forceinline__ void project_store(const ContactArrays& C, int e, const int4& inf, const ElemProj& r) {
    const Proj& p = r.p;
    C.borderM[e] = r.borderM; C.shiftM[e] = r.shiftM; C.segM[e] = r.segM;
    C.gapn[e] = p.gap;
    C.vgapn[e] = r.vgapn;
    C.gapt[e] = r.gapt;
    C.vgapt[e] = r.vgapt;
    C.xsi[e] = p.xsi; C.xn[e] = p.xn; C.yn[e] = p.yn; C.xt[e] = p.xt; C.yt[e] = p.yt;
    C.nM0[e] = r.nd0; C.nM1[e] = r.nd1; C.nM2[e] = r.nd2; C.nM3[e] = r.nd3;
    C.sM0[e] = r.sh0; C.sM1[e] = r.sh1; C.sM2[e] = r.sh2; C.sM3[e] = r.sh3;
    C.N0[e] = p.N0; C.N1[e] = p.N1; C.N2[e] = p.N2; C.N3[e] = p.N3;
    C.info[e] = inf;
}

// re-project the slave node, walk to the previous / next master segment on
// flags +-2/+-3, accumulate the tangential gap, relative velocities.  When
// both projections return +-3 the element keeps its own stored shape values.
__device__ inline void project_element(const DevModel& m, const ContactArrays& C, int e) {
    int4 inf, sinf;
    element_info_ld(m, C, e, inf, sinf);
    ElemProj r;
    project_core(m, C, e, inf, r);
    project_store(C, e, inf, r);
}

// DOF_TO_DISP gather of one node (Body::Update_kinematics, Code/Body.h:2176-2212)
// and its current position (Node::Update_current_positions, Code/Node.h:233-237)
__device__ inline void kinematics_node(const DevModel& m, int i) {
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

// force fields of a deformable node rebuilt from the last step's
// components (same expression as the update): Node::x_force and its parts
// (Code/Node.h:212-225); body forces m g (Body::Update_body_forces) unless
// with_body_forces is 0.  An inactive body's nodes keep their last total
// while their contact force goes on receiving sends (bn_acx, zero for
// active bodies).
__device__ inline void node_outputs(const DevModel& m, int j, int with_body_forces) {
    const int b = m.node_body[j];
    double cx = 0., cy = 0., dx = 0., dy = 0., nx = 0., ny = 0., rx = 0., ry = 0.;
    for (int e = m.bt_ptr[j]; e < m.bt_ptr[j + 1]; e++) {
        int bn = m.bt_bn[e];
        double f = m.bt_phi[e];
        cx += m.bn_pcx[bn] * f; cy += m.bn_pcy[bn] * f;
        dx += m.bn_pdx[bn] * f; dy += m.bn_pdy[bn] * f;
        nx += m.bn_pnx[bn] * f; ny += m.bn_pny[bn] * f;
        if (m.bn_acx) { rx += m.bn_acx[bn] * f; ry += m.bn_acy[bn] * f; }
    }
    double gx = m.gravity_x, gy = m.gravity_y;
    if (m.body_status[b] == 2 || !with_body_forces) { gx = 0.; gy = 0.; }
    double bx = m.mx[j] * gx, by = m.my[j] * gy;
    m.fcx[j] = m.bn_acx ? cx + rx : cx; m.fcy[j] = m.bn_acy ? cy + ry : cy;
    m.fdx[j] = dx; m.fdy[j] = dy; m.fnx[j] = nx; m.fny[j] = ny;
    m.fbx[j] = bx; m.fby[j] = by;
    const double lx = m.alid_on ? m.fax[j] : 0., ly = m.alid_on ? m.fay[j] : 0.;
    m.fx[j] = m.fix[j] + cx + 0. + bx + dx + nx + m.fkx[j] + lx;
    m.fy[j] = m.fiy[j] + cy + 0. + by + dy + ny + m.fky[j] + ly;
}

}  // namespace mel
