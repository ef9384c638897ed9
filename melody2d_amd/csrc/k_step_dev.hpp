// Device routines of the explicit step shared by the step kernels
// (k_step.hip) and the fused step (k_fused.hip).  Each routine names the
// reference function it replaces.  Compiled with -ffp-contract=off.
#pragma once
#include "mel_device.hpp"
#include "k_block.hpp"
#include "k_step.hpp"
#include "k_rigid_lists.hpp"
#include <climits>

namespace mel {


// ---------------------------------------------------------------------------
// Body::Update_borders (Code/Body.h:614-661): tributary length per slot.
// ---------------------------------------------------------------------------
__device__ inline double slot_length(const DevModel& m, int s) {
    int br = m.sl_border[s];
    int j = m.sl_j[s];
    int n = m.br_nslots[br];
    double P = m.period;
    int a0 = m.sl_n0[s], a1 = m.sl_n1[s], a2 = m.sl_n2[s];
    double x0 = m.xc[a0] + P * m.sl_s0[s], y0 = m.yc[a0];
    double x1 = m.xc[a1] + P * m.sl_s1[s], y1 = m.yc[a1];
    double x2 = m.xc[a2] + P * m.sl_s2[s], y2 = m.yc[a2];
    double l21 = __dsqrt_rn((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    double l10 = __dsqrt_rn((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
    bool simple = m.br_periodicity[br] == 1;
    if (j == 0 && simple) return l21 * 0.5;
    if (j == n - 1 && simple) return l10 * 0.5;
    return (l21 + l10) * 0.5;
}

// current tributary length of a slot: deformable borders from the current
// positions (the step recomputes them where they are used), rigid borders
// keep their stored value
__device__ inline double slot_len_now(const DevModel& m, int s) {
    int b = m.br_body[m.sl_border[s]];
    return m.body_rigid_flag[b] ? m.sl_len[s] : slot_length(m, s);
}


// ---------------------------------------------------------------------------
// Contact laws, Code/Model.h:81-335 (0.5 factors applied by the caller as in
// Code/Body.h:1703-1789)
// ---------------------------------------------------------------------------
__device__ inline double sqrt_r(double x) { return __dsqrt_rn(x); }
__device__ inline float sqrt_r(float x) { return __fsqrt_rn(x); }

// Written over the arithmetic type R: double in the FP64 form, float in the
// FP32-mixed form (gaps arrive rounded from FP64; a law that does not
// return-map the tangential gap or change the damage leaves the stored FP64
// values as they were: *_set flags).
template <class R>
__device__ inline void law_mohr_coulomb(R kn, R kt, R fric, R coh, R tens, R gapn, R& gapt, bool& gapt_set, R& Pn,
                                        R& Pt) {
    Pn = -kn * gapn;
    Pt = -kt * gapt;
    if (Pn < -tens) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
    else if (fabs(Pt) > coh + fric * Pn) {
        if (gapt > R(0)) Pt = -coh - fric * Pn;
        else Pt = coh + fric * Pn;
        gapt = -Pt / kt;
        gapt_set = true;
    }
}

template <class R>
__device__ inline void law_czm_linear(R kini, R Pnlim, R gapnlim, R Pnres, R NT, R gapn, R& gapt, bool& gapt_set, R& D,
                                      bool& D_set, R& Pn, R& Pt) {
    if (D >= R(1)) {
        Pn = -kini * gapn;
        Pt = -kini * gapt;
        if (Pn < -Pnres) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
        else if (fabs(Pt) > Pnres * NT) {
            Pt = gapt > R(0) ? -Pnres * NT : Pnres * NT;
            gapt = -Pt / kini;
            gapt_set = true;
        }
        return;
    }
    R gnmax = Pnlim / kini * (R(1) - D) + gapnlim * D;
    R gtmax = gnmax * NT;
    R kmax;
    bool n_in = gapn < gnmax, n_out = gapn > gnmax;
    bool t_in = fabs(gapt) < gtmax, t_out = fabs(gapt) > gtmax;
    if (n_in && t_in) {
        kmax = (Pnlim * (R(1) - D) + Pnres * D) / gnmax;
    } else if (n_out && t_in) {
        gnmax = gapn; gtmax = gnmax * NT;
        D = (gnmax - Pnlim / kini) / (gapnlim - Pnlim / kini);
        D_set = true;
        kmax = (Pnlim * (R(1) - D) + Pnres * D) / gnmax;
    } else if (n_in && t_out) {
        gtmax = fabs(gapt); gnmax = gtmax / NT;
        D = (gnmax - Pnlim / kini) / (gapnlim - Pnlim / kini);
        D_set = true;
        kmax = (Pnlim * (R(1) - D) + Pnres * D) / gnmax;
    } else if (n_out && t_out) {
        if (gapn / gnmax > fabs(gapt) / gtmax) { gnmax = gapn; gtmax = gnmax * NT; }
        else { gtmax = fabs(gapt); gnmax = gtmax / NT; }
        D = (gnmax - Pnlim / kini) / (gapnlim - Pnlim / kini);
        D_set = true;
        kmax = (Pnlim * (R(1) - D) + Pnres * D) / gnmax;
    } else {
        return;
    }
    if (gapn < R(0)) Pn = -kini * gapn;
    else Pn = -kmax * gapn;
    Pt = -kmax * gapt;
}

// Apply_CZM_fatigue (Code/Model.h:343-446): the linear cohesive law whose
// damage also grows with the traction increments above the fatigue
// threshold Pnfat; Pn, Pt enter as the element's previous tractions
template <class R>
__device__ inline void law_czm_fatigue(R kini, R Pnlim, R gapnres, R Pnres, R NT, R Pnfat, R Raten, R Ratet, R gapn,
                                       R& gapt, bool& gapt_set, R& D, R& Pn, R& Pt) {
    if (D < R(1)) {
        R gnmax = Pnlim / kini * (R(1) - D) + gapnres * D;
        R gtmax = gnmax * NT;
        R kmax;
        if (gapn < gnmax && fabs(gapt) < gtmax) {
            const R Pnm = Pn, Ptm = Pt;
            kmax = (Pnlim * (R(1) - D) + Pnres * D) / gnmax;
            Pn = gapn < R(0) ? -kini * gapn : -kmax * gapn;
            Pt = -kmax * gapt;
            const R Pnmax = Pnlim * (R(1) - D) + Pnres * D;
            if (fabs(Pn) > Pnfat && fabs(Pn) > fabs(Pnm))
                D = D + (fabs(Pn) - Pnfat) / (Pnmax - Pnfat) * Raten * (fabs(Pn) - fabs(Pnm));
            if (fabs(Pt) > NT * Pnfat && fabs(Pt) > fabs(Ptm))
                D = D + (fabs(Pt) - Pnfat * NT) / (NT * (Pnmax - Pnfat)) * Ratet * (fabs(Pt) - fabs(Ptm));
        } else {
            bool grew = true;
            if (gapn > gnmax && fabs(gapt) < gtmax) {
                gnmax = gapn;
                gtmax = gnmax * NT;
            } else if (gapn < gnmax && fabs(gapt) > gtmax) {
                gtmax = fabs(gapt);
                gnmax = gtmax / NT;
            } else if (gapn > gnmax && fabs(gapt) > gtmax) {
                if (gapn / gnmax > fabs(gapt) / gtmax) { gnmax = gapn; gtmax = gnmax * NT; }
                else { gtmax = fabs(gapt); gnmax = gtmax / NT; }
            } else {
                grew = false;     // on the elastic domain's boundary: tractions and damage unchanged
            }
            if (grew) {
                D = (gnmax - Pnlim / kini) / (gapnres - Pnlim / kini);
                kmax = (Pnlim * (R(1) - D) + Pnres * D) / gnmax;
                Pn = gapn < R(0) ? -kini * gapn : -kmax * gapn;
                Pt = -kmax * gapt;
            }
        }
    }
    if (D >= R(1)) {
        D = R(1);
        Pn = -kini * gapn;
        Pt = -kini * gapt;
        if (Pn < -Pnres) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
        else if (fabs(Pt) > Pnres * NT) {
            Pt = gapt > R(0) ? -Pnres * NT : Pnres * NT;
            gapt = -Pt / kini;
            gapt_set = true;
        }
    }
}

// Apply_Bonded_Mohr_Coulomb (Code/Model.h:453-506): damped Mohr-Coulomb with
// a bond (stiffness kbond in tension, its own friction, cohesion and tensile
// strength) that breaks for good (damage 1) when exceeded
template <class R>
__device__ inline void law_bonded_mc(R kn, R kt, R kbond, R fricbond, R cohbond, R tensbond, R fricfree, R cohfree,
                                     R tensfree, R damp, R mass, R length, R gapn, R vgapn, R& gapt, bool& gapt_set,
                                     R vgapt, R& D, R& Pn, R& Pt) {
    if (D >= R(1)) {
        Pn = -kn * gapn - damp * vgapn * R(2) * sqrt_r(mass * kn / length);
        if (Pn < -tensfree) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
        else Pt = -kt * gapt - damp * vgapt * R(2) * sqrt_r(mass * kt / length);
        if (fabs(Pt) > cohfree + fricfree * Pn) {
            Pt = gapt > R(0) ? -cohfree - fricfree * Pn : cohfree + fricfree * Pn;
            gapt = -Pt / kt;
            gapt_set = true;
        }
    } else {
        if (gapn < R(0)) Pn = -kn * gapn - damp * vgapn * R(2) * sqrt_r(mass * kn / length);
        else Pn = -kbond * gapn - damp * vgapn * R(2) * sqrt_r(mass * kbond / length);
        Pt = -kt * gapt - damp * vgapt * R(2) * sqrt_r(mass * kt / length);
        if (Pn < -tensbond) {
            D = R(1);
            if (Pn < -tensfree) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
        }
        if (fabs(Pt) > cohbond + fricbond * Pn) {
            D = R(1);
            Pt = gapt > R(0) ? -cohfree - fricfree * Pn : cohfree + fricfree * Pn;
            gapt = -Pt / kt;
            gapt_set = true;
        }
    }
}

// Update_contact_forces' law step for one element (Code/Body.h:1694-1840
// with the laws of Code/Model.h:81-335): tractions Pn, Pt and the damped
// law's work rate W from the element's gaps, in the arithmetic type R.
// gapt (return mapping) and D (CZM damage, nbint its first-use flag) are
// updated in place when the law changes them.
template <class R>
__device__ __forceinline__ void contact_law(const DevModel& m, int lt, const double* p, double gapn_d, double& gapt_d,
                                            double vgn_d, double vgt_d, double emass, double length0, double& D_d,
                                            int& nbint, double& P1_d, double& P2_d, double& Pn_o, double& Pt_o,
                                            double& W_o) {
    const R gapn = (R)gapn_d;
    R gapt = (R)gapt_d;
    bool gapt_set = false;
    R Pn = R(0), Pt = R(0), W = R(0);
    if (lt == 0) {                      // Frictionless
        Pn = gapn > R(0) ? R(0) : -(R)p[0] * gapn;
        Pt = R(0);
        gapt = R(0);
        gapt_set = true;
    } else if (lt == 1) {               // Cohesion
        R gtang = (R)p[2] * R(0.5), gtens = (R)p[3] * R(0.5);
        Pn = -(R)p[0] * gapn;
        Pt = -(R)p[1] * gapt;
        if (Pn < -gtens) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
        else if (fabs(Pt) > gtang) { Pt = gapt > R(0) ? -gtang : gtang; gapt = -Pt / (R)p[1]; gapt_set = true; }
    } else if (lt == 2) {               // MohrCoulomb
        law_mohr_coulomb<R>((R)p[0], (R)p[1], (R)p[2], (R)p[3] * R(0.5), (R)p[4] * R(0.5), gapn, gapt, gapt_set, Pn,
                            Pt);
    } else if (lt == 3) {               // DampedMohrCoulomb
        R kn = (R)p[0], kt = (R)p[1], fric = (R)p[2], coh = (R)p[3] * R(0.5), tens = (R)p[4] * R(0.5), damp = (R)p[5];
        R mass = (R)emass, L0 = (R)length0, vgn = (R)vgn_d, vgt = (R)vgt_d;
        Pn = R(0);
        if (gapn < R(0)) { Pn = -damp * vgn * R(2) * sqrt_r(mass * kn / L0); W += Pn * vgn; }
        Pn -= kn * gapn;
        if (Pn < -tens) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
        else {
            R Ptd = -damp * vgt * R(2) * sqrt_r(mass * kt / L0);
            Pt = -kt * gapt + Ptd;
            if (fabs(Pt) > coh + fric * Pn) {
                Pt = gapt > R(0) ? -coh - fric * Pn : coh + fric * Pn;
                gapt = -Pt / kt;
                gapt_set = true;
                W += Pt * vgt;
            } else W += Ptd * vgt;
        }
    } else if (lt == 4) {               // TwoSlopesMohrCoulomb
        R kn = (R)p[0], kt = (R)p[1], fric = (R)p[2], coh = (R)p[3] * R(0.5), tens = (R)p[4] * R(0.5), rati = (R)p[5];
        Pn = (R)vgn_d <= R(0) ? -kn * gapn : -kn * gapn * rati;
        if (Pn < -tens) { Pn = R(0); Pt = R(0); gapt = R(0); gapt_set = true; }
        else {
            Pt = -kt * gapt;
            if (fabs(Pt) > coh + fric * Pn) {
                Pt = gapt > R(0) ? -coh - fric * Pn : coh + fric * Pn;
                gapt = -Pt / kt;
                gapt_set = true;
            }
        }
    } else if (lt == 6) {               // CZMlinear
        R D = (R)D_d;
        bool D_set = false;
        if (nbint == 0) { D = R(1); D_set = true; nbint = 1; }
        law_czm_linear<R>((R)p[0], (R)p[1] * R(0.5), (R)p[2], (R)p[3] * R(0.5), (R)p[4], gapn, gapt, gapt_set, D, D_set,
                          Pn, Pt);
        if (D_set) D_d = (double)D;
    } else if (lt == 5 || lt == 7) {    // BondedMohrCoulomb, CZMfatigue: damage and the previous tractions
        R D = (R)D_d;
        Pn = (R)P1_d;
        Pt = (R)P2_d;
        if (nbint == 0) { D = R(1); nbint = 3; }
        if (lt == 5)
            law_bonded_mc<R>((R)p[0], (R)p[1], (R)p[2], (R)p[3], (R)p[4] * R(0.5), (R)p[5] * R(0.5), (R)p[6],
                             (R)p[7] * R(0.5), (R)p[8] * R(0.5), (R)p[9], (R)emass, (R)length0, gapn, (R)vgn_d, gapt,
                             gapt_set, (R)vgt_d, D, Pn, Pt);
        else
            law_czm_fatigue<R>((R)p[0], (R)p[1] * R(0.5), (R)p[2], (R)p[3] * R(0.5), (R)p[4], (R)p[5], (R)p[6],
                               (R)p[7], gapn, gapt, gapt_set, D, Pn, Pt);
        D_d = (double)D;
        P1_d = (double)Pn;
        P2_d = (double)Pt;
    }
    if (gapt_set) gapt_d = (double)gapt;
    Pn_o = (double)Pn;
    Pt_o = (double)Pt;
    W_o = (double)W;
}

// a point-force record (id r, force) stored at position pos of border node
// bn's bucket, and appended to it
__device__ inline void bucket_store(const DevModel& m, int bn, int pos, int r, double fx, double fy) {
    if (pos < BUCKET_CAP) {
        m.bk_rec[bn * BUCKET_CAP + pos] = r;
        m.bk_val[bn * BUCKET_CAP + pos] = make_double2(fx, fy);
    } else {
        atomicOr(&m.status_flags[0], 1);
    }
}

__device__ inline void bucket_append(const DevModel& m, int bn, int r, double fx, double fy) {
    bucket_store(m, bn, atomicAdd(&m.bk_cursor[bn], 1), r, fx, fy);
}

// a deformable slot's current tributary length from its stencil record
// (slot_length's expressions)
__device__ __forceinline__ double slot_length_stencil(const DevModel& m, int4 st) {
    const double P = m.period;
    const int w = st.w;
    const double x0 = m.xc[st.x] + P * ((w & 3) - 1), y0 = m.yc[st.x];
    const double x1 = m.xc[st.y] + P * (((w >> 2) & 3) - 1), y1 = m.yc[st.y];
    const double x2 = m.xc[st.z] + P * (((w >> 4) & 3) - 1), y2 = m.yc[st.z];
    const double l21 = __dsqrt_rn((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    const double l10 = __dsqrt_rn((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
    const int end = (w >> 6) & 3;
    if (end == 1) return l21 * 0.5;
    if (end == 2) return l10 * 0.5;
    return (l21 + l10) * 0.5;
}

// Body::Update_contacts (project_core) then Body::Update_contact_forces
// (Code/Body.h:1642-1941) minus the scatter, for one element: its force and
// its point-force records.  Everything the element reads is one load away
// from its own arrays (element_info), the slave's bucket slot is reserved
// before the projection, and the projected state stays in registers for the
// law.
__device__ __forceinline__ void contact_item(const DevModel& m, int e) {
    const ContactArrays& C = m.C;
    int4 inf, sinf;
    element_info_ld(m, C, e, inf, sinf);
    m.rec_bn[3 * e] = -1; m.rec_bn[3 * e + 1] = -1; m.rec_bn[3 * e + 2] = -1;
    for (int q = 0; q < 6; q++) m.rig_rec[6 * e + q] = 0.;
    const int fl = inf.z;
    if (fl & CI_INACTIVE_S) return;
    const int li = (fl & CI_LAW_MASK) - 1;
    // loads that depend only on the element's info
    int posS = 0;
    if (li >= 0 && !(fl & CI_RIGID_S)) posS = atomicAdd(&m.bk_cursor[inf.y], 1);
    const double emass = C.emass[e], int0 = C.int0[e];
    const int nbint = C.nbint[e];
    const int bS = C.bodyS[e], bM = C.bodyM[e];
    ElemProj r;
    project_core(m, C, e, inf, r);
    project_store(C, e, inf, r);
    if (li < 0) return;
    const int slot = inf.x;
    const double length = (fl & CI_RIGID_S) ? m.sl_len[slot] : slot_length_stencil(m, sinf);
    const double length0 = m.sl_len0[slot];
    double gapn = r.p.gap, gapt = r.gapt;
    const double* p = m.law_p + 12 * li;
    double Pn = 0., Pt = 0., W = 0.;
    const int lt = m.law_type[li];
    double D = int0;
    int nb = nbint;
    // the three-variable laws' previous tractions (zero for a fresh element,
    // whose law starts from a broken bond and does not read them)
    double P1 = 0., P2 = 0.;
    if ((lt == 5 || lt == 7) && nbint == 3) { P1 = C.int1[e]; P2 = C.int2[e]; }
    if (m.law_f32) contact_law<float>(m, lt, p, gapn, gapt, r.vgapn, r.vgapt, emass, length0, D, nb, P1, P2, Pn, Pt, W);
    else contact_law<double>(m, lt, p, gapn, gapt, r.vgapn, r.vgapt, emass, length0, D, nb, P1, P2, Pn, Pt, W);
    if (lt == 6 || lt == 5 || lt == 7) {
        if (nb != nbint) C.nbint[e] = nb;
        C.int0[e] = D;
        if (lt != 6) { C.int1[e] = P1; C.int2[e] = P2; }
    }
    const double Px = r.p.xn * Pn + r.p.xt * Pt;
    const double Py = r.p.yn * Pn + r.p.yt * Pt;
    const double L = m.law_fixed[li] ? length0 : length;
    const double Fx = Px * L, Fy = Py * L;
    W *= L * m.dt;
    C.gapn[e] = gapn; C.gapt[e] = gapt; C.len[e] = length; C.fx[e] = Fx; C.fy[e] = Fy;
    // Code/Body.h:1847-1848 add 0.5 W to the slave's and the master's
    // contact work; the element keeps its half until the works are flushed
    if (W != 0.) C.wacc[e] += 0.5 * W;
    if (!(fl & CI_RIGID_S)) {
        const int bn = inf.y;
        m.rec_bn[3 * e] = bn; m.rec_fx[3 * e] = Fx; m.rec_fy[3 * e] = Fy;
        bucket_store(m, bn, posS, 3 * e, Fx, Fy);
    } else {
        const double* R = m.body_rigid + bS * RIGID_FIELDS;
        double dx = r.xs - R[R_CUR], dy = r.ys - R[R_CUR + 1];
        m.rig_rec[6 * e] = Fx; m.rig_rec[6 * e + 1] = Fy; m.rig_rec[6 * e + 2] = -Fx * dy + Fy * dx;
    }
    if (!(fl & CI_RIGID_M)) {
        const int sl = inf.w + r.segM;
        const int bn1 = m.sl_bn[sl], bn2 = m.sl_bn[sl + 1];
        const double f1x = -Fx * r.p.N1, f1y = -Fy * r.p.N1, f2x = -Fx * r.p.N2, f2y = -Fy * r.p.N2;
        m.rec_bn[3 * e + 1] = bn1; m.rec_fx[3 * e + 1] = f1x; m.rec_fy[3 * e + 1] = f1y;
        m.rec_bn[3 * e + 2] = bn2; m.rec_fx[3 * e + 2] = f2x; m.rec_fy[3 * e + 2] = f2y;
        bucket_append(m, bn1, 3 * e + 1, f1x, f1y);
        bucket_append(m, bn2, 3 * e + 2, f2x, f2y);
    } else {
        const double* R = m.body_rigid + bM * RIGID_FIELDS;
        double dx = r.xs - R[R_CUR] - r.shiftM * m.period;
        double dy = r.ys - R[R_CUR + 1];
        m.rig_rec[6 * e + 3] = -Fx; m.rig_rec[6 * e + 4] = -Fy; m.rig_rec[6 * e + 5] = Fx * dy - Fy * dx;
    }
}

// point contact force per border node: records summed in record-id order
// (the bucket is emptied)
__device__ inline double2 bucket_total(const DevModel& m, int bn) {
    // the first BUCKET_FAST records are loaded together with the count (the
    // bucket's storage is always there; entries past the count are ignored)
    constexpr int BUCKET_FAST = 4;
    const int base = bn * BUCKET_CAP;
    int id[BUCKET_FAST];
    double2 v[BUCKET_FAST];
#pragma unroll
    for (int k = 0; k < BUCKET_FAST; k++) { id[k] = m.bk_rec[base + k]; v[k] = m.bk_val[base + k]; }
    const int n = min(m.bk_cursor[bn], BUCKET_CAP);
    double px = 0., py = 0.;
    if (n <= BUCKET_FAST) {
        // records in element order: sorting network on (id, force), unused
        // entries last
#pragma unroll
        for (int k = 0; k < BUCKET_FAST; k++)
            if (k >= n) id[k] = INT_MAX;
        auto ce = [&](int a, int b) {
            if (id[b] < id[a]) {
                int ti = id[a]; id[a] = id[b]; id[b] = ti;
                double2 tv = v[a]; v[a] = v[b]; v[b] = tv;
            }
        };
        ce(0, 1); ce(2, 3); ce(0, 2); ce(1, 3); ce(1, 2);
#pragma unroll
        for (int k = 0; k < BUCKET_FAST; k++)
            if (k < n) { px += v[k].x; py += v[k].y; }
    } else {
        int ids[BUCKET_CAP];
        for (int k = 0; k < n; k++) {
            int r = m.bk_rec[base + k];
            int q = k;
            while (q > 0 && ids[q - 1] > r) { ids[q] = ids[q - 1]; q--; }
            ids[q] = r;
        }
        for (int k = 0; k < n; k++) { px += m.rec_fx[ids[k]]; py += m.rec_fy[ids[k]]; }
    }
    m.bk_cursor[bn] = 0;
    return make_double2(px, py);
}

__device__ inline double2 bucket_sum_node(const DevModel& m, int bn) {
    const double2 p = bucket_total(m, bn);
    m.bn_pcx[bn] = p.x;
    m.bn_pcy[bn] = p.y;
    return p;
}

// a border node of an inactive body: what it received is added to what it
// received before (Send_contact_forces still reaches it, nothing resets it)
// and its point force of the last active step is kept
__device__ inline void bucket_drain_node(const DevModel& m, int bn) {
    const double2 p = bucket_total(m, bn);
    if (m.bn_acx) {
        m.bn_acx[bn] += p.x;
        m.bn_acy[bn] += p.y;
    }
}

// the contact point forces of a border node of body b
__device__ inline void bucket_gather(const DevModel& m, int bn, int b) {
    if (m.body_status[b] == 1) bucket_drain_node(m, bn);
    else bucket_sum_node(m, bn);
}

// ---------------------------------------------------------------------------
// Border::Update_bc (Code/Border.h:120-221): current BC values at time t
// ---------------------------------------------------------------------------
__device__ inline void bc_value(const DevModel& m, int type, int h0, int h1, double t, double& value, double& vel) {
    if (type == 1) return;
    if (type == 2) {
        value = 0.;
        for (int n = h0; n + 1 < h1; n++) {
            double t0 = m.hist_t[n], p0 = m.hist_v[n], t1 = m.hist_t[n + 1], p1 = m.hist_v[n + 1];
            if (t >= t1) value = value + (t1 - t0) * (p1 + p0) * 0.5;
            else {
                value = value + (t - t0) * (p0 + (p1 - p0) / (t1 - t0) * ((t - t0) * 0.5));
                vel = p0 + (p1 - p0) / (t1 - t0) * (t - t0);
                break;
            }
        }
    } else if (type == 3 || type == 4) {
        for (int n = h0; n + 1 < h1; n++) {
            double t0 = m.hist_t[n], p0 = m.hist_v[n], t1 = m.hist_t[n + 1], p1 = m.hist_v[n + 1];
            if ((t0 <= t) & (t <= t1)) { value = p0 + (p1 - p0) / (t1 - t0) * (t - t0); break; }
        }
    } else {
        value = 0.;
    }
}

__device__ inline void bc_eval_at(const DevModel& m, int br, double t) {
    const int* hp = m.br_hist_ptr + 2 * br;
    bc_value(m, m.br_xtype[br], hp[0], hp[1], t, m.br_xval[br], m.br_xvel[br]);
    bc_value(m, m.br_ytype[br], hp[1], hp[2], t, m.br_yval[br], m.br_yvel[br]);
}


// Body::Update_bc_forces deformable branch (Code/Body.h:672-815): point
// Neumann / Dirichlet values per border node (one slot per border node in
// practice; several slots are summed in slot order).
__device__ inline void bc_point_force_node(const DevModel& m, int bn) {
    double cnx = 0., cny = 0., cdx = 0., cdy = 0.;
    int node = m.bn_node[bn];
    int b = m.node_body[node];
    if (!m.body_rigid_flag[b] && m.body_status[b] != 1) {
        for (int q = m.bn_slot_ptr[bn]; q < m.bn_slot_ptr[bn + 1]; q++) {
            int s = m.bn_slots[q];
            int br = m.sl_border[s];
            int xt = m.br_xtype[br], yt = m.br_ytype[br];
            if (xt == 0 && yt == 0) continue;
            if (m.sl_j[s] >= m.br_nslots[br] - 1) continue;
            double L = slot_length(m, s);
            double px = 0., py = 0.;
            double tx = 0, ty = 0, tpx = 0, tpy = 0;
            if (xt == 4 || yt == 4) {
                int a = m.sl_n0[s], c = m.sl_n2[s];
                double x0 = m.X[a], y0 = m.Y[a], x2 = m.X[c], y2 = m.Y[c];
                double xp0 = m.xc[a], yp0 = m.yc[a], xp2 = m.xc[c], yp2 = m.yc[c];
                double invl = inv_sqrt((x2 - x0) * (x2 - x0) + (y2 - y0) * (y2 - y0));
                double invlp = inv_sqrt((xp2 - xp0) * (xp2 - xp0) + (yp2 - yp0) * (yp2 - yp0));
                tx = (x2 - x0) * invl; ty = (y2 - y0) * invl;
                tpx = (xp2 - xp0) * invlp; tpy = (yp2 - yp0) * invlp;
            }
            if (xt == 3) px = m.br_xval[br];
            else if (xt == 4) { px += m.br_xval[br] * (tx * tpx + ty * tpy); py += m.br_xval[br] * (tx * tpy - ty * tpx); }
            else if (xt == 2) px = m.br_xparam[2 * br] * (m.br_xval[br] - m.ux[node]) - m.br_xparam[2 * br + 1] * (m.br_xvel[br] - m.vx[node]);
            if (yt == 3) py = m.br_yval[br];
            else if (yt == 4) { px += m.br_yval[br] * (ty * tpx - tx * tpy); py += m.br_yval[br] * (ty * tpy + tx * tpx); }
            else if (yt == 2) py = m.br_yparam[2 * br] * (m.br_yval[br] - m.uy[node]) - m.br_yparam[2 * br + 1] * (m.br_yvel[br] - m.vy[node]);
            m.sl_xbp[s] = px; m.sl_ybp[s] = py;
            double a_nx = 0., a_ny = 0., a_dx = 0., a_dy = 0.;
            if (xt == 3) a_nx = px * L;
            else if (xt == 4) { a_nx = px * L; a_ny = py * L; }
            else if (xt == 2) a_dx = px * L;
            if (yt == 3) a_ny = py * L;
            else if (yt == 4) { a_nx = px * L; a_ny = py * L; }
            else if (yt == 2) a_dy = py * L;
            cnx += a_nx; cny += a_ny; cdx += a_dx; cdy += a_dy;
        }
    }
    m.bn_pnx[bn] = cnx; m.bn_pny[bn] = cny; m.bn_pdx[bn] = cdx; m.bn_pdy[bn] = cdy;
}


// ---------------------------------------------------------------------------
// rigid bodies: one thread per body for the serial parts
// ---------------------------------------------------------------------------
// Body::Update_bc_forces rigid branch (Code/Body.h:816-899) + Update_body_forces
// + the contact resultant (slave records and the serial sends of
// Code/Body.h:1948-1967), reduced in element order.  sc: LDS scratch of at
// least 6 * blockDim.x + 6 doubles.
struct RigidForces {
    double fcon[3], fneu[3], fdir[3], fbody[2];
};

// BC terms of run r (of BLOCK_THREADS runs of contiguous slots) of rigid
// border br (Code/Body.h:816-899, rigid branch of Update_bc_forces): Neumann
// and spring-dashpot forces of its slots and their moments about the body's
// current centre (xc0, yc0), added in slot order; rigid_forces_block adds
// the runs in slot order
__device__ inline void rigid_bc_run(const DevModel& m, int br, int r, double xc0, double yc0, double (&p6)[6]) {
    const int tid = r;
    const int xt = m.br_xtype[br], yt = m.br_ytype[br];
    const int ns = m.br_nslots[br];
    const int run = (ns + BLOCK_THREADS - 1) / BLOCK_THREADS;
    for (int q = 0; q < 6; q++) p6[q] = 0.;
    const int sb = m.br_slot0[br];
    const int j1 = min(ns, (tid + 1) * run);
    // batches of RUN_BATCH slots: their nodes, then the nodes' state,
    // loaded together (indices past the run repeat its last slot; their
    // terms add +0, which leaves the sums unchanged: they start at +0
    // and never become -0)
    constexpr int RUN_BATCH = 8;
    const bool spring_x = xt == 2, spring_y = yt == 2;
    const double xval = m.br_xval[br], yval = m.br_yval[br];
    for (int jb = tid * run; jb < j1; jb += RUN_BATCH) {
        int nd[RUN_BATCH];
        double Ls[RUN_BATCH], px[RUN_BATCH], py[RUN_BATCH];
#pragma unroll
        for (int u = 0; u < RUN_BATCH; u++) {
            const int jj = min(jb + u, j1 - 1);
            nd[u] = m.sl_node[sb + jj];
            Ls[u] = m.sl_len[sb + jj];
        }
#pragma unroll
        for (int u = 0; u < RUN_BATCH; u++) { px[u] = m.xc[nd[u]]; py[u] = m.yc[nd[u]]; }
#pragma unroll
        for (int u = 0; u < RUN_BATCH; u++) {
            const int node = nd[u];
            const double L = Ls[u];
            double fxn = 0., fyn = 0., fxd = 0., fyd = 0., bpx = 0., bpy = 0.;
            if (xt == 3 || xt == 4) { bpx = xval; fxn = bpx * L; }
            else if (spring_x) {
                bpx = m.br_xparam[2 * br] * (xval - m.ux[node]) + m.br_xparam[2 * br + 1] * (m.br_xvel[br] - m.vx[node]);
                fxd = bpx * L;
            }
            if (yt == 3 || yt == 4) { bpy = yval; fyn = bpy * L; }
            else if (spring_y) {
                bpy = m.br_yparam[2 * br] * (yval - m.uy[node]) + m.br_yparam[2 * br + 1] * (m.br_yvel[br] - m.vy[node]);
                fyd = bpy * L;
            }
            const double dx = px[u] - xc0, dy = py[u] - yc0;
            const bool in = jb + u < j1;
            if (in) { m.sl_xbp[sb + jb + u] = bpx; m.sl_ybp[sb + jb + u] = bpy; }
            p6[0] += in ? fxn : 0.; p6[1] += in ? fyn : 0.; p6[2] += in ? -fxn * dy + fyn * dx : 0.;
            p6[3] += in ? fxd : 0.; p6[4] += in ? fyd : 0.; p6[5] += in ? -fxd * dy + fyd * dx : 0.;
        }
    }
}

// bc_pre: the BC terms of the force launch (rbc_part), or nullptr to form them here
__device__ inline RigidForces rigid_forces_block(const DevModel& m, int rigid_index, double* sc,
                                                 const double* bc_pre) {
    RigidForces F{};
    const int b = m.rigid_ids[rigid_index];
    const int nc = *m.c_count;
    constexpr int nt = BLOCK_THREADS;   // workgroup size of every caller
    const int tid = threadIdx.x;
    double* red = sc;              // [nt][3], then [nt][6] for the BC runs
    double* fsum = sc + 6 * nt;    // [6]
    int* s_bcf = reinterpret_cast<int*>(sc + 6 * nt + 8 + RIGID_FIELDS + 8);   // [32] border gives BC forces
    // what depends only on the body, loaded with the contact records
    const int status = m.body_status[b];
    const int gb0 = m.border_off[b];
    const int nbr = m.body_nborders[b];
    const double* R = m.body_rigid + b * RIGID_FIELDS;
    const double xc0 = R[R_CUR], yc0 = R[R_CUR + 1];
    // only BC types 2 (spring-dashpot), 3 and 4 (pressure) give forces; the
    // others add zeros, which leave the sums unchanged.  One thread per
    // border reads its types.
    for (int k = tid; k < nbr && k < 32; k += nt) {
        const int xt = m.br_xtype[gb0 + k], yt = m.br_ytype[gb0 + k];
        s_bcf[k] = xt == 2 || xt == 3 || xt == 4 || yt == 2 || yt == 3 || yt == 4;
    }
    double sx = 0., sy = 0., sm = 0.;
    // the body's own elements (k_rigid_lists.hip): those it is the slave of,
    // then those it is the master of, in element order, strided over the
    // threads
    (void)nc;
    const int s0 = m.rs_range[2 * rigid_index], s1 = m.rs_range[2 * rigid_index + 1];
    const int q1 = min(m.rm_cnt[rigid_index], RIGID_LIST_CAP);
    const int* mlist = m.rm_list + (size_t)rigid_index * RIGID_LIST_CAP;
#pragma unroll 4
    for (int e = s0 + tid; e < s1; e += nt) {
        const double* r = m.rig_rec + 6 * e;
        sx += r[0]; sy += r[1]; sm += r[2];
    }
#pragma unroll 4
    for (int q = tid; q < q1; q += nt) {
        const double* r = m.rig_rec + 6 * mlist[q];
        sx += r[3]; sy += r[4]; sm += r[5];
    }
    if (m.n_rules > 0) {
        // the last sends of inactive bodies that named this body as master
        const int ng = min(m.gh_count[1], m.ghost_cap);
        for (int q = tid; q < ng; q += nt)
            if (m.gh_rb[q] == b) { sx += m.gh_rf[3 * q]; sy += m.gh_rf[3 * q + 1]; sm += m.gh_rf[3 * q + 2]; }
    }
    // an inactive body only adds what it received to its contact force
    // (the same for the whole workgroup)
    if (status == 1 && m.n_rules == 0) return F;
    red[3 * tid] = sx; red[3 * tid + 1] = sy; red[3 * tid + 2] = sm;
    __syncthreads();
    for (int w = nt / 2; w > 0; w >>= 1) {
        if (tid < w) { red[3 * tid] += red[3 * (tid + w)]; red[3 * tid + 1] += red[3 * (tid + w) + 1]; red[3 * tid + 2] += red[3 * (tid + w) + 2]; }
        __syncthreads();
    }
    if (tid == 0) { F.fcon[0] = red[0]; F.fcon[1] = red[1]; F.fcon[2] = red[2]; }
    if (status == 1) {
        if (tid == 0) {
            double* Rg = m.body_rigid + b * RIGID_FIELDS;
            for (int q = 0; q < 3; q++) Rg[R_FCON + q] += F.fcon[q];
        }
        return F;
    }
    // boundary conditions: every thread adds the terms of a contiguous run of
    // slots, then the runs are added in slot order (a two-level version of
    // the reference's serial loop)
    double* part = sc;
    double acc[6] = {0., 0., 0., 0., 0., 0.};
    for (int k = 0; k < nbr; k++) {
        if (k < 32 && !s_bcf[k]) continue;
        const int br = gb0 + k;
        const int xt = m.br_xtype[br], yt = m.br_ytype[br];
        const bool fx_bc = xt == 2 || xt == 3 || xt == 4, fy_bc = yt == 2 || yt == 3 || yt == 4;
        if (!fx_bc && !fy_bc) continue;
        double p6[6];
        if (bc_pre && m.br_rbc[br] >= 0) {
            const double* q6 = bc_pre + ((size_t)m.br_rbc[br] * BLOCK_THREADS + tid) * 6;
            for (int q = 0; q < 6; q++) p6[q] = q6[q];
        } else {
            rigid_bc_run(m, br, tid, xc0, yc0, p6);
        }
        __syncthreads();
        for (int q = 0; q < 6; q++) part[6 * tid + q] = p6[q];
        __syncthreads();
        if (tid < 6) {
            // runs in slot order; the loads of a batch of runs issue together
            double a = acc[tid];
#pragma unroll 16
            for (int r = 0; r < nt; r++) a += part[6 * r + tid];
            acc[tid] = a;
        }
    }
    if (tid < 6) fsum[tid] = acc[tid];
    __syncthreads();
    if (tid != 0) return F;
    for (int q = 0; q < 3; q++) { F.fneu[q] = fsum[q]; F.fdir[q] = fsum[3 + q]; }
    double gx = m.gravity_x, gy = m.gravity_y;
    if (m.body_status[b] == 2) { gx = 0.; gy = 0.; }
    F.fbody[0] = m.body_mass[b] * gx;
    F.fbody[1] = m.body_mass[b] * gy;
    return F;
}

// driven DOFs of Body::Update_bc (Code/Body.h:340-393) at time t
__device__ inline void driven(const DevModel& m, int h0, int h1, double t, double* d) {
    d[0] = 1.;
    d[1] = 0.;
    for (int n = h0; n + 1 < h1; n++) {
        double t0 = m.hist_t[n], p0 = m.hist_v[n], t1 = m.hist_t[n + 1], p1 = m.hist_v[n + 1];
        if (t >= t1) d[1] += (t1 - t0) * (p1 + p0) * 0.5;
        else {
            d[1] += (t - t0) * (p0 + (p1 - p0) / (t1 - t0) * ((t - t0) * 0.5));
            d[2] = (t - t0) * (p1 - p0) / (t1 - t0);
            d[3] = (p1 - p0) / (t1 - t0);
            break;
        }
    }
}

__device__ inline void rigid_driven_dofs(const DevModel& m, int b, double t, double* dv) {
    int gb0 = m.border_off[b];
    for (int k = 0; k < m.body_nborders[b]; k++) {
        int br = gb0 + k;
        const int* hp = m.br_hist_ptr + 2 * br;
        if (m.br_xtype[br] == 1) driven(m, hp[0], hp[1], t, dv);
        if (m.br_ytype[br] == 1) driven(m, hp[1], hp[2], t, dv + 4);
    }
}

// Sum_up_forces, Apply_Newton, Update_bc(t+dt), Apply_Euler for a rigid body
// (Code/Body.h:1984-2014, 2032-2037, 2071-2117)
// F: the body's forces from rigid_forces_block.  The body record is worked
// on in registers and stored back once (the contact work slot, which
// contact elements add to atomically, is left alone).
// rec: the body's record as it was when the step started (RIGID_FIELDS
// values), dvt: its driven DOFs at the end of the step (rigid_driven_dofs).
// the body's constants the step reads (loaded before its forces are known)
struct RigidScalars {
    int status, periodicity;
    double inv_mass, inv_inertia, mass;
};

__device__ inline RigidScalars rigid_scalars(const DevModel& m, int b) {
    RigidScalars r;
    r.status = m.body_status[b];
    r.periodicity = m.body_periodicity[b];
    r.inv_mass = m.body_inv_mass[b];
    r.inv_inertia = m.body_inv_inertia[b];
    r.mass = m.body_mass[b];
    return r;
}

__device__ inline void rigid_step_body(const DevModel& m, int b, const RigidForces& F, const double* rec,
                                       const double* dvt, const RigidScalars& rs) {
    if (rs.status == 1) return;
    double* Rg = m.body_rigid + b * RIGID_FIELDS;
    double R[RIGID_FIELDS];
#pragma unroll
    for (int k = 0; k < RIGID_FIELDS; k++) R[k] = rec[k];
    for (int q = 0; q < 3; q++) { R[R_FCON + q] = F.fcon[q]; R[R_FNEU + q] = F.fneu[q]; R[R_FDIR + q] = F.fdir[q]; }
    R[R_FBODY] = F.fbody[0];
    R[R_FBODY + 1] = F.fbody[1];
    double dt = m.dt;
    for (int q = 0; q < 3; q++)
        R[R_FTOT + q] = R[R_FCON + q] + R[R_FBODY + q] + R[R_FDIR + q] + R[R_FNEU + q] + R[R_FDAMP + q];
    R[R_ACC] = R[R_FTOT] * rs.inv_mass;
    R[R_ACC + 1] = R[R_FTOT + 1] * rs.inv_mass;
    R[R_ACC + 2] = R[R_FTOT + 2] * rs.inv_inertia;
    double nacc_x = R[R_ACC], nacc_y = R[R_ACC + 1];
    R[R_NODE_ACC] = nacc_x; R[R_NODE_ACC + 1] = nacc_y;   // node copies of Body::Apply_Newton (Code/Body.h:2038-2042)
    R[R_FDIR_NODE] = R[R_FDIR]; R[R_FDIR_NODE + 1] = R[R_FDIR + 1];   // copied to nodes by Code/Body.h:2005-2006
    double* dv = R + R_DRIVEN;
#pragma unroll
    for (int k = 0; k < 8; k++) dv[k] = dvt[k];
    double hx = R[R_VEL] + 0.5 * dt * R[R_ACC];
    double hy = R[R_VEL + 1] + 0.5 * dt * R[R_ACC + 1];
    double hr = R[R_VEL + 2] + 0.5 * dt * R[R_ACC + 2];
    R[R_WORK + W_BODY] += dt * (hx * R[R_FBODY] + hy * R[R_FBODY + 1] + hr * R[R_FBODY + 2]);
    R[R_WORK + W_DIRICHLET] += dt * (hx * R[R_FDIR] + hy * R[R_FDIR + 1] + hr * R[R_FDIR + 2]);
    R[R_WORK + W_NEUMANN] += dt * (hx * R[R_FNEU] + hy * R[R_FNEU + 1] + hr * R[R_FNEU + 2]);
    R[R_WORK + W_DAMPING] += dt * (hx * R[R_FDAMP] + hy * R[R_FDAMP + 1] + hr * R[R_FDAMP + 2]);
    double mass = rs.mass;
    if (dv[0] == 0.) {
        R[R_VEL] += dt * R[R_ACC];
        R[R_DISP] = R[R_DISP] + dt * R[R_VEL];
    } else if (dv[0] == 1.) {
        R[R_ACC] = dv[3]; R[R_VEL] = dv[2]; R[R_DISP] = dv[1];
        R[R_FDIR] = mass * R[R_ACC] - R[R_FTOT];
    }
    if (dv[4] == 0.) {
        R[R_VEL + 1] += dt * R[R_ACC + 1];
        R[R_DISP + 1] = R[R_DISP + 1] + dt * R[R_VEL + 1];
    } else if (dv[4] == 1.) {
        R[R_ACC + 1] = dv[7]; R[R_VEL + 1] = dv[6]; R[R_DISP + 1] = dv[5];
        R[R_FDIR + 1] = mass * R[R_ACC + 1] - R[R_FTOT + 1];
    }
    if (dv[0] == 0. && dv[4] == 0. && rs.periodicity != 2) {
        R[R_VEL + 2] += dt * R[R_ACC + 2];
        R[R_DISP + 2] = R[R_DISP + 2] + dt * R[R_VEL + 2];
    } else {
        R[R_VEL + 2] = 0.;
        R[R_DISP + 2] = 0.;
    }
    R[R_CUR] = R[R_XC] + R[R_DISP];
    R[R_CUR + 1] = R[R_YC] + R[R_DISP + 1];
    R[R_CUR + 2] = 0. + R[R_DISP + 2];
#pragma unroll
    for (int k = 0; k < RIGID_FIELDS; k++)
        if (k != R_WORK + W_CONTACT) Rg[k] = R[k];
}

// One rigid body by a workgroup: its record and driven DOFs (which do not
// depend on this step's forces) are staged in LDS while the resultants are
// formed (rigid_forces_block), then thread 0 steps the body.  sc: LDS
// scratch of RIGID_WG_LDS doubles.

// The driven DOFs at the end of the step of rigid body rigid_index (its
// record's previous values updated by rigid_driven_dofs), into dv[8].
__device__ inline void rigid_dv_end_of_step(const DevModel& m, int rigid_index, double* dv) {
    const int b = m.rigid_ids[rigid_index];
    const double* Rg = m.body_rigid + b * RIGID_FIELDS;
    for (int k = 0; k < 8; k++) dv[k] = Rg[R_DRIVEN + k];
    rigid_driven_dofs(m, b, *m.time + m.dt, dv);
}

// dv_pre: the driven DOFs formed earlier by rigid_dv_end_of_step (8 per
// rigid body), or nullptr to form them here
__device__ inline void rigid_body_wg(const DevModel& m, int rigid_index, double* sc, const double* dv_pre) {
    const int b = m.rigid_ids[rigid_index];
    double* s_rec = sc + 6 * BLOCK_THREADS + 8;
    double* s_dv = s_rec + RIGID_FIELDS;
    const double* Rg = m.body_rigid + b * RIGID_FIELDS;
    const RigidScalars rs = rigid_scalars(m, b);
    for (int k = threadIdx.x; k < RIGID_FIELDS; k += blockDim.x) s_rec[k] = Rg[k];
    if (dv_pre) {
        if (threadIdx.x < 8) s_dv[threadIdx.x] = dv_pre[8 * rigid_index + threadIdx.x];
    } else if (threadIdx.x == blockDim.x - 1) {
        double dv[8];
        rigid_dv_end_of_step(m, rigid_index, dv);
        for (int k = 0; k < 8; k++) s_dv[k] = dv[k];
    }
    const RigidForces F = rigid_forces_block(m, rigid_index, sc, dv_pre ? m.rbc_part : nullptr);   // its barriers publish s_rec, s_dv
    __syncthreads();
    if (threadIdx.x == 0) rigid_step_body(m, b, F, s_rec, s_dv, rs);
}

// the force launch's share of a rigid border's BC terms: item j is run
// j % BLOCK_THREADS of border rbc_border[j / BLOCK_THREADS]
__device__ inline void rigid_bc_item(const DevModel& m, int j) {
    const int k = j / BLOCK_THREADS, r = j - k * BLOCK_THREADS;
    const int br = m.rbc_border[k];
    const double* R = m.body_rigid + m.br_body[br] * RIGID_FIELDS;
    double p6[6];
    rigid_bc_run(m, br, r, R[R_CUR], R[R_CUR + 1], p6);
    double* o = m.rbc_part + (size_t)j * 6;
    for (int q = 0; q < 6; q++) o[q] = p6[q];
}

// rigid node kinematics and positions (Code/Body.h:2214-2226, 2281-2301);
// node force fields mirror the body totals as Code/Body.h:1999-2013 does.
// with_forces < 0: positions and velocities only, as the state a model is
// loaded with (Load_dynamic reads no node kinematics of rigid bodies)
__device__ inline void rigid_node_at(const DevModel& m, int i, const double* R, double c, double s,
                                     int with_forces, double Xi, double Yi) {
    double dx = Xi - R[R_XC], dy = Yi - R[R_YC];
    if (with_forces >= 0) {
        m.ux[i] = R[R_DISP] + dx * c - dy * s - dx;
        m.uy[i] = R[R_DISP + 1] + dx * s + dy * c - dy;
    }
    double xx = R[R_CUR] + dx * c - dy * s;
    double yy = R[R_CUR + 1] + dx * s + dy * c;
    m.xc[i] = xx;
    m.yc[i] = yy;
    double ex = xx - R[R_CUR], ey = yy - R[R_CUR + 1];
    m.vx[i] = -ey * R[R_VEL + 2] + R[R_VEL];
    m.vy[i] = ex * R[R_VEL + 2] + R[R_VEL + 1];
    if (with_forces > 0) {
        m.ax[i] = R[R_NODE_ACC]; m.ay[i] = R[R_NODE_ACC + 1];
        m.fcx[i] = R[R_FCON]; m.fcy[i] = R[R_FCON + 1];
        m.fbx[i] = R[R_FBODY]; m.fby[i] = R[R_FBODY + 1];
        m.fnx[i] = R[R_FNEU]; m.fny[i] = R[R_FNEU + 1];
        m.fkx[i] = R[R_FDAMP]; m.fky[i] = R[R_FDAMP + 1];
        m.fx[i] = R[R_FTOT]; m.fy[i] = R[R_FTOT + 1];
        // Body::Sum_up_forces copies the Dirichlet resultant before Apply_Euler
        // overwrites it for driven directions
        m.fdx[i] = R[R_FDIR_NODE]; m.fdy[i] = R[R_FDIR_NODE + 1];
    }
}

// what a rigid-body node needs besides its body's new state (loaded before
// the body has stepped)
struct RigidNodePre {
    int i, b;
    bool on;
    double X, Y;
};

__device__ inline RigidNodePre rigid_node_pre(const DevModel& m, int i) {
    RigidNodePre p;
    p.i = i;
    p.b = m.node_body[i];
    p.on = m.body_rigid_flag[p.b] && m.body_status[p.b] != 1;
    p.X = m.X[i];
    p.Y = m.Y[i];
    return p;
}

__device__ inline void rigid_node_apply(const DevModel& m, const RigidNodePre& p, int with_forces) {
    if (!p.on) return;
    const double* R = m.body_rigid + p.b * RIGID_FIELDS;
    double dr = R[R_DISP + 2] - 0.;
    rigid_node_at(m, p.i, R, cos(dr), sin(dr), with_forces, p.X, p.Y);
}

__device__ inline void rigid_node(const DevModel& m, int i, int with_forces) {
    rigid_node_apply(m, rigid_node_pre(m, i), with_forces);
}

// Initialize_contact_forces + Update_contacts + Update_contact_forces (records
// into buckets) for element i < nc, Update_bc_forces for deformable border
// node bc_bn[i - nc] after them
__device__ inline void contact_or_bc_item(const DevModel& m, int i, int nc, const int* bc_bn, int n_bc) {
    if (i < nc) {
        contact_item(m, i);
        return;
    }
    i -= nc;
    if (i < n_bc) bc_point_force_node(m, bc_bn[i]);
}

// Body::Update_borders (Code/Body.h:614-661) for the slots of one border node
__device__ inline void bn_lengths(const DevModel& m, int bn) {
    for (int q = m.bn_slot_ptr[bn]; q < m.bn_slot_ptr[bn + 1]; q++) {
        int s = m.bn_slots[q];
        m.sl_len[s] = slot_length(m, s);
    }
}


// The rest of the step per node block (a whole body): send of the contact
// point forces to the block's border nodes (bucket sums), Sum_up_forces +
// Apply_Newton + Apply_Euler with works for the owned nodes
// (Code/Body.h:1975-2068, Code/Node.h:212-245), Update_borders for the
// block's border slots (positions of the start of the step) and
// Update_kinematics / Update_current_positions for its border nodes.
// Workgroups past the node blocks move the rigid nodes, evaluate the BCs at
// t + dt and store the new time.
__device__ inline void node_block_finish(const DevModel& m, const BlockArrays& B, int blk) {
    __shared__ double red[BLOCK_THREADS / 64][6];
    const int t = threadIdx.x;
    const int body = B.blk_body[blk];
    const int nown = B.blk_nown[blk];
    const int* own = B.own_nodes + B.blk_own_off[blk];
    const int* bns = B.bn_list + B.blk_bn_off[blk];
    const int nbn = B.blk_bn_n[blk];        // 0 for blocks of split bodies
    if (m.body_status[body] == 1) {         // inactive body: no update (uniform over the block)
        for (int i = t; i < nbn; i += BLOCK_THREADS) bucket_drain_node(m, bns[i]);
        return;
    }
    for (int i = t; i < nbn; i += BLOCK_THREADS) bucket_sum_node(m, bns[i]);
    __syncthreads();
    double w[6] = {0., 0., 0., 0., 0., 0.};   // internal body dirichlet neumann damping alid
    double gx = m.gravity_x, gy = m.gravity_y;
    if (m.body_status[body] == 2) { gx = 0.; gy = 0.; }
    const double dt = m.dt;
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
        double apx = fx * m.imx[j], apy = fy * m.imy[j];
        double vpx = m.vpx[j], vpy = m.vpy[j];
        double hx = vpx + 0.5 * dt * apx, hy = vpy + 0.5 * dt * apy;
        w[0] += dt * (hx * ix + hy * iy);
        w[1] += dt * (hx * bx + hy * by);
        w[2] += dt * (hx * dx + hy * dy);
        w[3] += dt * (hx * nx + hy * ny);
        w[4] += dt * (hx * kx + hy * ky);
        w[5] += dt * (hx * lx + hy * ly);
        vpx = vpx + dt * apx;
        vpy = vpy + dt * apy;
        const double nux = m.upx[j] + dt * vpx, nuy = m.upy[j] + dt * vpy;
        if (!isfinite(nux) || !isfinite(nuy)) atomicOr(&m.status_flags[1], 1);
        m.upx[j] = nux;
        m.upy[j] = nuy;
        m.vpx[j] = vpx; m.vpy[j] = vpy;
        m.apx[j] = apx; m.apy[j] = apy;
    }
    // block work increments: fixed-order reduction (wavefront butterflies,
    // then the waves in order), accumulated with compensation across steps
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
            // Neumaier accumulation: acc[2s] sum, acc[2s+1] compensation
            int s = slot[q];
            double sum = acc[2 * s], x = v, nt = sum + x;
            acc[2 * s + 1] += fabs(sum) >= fabs(x) ? (sum - nt) + x : (x - nt) + sum;
            acc[2 * s] = nt;
        }
    }
    for (int i = t; i < nbn; i += BLOCK_THREADS) bn_lengths(m, bns[i]);
    __syncthreads();       // the block's new parameters and old-position lengths are in place
    for (int i = t; i < nbn; i += BLOCK_THREADS) kinematics_node(m, m.bn_node[bns[i]]);
}

// a slot's tributary-length inputs, loaded ahead of use (slot_length)
struct SlotPre {
    int j, n, simple;
    double x0, y0, x1, y1, x2, y2;
};

__device__ inline SlotPre slot_prefetch(const DevModel& m, int s) {
    SlotPre r;
    const int br = m.sl_border[s];
    r.j = m.sl_j[s];
    r.n = m.br_nslots[br];
    r.simple = m.br_periodicity[br] == 1;
    const double P = m.period;
    const int a0 = m.sl_n0[s], a1 = m.sl_n1[s], a2 = m.sl_n2[s];
    r.x0 = m.xc[a0] + P * m.sl_s0[s]; r.y0 = m.yc[a0];
    r.x1 = m.xc[a1] + P * m.sl_s1[s]; r.y1 = m.yc[a1];
    r.x2 = m.xc[a2] + P * m.sl_s2[s]; r.y2 = m.yc[a2];
    return r;
}

// the same from the slot's record (one load before the positions)
__device__ inline SlotPre slot_prefetch_rec(const DevModel& m, const BlockArrays& B, int s) {
    SlotPre r;
    const int4 a = B.sl_rec0[s], h = B.sl_rec1[s];
    r.j = a.w;
    r.n = h.w & ((1 << 30) - 1);
    r.simple = (h.w >> 30) & 1;
    const double P = m.period;
    r.x0 = m.xc[a.x] + P * h.x; r.y0 = m.yc[a.x];
    r.x1 = m.xc[a.y] + P * h.y; r.y1 = m.yc[a.y];
    r.x2 = m.xc[a.z] + P * h.z; r.y2 = m.yc[a.z];
    return r;
}

// slot_length from prefetched inputs (same expressions)
__device__ inline double slot_length_pre(const SlotPre& r) {
    double l21 = __dsqrt_rn((r.x2 - r.x1) * (r.x2 - r.x1) + (r.y2 - r.y1) * (r.y2 - r.y1));
    double l10 = __dsqrt_rn((r.x1 - r.x0) * (r.x1 - r.x0) + (r.y1 - r.y0) * (r.y1 - r.y0));
    if (r.j == 0 && r.simple) return l21 * 0.5;
    if (r.j == r.n - 1 && r.simple) return l10 * 0.5;
    return (l21 + l10) * 0.5;
}

// node_block_finish for a block that holds a whole body, as the two-launch
// step runs it: the same operations and summation orders, with the block's
// point forces and new nodal parameters staged in LDS so that the spreading
// to the nodes and the border-node kinematics read LDS instead of chasing
// global indices, and every load that does not depend on an earlier phase
// (node data, the slot stencils and old positions of the lengths, the
// DOF_TO_DISP rows of the kinematics) issued at the start.
//   fi     internal forces of the thread's owned nodes (o = t + q * BLOCK_THREADS)
//   sc     LDS scratch of at least 6 * (blk_bn_n + blk_nown) doubles
// Preconditions (checked when the blocks are built): the owned nodes are
// own[0] .. own[0] + nown - 1, every point-force spreading entry of an owned
// node names a border node of the block (bt_lbn is its index in the block's
// list) and every DOF_TO_DISP entry of a border node names an owned node.
constexpr int FIN_SLOTS = 2;     // slots of a border node with prefetched length inputs
constexpr int FIN_ROW = 12;      // DOF_TO_DISP entries of a border node prefetched
constexpr int FIN_BT = 6;        // point-force spreading entries of an owned node prefetched
__device__ inline void node_block_finish_lds(const DevModel& m, const BlockArrays& B, int blk,
                                             const double (&fi)[2][NODES_PER_THREAD], double* sc) {
    __shared__ double red[BLOCK_THREADS / 64][5];
    const int t = threadIdx.x;
    // the block's header (one load for everything the first loads need)
    const int4 h0 = B.blk_hdr[4 * blk], h1 = B.blk_hdr[4 * blk + 1], h3 = B.blk_hdr[4 * blk + 3];
    if (h1.z & BLK_INACTIVE) return;
    const int nown = h0.y;
    const int n0 = h0.x;
    const int* bns = B.bn_list + h3.z;
    const int nbn = h3.w;
    double* s_bnp = sc;                // [nbn][6]: contact, Dirichlet, Neumann point forces (x, y)
    double* s_up = sc + 6 * nbn;       // [nown][6]: new displacement, velocity, acceleration parameters

    // ---- loads issued up front ----
    // the owned nodes' inputs of the update
    double mx[NODES_PER_THREAD], my[NODES_PER_THREAD], imx[NODES_PER_THREAD], imy[NODES_PER_THREAD];
    double kx[NODES_PER_THREAD], ky[NODES_PER_THREAD], vpx[NODES_PER_THREAD], vpy[NODES_PER_THREAD];
    double upx[NODES_PER_THREAD], upy[NODES_PER_THREAD];
    int e0[NODES_PER_THREAD], e1[NODES_PER_THREAD];
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) {
        const int o = t + q * BLOCK_THREADS;
        const int j = n0 + (o < nown ? o : 0);
        mx[q] = m.mx[j]; my[q] = m.my[j]; imx[q] = m.imx[j]; imy[q] = m.imy[j];
        kx[q] = m.fkx[j]; ky[q] = m.fky[j]; vpx[q] = m.vpx[j]; vpy[q] = m.vpy[j];
        upx[q] = m.upx[j]; upy[q] = m.upy[j];
        e0[q] = m.bt_ptr[j]; e1[q] = o < nown ? m.bt_ptr[j + 1] : e0[q];
    }
    // the owned nodes' first point-force spreading entries
    int bl[NODES_PER_THREAD][FIN_BT];
    double bp[NODES_PER_THREAD][FIN_BT];
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++)
#pragma unroll
        for (int u = 0; u < FIN_BT; u++)
            if (e0[q] + u < e1[q]) { bl[q][u] = B.bt_lbn[e0[q] + u]; bp[q][u] = m.bt_phi[e0[q] + u]; }
    // the thread's first border node: its lengths' inputs (old positions)
    // and its kinematics row
    const bool mine = t < nbn;
    int bn_t = 0, sq0 = 0, sq1 = 0, knode = 0, k0 = 0, k1 = 0, slot_a = 0, slot_b = 0;
    SlotPre sp[FIN_SLOTS];
    int di[FIN_ROW];
    double dp[FIN_ROW];
    double kX = 0., kY = 0.;
    static_assert(FIN_SLOTS == 2, "the border-node record names two slots");
    if (mine) {
        const int4 r0 = B.fin_rec0[h3.z + t], r1 = B.fin_rec1[h3.z + t];
        const double2 rx = B.fin_recX[h3.z + t];
        bn_t = r0.x; knode = r0.y; k0 = r0.z; k1 = r0.w;
        sq0 = r1.x; sq1 = r1.y;
        slot_a = r1.z; slot_b = r1.w;
        kX = rx.x; kY = rx.y;
        if (sq0 < sq1) sp[0] = slot_prefetch_rec(m, B, r1.z);
        if (sq0 + 1 < sq1) sp[1] = slot_prefetch_rec(m, B, r1.w);
#pragma unroll
        for (int u = 0; u < FIN_ROW; u++)
            if (k0 + u < k1) { di[u] = m.d_idx[k0 + u]; dp[u] = m.d_phi[k0 + u]; }
    }

    // ---- sends: bucket sums, point forces staged ----
    for (int i = t; i < nbn; i += BLOCK_THREADS) {
        const int bn = i == t ? bn_t : bns[i];
        const double pdx = m.bn_pdx[bn], pdy = m.bn_pdy[bn], pnx = m.bn_pnx[bn], pny = m.bn_pny[bn];
        const double2 pc = bucket_sum_node(m, bn);
        double* q = s_bnp + 6 * i;
        q[0] = pc.x; q[1] = pc.y; q[2] = pdx; q[3] = pdy; q[4] = pnx; q[5] = pny;
    }
    __syncthreads();

    // ---- Sum_up_forces, Apply_Newton, Apply_Euler, works ----
    double w[5] = {0., 0., 0., 0., 0.};   // internal body dirichlet neumann damping
    double gx = m.gravity_x, gy = m.gravity_y;
    if (h1.z & BLK_NO_GRAVITY) { gx = 0.; gy = 0.; }
    const double dt = m.dt;
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) {
        const int o = t + q * BLOCK_THREADS;
        if (o >= nown) break;
        const int j = n0 + o;
        const double ix = fi[0][q], iy = fi[1][q];
        double cx = 0., cy = 0., dx = 0., dy = 0., nx = 0., ny = 0.;
#pragma unroll
        for (int u = 0; u < FIN_BT; u++) {
            if (e0[q] + u < e1[q]) {
                const double* p = s_bnp + 6 * bl[q][u];
                const double f = bp[q][u];
                cx += p[0] * f; cy += p[1] * f;
                dx += p[2] * f; dy += p[3] * f;
                nx += p[4] * f; ny += p[5] * f;
            }
        }
        for (int e = e0[q] + FIN_BT; e < e1[q]; e++) {
            const double* p = s_bnp + 6 * (int)B.bt_lbn[e];
            const double f = m.bt_phi[e];
            cx += p[0] * f; cy += p[1] * f;
            dx += p[2] * f; dy += p[3] * f;
            nx += p[4] * f; ny += p[5] * f;
        }
        const double bx = mx[q] * gx, by = my[q] * gy;
        const double fx = ix + cx + 0. + bx + dx + nx + kx[q] + 0.;
        const double fy = iy + cy + 0. + by + dy + ny + ky[q] + 0.;
        const double apx = fx * imx[q], apy = fy * imy[q];
        const double hx = vpx[q] + 0.5 * dt * apx, hy = vpy[q] + 0.5 * dt * apy;
        w[0] += dt * (hx * ix + hy * iy);
        w[1] += dt * (hx * bx + hy * by);
        w[2] += dt * (hx * dx + hy * dy);
        w[3] += dt * (hx * nx + hy * ny);
        w[4] += dt * (hx * kx[q] + hy * ky[q]);
        const double nvx = vpx[q] + dt * apx, nvy = vpy[q] + dt * apy;
        const double nux = upx[q] + dt * nvx, nuy = upy[q] + dt * nvy;
        if (!isfinite(nux) || !isfinite(nuy)) atomicOr(&m.status_flags[1], 1);
        m.upx[j] = nux; m.upy[j] = nuy;
        m.vpx[j] = nvx; m.vpy[j] = nvy;
        m.apx[j] = apx; m.apy[j] = apy;
        double* u = s_up + 6 * o;
        u[0] = nux; u[1] = nuy; u[2] = nvx; u[3] = nvy; u[4] = apx; u[5] = apy;
    }
#pragma unroll
    for (int q = 0; q < 5; q++) {
        double v = w[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((t & 63) == 0) red[t >> 6][q] = v;
    }

    // ---- Update_borders: lengths from the old positions ----
    if (mine) {
        if (sq0 < sq1) m.sl_len[slot_a] = slot_length_pre(sp[0]);
        if (sq0 + 1 < sq1) m.sl_len[slot_b] = slot_length_pre(sp[1]);
        for (int q = sq0 + FIN_SLOTS; q < sq1; q++) m.sl_len[m.bn_slots[q]] = slot_length(m, m.bn_slots[q]);
    }
    for (int i = t + BLOCK_THREADS; i < nbn; i += BLOCK_THREADS) bn_lengths(m, bns[i]);
    __syncthreads();       // works partials in place; lengths read the old positions before any moves
    if (t == 0) {
        const int slot[5] = {W_INTERNAL, W_BODY, W_DIRICHLET, W_NEUMANN, W_DAMPING};
        double* acc = B.blk_work + 14 * (size_t)blk;
#pragma unroll
        for (int q = 0; q < 5; q++) {
            double v = 0.;
            for (int wv = 0; wv < BLOCK_THREADS / 64; wv++) v += red[wv][q];
            int s = slot[q];
            double sum = acc[2 * s], x = v, nt = sum + x;
            acc[2 * s + 1] += fabs(sum) >= fabs(x) ? (sum - nt) + x : (x - nt) + sum;
            acc[2 * s] = nt;
        }
    }

    // ---- Update_kinematics and positions of the border nodes ----
    if (mine) {
        double dx = 0., dy = 0., vx = 0., vy = 0., ax = 0., ay = 0.;
#pragma unroll
        for (int u = 0; u < FIN_ROW; u++) {
            if (k0 + u < k1) {
                const double* up = s_up + 6 * (di[u] - n0);
                const double f = dp[u];
                dx += f * up[0]; dy += f * up[1];
                vx += f * up[2]; vy += f * up[3];
                ax += f * up[4]; ay += f * up[5];
            }
        }
        for (int k = k0 + FIN_ROW; k < k1; k++) {
            const double* up = s_up + 6 * (m.d_idx[k] - n0);
            const double f = m.d_phi[k];
            dx += f * up[0]; dy += f * up[1];
            vx += f * up[2]; vy += f * up[3];
            ax += f * up[4]; ay += f * up[5];
        }
        m.ux[knode] = dx; m.uy[knode] = dy; m.vx[knode] = vx; m.vy[knode] = vy; m.ax[knode] = ax; m.ay[knode] = ay;
        m.xc[knode] = kX + dx;
        m.yc[knode] = kY + dy;
    }
    for (int i = t + BLOCK_THREADS; i < nbn; i += BLOCK_THREADS) {
        const int node = m.bn_node[bns[i]];
        double dx = 0., dy = 0., vx = 0., vy = 0., ax = 0., ay = 0.;
        for (int k = m.d_ptr[node]; k < m.d_ptr[node + 1]; k++) {
            const double* up = s_up + 6 * (m.d_idx[k] - n0);
            const double f = m.d_phi[k];
            dx += f * up[0]; dy += f * up[1];
            vx += f * up[2]; vy += f * up[3];
            ax += f * up[4]; ay += f * up[5];
        }
        m.ux[node] = dx; m.uy[node] = dy; m.vx[node] = vx; m.vy[node] = vy; m.ax[node] = ax; m.ay[node] = ay;
        m.xc[node] = m.X[node] + dx;
        m.yc[node] = m.Y[node] + dy;
    }
}



}  // namespace mel
