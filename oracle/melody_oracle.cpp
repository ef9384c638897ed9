// ===========================================================================
// TEST INFRASTRUCTURE - the oracle.  CPU restatement (FP64, OpenMP) of the
// MELODY 2D explicit step that the HIP library is checked against.  Only the
// tests, the smoke check and the CPU-baseline leg of bench.py may load it;
// the product never does.
//
// It restates, over flat per-body arrays, the reference functions:
//   Euler_step                 Code/Solver.h:10-88
//   Solver_step (Adaptive_Euler, Mass_Scaling, Adaptive_Mass_Scaling)
//                              Code/Solver.h:98-486, with Body::Store /
//                              Restore, Apply_Euler_temporary,
//                              Update_kinematics_temporary, Compute_error,
//                              Compute_mass_scaling (Code/Body.h:2126-2264,
//                              2524-2659) and Code/Node.h:247-320
//   init sequence              Code/main.cpp:103-126, cadence :191-196
//   Body::Update_bc            Code/Body.h:335-395, Border::Update_bc Code/Border.h:120-321
//   Body::Update_borders       Code/Body.h:614-661
//   Body::Update_bc_forces     Code/Body.h:669-901
//   Body::Update_body_forces   Code/Body.h:909-929
//   Body::Update_internal_forces Code/Body.h:937-1049 (+ Code/Model.h:19-73)
//   Body::Update_damping_forces Code/Body.h:1057-1154
//   Body::Update_alid_forces   Code/Body.h:1163-1296 (+ Code/IO.h:1395-1584)
//   Body::Initialize_contact_forces Code/Body.h:1305-1333
//   Body::Update_contacts      Code/Body.h:1341-1634 (+ Code/Tools.h:106-219, 319-369)
//   Body::Update_contact_forces Code/Body.h:1642-1941 (+ Code/Model.h:81-335)
//   Body::Send_contact_forces  Code/Body.h:1948-1967
//   Sum_up_forces / Apply_Newton / Apply_Euler / Update_kinematics /
//   Update_current_positions   Code/Body.h:1975-2302, Code/Node.h:212-245
//   Update_proximity           Code/Proximity.h:8-473
//   Initialize_CZM             Code/Proximity.h:481-541
//   Update_contact_pressures   Code/Proximity.h:1014-1069
//   Body::Update_material (masses) Code/Body.h:2708-2783, 2877-2905
// keeping the reference's floating-point expression order and summation
// order (per-region accumulators, serial master-force sends), so that on the
// same input it reproduces the reference bit for bit.  Compiled with
// -ffp-contract=off.
//
// One documented deviation: when both projections of Update_contacts return
// flag +-3, the reference reuses shape values left over from the previous
// contact element of the loop (Code/Body.h:1349 locals); this restatement
// keeps the element's own stored shape values instead, as the device does.
// ===========================================================================
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>

#include "dump_format.hpp"
#include "model_io.hpp"

using namespace mel;

namespace {

struct State {
    // per node
    std::vector<double> xc, yc, ux, uy, vx, vy, ax, ay, upx, upy, vpx, vpy, apx, apy;
    std::vector<double> fix, fiy, fcx, fcy, fbx, fby, fdx, fdy, fnx, fny, fkx, fky, fx, fy, fax, fay;
    std::vector<double> regx, regy;           // n_regions x n_nodes
    std::vector<double> mx, my, imx, imy;
    // rigid body
    double cur[3] = {0, 0, 0}, disp[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, acc[3] = {0, 0, 0};
    double rfc[3] = {0, 0, 0}, rfb[3] = {0, 0, 0}, rfd[3] = {0, 0, 0}, rfn[3] = {0, 0, 0}, rfk[3] = {0, 0, 0},
           rf[3] = {0, 0, 0};
    double driven[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double mass = 0, inverse_mass = 0, inertia = 0, inverse_inertia = 0;
    double works[7] = {0, 0, 0, 0, 0, 0, 0};
    // per border
    std::vector<std::vector<double>> length, xcp, ycp;
    std::vector<std::vector<double>> xbp, ybp;    // per-slot BC pressure (Border::x_bc_pressure)
    std::vector<double> xbv, ybv, xbvel, ybvel;    // current bc value / velocity
    std::vector<double> pxmin, pxmax, pymin, pymax;
    std::vector<int> shmin, shmax;
    std::vector<int> neighbours;
    int nb_neighbours = 0;
    std::vector<ContactElement> contacts;
    std::vector<std::array<double, 3>> send_f;
    std::vector<std::array<int, 2>> send_to;
    // adaptive schemes: per node the temporary displacement parameters and
    // displacements, errors, inverse distance estimator, mass-scaling
    // factors and their increments, scaled masses (Code/Node.h:72-99);
    // per body the rigid temporaries, error and mass-scaling scalars
    std::vector<double> uptx, upty, utx, uty, xerr, yerr, enorm, ide, fsx, fsy, dfsx, dfsy, fsn, msm;
    double disp_t[3] = {0, 0, 0}, vel_t[3] = {0, 0, 0};
    double total_error = 0, max_error = 0;
    int node_for_max = 0;
    double mass_ms = 0, factor_ms = 1, delta_factor_ms = 0, max_factor_ms = 0;
    double damage = 0;                        // Body::damage (Code/Body.h:2667-2700)
};

struct Sim {
    Model m;
    std::vector<State> s;
    double time = 0;
    int iteration = 0;
    double next_contact_update = 0;
    std::vector<double> monitor_row;          // Monitoring's row of the last step
};

// --------------------------------------------------------------------------
// geometry (Code/Tools.h)
// --------------------------------------------------------------------------

// Linear branch of Closest_2_segment, Code/Tools.h:106-219.  Outputs are
// only written where the reference writes them.
void closest_linear(double Xs, double Ys, double X0, double Y0, double X1, double Y1, double X2, double Y2,
                    double X3, double Y3, double& Gap, double& Xsi, double& Xnorm, double& Ynorm, double& Xtan,
                    double& Ytan, double& N0, double& N1, double& N2, double& N3, int& flag) {
    double il = std::pow((X2 - X1) * (X2 - X1) + (Y2 - Y1) * (Y2 - Y1), -0.5);
    double ilm = std::pow((X1 - X0) * (X1 - X0) + (Y1 - Y0) * (Y1 - Y0), -0.5);
    double ilp = std::pow((X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2), -0.5);
    Xtan = (X2 - X1) * il;
    Ytan = (Y2 - Y1) * il;
    Xnorm = Ytan;
    Ynorm = -Xtan;
    double tmx = (X1 - X0) * ilm, tmy = (Y1 - Y0) * ilm;
    double tpx = (X3 - X2) * ilp, tpy = (Y3 - Y2) * ilp;
    bool after_start = ((Xs - X1) * (tmx + Xtan) + (Ys - Y1) * (tmy + Ytan)) *
                           (Xtan * (tmx + Xtan) + Ytan * (tmy + Ytan)) >= 0.;
    if (!after_start) {
        flag = -3;
        Xsi = -1.;
        return;
    }
    bool before_end = ((Xs - X2) * (Xtan + tpx) + (Ys - Y2) * (Ytan + tpy)) *
                          (Xtan * (Xtan + tpx) + Ytan * (Ytan + tpy)) <= 0.;
    if (!before_end) {
        flag = 3;
        Xsi = 1.;
        return;
    }
    Xsi = -1. + 2. * ((X2 - X1) * (Xs - X1) + (Y2 - Y1) * (Ys - Y1)) * il * il;
    double xcl, ycl;
    if ((Xsi < 1.) && (Xsi > -1.)) {
        Xnorm = Ytan; Ynorm = -Xtan;
        xcl = 0.5 * (X1 + X2) + Xsi * 0.5 * (X2 - X1);
        ycl = 0.5 * (Y1 + Y2) + Xsi * 0.5 * (Y2 - Y1);
        Gap = Xnorm * (Xs - xcl) + Ynorm * (Ys - ycl);
        N0 = 0.; N1 = 0.5 * (1. - Xsi); N2 = 0.5 * (1. + Xsi); N3 = 0.;
        flag = 0;
    } else if (Xsi == -1.) {
        Xnorm = Ytan; Ynorm = -Xtan;
        Gap = Xnorm * (Xs - X1) + Ynorm * (Ys - Y1);
        N0 = 0.; N1 = 1.; N2 = 0.; N3 = 0.;
        flag = -1;
    } else if (Xsi < -1.) {
        double idm = std::pow((Xs - X1) * (Xs - X1) + (Ys - Y1) * (Ys - Y1), -0.5);
        Xnorm = (Xs - X1) * idm;
        Ynorm = (Ys - Y1) * idm;
        if (-Ynorm * (Xtan + tmx) + Xnorm * (Ytan + tmy) < 0.) { Xnorm = -Xnorm; Ynorm = -Ynorm; }
        Xtan = -Ynorm; Ytan = Xnorm;
        Gap = Xnorm * (Xs - X1) + Ynorm * (Ys - Y1);
        N0 = 0.; N1 = 1.; N2 = 0.; N3 = 0.;
        Xsi = -1.;
        flag = -2;
    } else if (Xsi == 1.) {
        Xnorm = Ytan; Ynorm = -Xtan;
        Gap = Xnorm * (Xs - X2) + Ynorm * (Ys - Y2);
        N0 = 0.; N1 = 0.; N2 = 1.; N3 = 0.;
        flag = 1;
    } else if (Xsi > 1.) {
        double idp = std::pow((Xs - X2) * (Xs - X2) + (Ys - Y2) * (Ys - Y2), -0.5);
        Xnorm = (Xs - X2) * idp;
        Ynorm = (Ys - Y2) * idp;
        if (-Ynorm * (Xtan + tpx) + Xnorm * (Ytan + tpy) < 0.) { Xnorm = -Xnorm; Ynorm = -Ynorm; }
        Xtan = -Ynorm; Ytan = Xnorm;
        Gap = Xnorm * (Xs - X2) + Ynorm * (Ys - Y2);
        N0 = 0.; N1 = 0.; N2 = 1.; N3 = 0.;
        Xsi = 1.;
        flag = 2;
    }
}

// Closest_2_segment_self, Code/Tools.h:319-369: the interior case writes the
// projection; the flag always ends at -3.
void closest_self(double Xs, double Ys, double X1, double Y1, double X2, double Y2, double& Gap, double& Xsi,
                  double& Xnorm, double& Ynorm, double& Xtan, double& Ytan, double& N0, double& N1, double& N2,
                  double& N3, int& flag) {
    double il = std::pow((X2 - X1) * (X2 - X1) + (Y2 - Y1) * (Y2 - Y1), -0.5);
    Xtan = (X2 - X1) * il;
    Ytan = (Y2 - Y1) * il;
    if (((Xs - X1) * Xtan + (Ys - Y1) * Ytan) * (Xtan * Xtan + Ytan * Ytan) >= 0.) {
        if (((Xs - X2) * Xtan + (Ys - Y2) * Ytan) * (Xtan * Xtan + Ytan * Ytan) <= 0.) {
            Xsi = -1. + 2. * ((X2 - X1) * (Xs - X1) + (Y2 - Y1) * (Ys - Y1)) * il * il;
            if ((Xsi < 1.) && (Xsi > -1.)) {
                Xnorm = Ytan; Ynorm = -Xtan;
                double xcl = 0.5 * (X1 + X2) + Xsi * 0.5 * (X2 - X1);
                double ycl = 0.5 * (Y1 + Y2) + Xsi * 0.5 * (Y2 - Y1);
                Gap = Xnorm * (Xs - xcl) + Ynorm * (Ys - ycl);
                N0 = 0.; N1 = 0.5 * (1. - Xsi); N2 = 0.5 * (1. + Xsi); N3 = 0.;
            }
        } else {
            Xsi = 1.;
        }
    } else {
        Xsi = -1.;
    }
    flag = -3;
}

// Detect_Segment_Proximity, Code/Proximity.h:8-27
int detect_segment(double Xs, double Ys, double X1, double Y1, double X2, double Y2, double d) {
    d = d * d;
    if ((Xs - X1) * (Xs - X1) + (Ys - Y1) * (Ys - Y1) < d) return 1;
    if ((Xs - X2) * (Xs - X2) + (Ys - Y2) * (Ys - Y2) < d) return 1;
    double dx = X2 - X1, dy = Y2 - Y1;
    double a = dx * dx + dy * dy;
    double xsi = -1. + 2. * (dx * (Xs - X1) + dy * (Ys - Y1)) / a;
    if ((xsi < 1.) && (xsi > -1.)) {
        double b = dy * (Xs - 0.5 * (X1 + X2) - xsi * 0.5 * dx) - dx * (Ys - 0.5 * (Y1 + Y2) - xsi * 0.5 * dy);
        if (b * b < a * d) return 1;
    }
    return 0;
}

// --------------------------------------------------------------------------
// boundary conditions
// --------------------------------------------------------------------------

// Border::Update_bc, Code/Border.h:120-221, for one direction
void border_bc_value(const BorderBC& bc, double t, double& value, double& velocity) {
    if (bc.type == BC_DRIVEN) return;
    if (bc.type == BC_SOFT) {
        value = 0.;
        for (size_t n = 0; n + 1 < bc.hist.t.size(); n++) {
            double t0 = bc.hist.t[n], p0 = bc.hist.v[n], t1 = bc.hist.t[n + 1], p1 = bc.hist.v[n + 1];
            if (t >= t1) value = value + (t1 - t0) * (p1 + p0) * 0.5;
            else {
                value = value + (t - t0) * (p0 + (p1 - p0) / (t1 - t0) * ((t - t0) * 0.5));
                velocity = p0 + (p1 - p0) / (t1 - t0) * (t - t0);
                break;
            }
        }
    } else if (bc.type == BC_ORIENTED || bc.type == BC_FOLLOWING) {
        for (size_t n = 0; n + 1 < bc.hist.t.size(); n++) {
            double t0 = bc.hist.t[n], p0 = bc.hist.v[n], t1 = bc.hist.t[n + 1], p1 = bc.hist.v[n + 1];
            if ((t0 <= t) & (t <= t1)) {
                value = p0 + (p1 - p0) / (t1 - t0) * (t - t0);
                break;
            }
        }
    } else {
        value = 0.;
    }
}

// driven displacement / velocity increments of Body::Update_bc, Code/Body.h:340-393
void driven_dof(const History& h, double t, double* d) {
    d[0] = 1.;
    d[1] = 0.;
    for (size_t n = 0; n + 1 < h.t.size(); n++) {
        double t0 = h.t[n], p0 = h.v[n], t1 = h.t[n + 1], p1 = h.v[n + 1];
        if (t >= t1) d[1] += (t1 - t0) * (p1 + p0) * 0.5;
        else {
            d[1] += (t - t0) * (p0 + (p1 - p0) / (t1 - t0) * ((t - t0) * 0.5));
            d[2] = (t - t0) * (p1 - p0) / (t1 - t0);
            d[3] = (p1 - p0) / (t1 - t0);
            break;
        }
    }
}

void update_bc(Sim& S, int bi, double t) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    for (size_t n = 0; n < b.borders.size(); n++) {
        border_bc_value(b.borders[n].xbc, t, s.xbv[n], s.xbvel[n]);
        border_bc_value(b.borders[n].ybc, t, s.ybv[n], s.ybvel[n]);
        if (b.borders[n].xbc.type == BC_DRIVEN) driven_dof(b.borders[n].xbc.hist, t, s.driven);
        if (b.borders[n].ybc.type == BC_DRIVEN) driven_dof(b.borders[n].ybc.hist, t, s.driven + 4);
    }
}

void update_borders(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    if (b.rigid) return;
    State& s = S.s[bi];
    double period = S.m.xmax_period - S.m.xmin_period;
    for (size_t i = 0; i < b.borders.size(); i++) {
        const Border& B = b.borders[i];
        int n = (int)B.nodes.size();
        for (int j = 0; j < n; j++) {
            double x0 = s.xc[B.node0[j]] + period * B.shift0[j], y0 = s.yc[B.node0[j]];
            double x1 = s.xc[B.node1[j]] + period * B.shift1[j], y1 = s.yc[B.node1[j]];
            double x2 = s.xc[B.node2[j]] + period * B.shift2[j], y2 = s.yc[B.node2[j]];
            double l21 = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
            double l10 = std::sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
            double L;
            if (j == 0 && B.periodicity == PER_SIMPLE) L = l21 * 0.5;
            else if (j == n - 1 && B.periodicity == PER_SIMPLE) L = l10 * 0.5;
            else L = (l21 + l10) * 0.5;
            s.length[i][j] = L;
        }
    }
}

void update_bc_forces(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    if (!b.rigid) {
        std::fill(s.fnx.begin(), s.fnx.end(), 0.);
        std::fill(s.fny.begin(), s.fny.end(), 0.);
        std::fill(s.fdx.begin(), s.fdx.end(), 0.);
        std::fill(s.fdy.begin(), s.fdy.end(), 0.);
        for (size_t i = 0; i < b.borders.size(); i++) {
            const Border& B = b.borders[i];
            int xt = B.xbc.type, yt = B.ybc.type;
            std::fill(s.xbp[i].begin(), s.xbp[i].end(), 0.);
            std::fill(s.ybp[i].begin(), s.ybp[i].end(), 0.);
            if (xt == BC_NONE && yt == BC_NONE) continue;
            int nn = (int)B.nodes.size();
            for (int n = 0; n < nn - 1; n++) {
                int node = B.nodes[n];
                double L = s.length[i][n];
                double px = 0., py = 0.;
                auto follow = [&](double& tx, double& ty, double& tpx, double& tpy) {
                    int a = B.node0[n], c = B.node2[n];
                    double x0 = b.X[a], y0 = b.Y[a], x2 = b.X[c], y2 = b.Y[c];
                    double xp0 = s.xc[a], yp0 = s.yc[a], xp2 = s.xc[c], yp2 = s.yc[c];
                    double invl = std::pow((x2 - x0) * (x2 - x0) + (y2 - y0) * (y2 - y0), -0.5);
                    double invlp = std::pow((xp2 - xp0) * (xp2 - xp0) + (yp2 - yp0) * (yp2 - yp0), -0.5);
                    tx = (x2 - x0) * invl; ty = (y2 - y0) * invl;
                    tpx = (xp2 - xp0) * invlp; tpy = (yp2 - yp0) * invlp;
                };
                if (xt == BC_ORIENTED) px = s.xbv[i];
                else if (xt == BC_FOLLOWING) {
                    double tx, ty, tpx, tpy;
                    follow(tx, ty, tpx, tpy);
                    px += s.xbv[i] * (tx * tpx + ty * tpy);
                    py += s.xbv[i] * (tx * tpy - ty * tpx);
                } else if (xt == BC_SOFT) {
                    px = B.xbc.stiffness * (s.xbv[i] - s.ux[node]) - B.xbc.damping * (s.xbvel[i] - s.vx[node]);
                }
                if (yt == BC_ORIENTED) py = s.ybv[i];
                else if (yt == BC_FOLLOWING) {
                    double tx, ty, tpx, tpy;
                    follow(tx, ty, tpx, tpy);
                    px += s.ybv[i] * (ty * tpx - tx * tpy);
                    py += s.ybv[i] * (ty * tpy + tx * tpx);
                } else if (yt == BC_SOFT) {
                    py = B.ybc.stiffness * (s.ybv[i] - s.uy[node]) - B.ybc.damping * (s.ybvel[i] - s.vy[node]);
                }
                s.xbp[i][n] = px;
                s.ybp[i][n] = py;
                for (int k = b.dptr[node]; k < b.dptr[node + 1]; k++) {
                    double sf = b.dphi[k];
                    double fxn = 0., fyn = 0., fxd = 0., fyd = 0.;
                    if (xt == BC_ORIENTED) fxn = px * L * sf;
                    else if (xt == BC_FOLLOWING) { fxn = px * L * sf; fyn = py * L * sf; }
                    else if (xt == BC_SOFT) fxd = px * L * sf;
                    if (yt == BC_ORIENTED) fyn = py * L * sf;
                    else if (yt == BC_FOLLOWING) { fxn = px * L * sf; fyn = py * L * sf; }
                    else if (yt == BC_SOFT) fyd = py * L * sf;
                    int j = b.didx[k];
                    s.fnx[j] += fxn; s.fny[j] += fyn; s.fdx[j] += fxd; s.fdy[j] += fyd;
                }
            }
        }
    } else {
        for (int q = 0; q < 3; q++) { s.rfn[q] = 0.; s.rfd[q] = 0.; }
        for (size_t i = 0; i < b.borders.size(); i++) {
            const Border& B = b.borders[i];
            int xt = B.xbc.type, yt = B.ybc.type;
            std::fill(s.xbp[i].begin(), s.xbp[i].end(), 0.);
            std::fill(s.ybp[i].begin(), s.ybp[i].end(), 0.);
            if (xt == BC_NONE && yt == BC_NONE) continue;
            int nn = (int)B.nodes.size();
            for (int n = 0; n < nn; n++) {
                int node = B.nodes[n];
                double L = s.length[i][n];
                double fxn = 0., fyn = 0., fxd = 0., fyd = 0.;
                double px = 0., py = 0.;
                if (xt == BC_ORIENTED || xt == BC_FOLLOWING) { px = s.xbv[i]; fxn = px * L; }
                else if (xt == BC_SOFT) {
                    px = B.xbc.stiffness * (s.xbv[i] - s.ux[node]) + B.xbc.damping * (s.xbvel[i] - s.vx[node]);
                    fxd = px * L;
                }
                if (yt == BC_ORIENTED || yt == BC_FOLLOWING) { py = s.ybv[i]; fyn = py * L; }
                else if (yt == BC_SOFT) {
                    py = B.ybc.stiffness * (s.ybv[i] - s.uy[node]) + B.ybc.damping * (s.ybvel[i] - s.vy[node]);
                    fyd = py * L;
                }
                s.xbp[i][n] = px;
                s.ybp[i][n] = py;
                double dx = s.xc[node] - s.cur[0], dy = s.yc[node] - s.cur[1];
                s.rfn[0] += fxn; s.rfn[1] += fyn; s.rfn[2] += -fxn * dy + fyn * dx;
                s.rfd[0] += fxd; s.rfd[1] += fyd; s.rfd[2] += -fxd * dy + fyd * dx;
            }
        }
    }
}

void update_body_forces(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    double gx = S.m.gravity_x, gy = S.m.gravity_y;
    if (b.status == ST_NOGRAVITY) { gx = 0.; gy = 0.; }
    if (!b.rigid) {
        for (int i = 0; i < b.n_nodes; i++) { s.fbx[i] = s.mx[i] * gx; s.fby[i] = s.my[i] * gy; }
    } else {
        s.rfb[0] = s.mass * gx;
        s.rfb[1] = s.mass * gy;
    }
}

void update_damping_forces(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    if (b.rigid) return;
    State& s = S.s[bi];
    std::fill(s.fkx.begin(), s.fkx.end(), 0.);
    std::fill(s.fky.begin(), s.fky.end(), 0.);
    for (size_t i = 0; i < b.mat_damping.size(); i++) {
        int dof0 = b.mat_dof[2 * i], dof1 = b.mat_dof[2 * i + 1];
        double d = b.mat_damping[i];
        int n0 = dof0 / 2, r0 = dof0 - 2 * n0, n1 = dof1 / 2, r1 = dof1 - 2 * n1;
        if (r0 == 0 && r1 == 0) {
            s.fkx[n0] -= s.vpx[n1] * d;
            if (dof0 != dof1) s.fkx[n1] -= s.vpx[n0] * d;
        } else if (r0 == 0 && r1 == 1) {
            s.fkx[n0] -= s.vpy[n1] * d;
            s.fky[n1] -= s.vpx[n0] * d;
        } else if (r0 == 1 && r1 == 0) {
            s.fky[n0] -= s.vpx[n1] * d;
            s.fkx[n1] -= s.vpy[n0] * d;
        } else {
            s.fky[n0] -= s.vpy[n1] * d;
            if (dof0 != dof1) s.fky[n1] -= s.vpy[n0] * d;
        }
    }
}

// --------------------------------------------------------------------------
// internal forces (Code/Body.h:937-1049, Code/Model.h:19-73)
// --------------------------------------------------------------------------
// Body::Update_alid_forces (Code/Body.h:1163-1296): mass-proportional
// terms from the physical velocities, then the stiffness-proportional terms
// of the matrix entries in entry order, relative to the layers' imposed
// velocity at time t
void update_alid_forces(Sim& S, int bi, double t) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    if (b.rigid || !b.alid.on) return;
    auto pv = [&](const History& h, bool& on) {
        on = !h.t.empty();
        for (size_t n = 0; n + 1 < h.t.size(); n++) {
            double t0 = h.t[n], p0 = h.v[n], t1 = h.t[n + 1], p1 = h.v[n + 1];
            if ((t0 <= t) & (t <= t1)) return p0 + (p1 - p0) / (t1 - t0) * (t - t0);
        }
        return 0.;
    };
    bool onx, ony;
    const double pvx = pv(b.alid.xh, onx), pvy = pv(b.alid.yh, ony);
    for (int i = 0; i < b.n_nodes; i++) {
        const double beta = b.alid_beta[i];
        s.fax[i] = onx ? -beta * s.mx[i] * (s.vx[i] - pvx) : 0.;
        s.fay[i] = ony ? -beta * s.my[i] * (s.vy[i] - pvy) : 0.;
    }
    for (size_t e = 0; e < b.mat_stiffness.size(); e++) {
        const int dof0 = b.mat_dof[2 * e], dof1 = b.mat_dof[2 * e + 1];
        const int n0 = dof0 / 2, r0 = dof0 - n0 * 2, n1 = dof1 / 2, r1 = dof1 - n1 * 2;
        const double d = b.mat_stiffness[e] / (1. / b.alid_alpha[n0] + 1. / b.alid_alpha[n1]);
        const double vp0[2] = {s.vpx[n0] - pvx, s.vpy[n0] - pvy}, vp1[2] = {s.vpx[n1] - pvx, s.vpy[n1] - pvy};
        double* f0 = r0 == 0 ? &s.fax[n0] : &s.fay[n0];
        double* f1 = r1 == 0 ? &s.fax[n1] : &s.fay[n1];
        const bool on0 = r0 == 0 ? onx : ony, on1 = r1 == 0 ? onx : ony;
        if (on0) *f0 += -vp1[r1] * d;
        if (dof0 != dof1 && on1) *f1 += -vp0[r0] * d;
    }
}

void update_internal_forces(Sim& S, int bi, int region) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    const Material& mat = S.m.materials[b.material];
    int G = b.n_gauss, R = b.n_regions;
    int per = (int)std::ceil((double)G / (double)R);
    double* rx = &s.regx[(size_t)region * b.n_nodes];
    double* ry = &s.regy[(size_t)region * b.n_nodes];
    for (int g = region * per; g < std::min(G, (region + 1) * per); g++) {
        int k0 = b.gptr[g], k1 = b.gptr[g + 1];
        double c0;
        if (mat.type == MAT_ELASTIC_LINEAR) {
            double E11 = 0., E12 = 0., E22 = 0.;
            for (int k = k0; k < k1; k++) {
                int j = b.gidx[k];
                E11 += b.gphix[k] * s.upx[j];
                E12 += b.gphix[k] * s.upy[j] + b.gphiy[k] * s.upx[j];
                E22 += b.gphiy[k] * s.upy[j];
            }
            E12 = 0.5 * E12;
            double Mu = mat.p[3], Lambda = mat.p[4];
            double J = 1. + E11 + E22;
            double S33 = Lambda * (J - 1.);
            double Mu2 = 2 * Mu;
            double S11 = S33 + Mu2 * E11, S22 = S33 + Mu2 * E22, S12 = Mu2 * E12;
            c0 = b.gw[g] * b.gjac[g];
            for (int k = k0; k < k1; k++) {
                int j = b.gidx[k];
                rx[j] -= c0 * (b.gphix[k] * S11 + b.gphiy[k] * S12);
                ry[j] -= c0 * (b.gphiy[k] * S22 + b.gphix[k] * S12);
            }
        } else {
            double F11 = 1., F12 = 0., F21 = 0., F22 = 1.;
            for (int k = k0; k < k1; k++) {
                int j = b.gidx[k];
                F11 += b.gphix[k] * s.upx[j];
                F12 += b.gphiy[k] * s.upx[j];
                F21 += b.gphix[k] * s.upy[j];
                F22 += b.gphiy[k] * s.upy[j];
            }
            double Mu = mat.p[3], Kappa = mat.p[4];
            double F11F12 = F11 * F11 + F12 * F12;
            double F21F22 = F21 * F21 + F22 * F22;
            double J = F11 * F22 - F12 * F21;
            double invJ = 1. / J;
            double K = Kappa * (J - 1.);
            double MuJ53 = Mu * std::pow(J, -1.666666666667);
            double trB3 = 0.333333333333 * (F11F12 + F21F22 + 1.);
            double sig11 = K + MuJ53 * (F11F12 - trB3);
            double sig22 = K + MuJ53 * (F21F22 - trB3);
            double sig12 = MuJ53 * (F11 * F21 + F12 * F22);
            double S11 = (F12 * (F12 * sig22 - F22 * sig12) - F22 * (F12 * sig12 - F22 * sig11)) * invJ;
            double S22 = (F11 * (F11 * sig22 - F21 * sig12) - F21 * (F11 * sig12 - F21 * sig11)) * invJ;
            double S12 = (F21 * (F12 * sig12 - F22 * sig11) - F11 * (F12 * sig22 - F22 * sig12)) * invJ;
            c0 = b.gw[g] * b.gjac[g];
            for (int k = k0; k < k1; k++) {
                int j = b.gidx[k];
                rx[j] -= c0 * (b.gphix[k] * (F11 * S11 + F12 * S12) + b.gphiy[k] * (F12 * S22 + F11 * S12));
                ry[j] -= c0 * (b.gphix[k] * (F21 * S11 + F22 * S12) + b.gphiy[k] * (F22 * S22 + F21 * S12));
            }
        }
    }
}

// --------------------------------------------------------------------------
// contacts
// --------------------------------------------------------------------------
void initialize_contact_forces(Sim& S, int bi) {
    State& s = S.s[bi];
    if (!S.m.bodies[bi].rigid) {
        std::fill(s.fcx.begin(), s.fcx.end(), 0.);
        std::fill(s.fcy.begin(), s.fcy.end(), 0.);
    } else {
        s.rfc[0] = s.rfc[1] = s.rfc[2] = 0.;
    }
    s.send_f.clear();
    s.send_to.clear();
}

void stencil(const Border& B, int seg, int* nd, int* sh) {
    nd[0] = B.node0[seg]; nd[1] = B.node1[seg]; nd[2] = B.node2[seg]; nd[3] = B.node3[seg];
    sh[0] = B.shift0[seg]; sh[1] = B.shift1[seg]; sh[2] = B.shift2[seg]; sh[3] = B.shift3[seg];
}

void update_contacts(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    double period = S.m.xmax_period - S.m.xmin_period;
    for (ContactElement& c : s.contacts) {
        int bodyM = c.bodyM, borderM = c.borderM, shiftM = c.shiftM, segmentM = c.segmentM;
        const Body& BM = S.m.bodies[bodyM];
        const State& SM = S.s[bodyM];
        int per = BM.borders[borderM].periodicity;
        double gapn = c.gapn, gapt = c.gapt, xsi = c.xsi, xn = c.xnorm, yn = c.ynorm, xt = c.xtan, yt = c.ytan;
        double N[4] = {c.shapeM[0], c.shapeM[1], c.shapeM[2], c.shapeM[3]};
        int nd[4], sh[4];
        stencil(BM.borders[borderM], segmentM, nd, sh);
        double xsi_prev = xsi;
        double xs = s.xc[c.nodeS], ys = s.yc[c.nodeS];
        double X[4], Y[4];
        auto coords = [&]() {
            for (int k = 0; k < 4; k++) { X[k] = SM.xc[nd[k]] + period * (shiftM + sh[k]); Y[k] = SM.yc[nd[k]]; }
        };
        int flag = 0;
        auto project = [&]() {
            if (bi == bodyM && (nd[0] == c.nodeS || nd[3] == c.nodeS))
                closest_self(xs, ys, X[1], Y[1], X[2], Y[2], gapn, xsi, xn, yn, xt, yt, N[0], N[1], N[2], N[3], flag);
            else
                closest_linear(xs, ys, X[0], Y[0], X[1], Y[1], X[2], Y[2], X[3], Y[3], gapn, xsi, xn, yn, xt, yt,
                               N[0], N[1], N[2], N[3], flag);
        };
        coords();
        project();
        if (flag == -2 || flag == -3) {
            gapt = gapt + 0.5 * (-1. - xsi_prev) * std::sqrt((X[2] - X[1]) * (X[2] - X[1]) + (Y[2] - Y[1]) * (Y[2] - Y[1]));
            if (per == PER_SIMPLE) {
                if (segmentM == 0) {
                    borderM = BM.borders[borderM].before;
                    segmentM = (int)BM.borders[borderM].nodes.size() - 2;
                } else segmentM--;
            } else if (per == PER_PERIODIC) {
                if (segmentM == 0) {
                    if (borderM == 0) shiftM--;
                    else if (borderM == 1) shiftM++;
                    segmentM = (int)BM.borders[borderM].nodes.size() - 2;
                } else segmentM--;
            } else {
                if (segmentM == 0) segmentM = (int)BM.borders[borderM].nodes.size() - 2;
                else segmentM--;
            }
            stencil(BM.borders[borderM], segmentM, nd, sh);
            coords();
            xsi = 1.;
            project();
            gapt = gapt + 0.5 * (-1. + xsi) * std::sqrt((X[2] - X[1]) * (X[2] - X[1]) + (Y[2] - Y[1]) * (Y[2] - Y[1]));
            c.bodyM = bodyM; c.borderM = borderM; c.shiftM = shiftM; c.segmentM = segmentM;
        } else if (flag == 2 || flag == 3) {
            gapt = gapt + 0.5 * (-1. + xsi_prev) * std::sqrt((X[2] - X[1]) * (X[2] - X[1]) + (Y[2] - Y[1]) * (Y[2] - Y[1]));
            int last = (int)BM.borders[borderM].nodes.size() - 2;
            if (per == PER_SIMPLE) {
                if (segmentM == last) { borderM = BM.borders[borderM].after; segmentM = 0; }
                else segmentM++;
            } else if (per == PER_PERIODIC) {
                if (segmentM == last) {
                    if (borderM == 0) shiftM++;
                    else if (borderM == 1) shiftM--;
                    segmentM = 0;
                } else segmentM++;
            } else {
                if (segmentM == last) segmentM = 0;
                else segmentM++;
            }
            stencil(BM.borders[borderM], segmentM, nd, sh);
            coords();
            xsi = -1.;
            project();
            gapt = gapt + 0.5 * (1. + xsi) * std::sqrt((X[2] - X[1]) * (X[2] - X[1]) + (Y[2] - Y[1]) * (Y[2] - Y[1]));
            c.bodyM = bodyM; c.borderM = borderM; c.shiftM = shiftM; c.segmentM = segmentM;
        } else {
            gapt = gapt + 0.5 * (xsi - xsi_prev) * std::sqrt((X[2] - X[1]) * (X[2] - X[1]) + (Y[2] - Y[1]) * (Y[2] - Y[1]));
        }
        double vgx = s.vx[c.nodeS] - (N[0] * SM.vx[nd[0]] + N[1] * SM.vx[nd[1]] + N[2] * SM.vx[nd[2]] + N[3] * SM.vx[nd[3]]);
        double vgy = s.vy[c.nodeS] - (N[0] * SM.vy[nd[0]] + N[1] * SM.vy[nd[1]] + N[2] * SM.vy[nd[2]] + N[3] * SM.vy[nd[3]]);
        c.gapn = gapn;
        c.vgapn = vgx * xn + vgy * yn;
        c.gapt = gapt;
        c.vgapt = vgx * xt + vgy * yt;
        c.xsi = xsi; c.xnorm = xn; c.ynorm = yn; c.xtan = xt; c.ytan = yt;
        for (int k = 0; k < 4; k++) { c.nodeM[k] = nd[k]; c.shapeM[k] = N[k]; c.shiftMk[k] = sh[k]; }
    }
}

// contact laws, Code/Model.h:81-335 (the 0.5 factors are applied by the caller)
void law_frictionless(double kn, double gapn, double& gapt, double& Pn, double& Pt) {
    if (gapn > 0.) Pn = 0.;
    else Pn = -kn * gapn;
    Pt = 0.;
    gapt = 0.;
}
void law_cohesion(double kn, double kt, double gtang, double gtens, double gapn, double& gapt, double& Pn, double& Pt) {
    Pn = -kn * gapn;
    Pt = -kt * gapt;
    if (Pn < -gtens) { Pn = 0.; Pt = 0.; gapt = 0.; }
    else if (std::abs(Pt) > gtang) {
        Pt = gapt > 0. ? -gtang : gtang;
        gapt = -Pt / kt;
    }
}
void law_mohr_coulomb(double kn, double kt, double fric, double coh, double tens, double gapn, double& gapt,
                      double& Pn, double& Pt) {
    Pn = -kn * gapn;
    Pt = -kt * gapt;
    if (Pn < -tens) { Pn = 0.; Pt = 0.; gapt = 0.; }
    else if (std::abs(Pt) > coh + fric * Pn) {
        if (gapt > 0.) Pt = -coh - fric * Pn;
        else Pt = coh + fric * Pn;
        gapt = -Pt / kt;
    }
}
void law_damped_mc(double kn, double kt, double fric, double coh, double tens, double damp, double mass, double length,
                   double gapn, double vgapn, double& gapt, double vgapt, double& Pn, double& Pt, double& W) {
    Pn = 0.;
    if (gapn < 0.) {
        Pn = -damp * vgapn * 2. * std::sqrt(mass * kn / length);
        W += Pn * vgapn;
    }
    Pn -= kn * gapn;
    if (Pn < -tens) { Pn = 0.; Pt = 0.; gapt = 0.; }
    else {
        double Ptd = -damp * vgapt * 2. * std::sqrt(mass * kt / length);
        Pt = -kt * gapt + Ptd;
        if (std::abs(Pt) > coh + fric * Pn) {
            if (gapt > 0.) Pt = -coh - fric * Pn;
            else Pt = coh + fric * Pn;
            gapt = -Pt / kt;
            W += Pt * vgapt;
        } else W += Ptd * vgapt;
    }
}
void law_two_slopes(double kn, double kt, double fric, double coh, double tens, double rati, double gapn, double vgapn,
                    double& gapt, double& Pn, double& Pt) {
    if (vgapn <= 0.) Pn = -kn * gapn;
    else Pn = -kn * gapn * rati;
    if (Pn < -tens) { Pn = 0.; Pt = 0.; gapt = 0.; }
    else {
        Pt = -kt * gapt;
        if (std::abs(Pt) > coh + fric * Pn) {
            if (gapt > 0.) Pt = -coh - fric * Pn;
            else Pt = coh + fric * Pn;
            gapt = -Pt / kt;
        }
    }
}
void law_czm_linear(double kini, double Pnlim, double gapnlim, double Pnres, double NT, double gapn, double& gapt,
                    double& D, double& Pn, double& Pt) {
    if (D >= 1.) {
        Pn = -kini * gapn;
        Pt = -kini * gapt;
        if (Pn < -Pnres) { Pn = 0.; Pt = 0.; gapt = 0.; }
        else if (std::abs(Pt) > Pnres * NT) {
            Pt = gapt > 0. ? -Pnres * NT : Pnres * NT;
            gapt = -Pt / kini;
        }
        return;
    }
    double gnmax = Pnlim / kini * (1. - D) + gapnlim * D;
    double gtmax = gnmax * NT;
    double kmax;
    bool n_in = gapn < gnmax, n_out = gapn > gnmax;
    bool t_in = std::abs(gapt) < gtmax, t_out = std::abs(gapt) > gtmax;
    if (n_in && t_in) {
        kmax = (Pnlim * (1. - D) + Pnres * D) / gnmax;
    } else if (n_out && t_in) {
        gnmax = gapn;
        gtmax = gnmax * NT;
        D = (gnmax - Pnlim / kini) / (gapnlim - Pnlim / kini);
        kmax = (Pnlim * (1. - D) + Pnres * D) / gnmax;
    } else if (n_in && t_out) {
        gtmax = std::abs(gapt);
        gnmax = gtmax / NT;
        D = (gnmax - Pnlim / kini) / (gapnlim - Pnlim / kini);
        kmax = (Pnlim * (1. - D) + Pnres * D) / gnmax;
    } else if (n_out && t_out) {
        if (gapn / gnmax > std::abs(gapt) / gtmax) { gnmax = gapn; gtmax = gnmax * NT; }
        else { gtmax = std::abs(gapt); gnmax = gtmax / NT; }
        D = (gnmax - Pnlim / kini) / (gapnlim - Pnlim / kini);
        kmax = (Pnlim * (1. - D) + Pnres * D) / gnmax;
    } else {
        return;  // on the boundary of the elastic domain the reference leaves Pn, Pt untouched
    }
    if (gapn < 0.) Pn = -kini * gapn;
    else Pn = -kmax * gapn;
    Pt = -kmax * gapt;
}

// Apply_CZM_fatigue, Code/Model.h:343-446 (Pn, Pt enter as the previous tractions)
void law_czm_fatigue(double kini, double Pnlim, double gapnres, double Pnres, double NT, double Pnfat, double Raten,
                     double Ratet, double gapn, double& gapt, double& D, double& Pn, double& Pt) {
    if (D < 1.) {
        double gnmax = Pnlim / kini * (1. - D) + gapnres * D;
        double gtmax = gnmax * NT;
        double kmax;
        if (gapn < gnmax && std::abs(gapt) < gtmax) {
            double Pnm = Pn, Ptm = Pt;
            kmax = (Pnlim * (1. - D) + Pnres * D) / gnmax;
            if (gapn < 0.) Pn = -kini * gapn;
            else Pn = -kmax * gapn;
            Pt = -kmax * gapt;
            double Pnmax = Pnlim * (1. - D) + Pnres * D;
            if (std::abs(Pn) > Pnfat && std::abs(Pn) > std::abs(Pnm)) {
                double ratio = (std::abs(Pn) - Pnfat) / (Pnmax - Pnfat);
                D = D + ratio * Raten * (std::abs(Pn) - std::abs(Pnm));
            }
            if (std::abs(Pt) > NT * Pnfat && std::abs(Pt) > std::abs(Ptm)) {
                double ratio = (std::abs(Pt) - Pnfat * NT) / (NT * (Pnmax - Pnfat));
                D = D + ratio * Ratet * (std::abs(Pt) - std::abs(Ptm));
            }
        } else {
            bool grew = true;
            if (gapn > gnmax && std::abs(gapt) < gtmax) { gnmax = gapn; gtmax = gnmax * NT; }
            else if (gapn < gnmax && std::abs(gapt) > gtmax) { gtmax = std::abs(gapt); gnmax = gtmax / NT; }
            else if (gapn > gnmax && std::abs(gapt) > gtmax) {
                if (gapn / gnmax > std::abs(gapt) / gtmax) { gnmax = gapn; gtmax = gnmax * NT; }
                else { gtmax = std::abs(gapt); gnmax = gtmax / NT; }
            } else grew = false;
            if (grew) {
                D = (gnmax - Pnlim / kini) / (gapnres - Pnlim / kini);
                kmax = (Pnlim * (1. - D) + Pnres * D) / gnmax;
                if (gapn < 0.) Pn = -kini * gapn;
                else Pn = -kmax * gapn;
                Pt = -kmax * gapt;
            }
        }
    }
    if (D >= 1.) {
        D = 1.;
        Pn = -kini * gapn;
        Pt = -kini * gapt;
        if (Pn < -Pnres) { Pn = 0.; Pt = 0.; gapt = 0.; }
        else if (std::abs(Pt) > Pnres * NT) {
            Pt = gapt > 0. ? -Pnres * NT : Pnres * NT;
            gapt = -Pt / kini;
        }
    }
}

// Apply_Bonded_Mohr_Coulomb, Code/Model.h:453-506
void law_bonded_mc(double kn, double kt, double kbond, double fricbond, double cohbond, double tensbond,
                   double fricfree, double cohfree, double tensfree, double damp, double mass, double length,
                   double gapn, double vgapn, double& gapt, double vgapt, double& D, double& Pn, double& Pt) {
    if (D >= 1.) {
        Pn = -kn * gapn - damp * vgapn * 2. * std::sqrt(mass * kn / length);
        if (Pn < -tensfree) { Pn = 0.; Pt = 0.; gapt = 0.; }
        else Pt = -kt * gapt - damp * vgapt * 2. * std::sqrt(mass * kt / length);
        if (std::abs(Pt) > cohfree + fricfree * Pn) {
            Pt = gapt > 0. ? -cohfree - fricfree * Pn : cohfree + fricfree * Pn;
            gapt = -Pt / kt;
        }
    } else {
        if (gapn < 0.) Pn = -kn * gapn - damp * vgapn * 2. * std::sqrt(mass * kn / length);
        else Pn = -kbond * gapn - damp * vgapn * 2. * std::sqrt(mass * kbond / length);
        Pt = -kt * gapt - damp * vgapt * 2. * std::sqrt(mass * kt / length);
        if (Pn < -tensbond) {
            D = 1.;
            if (Pn < -tensfree) { Pn = 0.; Pt = 0.; gapt = 0.; }
        }
        if (std::abs(Pt) > cohbond + fricbond * Pn) {
            D = 1.;
            Pt = gapt > 0. ? -cohfree - fricfree * Pn : cohfree + fricfree * Pn;
            gapt = -Pt / kt;
        }
    }
}

void update_contact_forces(Sim& S, int bi, double dt) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    double period = S.m.xmax_period - S.m.xmin_period;
    double Pn = 0., Pt = 0.;   // carried between elements as in Code/Body.h:1647
    for (ContactElement& c : s.contacts) {
        double W = 0.;
        double length = s.length[c.borderS][c.border_nodeS];
        double initial_length = b.borders[c.borderS].initial_length[c.border_nodeS];
        double gapn = c.gapn, gapt = c.gapt;
        int bodyM = c.bodyM;
        int li = S.m.law_index(b.material, S.m.bodies[bodyM].material);
        if (li < 0) continue;
        const ContactLaw& law = S.m.laws[li];
        const std::vector<double>& p = law.p;
        switch (law.type) {
            case LAW_FRICTIONLESS: law_frictionless(p[0], gapn, gapt, Pn, Pt); break;
            case LAW_COHESION: law_cohesion(p[0], p[1], p[2] * 0.5, p[3] * 0.5, gapn, gapt, Pn, Pt); break;
            case LAW_MOHR_COULOMB: law_mohr_coulomb(p[0], p[1], p[2], p[3] * 0.5, p[4] * 0.5, gapn, gapt, Pn, Pt); break;
            case LAW_DAMPED_MOHR_COULOMB:
                law_damped_mc(p[0], p[1], p[2], p[3] * 0.5, p[4] * 0.5, p[5], c.effective_mass, initial_length, gapn,
                              c.vgapn, gapt, c.vgapt, Pn, Pt, W);
                break;
            case LAW_TWO_SLOPES:
                law_two_slopes(p[0], p[1], p[2], p[3] * 0.5, p[4] * 0.5, p[5], gapn, c.vgapn, gapt, Pn, Pt);
                break;
            case LAW_CZM_LINEAR: {
                if (c.nb_internal == 0) { c.nb_internal = 1; c.internal.assign(1, 1.); }
                double D = c.internal[0];
                law_czm_linear(p[0], p[1] * 0.5, p[2], p[3] * 0.5, p[4], gapn, gapt, D, Pn, Pt);
                c.internal[0] = D;
                break;
            }
            case LAW_BONDED_MOHR_COULOMB:
            case LAW_CZM_FATIGUE: {
                // Code/Body.h:1745-1810: damage and the previous tractions from
                // the element (a fresh element starts from a broken bond, and its
                // law then does not read the tractions)
                double D = c.internal.empty() ? 0. : c.internal[0];
                if (c.internal.size() >= 3) { Pn = c.internal[1]; Pt = c.internal[2]; }
                if (c.nb_internal == 0) { D = 1.; c.nb_internal = 3; c.internal = {1., 0., 0.}; }
                if (c.internal.size() < 3) c.internal.resize(3, 0.);
                if (law.type == LAW_BONDED_MOHR_COULOMB)
                    law_bonded_mc(p[0], p[1], p[2], p[3], p[4] * 0.5, p[5] * 0.5, p[6], p[7] * 0.5, p[8] * 0.5, p[9],
                                  c.effective_mass, initial_length, gapn, c.vgapn, gapt, c.vgapt, D, Pn, Pt);
                else
                    law_czm_fatigue(p[0], p[1] * 0.5, p[2], p[3] * 0.5, p[4], p[5], p[6], p[7], gapn, gapt, D, Pn, Pt);
                c.internal[0] = D;
                c.internal[1] = Pn;
                c.internal[2] = Pt;
                break;
            }
            default: break;
        }
        double Px = c.xnorm * Pn + c.xtan * Pt;
        double Py = c.ynorm * Pn + c.ytan * Pt;
        double Fx, Fy;
        if (law.fixed_length) { Fx = Px * initial_length; Fy = Py * initial_length; W *= initial_length * dt; }
        else { Fx = Px * length; Fy = Py * length; W *= length * dt; }
        c.gapn = gapn; c.gapt = gapt; c.length = length; c.fx = Fx; c.fy = Fy;
        s.works[1] += 0.5 * W;
        S.s[bodyM].works[1] += 0.5 * W;
        if (!b.rigid) {
            for (int k = b.dptr[c.nodeS]; k < b.dptr[c.nodeS + 1]; k++) {
                s.fcx[b.didx[k]] += Fx * b.dphi[k];
                s.fcy[b.didx[k]] += Fy * b.dphi[k];
            }
        } else {
            double dx = s.xc[c.nodeS] - s.cur[0], dy = s.yc[c.nodeS] - s.cur[1];
            s.rfc[0] += Fx;
            s.rfc[1] += Fy;
            s.rfc[2] += -Fx * dy + Fy * dx;
        }
        const Body& BM = S.m.bodies[bodyM];
        if (!BM.rigid) {
            for (int q = 0; q < 4; q++) {
                int nm = c.nodeM[q];
                for (int k = BM.dptr[nm]; k < BM.dptr[nm + 1]; k++) {
                    s.send_f.push_back({-Fx * c.shapeM[q] * BM.dphi[k], -Fy * c.shapeM[q] * BM.dphi[k], 0.});
                    s.send_to.push_back({bodyM, BM.didx[k]});
                }
            }
        } else {
            double dx = s.xc[c.nodeS] - S.s[bodyM].cur[0] - c.shiftM * period;
            double dy = s.yc[c.nodeS] - S.s[bodyM].cur[1];
            s.send_f.push_back({-Fx, -Fy, Fx * dy - Fy * dx});
            s.send_to.push_back({bodyM, -1});
        }
    }
}

void send_contact_forces(Sim& S, int bi) {
    State& s = S.s[bi];
    for (size_t j = 0; j < s.send_to.size(); j++) {
        State& t = S.s[s.send_to[j][0]];
        int n = s.send_to[j][1];
        if (n == -1) { t.rfc[0] += s.send_f[j][0]; t.rfc[1] += s.send_f[j][1]; t.rfc[2] += s.send_f[j][2]; }
        else { t.fcx[n] += s.send_f[j][0]; t.fcy[n] += s.send_f[j][1]; }
    }
}

// --------------------------------------------------------------------------
// integration
// --------------------------------------------------------------------------
// Sum_up_forces + Apply_Newton (Code/Node.h:212-231, Code/Body.h:1975-2043):
// accelerations divided by the mass-scaling factors (1 unless a
// mass-scaling scheme changed them)
void sum_newton(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    int n = b.n_nodes;
    if (!b.rigid) {
        // Sum_up_forces (Code/Node.h:212-225)
        for (int i = 0; i < n; i++) {
            double ix = 0., iy = 0.;
            for (int r = 0; r < b.n_regions; r++) {
                ix += s.regx[(size_t)r * n + i];
                iy += s.regy[(size_t)r * n + i];
                s.regx[(size_t)r * n + i] = 0.;
                s.regy[(size_t)r * n + i] = 0.;
            }
            s.fix[i] = ix; s.fiy[i] = iy;
            s.fx[i] = ix + s.fcx[i] + 0. + s.fbx[i] + s.fdx[i] + s.fnx[i] + s.fkx[i] + s.fax[i];
            s.fy[i] = iy + s.fcy[i] + 0. + s.fby[i] + s.fdy[i] + s.fny[i] + s.fky[i] + s.fay[i];
        }
        // Apply_Newton (Code/Node.h:227-231)
        for (int i = 0; i < n; i++) { s.apx[i] = s.fx[i] * s.imx[i] / s.fsx[i]; s.apy[i] = s.fy[i] * s.imy[i] / s.fsy[i]; }
        for (int i = 0; i < n; i++) { s.ax[i] = 0.; s.ay[i] = 0.; }
    } else {
        for (int q = 0; q < 3; q++) s.rf[q] = s.rfc[q] + s.rfb[q] + s.rfd[q] + s.rfn[q] + s.rfk[q];
        s.acc[0] = s.rf[0] * s.inverse_mass / s.factor_ms;
        s.acc[1] = s.rf[1] * s.inverse_mass / s.factor_ms;
        s.acc[2] = s.rf[2] * s.inverse_inertia / s.factor_ms;
        for (int i = 0; i < n; i++) {
            s.fcx[i] = s.rfc[0]; s.fcy[i] = s.rfc[1];
            s.fbx[i] = s.rfb[0]; s.fby[i] = s.rfb[1];
            s.fdx[i] = s.rfd[0]; s.fdy[i] = s.rfd[1];
            s.fnx[i] = s.rfn[0]; s.fny[i] = s.rfn[1];
            s.fkx[i] = s.rfk[0]; s.fky[i] = s.rfk[1];
            s.fx[i] = s.rf[0]; s.fy[i] = s.rf[1];
            s.ax[i] = s.acc[0]; s.ay[i] = s.acc[1];
        }
    }
}

// Apply_Euler (Code/Body.h:2051-2118) + Update_kinematics +
// Update_current_positions (Code/Body.h:2176-2302)
void euler_kinematics(Sim& S, int bi, double dt) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    int n = b.n_nodes;
    if (!b.rigid) {
        for (int i = 0; i < n; i++) {
            double vx = s.vpx[i] + 0.5 * dt * s.apx[i];
            double vy = s.vpy[i] + 0.5 * dt * s.apy[i];
            s.works[0] += dt * (vx * s.fix[i] + vy * s.fiy[i]);
            s.works[2] += dt * (vx * s.fbx[i] + vy * s.fby[i]);
            s.works[3] += dt * (vx * s.fdx[i] + vy * s.fdy[i]);
            s.works[4] += dt * (vx * s.fnx[i] + vy * s.fny[i]);
            s.works[5] += dt * (vx * s.fkx[i] + vy * s.fky[i]);
            s.works[6] += dt * (vx * s.fax[i] + vy * s.fay[i]);
            s.vpx[i] = s.vpx[i] + dt * s.apx[i];
            s.vpy[i] = s.vpy[i] + dt * s.apy[i];
            s.upx[i] = s.upx[i] + dt * s.vpx[i];
            s.upy[i] = s.upy[i] + dt * s.vpy[i];
        }
        // Update_kinematics (Code/Body.h:2176-2212)
        for (int i = 0; i < n; i++) {
            double dx = 0., dy = 0., vx = 0., vy = 0., ax = 0., ay = 0.;
            for (int k = b.dptr[i]; k < b.dptr[i + 1]; k++) {
                int j = b.didx[k];
                double f = b.dphi[k];
                dx += f * s.upx[j]; dy += f * s.upy[j];
                vx += f * s.vpx[j]; vy += f * s.vpy[j];
                ax += f * s.apx[j]; ay += f * s.apy[j];
            }
            s.ux[i] = dx; s.uy[i] = dy; s.vx[i] = vx; s.vy[i] = vy; s.ax[i] = ax; s.ay[i] = ay;
        }
        for (int i = 0; i < n; i++) { s.xc[i] = b.X[i] + s.ux[i]; s.yc[i] = b.Y[i] + s.uy[i]; }
    } else {
        double vx = s.vel[0] + 0.5 * dt * s.acc[0];
        double vy = s.vel[1] + 0.5 * dt * s.acc[1];
        double vr = s.vel[2] + 0.5 * dt * s.acc[2];
        s.works[2] += dt * (vx * s.rfb[0] + vy * s.rfb[1] + vr * s.rfb[2]);
        s.works[3] += dt * (vx * s.rfd[0] + vy * s.rfd[1] + vr * s.rfd[2]);
        s.works[4] += dt * (vx * s.rfn[0] + vy * s.rfn[1] + vr * s.rfn[2]);
        s.works[5] += dt * (vx * s.rfk[0] + vy * s.rfk[1] + vr * s.rfk[2]);
        if (s.driven[0] == 0.) {
            s.vel[0] += dt * s.acc[0];
            s.disp[0] = s.disp[0] + dt * s.vel[0];
        } else if (s.driven[0] == 1.) {
            s.acc[0] = s.driven[3];
            s.vel[0] = s.driven[2];
            s.disp[0] = s.driven[1];
            s.rfd[0] = s.mass * s.acc[0] - s.rf[0];
        }
        if (s.driven[4] == 0.) {
            s.vel[1] += dt * s.acc[1];
            s.disp[1] = s.disp[1] + dt * s.vel[1];
        } else if (s.driven[4] == 1.) {
            s.acc[1] = s.driven[7];
            s.vel[1] = s.driven[6];
            s.disp[1] = s.driven[5];
            s.rfd[1] = s.mass * s.acc[1] - s.rf[1];
        }
        if (s.driven[0] == 0. && s.driven[4] == 0. && b.periodicity != PER_PERIODIC) {
            s.vel[2] += dt * s.acc[2];
            s.disp[2] = s.disp[2] + dt * s.vel[2];
        } else {
            s.vel[2] = 0.;
            s.disp[2] = 0.;
        }
        // Update_kinematics rigid branch then Update_current_positions (Code/Body.h:2214-2226, 2281-2301)
        double dr = s.disp[2] - 0.;
        for (int i = 0; i < n; i++) {
            double dx = b.X[i] - b.xc, dy = b.Y[i] - b.yc;
            s.ux[i] = s.disp[0] + dx * std::cos(dr) - dy * std::sin(dr) - dx;
            s.uy[i] = s.disp[1] + dx * std::sin(dr) + dy * std::cos(dr) - dy;
        }
        s.cur[0] = b.xc + s.disp[0];
        s.cur[1] = b.yc + s.disp[1];
        s.cur[2] = 0. + s.disp[2];
        for (int i = 0; i < n; i++) {
            double dx = b.X[i] - b.xc, dy = b.Y[i] - b.yc;
            s.xc[i] = s.cur[0] + dx * std::cos(dr) - dy * std::sin(dr);
            s.yc[i] = s.cur[1] + dx * std::sin(dr) + dy * std::cos(dr);
            double ex = s.xc[i] - s.cur[0], ey = s.yc[i] - s.cur[1];
            s.vx[i] = -ey * s.vel[2] + s.vel[0];
            s.vy[i] = ex * s.vel[2] + s.vel[1];
        }
    }
}

void sum_newton_euler_kinematics(Sim& S, int bi, double t, double dt) {
    sum_newton(S, bi);
    update_bc(S, bi, t + dt);
    euler_kinematics(S, bi, dt);
}

void rigid_positions(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    double dr = s.disp[2] - 0.;
    s.cur[0] = b.xc + s.disp[0];
    s.cur[1] = b.yc + s.disp[1];
    s.cur[2] = 0. + s.disp[2];
    for (int i = 0; i < b.n_nodes; i++) {
        double dx = b.X[i] - b.xc, dy = b.Y[i] - b.yc;
        s.xc[i] = s.cur[0] + dx * std::cos(dr) - dy * std::sin(dr);
        s.yc[i] = s.cur[1] + dx * std::sin(dr) + dy * std::cos(dr);
        double ex = s.xc[i] - s.cur[0], ey = s.yc[i] - s.cur[1];
        s.vx[i] = -ey * s.vel[2] + s.vel[0];
        s.vy[i] = ex * s.vel[2] + s.vel[1];
    }
}

void update_contact_pressures(Sim& S) {
    for (auto& s : S.s)
        for (size_t j = 0; j < s.xcp.size(); j++) {
            std::fill(s.xcp[j].begin(), s.xcp[j].end(), 0.);
            std::fill(s.ycp[j].begin(), s.ycp[j].end(), 0.);
        }
    for (size_t i = 0; i < S.s.size(); i++)
        for (const ContactElement& c : S.s[i].contacts) {
            double il = 1. / c.length;
            State& t = S.s[c.bodyM];
            S.s[i].xcp[c.borderS][c.border_nodeS] += c.fx * il;
            S.s[i].ycp[c.borderS][c.border_nodeS] += c.fy * il;
            t.xcp[c.borderM][c.segmentM] -= c.fx * il * c.shapeM[1];
            t.ycp[c.borderM][c.segmentM] -= c.fy * il * c.shapeM[1];
            t.xcp[c.borderM][c.segmentM + 1] -= c.fx * il * c.shapeM[2];
            t.ycp[c.borderM][c.segmentM + 1] -= c.fy * il * c.shapeM[2];
        }
}

// --------------------------------------------------------------------------
// proximity (Code/Proximity.h:48-473)
// --------------------------------------------------------------------------
void update_proximity(Sim& S) {
    Model& m = S.m;
    int NB = (int)m.bodies.size();
    double period = m.xmax_period - m.xmin_period;
    double invperiod = 1. / period;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < NB; i++) {
        State& s = S.s[i];
        Body& b = m.bodies[i];
        s.neighbours.clear();
        s.nb_neighbours = 0;
        if (b.status == ST_INACTIVE) continue;
        for (size_t j = 0; j < b.borders.size(); j++) {
            double xmin = 1.e100, xmax = -1.e100, ymin = 1.e100, ymax = -1.e100;
            double d = b.detection_distance;
            for (int nd : b.borders[j].nodes) {
                double x = s.xc[nd], y = s.yc[nd];
                if (x - d < xmin) xmin = x - d;
                if (x + d > xmax) xmax = x + d;
                if (y - d < ymin) ymin = y - d;
                if (y + d > ymax) ymax = y + d;
            }
            s.pxmin[j] = xmin; s.pxmax[j] = xmax; s.pymin[j] = ymin; s.pymax[j] = ymax;
            s.shmin[j] = (int)std::floor((xmin - m.xmin_period) * invperiod);
            s.shmax[j] = (int)std::floor((xmax - m.xmin_period) * invperiod);
        }
    }
#pragma omp parallel for schedule(dynamic)
    for (int b1 = 0; b1 < NB; b1++) {
        if (m.bodies[b1].status == ST_INACTIVE) continue;
        State& s1 = S.s[b1];
        for (size_t r1 = 0; r1 < m.bodies[b1].borders.size(); r1++) {
            for (int b2 = 0; b2 < NB; b2++) {
                if (m.bodies[b2].status == ST_INACTIVE || b1 == b2) continue;
                const State& s2 = S.s[b2];
                for (size_t r2 = 0; r2 < m.bodies[b2].borders.size(); r2++) {
                    if ((s2.pymax[r2] < s1.pymin[r1]) || (s1.pymax[r1] < s2.pymin[r2])) continue;
                    for (int sh = s1.shmin[r1] - s2.shmax[r2]; sh <= s1.shmax[r1] - s2.shmin[r2]; sh++) {
                        if ((s2.pxmax[r2] + sh * period > s1.pxmin[r1]) && (s2.pxmin[r2] + sh * period < s1.pxmax[r1])) {
                            bool exists = false;
                            for (int k = 0; k < s1.nb_neighbours; k++)
                                if (s1.neighbours[3 * k] == b2 && s1.neighbours[3 * k + 1] == (int)r2 &&
                                    s1.neighbours[3 * k + 2] == sh) { exists = true; break; }
                            if (!exists) {
                                s1.neighbours.push_back(b2);
                                s1.neighbours.push_back((int)r2);
                                s1.neighbours.push_back(sh);
                                s1.nb_neighbours++;
                            }
                        }
                    }
                }
            }
        }
    }
#pragma omp parallel for schedule(dynamic)
    for (int bS = 0; bS < NB; bS++) {
        State& sS = S.s[bS];
        const Body& BS = m.bodies[bS];
        std::vector<ContactElement> out;
        ContactElement fresh;
        for (size_t rS = 0; rS < BS.borders.size(); rS++) {
            const Border& brS = BS.borders[rS];
            int nS = (int)brS.nodes.size();
            for (int nb = 0; nb < sS.nb_neighbours; nb++) {
                int bM = sS.neighbours[3 * nb], rM = sS.neighbours[3 * nb + 1], shM = sS.neighbours[3 * nb + 2];
                const Body& BM = m.bodies[bM];
                const State& sM = S.s[bM];
                const Border& brM = BM.borders[rM];
                int nM = (int)brM.nodes.size();
                double emass;
                if (sM.mass == 0.) emass = 0.;
                else if (sS.mass == 0.) emass = 0.;
                else emass = 1. / (sM.inverse_mass + sS.inverse_mass);
                for (int jS = 0; jS < nS; jS++) {
                    int nodeS = brS.nodes[jS];
                    double len = sS.length[rS][jS];
                    double Xs = sS.xc[nodeS], Ys = sS.yc[nodeS];
                    for (int seg = 0; seg < nM - 1; seg++) {
                        int n1 = brM.node1[seg], n2 = brM.node2[seg];
                        double X1 = sM.xc[n1] + period * (shM + brM.shift1[seg]), Y1 = sM.yc[n1];
                        double X2 = sM.xc[n2] + period * (shM + brM.shift2[seg]), Y2 = sM.yc[n2];
                        if (detect_segment(Xs, Ys, X1, Y1, X2, Y2, BS.detection_distance) == 0) continue;
                        int n0 = brM.node0[seg], n3 = brM.node3[seg];
                        double X0 = sM.xc[n0] + period * (shM + brM.shift0[seg]), Y0 = sM.yc[n0];
                        double X3 = sM.xc[n3] + period * (shM + brM.shift3[seg]), Y3 = sM.yc[n3];
                        double gap = 0, xsi = 0., xn = 0, yn = 0, xt = 0, yt = 0, N[4] = {0, 0, 0, 0};
                        int flag = 0;
                        closest_linear(Xs, Ys, X0, Y0, X1, Y1, X2, Y2, X3, Y3, gap, xsi, xn, yn, xt, yt, N[0], N[1],
                                       N[2], N[3], flag);
                        if ((flag == -2 || flag == -1 || flag == 0) && (gap > -BS.contact_distance)) {
                            bool found = false;
                            for (const ContactElement& o : sS.contacts)
                                if (o.borderS == (int)rS && o.border_nodeS == jS && o.bodyM == bM && o.borderM == rM &&
                                    std::abs(o.segmentM - seg) <= 1) {
                                    fresh = o;
                                    found = true;
                                    break;
                                }
                            if (!found) {
                                fresh.borderS = (int)rS; fresh.border_nodeS = jS; fresh.nodeS = nodeS;
                                fresh.bodyM = bM; fresh.borderM = rM; fresh.shiftM = shM; fresh.segmentM = seg;
                                fresh.nodeM[0] = n0; fresh.nodeM[1] = n1; fresh.nodeM[2] = n2; fresh.nodeM[3] = n3;
                                fresh.shiftMk[0] = brM.shift0[seg]; fresh.shiftMk[1] = brM.shift1[seg];
                                fresh.shiftMk[2] = brM.shift2[seg]; fresh.shiftMk[3] = brM.shift3[seg];
                                for (int q = 0; q < 4; q++) fresh.shapeM[q] = N[q];
                                fresh.gapn = gap; fresh.gapt = 0.; fresh.xsi = xsi;
                                fresh.xnorm = xn; fresh.ynorm = yn; fresh.xtan = xt; fresh.ytan = yt;
                                fresh.effective_mass = emass; fresh.length = len; fresh.fx = 0.; fresh.fy = 0.;
                                fresh.nb_internal = 0; fresh.internal.assign(1, 0.);
                            }
                            out.push_back(fresh);
                        }
                    }
                }
            }
        }
        sS.contacts.swap(out);
    }
    if (m.flags[F_NO_MONITORING] == 0) {
#pragma omp parallel for schedule(dynamic)
        for (int bS = 0; bS < NB; bS++) {
            const Body& BS = m.bodies[bS];
            if (BS.rigid || BS.periodicity == PER_PERIODIC) continue;
            State& sS = S.s[bS];
            std::vector<ContactElement> out = sS.contacts;
            size_t n_old = sS.contacts.size();
            ContactElement fresh;
            double emass = sS.mass;
            for (size_t rS = 0; rS < BS.borders.size(); rS++) {
                const Border& brS = BS.borders[rS];
                for (int jS = 0; jS < (int)brS.nodes.size(); jS++) {
                    int nodeS = brS.nodes[jS];
                    double len = sS.length[rS][jS];
                    double Xs = sS.xc[nodeS], Ys = sS.yc[nodeS];
                    for (size_t rM = 0; rM < BS.borders.size(); rM++) {
                        const Border& brM = BS.borders[rM];
                        for (int seg = 0; seg < (int)brM.nodes.size() - 1; seg++) {
                            int n1 = brM.node1[seg], n2 = brM.node2[seg];
                            if (n1 == nodeS || n2 == nodeS) continue;
                            double X1 = sS.xc[n1], Y1 = sS.yc[n1], X2 = sS.xc[n2], Y2 = sS.yc[n2];
                            if (detect_segment(Xs, Ys, X1, Y1, X2, Y2, BS.detection_distance) == 0) continue;
                            int n0 = brM.node0[seg], n3 = brM.node3[seg];
                            double X0 = sS.xc[n0], Y0 = sS.yc[n0], X3 = sS.xc[n3], Y3 = sS.yc[n3];
                            double gap = 0, xsi = 0., xn = 0, yn = 0, xt = 0, yt = 0, N[4] = {0, 0, 0, 0};
                            int flag = 0;
                            if (n0 == nodeS || n3 == nodeS)
                                closest_self(Xs, Ys, X1, Y1, X2, Y2, gap, xsi, xn, yn, xt, yt, N[0], N[1], N[2], N[3], flag);
                            else
                                closest_linear(Xs, Ys, X0, Y0, X1, Y1, X2, Y2, X3, Y3, gap, xsi, xn, yn, xt, yt, N[0],
                                               N[1], N[2], N[3], flag);
                            if (flag == 0 && gap > -BS.contact_distance) {
                                bool found = false;
                                for (size_t k = 0; k < n_old; k++) {
                                    const ContactElement& o = sS.contacts[k];
                                    if (o.borderS == (int)rS && o.border_nodeS == jS && o.bodyM == bS &&
                                        o.borderM == (int)rM && std::abs(o.segmentM - seg) <= 1.) {
                                        fresh = o;
                                        found = true;
                                        break;
                                    }
                                }
                                if (!found) {
                                    fresh.borderS = (int)rS; fresh.border_nodeS = jS; fresh.nodeS = nodeS;
                                    fresh.bodyM = bS; fresh.borderM = (int)rM; fresh.shiftM = 0; fresh.segmentM = seg;
                                    fresh.nodeM[0] = n0; fresh.nodeM[1] = n1; fresh.nodeM[2] = n2; fresh.nodeM[3] = n3;
                                    for (int q = 0; q < 4; q++) { fresh.shiftMk[q] = 0; fresh.shapeM[q] = N[q]; }
                                    fresh.gapn = gap; fresh.gapt = 0.; fresh.xsi = xsi;
                                    fresh.xnorm = xn; fresh.ynorm = yn; fresh.xtan = xt; fresh.ytan = yt;
                                    fresh.effective_mass = emass; fresh.length = len; fresh.fx = 0.; fresh.fy = 0.;
                                    fresh.nb_internal = 0; fresh.internal.assign(1, 0.);
                                }
                                out.push_back(fresh);
                            }
                        }
                    }
                }
            }
            sS.contacts.swap(out);
        }
    }
}

void initialize_czm(Sim& S) {
    for (size_t i = 0; i < S.s.size(); i++) update_contacts(S, (int)i);
    S.m.flags[F_INITIALIZE_CZM] = 0;
    for (size_t bS = 0; bS < S.s.size(); bS++) {
        for (ContactElement& c : S.s[bS].contacts) {
            const std::string& mS = S.m.materials[S.m.bodies[bS].material].name;
            const std::string& mM = S.m.materials[S.m.bodies[c.bodyM].material].name;
            const ContactLaw* law = nullptr;
            for (const ContactLaw& l : S.m.laws)
                if ((l.m1 == mS && l.m2 == mM) || (l.m2 == mS && l.m1 == mM)) { law = &l; break; }
            if (law && law->type == LAW_CZM_LINEAR) {
                c.nb_internal = 1;
                c.internal.assign(1, c.gapn < law->p[5] ? 0. : 1.);
            } else if (law && (law->type == LAW_CZM_FATIGUE || law->type == LAW_BONDED_MOHR_COULOMB)) {
                const double gapinit = law->p[law->type == LAW_CZM_FATIGUE ? 8 : 10];
                c.nb_internal = 3;
                c.internal = {c.gapn < gapinit ? 0. : 1., 0., 0.};
            }
        }
    }
}

// Body::Update_material mass part (Code/Body.h:2751-2783) and rigid part (:2877-2905)
void update_material(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    double rho = S.m.materials[b.material].p[0];
    if (!b.rigid) {
        if (S.m.flags[F_UPDATE_MASS]) {
            s.mass = 0.;
            std::fill(s.mx.begin(), s.mx.end(), 0.);
            std::fill(s.my.begin(), s.my.end(), 0.);
            for (int g = 0; g < b.n_gauss; g++)
                for (int j = b.gptr[g]; j < b.gptr[g + 1]; j++)
                    for (int k = b.gptr[g]; k < b.gptr[g + 1]; k++) {
                        double dm = b.gw[g] * b.gjac[g] * b.gphi[j] * b.gphi[k];
                        s.mx[b.gidx[j]] += dm;
                        s.my[b.gidx[j]] += dm;
                    }
            for (int i = 0; i < b.n_nodes; i++) {
                s.mx[i] *= rho;
                s.my[i] *= rho;
                s.imx[i] = 1. / s.mx[i];
                s.imy[i] = 1. / s.my[i];
                s.mass += std::sqrt(s.mx[i] * s.mx[i] + s.my[i] * s.my[i]);
                s.inverse_mass = 1 / s.mass;
            }
        }
        // Rayleigh damping rebuilt from the stored stiffness (Code/Body.h:2862-2875)
        if (S.m.flags[F_UPDATE_DAMPING]) {
            double alpha = S.m.materials[b.material].p[1], beta = S.m.materials[b.material].p[2];
            for (size_t i = 0; i < b.mat_damping.size(); i++) {
                if (b.mat_dof[2 * i] == b.mat_dof[2 * i + 1])
                    b.mat_damping[i] = alpha * b.mat_stiffness[i] + beta * s.mx[b.mat_dof[2 * i] / 2];
                else
                    b.mat_damping[i] = alpha * b.mat_stiffness[i];
            }
        }
    } else {
        double mass = 0., inertia = 0.;
        for (size_t c = 0; c < b.cells.size() / 3; c++) {
            int n0 = b.cells[3 * c], n1 = b.cells[3 * c + 1], n2 = b.cells[3 * c + 2];
            double x0 = b.X[n0], y0 = b.Y[n0], x1 = b.X[n1], y1 = b.Y[n1], x2 = b.X[n2], y2 = b.Y[n2];
            double xc = (x0 + x1 + x2) / 3., yc = (y0 + y1 + y2) / 3.;
            auto surf = [](double a0, double b0, double a1, double b1, double a2, double b2) {
                return 0.5 * std::abs((a1 - a0) * (b2 - b0) - (a2 - a0) * (b1 - b0));
            };
            double sf = surf(x0, y0, x1, y1, x2, y2);
            mass += sf;
            inertia += sf * ((xc - b.xc) * (xc - b.xc) + (yc - b.yc) * (yc - b.yc));
            // Triangle_inertia (Code/Tools.h:390-413)
            double ti = 0.;
            double px[3] = {x0, x1, x2}, py[3] = {y0, y1, y2};
            for (int e = 0; e < 3; e++) {
                double ax = px[e], ay = py[e], bx = px[(e + 1) % 3], by = py[(e + 1) % 3];
                double area = surf(xc, yc, ax, ay, bx, by);
                double PP = (ax - xc) * (ax - xc) + (ay - yc) * (ay - yc);
                double PQ = (ax - xc) * (bx - xc) + (ay - yc) * (by - yc);
                double QQ = (bx - xc) * (bx - xc) + (by - yc) * (by - yc);
                ti += area * (PP + PQ + QQ);
            }
            inertia += std::abs(ti / 6.);
        }
        s.mass = mass * rho;
        s.inertia = inertia * rho;
        s.inverse_mass = 1. / s.mass;
        s.inverse_inertia = 1. / s.inertia;
    }
}

// --------------------------------------------------------------------------
// set-up, step, dump
// --------------------------------------------------------------------------
void setup(Sim& S) {
    Model& m = S.m;
    S.s.assign(m.bodies.size(), State());
    for (size_t bi = 0; bi < m.bodies.size(); bi++) {
        Body& b = m.bodies[bi];
        State& s = S.s[bi];
        int n = b.n_nodes;
        auto z = [&](std::vector<double>& v) { v.assign(n, 0.); };
        for (auto* v : {&s.xc, &s.yc, &s.fix, &s.fiy, &s.fcx, &s.fcy, &s.fbx, &s.fby, &s.fdx, &s.fdy, &s.fnx,
                        &s.fny, &s.fkx, &s.fky, &s.fx, &s.fy, &s.fax, &s.fay})
            z(*v);
        s.ux = b.ux; s.uy = b.uy; s.vx = b.vx; s.vy = b.vy; s.ax = b.ax; s.ay = b.ay;
        s.upx = b.upx; s.upy = b.upy; s.vpx = b.vpx; s.vpy = b.vpy; s.apx = b.apx; s.apy = b.apy;
        s.mx = b.mx; s.my = b.my; s.imx = b.imx; s.imy = b.imy;
        s.regx.assign((size_t)std::max(b.n_regions, 1) * n, 0.);
        s.regy.assign((size_t)std::max(b.n_regions, 1) * n, 0.);
        for (int q = 0; q < 7; q++) s.works[q] = b.works[q];
        size_t nbr = b.borders.size();
        s.length.resize(nbr); s.xcp.resize(nbr); s.ycp.resize(nbr); s.xbp.resize(nbr); s.ybp.resize(nbr);
        for (size_t k = 0; k < nbr; k++) {
            s.length[k] = b.borders[k].initial_length;
            s.xcp[k].assign(b.borders[k].nodes.size(), 0.);
            s.ycp[k].assign(b.borders[k].nodes.size(), 0.);
            s.xbp[k].assign(b.borders[k].nodes.size(), 0.);
            s.ybp[k].assign(b.borders[k].nodes.size(), 0.);
        }
        s.xbv.assign(nbr, 0.); s.ybv.assign(nbr, 0.); s.xbvel.assign(nbr, 0.); s.ybvel.assign(nbr, 0.);
        s.pxmin.assign(nbr, 0.); s.pxmax.assign(nbr, 0.); s.pymin.assign(nbr, 0.); s.pymax.assign(nbr, 0.);
        s.shmin.assign(nbr, 0); s.shmax.assign(nbr, 0);
        for (auto* v : {&s.uptx, &s.upty, &s.utx, &s.uty, &s.xerr, &s.yerr, &s.enorm, &s.msm}) z(*v);
        // Node constructor values (Code/Node.h:186-200) and the inverse
        // distance estimator of Load_static (Code/IO.h:1620, 1819)
        s.fsx.assign(n, 1.); s.fsy.assign(n, 1.); s.fsn.assign(n, 1.);
        s.dfsx.assign(n, 1.); s.dfsy.assign(n, 1.);
        s.ide.assign(n, 0.);
        for (int i = 0; i < n; i++) {
            if (b.rigid) s.ide[i] = 1. / b.nodal_distance;
            else if (!b.dptr.empty())
                s.ide[i] = 1. / (b.domain[i] * std::sqrt(3.1415927 / 4. / (b.dptr[i + 1] - b.dptr[i])));
        }
        s.contacts = b.contacts;
        s.neighbours = b.neighbours;
        s.nb_neighbours = (int)b.neighbours.size() / 3;
        if (!b.rigid) {
            s.mass = 0.;
            s.inverse_mass = 0.;
            for (int i = 0; i < n; i++) { s.xc[i] = b.X[i] + s.ux[i]; s.yc[i] = b.Y[i] + s.uy[i]; }
        } else {
            s.mass = b.mass; s.inertia = b.inertia; s.inverse_mass = b.inverse_mass; s.inverse_inertia = b.inverse_inertia;
            for (int q = 0; q < 3; q++) { s.cur[q] = b.rigid_cur[q]; s.disp[q] = b.rigid_disp[q]; s.vel[q] = b.rigid_vel[q]; }
            rigid_positions(S, (int)bi);
        }
    }
    S.time = m.time;
    S.iteration = m.number_iteration;
    S.next_contact_update = m.next_contact_update;
    // Code/main.cpp:103-126
    int NB = (int)m.bodies.size();
    for (int i = 0; i < NB; i++) update_borders(S, i);
    for (int i = 0; i < NB; i++) update_bc(S, i, S.time);
    update_proximity(S);
    if (m.flags[F_INITIALIZE_CZM] == 1) initialize_czm(S);
    for (int i = 0; i < NB; i++) update_material(S, i);
    update_proximity(S);
    if (m.flags[F_INITIALIZE_CZM] == 1) initialize_czm(S);
}

void euler_step(Sim& S) {
    Model& m = S.m;
    int NB = (int)m.bodies.size();
    double dt = m.dt;
#pragma omp parallel
    {
#pragma omp for schedule(dynamic)
        for (int i = 0; i < NB; i++) {
            if (m.bodies[i].status == ST_INACTIVE) continue;
            initialize_contact_forces(S, i);
            update_borders(S, i);
            update_contacts(S, i);
        }
#pragma omp for schedule(dynamic)
        for (int i = 0; i < NB; i++) {
            if (m.bodies[i].status == ST_INACTIVE) continue;
            update_bc_forces(S, i);
            update_body_forces(S, i);
            update_damping_forces(S, i);
            update_alid_forces(S, i, S.time);
        }
    }
    // contact forces: the reference runs them per body in parallel; the
    // damped-law work update of another body is a data race there, so the
    // restatement runs this loop in body order.
    for (int i = 0; i < NB; i++)
        if (m.bodies[i].status != ST_INACTIVE) update_contact_forces(S, i, dt);
    for (int i = 0; i < NB; i++) send_contact_forces(S, i);
    std::vector<std::pair<int, int>> regions;
    for (int i = 0; i < NB; i++)
        if (!m.bodies[i].rigid)
            for (int r = 0; r < m.bodies[i].n_regions; r++) regions.push_back({i, r});
#pragma omp parallel
    {
#pragma omp for schedule(dynamic)
        for (int k = 0; k < (int)regions.size(); k++) {
            if (m.bodies[regions[k].first].status == ST_INACTIVE) continue;
            update_internal_forces(S, regions[k].first, regions[k].second);
        }
#pragma omp for schedule(dynamic)
        for (int i = 0; i < NB; i++) {
            if (m.bodies[i].status == ST_INACTIVE) continue;
            sum_newton_euler_kinematics(S, i, S.time, dt);
        }
    }
    S.time += dt;
    update_contact_pressures(S);
}

// --------------------------------------------------------------------------
// adaptive schemes (Code/Solver.h:98-486)
// --------------------------------------------------------------------------

// Body::Apply_Euler_temporary + Update_kinematics_temporary
// (Code/Body.h:2126-2264, Code/Node.h:247-251)
void euler_temporary(Sim& S, int bi, double dt) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    const int n = b.n_nodes;
    if (!b.rigid) {
        for (int i = 0; i < n; i++) {
            s.uptx[i] = s.upx[i] + dt * (s.vpx[i] + dt * s.apx[i]);
            s.upty[i] = s.upy[i] + dt * (s.vpy[i] + dt * s.apy[i]);
        }
        for (int i = 0; i < n; i++) {
            double dx = 0., dy = 0.;
            for (int k = b.dptr[i]; k < b.dptr[i + 1]; k++) {
                dx += b.dphi[k] * s.uptx[b.didx[k]];
                dy += b.dphi[k] * s.upty[b.didx[k]];
            }
            s.utx[i] = dx; s.uty[i] = dy;
        }
        return;
    }
    if (s.driven[0] == 0.) {
        s.vel_t[0] = s.vel[0] + dt * s.acc[0];
        s.disp_t[0] = s.disp[0] + dt * s.vel_t[0];
    } else if (s.driven[0] == 1.) {
        s.vel_t[0] = s.driven[2];
        s.disp_t[0] = s.driven[1];
    }
    if (s.driven[4] == 0.) {
        s.vel_t[1] = s.vel[1] + dt * s.acc[1];
        s.disp_t[1] = s.disp[1] + dt * s.vel_t[1];
    } else if (s.driven[4] == 1.) {
        s.vel_t[1] = s.driven[6];
        s.disp_t[1] = s.driven[5];
    }
    if (s.driven[0] == 0. && s.driven[4] == 0. && b.periodicity != PER_PERIODIC) {
        s.vel_t[2] = s.vel[2] + dt * s.acc[2];
        s.disp_t[2] = s.disp[2] + dt * s.vel_t[2];
    } else {
        s.vel_t[2] = 0.;
        s.disp_t[2] = 0.;
    }
    double dr = s.disp_t[2] - 0.;
    for (int i = 0; i < n; i++) {
        double dx = b.X[i] - b.xc, dy = b.Y[i] - b.yc;
        s.utx[i] = s.disp_t[0] + dx * std::cos(dr) - dy * std::sin(dr) - dx;
        s.uty[i] = s.disp_t[1] + dx * std::sin(dr) + dy * std::cos(dr) - dy;
    }
}

// Body::Compute_error (Code/Body.h:2524-2551, Code/Node.h:253-272)
void compute_error(Sim& S, int bi) {
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    s.total_error = 0.;
    if (!b.rigid) {
        s.max_error = 0.;
        s.node_for_max = 0;
        for (int i = 0; i < b.n_nodes; i++) {
            s.xerr[i] = s.driven[0] == 1. ? 0. : s.ux[i] - s.utx[i];
            s.yerr[i] = s.driven[4] == 1. ? 0. : s.uy[i] - s.uty[i];
            s.enorm[i] = std::sqrt(s.xerr[i] * s.xerr[i] + s.yerr[i] * s.yerr[i]) * s.ide[i];
            s.total_error += s.enorm[i];
            if (s.enorm[i] == 0.) s.enorm[i] = 1.e-16;
            if (s.enorm[i] > s.max_error) { s.max_error = s.enorm[i]; s.node_for_max = i; }
        }
        return;
    }
    if (s.driven[0] != 1.) s.total_error += (s.disp[0] - s.disp_t[0]) * (s.disp[0] - s.disp_t[0]);
    if (s.driven[4] != 1.) s.total_error += (s.disp[1] - s.disp_t[1]) * (s.disp[1] - s.disp_t[1]);
    s.total_error = 0.1 * std::sqrt(s.total_error) / b.nodal_distance;
    if (s.total_error == 0.) s.total_error = 1.e-16;
    s.max_error = s.total_error;
    s.node_for_max = -1;
    for (int i = 0; i < b.n_nodes; i++) s.enorm[i] = s.total_error;
}

// Body::Compute_mass_scaling (Code/Body.h:2557-2585, Code/Node.h:274-320)
void compute_mass_scaling(Sim& S, int bi) {
    const Model& m = S.m;
    Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    const double thr = m.target_error * m.error_factor_mass_scaling;
    s.mass_ms = 0.;
    if (!b.rigid) {
        s.max_factor_ms = 0.;
        auto one = [&](double err, double ide, double& f, double& df) {
            if (std::fabs(err * ide) >= thr) {
                df = std::pow((std::fabs(err * ide) - thr), m.control_parameter_mass_scaling);
                f += df;
                if (f > m.max_mass_scaling) f = m.max_mass_scaling;
            } else {
                if (f != 1) f *= m.decrease_factor_mass_scaling;
                if (f < 1) f = 1;
            }
        };
        for (int i = 0; i < b.n_nodes; i++) {
            one(s.xerr[i], s.ide[i], s.fsx[i], s.dfsx[i]);
            one(s.yerr[i], s.ide[i], s.fsy[i], s.dfsy[i]);
            s.fsn[i] = (s.fsx[i] + s.fsy[i]) * 0.5;
            if (s.fsn[i] > s.max_factor_ms) s.max_factor_ms = s.fsn[i];
            s.msm[i] = std::sqrt(s.mx[i] * s.mx[i] * s.fsx[i] * s.fsx[i] + s.my[i] * s.my[i] * s.fsy[i] * s.fsy[i]);
            s.mass_ms += s.msm[i];
        }
        return;
    }
    if (s.total_error >= thr) {
        s.delta_factor_ms = std::pow((s.total_error - thr), m.control_parameter_mass_scaling);
        s.factor_ms += s.delta_factor_ms;
        if (s.factor_ms >= m.max_mass_scaling) s.factor_ms = m.max_mass_scaling;
    } else {
        if (s.factor_ms != 1) s.factor_ms *= m.decrease_factor_mass_scaling;
        if (s.factor_ms < 1) s.factor_ms = 1;
    }
    s.mass_ms = s.mass * s.factor_ms;
}

// Body::Restore (Code/Body.h:2625-2652): the stored nodes, borders, contact
// elements, works, mass-scaling increment and rigid current / displacement /
// velocity come back; the rest of the body (driven DOFs, accelerations,
// resultants, its mass-scaling factor, error scalars) stays as it is
void restore(State& live, const State& st) {
    State keep = live;
    live = st;
    for (int q = 0; q < 3; q++) {
        live.acc[q] = keep.acc[q]; live.rfc[q] = keep.rfc[q]; live.rfb[q] = keep.rfb[q];
        live.rfd[q] = keep.rfd[q]; live.rfn[q] = keep.rfn[q]; live.rfk[q] = keep.rfk[q]; live.rf[q] = keep.rf[q];
        live.disp_t[q] = keep.disp_t[q]; live.vel_t[q] = keep.vel_t[q];
    }
    for (int q = 0; q < 8; q++) live.driven[q] = keep.driven[q];
    live.total_error = keep.total_error; live.max_error = keep.max_error; live.node_for_max = keep.node_for_max;
    live.mass_ms = keep.mass_ms; live.factor_ms = keep.factor_ms; live.max_factor_ms = keep.max_factor_ms;
    live.neighbours = keep.neighbours; live.nb_neighbours = keep.nb_neighbours;
}

struct SolverScalars {
    bool stepped = false;
    int neval = 0;
    double max_error = 0, mean_error = 0, total_mass = 0, max_mass = 0;
};
SolverScalars g_solver;

// one half of an attempt: contacts and forces with the half step's law
// time step, then the half's end (Code/Solver.h:155-214 and 218-279)
void solver_half(Sim& S, int half, double dt) {
    Model& m = S.m;
    const int NB = (int)m.bodies.size();
    const double t = S.time;
#pragma omp parallel
    {
#pragma omp for schedule(dynamic)
        for (int i = 0; i < NB; i++) {
            if (m.bodies[i].status == ST_INACTIVE) continue;
            initialize_contact_forces(S, i);
            if (half == 2) update_bc(S, i, t + dt * 0.5);
            update_borders(S, i);
            update_contacts(S, i);
        }
#pragma omp for schedule(dynamic)
        for (int i = 0; i < NB; i++) {
            if (m.bodies[i].status == ST_INACTIVE) continue;
            update_bc_forces(S, i);
            update_body_forces(S, i);
            update_damping_forces(S, i);
            update_alid_forces(S, i, half == 1 ? t : t + dt * 0.5);
        }
    }
    for (int i = 0; i < NB; i++)
        if (m.bodies[i].status != ST_INACTIVE) update_contact_forces(S, i, dt * 0.5);
    for (int i = 0; i < NB; i++) send_contact_forces(S, i);
    std::vector<std::pair<int, int>> regions;
    for (int i = 0; i < NB; i++)
        if (!m.bodies[i].rigid)
            for (int r = 0; r < m.bodies[i].n_regions; r++) regions.push_back({i, r});
    const bool ms = m.scheme == SCH_MASS_SCALING || m.scheme == SCH_ADAPTIVE_MASS_SCALING;
#pragma omp parallel
    {
#pragma omp for schedule(dynamic)
        for (int k = 0; k < (int)regions.size(); k++) {
            if (m.bodies[regions[k].first].status == ST_INACTIVE) continue;
            update_internal_forces(S, regions[k].first, regions[k].second);
        }
#pragma omp for schedule(dynamic)
        for (int i = 0; i < NB; i++) {
            if (m.bodies[i].status == ST_INACTIVE) continue;
            sum_newton(S, i);
            if (half == 1) {
                euler_temporary(S, i, dt);
                euler_kinematics(S, i, dt * 0.5);
            } else {
                update_bc(S, i, t + dt);
                euler_kinematics(S, i, dt * 0.5);
                compute_error(S, i);
                if (ms) compute_mass_scaling(S, i);
            }
        }
    }
}

// Solver_step: attempts from the stored state until one is accepted; the
// time step S.m.dt is updated after every attempt
void solver_step(Sim& S) {
    Model& m = S.m;
    const int NB = (int)m.bodies.size();
    std::vector<State> stored = S.s;     // Body::Store (Code/Body.h:2593-2617)
    g_solver.neval = 0;
    g_solver.stepped = true;
    for (;;) {
        g_solver.neval++;
        double& dt = m.dt;
        solver_half(S, 1, dt);
        solver_half(S, 2, dt);
        double mean_error = 0., max_error = 0.;
        int nnodes = 0;
        for (int i = 0; i < NB; i++) {
            if (m.bodies[i].status == ST_INACTIVE) continue;
            mean_error += S.s[i].total_error;
            nnodes += m.bodies[i].n_nodes;
            if (S.s[i].max_error > max_error) max_error = S.s[i].max_error;
        }
        mean_error /= nnodes;
        g_solver.max_error = max_error;
        g_solver.mean_error = mean_error;
        if (m.scheme == SCH_MASS_SCALING || m.scheme == SCH_ADAPTIVE_MASS_SCALING) {
            double total_mass = 0., max_mass = 0.;
            for (int i = 0; i < NB; i++) {
                if (m.bodies[i].status == ST_INACTIVE) continue;
                total_mass += S.s[i].mass_ms;
                if (S.s[i].mass_ms * S.s[i].inverse_mass > max_mass) max_mass = S.s[i].mass_ms * S.s[i].inverse_mass;
            }
            g_solver.total_mass = total_mass;
            g_solver.max_mass = max_mass;
        }
        const double cup = m.contact_update_period;
        bool fail;
        double correction1;
        if (m.scheme == SCH_ADAPTIVE_EULER) {
            fail = max_error > m.target_error * m.accepted_ratio;
            correction1 = max_error / m.target_error;
        } else if (m.scheme == SCH_MASS_SCALING) {
            fail = max_error > m.target_error * m.accepted_ratio;
            correction1 = max_error * m.inv_target_error;
        } else {
            fail = g_solver.max_mass > m.accepted_ratio + 0.05;
            correction1 = fail ? g_solver.max_mass : 0.;
        }
        const double correction2 = 1. / m.accepted_ratio;
        if (fail)
            for (int i = 0; i < NB; i++) restore(S.s[i], stored[i]);
        else
            S.time += dt;
        if (m.scheme == SCH_ADAPTIVE_MASS_SCALING && !fail) dt = dt / std::pow(correction2, m.control_parameter);
        else if (correction1 > correction2) dt = dt / std::pow(correction1, m.control_parameter);
        else dt = dt / std::pow(correction2, m.control_parameter);
        if (dt > cup) dt = cup;
        if (!fail) {
            update_contact_pressures(S);
            return;
        }
    }
}

// Per-body energy scalars {kinetic energy from physical velocities and
// nodal masses, kinetic energy from velocity parameters (rigid bodies: from
// the body's velocity, mass and inertia), boundary work of the last step}:
// the expressions of the MONITOR_ENERGY block, Code/Monitoring.h:508-537,
// and of monitor type 9, Code/Monitoring.h:395-412, in their summation order.
std::vector<double> body_energy(const Sim& S, int bi) {
    const Body& b = S.m.bodies[bi];
    const State& s = S.s[bi];
    double ke = 0., kp = 0., dw = 0.;
    for (size_t j = 0; j < b.borders.size(); j++) {
        const Border& B = b.borders[j];
        for (size_t k = 0; k < B.nodes.size(); k++) {
            int node = B.nodes[k];
            dw += (s.xbp[j][k] * s.vx[node] + s.ybp[j][k] * s.vy[node]) * s.length[j][k] * S.m.dt;
        }
    }
    for (int j = 0; j < b.n_nodes; j++) {
        ke += 0.5 * s.vx[j] * s.vx[j] * s.mx[j];
        ke += 0.5 * s.vy[j] * s.vy[j] * s.my[j];
    }
    if (!b.rigid) {
        for (int j = 0; j < b.n_nodes; j++) {
            kp += 0.5 * s.vpx[j] * s.vpx[j] * s.mx[j];
            kp += 0.5 * s.vpy[j] * s.vpy[j] * s.my[j];
        }
    } else {
        kp += 0.5 * s.vel[0] * s.vel[0] * s.mass;
        kp += 0.5 * s.vel[1] * s.vel[1] * s.mass;
        kp += 0.5 * s.vel[2] * s.vel[2] * s.inertia;
    }
    return {ke, kp, dw};
}

// Body::Compute_nodal_stresses (Code/Body.h:2310-2519): stresses, strains,
// principal and equivalent stresses at every node of body bi from the
// derivatives of its DOF_TO_DISP row (its own nodes' displacement
// parameters), with the materials of Code/Model.h:19-73; 19 node fields one
// after another (SigmaI, SigmaII, SigmaIII, jacobian, Sxx, Syy, Sxy, Szz,
// Sigmaxx, Sigmayy, Sigmaxy, Sigmazz, SigmaTresca, SigmaVM, SigmaSph, Exx,
// Eyy, Exy, NormE).  Rigid bodies: zeros, jacobian 1.
std::vector<double> nodal_stresses(const Sim& S, int bi) {
    const Body& b = S.m.bodies[bi];
    const State& s = S.s[bi];
    const int n = b.n_nodes;
    std::vector<double> out(19 * (size_t)n, 0.);
    auto put = [&](int f, int i, double v) { out[(size_t)f * n + i] = v; };
    if (b.rigid) {
        for (int i = 0; i < n; i++) put(3, i, 1.);
        return out;
    }
    const Material& mat = S.m.materials[b.material];
    for (int i = 0; i < n; i++) {
        double J = 0, S11 = 0, S12 = 0, S22 = 0, S33 = 0;
        double Sxx = 0, Syy = 0, Sxy = 0, Szz = 0, Exx = 0, Eyy = 0, Exy = 0, NormE = 0;
        if (mat.type == MAT_ELASTIC_LINEAR) {
            double E11 = 0, E12 = 0, E22 = 0;
            for (int k = b.dptr[i]; k < b.dptr[i + 1]; k++) {
                const int j = b.didx[k];
                E11 += b.dphix[k] * s.upx[j];
                E12 += b.dphix[k] * s.upy[j] + b.dphiy[k] * s.upx[j];
                E22 += b.dphiy[k] * s.upy[j];
            }
            E12 = 0.5 * E12;
            const double Mu2 = 2 * mat.p[3], Lambda = mat.p[4];
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
            for (int k = b.dptr[i]; k < b.dptr[i + 1]; k++) {
                const int j = b.didx[k];
                F11 += b.dphix[k] * s.upx[j];
                F12 += b.dphiy[k] * s.upx[j];
                F21 += b.dphix[k] * s.upy[j];
                F22 += b.dphiy[k] * s.upy[j];
            }
            const double Mu = mat.p[3], Kappa = mat.p[4];
            const double F11F12 = F11 * F11 + F12 * F12, F21F22 = F21 * F21 + F22 * F22;
            J = F11 * F22 - F12 * F21;
            const double invJ = 1. / J, K = Kappa * (J - 1.);
            const double MuJ53 = Mu * std::pow(J, -1.666666666667);
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
        NormE = std::pow(Exx * Exx + Eyy * Eyy + 2. * Exy * Exy, 0.5);
        const double root = std::sqrt((Sxx + Syy) * (Sxx + Syy) - 4. * (Sxx * Syy - Sxy * Sxy));
        const double A = 0.5 * (Sxx + Syy + root), Bv = 0.5 * (Sxx + Syy - root);
        double s1, s2, s3;
        if (A > Bv && A > Szz) { s1 = A; if (Bv > Szz) { s2 = Bv; s3 = Szz; } else { s2 = Szz; s3 = Bv; } }
        else if (Bv > A && Bv > Szz) { s1 = Bv; if (A > Szz) { s2 = A; s3 = Szz; } else { s2 = Szz; s3 = A; } }
        else { s1 = Szz; if (Bv > A) { s2 = Bv; s3 = A; } else { s2 = A; s3 = Bv; } }
        const double vm = std::sqrt((Sxx - Syy) * (Sxx - Syy) + (Sxx - Szz) * (Sxx - Szz) + (Szz - Syy) * (Szz - Syy) +
                                    6. * Sxy * Sxy) * 0.70710678118655;
        const double v[19] = {s1, s2, s3, J, S11, S22, S12, S33, Sxx, Syy, Sxy, Szz, s1 - s3, vm,
                              0.333333333 * (Sxx + Syy + Szz), Exx, Eyy, Exy, NormE};
        for (int f = 0; f < 19; f++) put(f, i, v[f]);
    }
    return out;
}

void dump(Sim& S, const std::string& path) {
    DumpWriter w(path);
    w.scalar("time", S.time);
    w.scalar("iteration", (double)S.iteration);
    w.scalar("n_bodies", (double)S.s.size());
    const bool adaptive = S.m.scheme != SCH_EULER;
    const bool ms = S.m.scheme == SCH_MASS_SCALING || S.m.scheme == SCH_ADAPTIVE_MASS_SCALING;
    if (adaptive) w.scalar("dt", S.m.dt);
    if (adaptive && g_solver.stepped) {
        w.scalar("neval", (double)g_solver.neval);
        w.scalar("max_error", g_solver.max_error);
        w.scalar("mean_error", g_solver.mean_error);
        if (ms) {
            w.scalar("total_mass", g_solver.total_mass);
            w.scalar("max_mass", g_solver.max_mass);
        }
    }
    if (!S.m.monitored.empty() && !S.m.flags[F_NO_MONITORING] && S.iteration > 0) w.f64("monitor", S.monitor_row);
    for (size_t bi = 0; bi < S.s.size(); bi++) {
        const State& s = S.s[bi];
        const Body& b = S.m.bodies[bi];
        std::string p = "b" + std::to_string(bi) + ".";
        w.f64(p + "x_current", s.xc); w.f64(p + "y_current", s.yc);
        w.f64(p + "x_disp", s.ux); w.f64(p + "y_disp", s.uy);
        w.f64(p + "x_vel", s.vx); w.f64(p + "y_vel", s.vy);
        w.f64(p + "x_acc", s.ax); w.f64(p + "y_acc", s.ay);
        w.f64(p + "x_disp_param", b.rigid ? std::vector<double>(b.n_nodes, 0.) : s.upx);
        w.f64(p + "y_disp_param", b.rigid ? std::vector<double>(b.n_nodes, 0.) : s.upy);
        w.f64(p + "x_vel_param", b.rigid ? std::vector<double>(b.n_nodes, 0.) : s.vpx);
        w.f64(p + "y_vel_param", b.rigid ? std::vector<double>(b.n_nodes, 0.) : s.vpy);
        w.f64(p + "x_acc_param", b.rigid ? std::vector<double>(b.n_nodes, 0.) : s.apx);
        w.f64(p + "y_acc_param", b.rigid ? std::vector<double>(b.n_nodes, 0.) : s.apy);
        w.f64(p + "x_internal_force", s.fix); w.f64(p + "y_internal_force", s.fiy);
        w.f64(p + "x_contact_force", s.fcx); w.f64(p + "y_contact_force", s.fcy);
        w.f64(p + "x_body_force", s.fbx); w.f64(p + "y_body_force", s.fby);
        w.f64(p + "x_dirichlet_force", s.fdx); w.f64(p + "y_dirichlet_force", s.fdy);
        w.f64(p + "x_neumann_force", s.fnx); w.f64(p + "y_neumann_force", s.fny);
        w.f64(p + "x_damping_force", s.fkx); w.f64(p + "y_damping_force", s.fky);
        w.f64(p + "x_force", s.fx); w.f64(p + "y_force", s.fy);
        w.f64(p + "x_mass", s.mx); w.f64(p + "y_mass", s.my);
        w.f64(p + "x_alid_force", s.fax); w.f64(p + "y_alid_force", s.fay);
        w.f64(p + "rigid", {s.cur[0], s.cur[1], s.cur[2], s.disp[0], s.disp[1], s.disp[2], s.vel[0], s.vel[1],
                            s.vel[2], s.acc[0], s.acc[1], s.acc[2], s.rfc[0], s.rfc[1], s.rfc[2], s.rf[0], s.rf[1],
                            s.rf[2], s.mass, s.inertia, s.rfd[0], s.rfd[1], s.rfd[2],
                            s.rfn[0], s.rfn[1], s.rfn[2]});
        w.f64(p + "works", std::vector<double>(s.works, s.works + 7));
        w.scalar(p + "mass", s.mass);
        w.f64(p + "energy", body_energy(S, (int)bi));
        if (!S.m.deactivated.empty() || !S.m.monitored.empty()) w.f64(p + "status", {(double)b.status, s.damage});
        if (adaptive && g_solver.stepped) {
            std::vector<double> er = {s.total_error, s.max_error, (double)s.node_for_max};
            if (ms) { er.push_back(s.mass_ms); er.push_back(s.factor_ms); }
            w.f64(p + "error", er);
            if (ms) { w.f64(p + "x_msf", s.fsx); w.f64(p + "y_msf", s.fsy); }
        }
        w.i32(p + "neighbours", s.neighbours);
        std::vector<int> ci;
        std::vector<double> cd;
        for (const ContactElement& c : s.contacts) {
            int ints[] = {c.borderS, c.border_nodeS, c.nodeS, c.bodyM, c.borderM, c.shiftM, c.segmentM,
                          c.nodeM[0], c.nodeM[1], c.nodeM[2], c.nodeM[3], c.shiftMk[0], c.shiftMk[1], c.shiftMk[2],
                          c.shiftMk[3], c.nb_internal};
            ci.insert(ci.end(), ints, ints + 16);
            double ds[] = {c.shapeM[0], c.shapeM[1], c.shapeM[2], c.shapeM[3], c.gapn, c.vgapn, c.gapt, c.vgapt,
                           c.xsi, c.xnorm, c.ynorm, c.xtan, c.ytan, c.effective_mass, c.length, c.fx, c.fy,
                           c.internal.empty() ? 0. : c.internal[0]};
            cd.insert(cd.end(), ds, ds + 18);
        }
        w.i32(p + "contacts_i", ci);
        w.f64(p + "contacts_d", cd);
        std::vector<double> cint;
        for (const ContactElement& c : s.contacts)
            for (size_t k = 0; k < 3; k++) cint.push_back(k < c.internal.size() ? c.internal[k] : 0.);
        w.f64(p + "contacts_int", cint);
        if (std::getenv("MEL_DUMP_STRESS")) w.f64(p + "stress", nodal_stresses(S, (int)bi));
        for (size_t k = 0; k < b.borders.size(); k++) {
            std::string q = p + "border" + std::to_string(k) + ".";
            w.f64(q + "length", s.length[k]);
            w.f64(q + "x_contact_pressure", s.xcp[k]);
            w.f64(q + "y_contact_pressure", s.ycp[k]);
        }
    }
}

// --------------------------------------------------------------------------
// deactivation (Update_status, Code/Proximity.h:550-1007)
// --------------------------------------------------------------------------

// Body::Update_damage (Code/Body.h:2667-2700): per border node the smallest
// first internal variable of its contact elements that have one (1 when
// none), averaged over the border nodes; never decreases
void update_damage(Sim& S, int bi) {
    const Body& b = S.m.bodies[bi];
    State& s = S.s[bi];
    double total_nodes = 0., damage = 0.;
    const double previous = s.damage;
    for (size_t i = 0; i < b.borders.size(); i++) {
        const int nbn = (int)b.borders[i].nodes.size();
        total_nodes += nbn;
        for (int j = 0; j < nbn; j++) {
            double d = 1.;
            for (const ContactElement& c : s.contacts)
                if (c.borderS == (int)i && c.border_nodeS == j && c.nb_internal > 0 && c.internal[0] < d)
                    d = c.internal[0];
            damage += d;
        }
    }
    damage = damage / total_nodes;
    if (damage < previous) damage = previous;
    s.damage = damage;
}

// The rules in order.  x, y, r carry over from one rule to the next, as the
// reference's locals do (a deformable body's r, a force kind without a
// branch).  Types 0-3 read a node's or the body's position, displacement,
// velocity, acceleration (body level: x_mass-weighted means over the nodes
// of a deformable body, the body's own values for a rigid one); type 4 a
// node's force of one kind, or over the nodes (the reference tests the type
// and node count of body [2], the force kind, and reads the nodes of body
// [3]); type 8 the body's damage.  A rule that holds sets the body inactive.
void update_status(Sim& S) {
    Model& m = S.m;
    double x = 0., y = 0., r = 0.;
    auto compare = [](int cmp, double v, double thr) {
        return (cmp == 0 && v <= thr) || (cmp == 1 && v == thr) || (cmp == 2 && v >= thr);
    };
    for (const std::vector<double>& D : m.deactivated) {
        const int type = (int)D[0];
        if (type >= 0 && type <= 3) {
            const int bi = (int)D[2];
            const Body& b = m.bodies[bi];
            const State& s = S.s[bi];
            const std::vector<double>* nx[4] = {&s.xc, &s.ux, &s.vx, &s.ax};
            const std::vector<double>* ny[4] = {&s.yc, &s.uy, &s.vy, &s.ay};
            const double* rb[4] = {s.cur, s.disp, s.vel, s.acc};
            if (D[3] >= 0) {
                const int n = (int)D[3];
                x = (*nx[type])[n];
                y = (*ny[type])[n];
            } else if (!b.rigid) {
                double mass = 0.;
                x = 0.;
                y = 0.;
                for (int j = 0; j < b.n_nodes; j++) {
                    x += (*nx[type])[j] * s.mx[j];
                    y += (*ny[type])[j] * s.my[j];
                    mass += s.mx[j];
                }
                x /= mass;
                y /= mass;
            } else {
                x = rb[type][0];
                y = rb[type][1];
                r = rb[type][2];
            }
            const int comp = (int)D[1];
            double v;
            if (type == 0) v = comp == 0 ? x : (comp == 1 ? y : r);
            else v = comp == 0 ? x : (comp == 1 ? y : (comp == 2 ? std::pow(x * x + y * y, 0.5) : r));
            if (compare((int)D[4], v, D[5])) m.bodies[bi].status = ST_INACTIVE;
        } else if (type == 4) {
            const int kind = (int)D[2], bi = (int)D[3];
            const State& s = S.s[bi];
            const std::vector<double>* fx[7] = {&s.fx, &s.fix, &s.fcx, &s.fbx, &s.fdx, &s.fnx, &s.fkx};
            const std::vector<double>* fy[7] = {&s.fy, &s.fiy, &s.fcy, &s.fby, &s.fdy, &s.fny, &s.fky};
            if (D[4] >= 0) {
                const int n = (int)D[4];
                if (kind <= 6) { x = (*fx[kind])[n]; y = (*fy[kind])[n]; }
            } else if (!m.bodies[kind].rigid) {
                x = 0.;
                y = 0.;
                for (int j = 0; j < m.bodies[kind].n_nodes; j++)
                    if (kind <= 6) { x += (*fx[kind])[j]; y += (*fy[kind])[j]; }
            } else {
                // the rigid branch has no internal (1) or ALID (7) kind
                const double* rf[7] = {s.rf, nullptr, s.rfc, s.rfb, s.rfd, s.rfn, s.rfk};
                if (kind <= 6 && rf[kind]) { x = rf[kind][0]; y = rf[kind][1]; r = rf[kind][2]; }
            }
            const int comp = (int)D[1];
            const double v = comp == 0 ? x : (comp == 1 ? y : (comp == 2 ? std::pow(x * x + y * y, 0.5) : r));
            if (compare((int)D[5], v, D[6])) m.bodies[bi].status = ST_INACTIVE;
        } else if (type == 8) {
            const int bi = (int)D[1];
            update_damage(S, bi);
            if (compare((int)D[2], S.s[bi].damage, D[3])) m.bodies[bi].status = ST_INACTIVE;
        }
    }
}

// --------------------------------------------------------------------------
// Monitoring (Code/Monitoring.h:8-559): one row per step
// --------------------------------------------------------------------------
// The row: iteration, time, time step, and the adaptive schemes' neval,
// max_error, mean_error (locals the reference leaves uninitialised under
// Euler_step: 0 here), then one value per entry that yields one, then with
// MONITOR_ENERGY the boundary work, elastic energy (the laws set their
// energy to 0, Code/Model.h:28, 73) and kinetic energy.  x, y, r carry over
// between entries as the reference's locals do (they start uninitialised
// there; 0 here).
void monitor(Sim& S) {
    Model& m = S.m;
    const int NB = (int)m.bodies.size();
    std::vector<double> row = {(double)S.iteration, S.time, m.dt, 0., 0., 0.};
    if (m.scheme != SCH_EULER) { row[3] = g_solver.neval; row[4] = g_solver.max_error; row[5] = g_solver.mean_error; }
    double x = 0., y = 0., r = 0.;
    for (const std::vector<double>& D : m.monitored) {
        const int type = (int)D[0];
        if (type >= 0 && type <= 3) {
            const int bi = (int)D[2];
            const Body& b = m.bodies[bi];
            const State& s = S.s[bi];
            const std::vector<double>* nx[4] = {&s.xc, &s.ux, &s.vx, &s.ax};
            const std::vector<double>* ny[4] = {&s.yc, &s.uy, &s.vy, &s.ay};
            const double* rb[4] = {s.cur, s.disp, s.vel, s.acc};
            if (D[3] >= 0) {
                x = (*nx[type])[(int)D[3]];
                y = (*ny[type])[(int)D[3]];
            } else if (!b.rigid) {
                double mass = 0.;
                x = 0.;
                y = 0.;
                for (int j = 0; j < b.n_nodes; j++) {
                    x += (*nx[type])[j] * s.mx[j];
                    y += (*ny[type])[j] * s.my[j];
                    mass += s.mx[j];
                }
                x /= mass;
                y /= mass;
            } else {
                x = rb[type][0];
                y = rb[type][1];
                r = rb[type][2];
            }
            const int c = (int)D[1];
            if (type == 0) {
                if (c == 0) row.push_back(x);
                else if (c == 1) row.push_back(y);
                else if (c == 2) row.push_back(r);
            } else {
                if (c == 0) row.push_back(x);
                else if (c == 1) row.push_back(y);
                else if (c == 2) row.push_back(std::pow(x * x + y * y, 0.5));
                else if (c == 3) row.push_back(r);
            }
        } else if (type == 4) {
            const int kind = (int)D[2], bi = (int)D[3];
            const Body& b = m.bodies[bi];
            const State& s = S.s[bi];
            const std::vector<double>* fx[8] = {&s.fx, &s.fix, &s.fcx, &s.fbx, &s.fdx, &s.fnx, &s.fkx, &s.fax};
            const std::vector<double>* fy[8] = {&s.fy, &s.fiy, &s.fcy, &s.fby, &s.fdy, &s.fny, &s.fky, &s.fay};
            if (D[4] >= 0) {
                if (kind <= 7) { x = (*fx[kind])[(int)D[4]]; y = (*fy[kind])[(int)D[4]]; }
            } else if (!b.rigid) {
                x = 0.;
                y = 0.;
                for (int j = 0; j < m.bodies[kind].n_nodes; j++)
                    if (kind <= 7) { x += (*fx[kind])[j]; y += (*fy[kind])[j]; }
            } else {
                const double* rf[8] = {s.rf, nullptr, s.rfc, s.rfb, s.rfd, s.rfn, s.rfk, nullptr};
                if (kind <= 7 && rf[kind]) { x = rf[kind][0]; y = rf[kind][1]; r = rf[kind][2]; }
            }
            const int c = (int)D[1];
            if (c == 0) row.push_back(x);
            else if (c == 1) row.push_back(y);
            else if (c == 2) row.push_back(std::pow(x * x + y * y, 0.5));
            else if (c == 3) row.push_back(r);
        } else if (type == 5) {
            row.push_back(1.);          // Node::jacobian until Compute_nodal_stresses runs
        } else if (type == 6) {
            if ((int)D[1] <= 8) row.push_back(0.);    // Node::Sigma* until Compute_nodal_stresses runs
        } else if (type == 7) {
            const int bi = (int)D[2], c = (int)D[1];
            for (const ContactElement& e : S.s[bi].contacts) {
                if (e.borderS == (int)D[3] && e.border_nodeS == (int)D[4]) {
                    const double v[6] = {e.gapn, e.gapt, e.xsi, e.xnorm, e.ynorm, e.internal.empty() ? 0. : e.internal[0]};
                    if (c >= 0 && c <= 5) row.push_back(v[c]);
                    break;
                }
            }
        } else if (type == 8) {
            update_damage(S, (int)D[1]);
            row.push_back(S.s[(int)D[1]].damage);
        } else if (type == 9) {
            x = 0.;
            const int c = (int)D[1], bsel = (int)D[2];
            auto kinetic = [&](int n) {
                const Body& b = m.bodies[n];
                const State& s = S.s[n];
                if (!b.rigid) {
                    for (int j = 0; j < b.n_nodes; j++) {
                        x += 0.5 * s.vpx[j] * s.vpx[j] * s.mx[j];
                        x += 0.5 * s.vpy[j] * s.vpy[j] * s.my[j];
                    }
                } else {
                    x += 0.5 * s.vel[0] * s.vel[0] * s.mass;
                    x += 0.5 * s.vel[1] * s.vel[1] * s.mass;
                    x += 0.5 * s.vel[2] * s.vel[2] * s.inertia;
                }
            };
            // the elastic energies are zero (the laws' energy is 0)
            if (bsel >= 0 && c == 0) kinetic(bsel);
            else if (bsel == -1 && c == 0)
                for (int n = 0; n < NB; n++) kinetic(n);
            row.push_back(x);
        } else if (type == 10) {
            const int c = (int)D[1], bsel = (int)D[2];
            if (bsel >= 0) {
                if (c >= 0 && c <= 6) row.push_back(S.s[bsel].works[c]);
            } else if (bsel == -1) {
                x = 0.;
                for (int n = 0; n < NB; n++)
                    if (c >= 0 && c <= 6) x += S.s[n].works[c];
                row.push_back(x);
            }
        }
    }
    if (m.flags[F_MONITOR_ENERGY] == 1) {
        double dwork = 0., elastic = 0., kinetic = 0.;
        for (int i = 0; i < NB; i++) {
            const Body& b = m.bodies[i];
            const State& s = S.s[i];
            for (size_t j = 0; j < b.borders.size(); j++)
                for (size_t k = 0; k < b.borders[j].nodes.size(); k++) {
                    const int node = b.borders[j].nodes[k];
                    dwork += (s.xbp[j][k] * s.vx[node] + s.ybp[j][k] * s.vy[node]) * s.length[j][k] * m.dt;
                }
            for (int j = 0; j < b.n_nodes; j++) {
                kinetic += 0.5 * s.vx[j] * s.vx[j] * s.mx[j];
                kinetic += 0.5 * s.vy[j] * s.vy[j] * s.my[j];
            }
        }
        row.push_back(dwork);
        row.push_back(elastic);
        row.push_back(kinetic);
    }
    S.monitor_row = row;
}

int run(Sim& S, int nsteps, const char* out_dir, const int* dump_steps, int n_dump, double* seconds) {
    auto want = [&](int st) {
        for (int k = 0; k < n_dump; k++)
            if (dump_steps[k] == st) return true;
        return false;
    };
    char name[64];
    if (out_dir && want(0)) {
        std::snprintf(name, sizeof name, "/step_%06d.bin", 0);
        dump(S, std::string(out_dir) + name);
    }
    auto t0 = std::chrono::steady_clock::now();
    for (int st = 1; st <= nsteps; st++) {
        S.iteration++;
        if (S.m.scheme == SCH_EULER) euler_step(S);
        else solver_step(S);
        // Monitoring (Code/main.cpp:172-176), then the contact update
        if (!S.m.flags[F_NO_MONITORING]) monitor(S);
        if (S.time > S.next_contact_update - 1.e-6 * S.m.dt) {
            S.next_contact_update = S.next_contact_update + S.m.contact_update_period;
            update_proximity(S);
            update_status(S);
        }
        if (out_dir && want(st)) {
            std::snprintf(name, sizeof name, "/step_%06d.bin", st);
            dump(S, std::string(out_dir) + name);
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// the oracle's state written in the DYNAMIC_nnnnn layout by the library's
// writer (Write_dynamic, Code/IO.h:2351-2636)
bool write_state(Sim& S, const std::string& path, std::string& err) {
    Model m = S.m;
    m.time = S.time;
    m.number_iteration = S.iteration;
    m.next_contact_update = S.next_contact_update;
    for (size_t bi = 0; bi < m.bodies.size(); bi++) {
        Body& b = m.bodies[bi];
        const State& s = S.s[bi];
        b.ux = s.ux; b.uy = s.uy; b.vx = s.vx; b.vy = s.vy; b.ax = s.ax; b.ay = s.ay;
        b.upx = s.upx; b.upy = s.upy; b.vpx = s.vpx; b.vpy = s.vpy; b.apx = s.apx; b.apy = s.apy;
        for (int q = 0; q < 3; q++) { b.rigid_cur[q] = s.cur[q]; b.rigid_disp[q] = s.disp[q]; b.rigid_vel[q] = s.vel[q]; }
        for (int q = 0; q < 7; q++) b.works[q] = s.works[q];
        if (!b.rigid) {
            b.forces_out.assign(16 * (size_t)b.n_nodes, 0.);
            for (int i = 0; i < b.n_nodes; i++) {
                const double v[16] = {s.fix[i], s.fiy[i], s.fcx[i], s.fcy[i], 0., 0., s.fbx[i], s.fby[i],
                                      s.fdx[i], s.fdy[i], s.fnx[i], s.fny[i], s.fkx[i], s.fky[i], s.fx[i], s.fy[i]};
                for (int k = 0; k < 16; k++) b.forces_out[16 * (size_t)i + k] = v[k];
            }
        } else {
            const double* r[6] = {s.rfc, s.rfb, s.rfd, s.rfn, s.rfk, s.rf};
            for (int q = 0; q < 6; q++)
                for (int k = 0; k < 3; k++) b.rigid_forces_out[3 * q + k] = r[q][k];
        }
        b.neighbours = s.neighbours;
        b.contacts = s.contacts;
        for (size_t j = 0; j < b.borders.size(); j++) {
            b.borders[j].length_out = s.length[j];
            b.borders[j].xcp_out = s.xcp[j];
            b.borders[j].ycp_out = s.ycp[j];
            b.borders[j].xbp_out = s.xbp[j];
            b.borders[j].ybp_out = s.ybp[j];
        }
    }
    return write_dynamic_text(path, m, err);
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// Load the model (dir of ASCII files when save_index >= 0, else a binary
// container at `path`), run the initialisation and `nsteps` Euler steps with
// the contact-update cadence, dumping the listed steps into out_dir (may be
// null).  Returns 0 on success.
int oracle_run(const char* path, int save_index, int nsteps, const char* out_dir, const int* dump_steps,
               int n_dump, int nthreads, double* seconds) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    Sim S;
    std::string err;
    bool ok = save_index >= 0 ? load_model_files(path, save_index, S.m, err) : load_model_binary(path, S.m, err);
    if (!ok) { g_err = err; return 1; }
    for (auto& b : S.m.bodies)
        if (S.m.flags[F_UPDATE_STIFFNESS] && !b.mat_dof.empty()) { g_err = "UPDATE_STIFFNESS_MATRIX is not supported"; return 1; }
    setup(S);
    g_solver = SolverScalars();
    return run(S, nsteps, out_dir, dump_steps, n_dump, seconds);
}

// Load and initialise as oracle_run, run `nsteps` steps, then write the state
// to `dynamic_path` in the DYNAMIC_nnnnn layout.  Returns 0 on success.
int oracle_write_after(const char* path, int save_index, int nsteps, const char* dynamic_path, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    Sim S;
    std::string err;
    bool ok = save_index >= 0 ? load_model_files(path, save_index, S.m, err) : load_model_binary(path, S.m, err);
    if (!ok) { g_err = err; return 1; }
    setup(S);
    g_solver = SolverScalars();
    run(S, nsteps, nullptr, nullptr, 0, nullptr);
    if (!write_state(S, dynamic_path, err)) { g_err = err; return 1; }
    return 0;
}

}  // extern "C"
