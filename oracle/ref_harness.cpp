// Test infrastructure (never shipped, never linked into the product).
//
// Drives the reference MELODY 2D solver, compiled from its own headers under
// /root/reference/Code, on a directory of model files and dumps FP64 state in
// the record format of dump_format.hpp.  The sequence is exactly the
// reference driver's: Load_static + Load_dynamic, the initialisation block of
// Code/main.cpp:103-126, then for each iteration Euler_step
// (Code/main.cpp:148-154) followed by the contact-update cadence of
// Code/main.cpp:191-196.  No graphics, monitoring or saves are run; they do
// not feed back into the step.
//
// usage: melody_ref <model_dir> <nsteps> <out_dir> <dump_steps,comma,list> [save_index]
// (save_index: the DYNAMIC_<save_index>.asc file to start from, default 0)
#include <iostream>
#include <fstream>
#include <iomanip>
#include <vector>
#include <string>
#include <cmath>
#include <algorithm>
#include <sstream>
#include <omp.h>
#include <unistd.h>
using namespace std;
#include "Tools.h"
#include "Material.h"
#include "Contact_Law.h"
#include "Node.h"
#include "Gauss.h"
#include "Border.h"
#include "Contact_Element.h"
#include "Body.h"
#include "Spy.h"
#include "IO.h"
#include "Solver.h"
#include "Proximity.h"
#include "Graphic.h"
#include "Monitoring.h"

#include "dump_format.hpp"

// Per-body energy scalars with the expressions of the reference's monitors:
// kinetic energy from physical velocities and nodal masses and the boundary
// work increment of the MONITOR_ENERGY block (Code/Monitoring.h:508-537), and
// kinetic energy from velocity parameters (monitor type 9,
// Code/Monitoring.h:395-412; rigid bodies from their own velocity, mass and
// inertia).
static vector<double> body_energy(Body& B, double Deltat) {
    double ke = 0., kp = 0., dw = 0.;
    for (int j = 0; j < B.nb_borders; j++)
        for (int k = 0; k < B.borders[j].number_border_nodes; k++) {
            double px = B.borders[j].x_bc_pressure[k];
            double py = B.borders[j].y_bc_pressure[k];
            double length = B.borders[j].length[k];
            double vx = B.nodes[B.borders[j].border_nodes[k]].x_velocity;
            double vy = B.nodes[B.borders[j].border_nodes[k]].y_velocity;
            dw += (px * vx + py * vy) * length * Deltat;
        }
    for (int j = 0; j < B.nb_nodes; j++) {
        ke += 0.5 * B.nodes[j].x_velocity * B.nodes[j].x_velocity * B.nodes[j].x_mass;
        ke += 0.5 * B.nodes[j].y_velocity * B.nodes[j].y_velocity * B.nodes[j].y_mass;
    }
    if (B.type == "deformable") {
        for (int j = 0; j < B.nb_nodes; j++) {
            kp += 0.5 * B.nodes[j].x_velocity_parameter * B.nodes[j].x_velocity_parameter * B.nodes[j].x_mass;
            kp += 0.5 * B.nodes[j].y_velocity_parameter * B.nodes[j].y_velocity_parameter * B.nodes[j].y_mass;
        }
    } else if (B.type == "rigid") {
        kp += 0.5 * B.x_velocity * B.x_velocity * B.mass;
        kp += 0.5 * B.y_velocity * B.y_velocity * B.mass;
        kp += 0.5 * B.r_velocity * B.r_velocity * B.inertia;
    }
    return {ke, kp, dw};
}

// Records to write: all, or (MEL_DUMP_FIELDS set) only the scalar records
// and the per-body records whose names are listed, comma-separated.
static bool keep_record(const string& name) {
    static const char* env = getenv("MEL_DUMP_FIELDS");
    if (!env || !*env) return true;
    size_t dot = name.find('.');
    string key = dot == string::npos ? name : name.substr(dot + 1);
    if (key.compare(0, 6, "border") == 0 && key.find('.') != string::npos) key = key.substr(key.find('.') + 1);
    string list = string(",") + env + ",";
    return dot == string::npos || list.find("," + key + ",") != string::npos;
}

struct FilteredWriter {
    DumpWriter w;
    explicit FilteredWriter(const string& path) : w(path) {}
    bool ok() const { return w.ok(); }
    void f64(const string& n, const vector<double>& v) { if (keep_record(n)) w.f64(n, v); }
    void i32(const string& n, const vector<int>& v) { if (keep_record(n)) w.i32(n, v); }
    void scalar(const string& n, double x) { if (keep_record(n)) w.scalar(n, x); }
};

// materials of the run, for the nodal stresses dumped when MEL_DUMP_STRESS is set
static vector<Material>* g_materials = nullptr;

// the adaptive schemes' step control (Code/Solver.h:98-486), dumped when
// the scheme is not Euler
struct SolverScalars {
    bool adaptive = false, mass_scaling = false, stepped = false;
    int neval = 0;
    double max_error = 0, mean_error = 0, total_mass = 0, max_mass = 0;
};
static SolverScalars g_solver;
// the model has DEACTIVATION rules or MONITORING entries: the bodies'
// status and damage are dumped
static bool g_has_rules = false;
// the model has MONITORING entries: the last step's Monitoring row is dumped
static bool g_has_monitor = false;
static vector<double> g_monitor_row;

static void dump_state(const string& path, vector<Body>& Bodies, double Time, int iteration, double Deltat) {
    FilteredWriter w(path);
    if (!w.ok()) { cerr << "cannot write " << path << endl; exit(2); }
    w.scalar("time", Time);
    w.scalar("iteration", (double)iteration);
    w.scalar("n_bodies", (double)Bodies.size());
    if (g_has_monitor && iteration > 0) w.f64("monitor", g_monitor_row);
    if (g_solver.adaptive) w.scalar("dt", Deltat);
    if (g_solver.adaptive && g_solver.stepped) {
        w.scalar("neval", (double)g_solver.neval);
        w.scalar("max_error", g_solver.max_error);
        w.scalar("mean_error", g_solver.mean_error);
        if (g_solver.mass_scaling) {
            w.scalar("total_mass", g_solver.total_mass);
            w.scalar("max_mass", g_solver.max_mass);
        }
    }
    for (size_t b = 0; b < Bodies.size(); b++) {
        Body& B = Bodies[b];
        string p = "b" + to_string(b) + ".";
        int n = B.nb_nodes;
        vector<double> v(n);
#define NODE_FIELD(NAME, EXPR) for (int i = 0; i < n; i++) { Node& N = B.nodes[i]; (void)N; v[i] = (EXPR); } w.f64(p + NAME, v);
        NODE_FIELD("x_current", N.x_current)
        NODE_FIELD("y_current", N.y_current)
        NODE_FIELD("x_disp", N.x_displacement)
        NODE_FIELD("y_disp", N.y_displacement)
        NODE_FIELD("x_vel", N.x_velocity)
        NODE_FIELD("y_vel", N.y_velocity)
        NODE_FIELD("x_acc", N.x_acceleration)
        NODE_FIELD("y_acc", N.y_acceleration)
        NODE_FIELD("x_disp_param", N.x_displacement_parameter)
        NODE_FIELD("y_disp_param", N.y_displacement_parameter)
        NODE_FIELD("x_vel_param", N.x_velocity_parameter)
        NODE_FIELD("y_vel_param", N.y_velocity_parameter)
        NODE_FIELD("x_acc_param", N.x_acceleration_parameter)
        NODE_FIELD("y_acc_param", N.y_acceleration_parameter)
        NODE_FIELD("x_internal_force", N.x_internal_force)
        NODE_FIELD("y_internal_force", N.y_internal_force)
        NODE_FIELD("x_contact_force", N.x_contact_force)
        NODE_FIELD("y_contact_force", N.y_contact_force)
        NODE_FIELD("x_body_force", N.x_body_force)
        NODE_FIELD("y_body_force", N.y_body_force)
        NODE_FIELD("x_dirichlet_force", N.x_dirichlet_force)
        NODE_FIELD("y_dirichlet_force", N.y_dirichlet_force)
        NODE_FIELD("x_neumann_force", N.x_neumann_force)
        NODE_FIELD("y_neumann_force", N.y_neumann_force)
        NODE_FIELD("x_damping_force", N.x_damping_force)
        NODE_FIELD("y_damping_force", N.y_damping_force)
        NODE_FIELD("x_force", N.x_force)
        NODE_FIELD("y_force", N.y_force)
        NODE_FIELD("x_mass", N.x_mass)
        NODE_FIELD("y_mass", N.y_mass)
        NODE_FIELD("x_alid_force", N.x_alid_force)
        NODE_FIELD("y_alid_force", N.y_alid_force)
        w.f64(p + "rigid", {B.x_current, B.y_current, B.r_current, B.x_displacement, B.y_displacement,
                           B.r_displacement, B.x_velocity, B.y_velocity, B.r_velocity, B.x_acceleration,
                           B.y_acceleration, B.r_acceleration, B.x_contact_force, B.y_contact_force,
                           B.r_contact_force, B.x_force, B.y_force, B.r_force, B.mass, B.inertia,
                           B.x_dirichlet_force, B.y_dirichlet_force, B.r_dirichlet_force,
                           B.x_neumann_force, B.y_neumann_force, B.r_neumann_force});
        w.f64(p + "works", {B.internal_work, B.contact_work, B.body_work, B.dirichlet_work,
                           B.neumann_work, B.damping_work, B.alid_work});
        w.scalar(p + "mass", B.mass);
        w.f64(p + "energy", body_energy(B, Deltat));
        if (g_has_rules) w.f64(p + "status", {B.status == "inactive" ? 1. : (B.status == "nogravity" ? 2. : 0.), B.damage});
        if (g_solver.adaptive && g_solver.stepped) {
            // Body::Compute_error / Compute_mass_scaling results (Code/Body.h:2524-2585),
            // once a step has set them (the body's mass_mass_scaling has no initial value)
            vector<double> er = {B.total_error, B.max_error, (double)B.node_for_max_error};
            if (g_solver.mass_scaling) { er.push_back(B.mass_mass_scaling); er.push_back(B.factor_mass_scaling); }
            w.f64(p + "error", er);
            if (g_solver.mass_scaling) {
                NODE_FIELD("x_msf", N.x_factor_mass_scaling)
                NODE_FIELD("y_msf", N.y_factor_mass_scaling)
            }
        }
        vector<int> nb;
        for (int k = 0; k < B.nb_neighbours; k++)
            for (int c = 0; c < 3; c++) nb.push_back(B.neighbours[k][c]);
#undef NODE_FIELD
        w.i32(p + "neighbours", nb);
        vector<int> ci;
        vector<double> cd;
        for (int k = 0; k < B.nb_contact_elements; k++) {
            Contact_element& C = B.contact_elements[k];
            int ints[] = {C.borderS, C.border_nodeS, C.nodeS, C.bodyM, C.borderM, C.shiftM, C.segmentM,
                          C.nodeM0, C.nodeM1, C.nodeM2, C.nodeM3, C.shiftM0, C.shiftM1, C.shiftM2,
                          C.shiftM3, C.nb_internal};
            ci.insert(ci.end(), ints, ints + 16);
            double ds[] = {C.shapeM0, C.shapeM1, C.shapeM2, C.shapeM3, C.gapn, C.vgapn, C.gapt, C.vgapt,
                           C.xsi, C.xnorm, C.ynorm, C.xtan, C.ytan, C.effective_mass, C.length, C.fx,
                           C.fy, C.internal.empty() ? 0.0 : C.internal[0]};
            cd.insert(cd.end(), ds, ds + 18);
        }
        w.i32(p + "contacts_i", ci);
        w.f64(p + "contacts_d", cd);
        vector<double> cint;
        for (int k = 0; k < B.nb_contact_elements; k++) {
            Contact_element& C = B.contact_elements[k];
            // a fresh element's internal vector holds one value; the
            // three-variable laws make it three
            for (int q = 0; q < 3; q++) cint.push_back(q < (int)C.internal.size() ? C.internal[q] : 0.0);
        }
        w.f64(p + "contacts_int", cint);
        if (getenv("MEL_DUMP_STRESS") && g_materials) {
            // Body::Compute_nodal_stresses (Code/Body.h:2310-2519), its 19
            // node fields one after another
            B.Compute_nodal_stresses((int)g_materials->size(), *g_materials);
            vector<double> st;
            st.reserve(19 * (size_t)n);
#define STRESS_FIELD(EXPR) for (int i = 0; i < n; i++) { Node& N = B.nodes[i]; st.push_back(EXPR); }
            STRESS_FIELD(N.SigmaI) STRESS_FIELD(N.SigmaII) STRESS_FIELD(N.SigmaIII) STRESS_FIELD(N.jacobian)
            STRESS_FIELD(N.Sxx) STRESS_FIELD(N.Syy) STRESS_FIELD(N.Sxy) STRESS_FIELD(N.Szz)
            STRESS_FIELD(N.Sigmaxx) STRESS_FIELD(N.Sigmayy) STRESS_FIELD(N.Sigmaxy) STRESS_FIELD(N.Sigmazz)
            STRESS_FIELD(N.SigmaTresca) STRESS_FIELD(N.SigmaVM) STRESS_FIELD(N.SigmaSph)
            STRESS_FIELD(N.Exx) STRESS_FIELD(N.Eyy) STRESS_FIELD(N.Exy) STRESS_FIELD(N.NormE)
#undef STRESS_FIELD
            w.f64(p + "stress", st);
        }
        for (int k = 0; k < B.nb_borders; k++) {
            string q = p + "border" + to_string(k) + ".";
            w.f64(q + "length", B.borders[k].length);
            w.f64(q + "x_contact_pressure", B.borders[k].x_contact_pressure);
            w.f64(q + "y_contact_pressure", B.borders[k].y_contact_pressure);
        }
    }
}

int main(int argc, char** argv) {
    if (argc < 5) {
        cerr << "usage: melody_ref <model_dir> <nsteps> <out_dir> <dump_steps>" << endl;
        return 1;
    }
    string out_dir = argv[3];
    if (out_dir[0] != '/') {
        char buf[4096];
        if (!getcwd(buf, sizeof buf)) return 1;
        out_dir = string(buf) + "/" + out_dir;
    }
    int nsteps = atoi(argv[2]);
    vector<int> dump_steps;
    {
        stringstream ss(argv[4]);
        string tok;
        while (getline(ss, tok, ',')) if (!tok.empty()) dump_steps.push_back(atoi(tok.c_str()));
    }
    if (chdir(argv[1]) != 0) { cerr << "cannot enter " << argv[1] << endl; return 1; }
    // silence the reference's console output
    ofstream devnull("/dev/null");
    streambuf* old = cout.rdbuf(devnull.rdbuf());

    string Simulation_name;
    int Nb_materials;
    vector<Material> Materials;
    int Nb_contact_laws;
    vector<Contact_law> Contact_laws;
    string Solver;
    double Tini, Deltat, Tend, Time;
    double Target_error, Inv_Target_error, Control_parameter, Accepted_ratio;
    double Max_mass_scaling, Control_parameter_mass_scaling, Error_factor_mass_scaling, Decrease_factor_mass_scaling;
    double Save_period, Print_period, Contact_update_period;
    double Next_save, Next_print, Next_contact_update;
    int Number_save(argc > 5 ? atoi(argv[5]) : 0), Number_print, Number_iteration;
    double Xmin_period, Xmax_period, Penalty = 0;
    double Xgravity, Ygravity;
    int Activate_plot;
    double Xmin_plot, Xmax_plot, Ymin_plot, Ymax_plot;
    int Nb_bodies;
    vector<Body> Bodies;
    int Nb_monitored = 0;
    vector<vector<double>> Monitored;
    int Nb_deactivated = 0;
    vector<vector<double>> Deactivated;
    int Nb_spies;
    vector<Spy> Spies;
    int Nb_regions = 0;
    vector<vector<int>> Regions;
    vector<int> flags(11);
    vector<int> To_Plot(41);
    vector<vector<int>> Contacts_Table;

    g_materials = &Materials;
    Load_static(Simulation_name, Nb_materials, Materials, Nb_contact_laws, Contact_laws, Contacts_Table,
                Solver, Tini, Deltat, Tend, Target_error, Inv_Target_error, Control_parameter, Accepted_ratio,
                Max_mass_scaling, Control_parameter_mass_scaling, Error_factor_mass_scaling,
                Decrease_factor_mass_scaling, Save_period, Print_period, Contact_update_period, Xmin_period,
                Xmax_period, Penalty, Xgravity, Ygravity, Activate_plot, Xmin_plot, Xmax_plot, Ymin_plot,
                Ymax_plot, Nb_monitored, Monitored, Nb_deactivated, Deactivated, Nb_spies, Spies, Nb_regions,
                Regions, Nb_bodies, Bodies, To_Plot);
    Load_dynamic(Number_save, Number_print, Time, Number_iteration, Deltat, Solver, Next_save, Next_print,
                 Next_contact_update, Xmin_period, Xmax_period, Bodies, flags);
    g_has_rules = Nb_deactivated > 0 || Nb_monitored > 0;
    g_has_monitor = Nb_monitored > 0 && flags[8] == 0;
    vector<vector<double>> monitoring;

    // initialisation block, Code/main.cpp:103-126
    for (int i = 0; i < Nb_bodies; i++) Bodies[i].Update_borders(Xmin_period, Xmax_period);
    for (int i = 0; i < Nb_bodies; i++) Bodies[i].Update_bc(Time);
    Update_proximity(Nb_bodies, Bodies, Xmin_period, Xmax_period, flags);
    if (flags[3] == 1)
        Initialize_CZM(Nb_bodies, Bodies, Nb_contact_laws, Contact_laws, flags, Xmin_period, Xmax_period);
#pragma omp parallel
    {
#pragma omp for schedule(dynamic)
        for (int i = 0; i < Nb_bodies; i++) Bodies[i].Update_material(Nb_materials, Materials, flags);
    }
    Update_proximity(Nb_bodies, Bodies, Xmin_period, Xmax_period, flags);
    if (flags[3] == 1)
        Initialize_CZM(Nb_bodies, Bodies, Nb_contact_laws, Contact_laws, flags, Xmin_period, Xmax_period);

    g_solver.adaptive = Solver != "Euler";
    g_solver.mass_scaling = Solver == "Mass_Scaling" || Solver == "Adaptive_Mass_Scaling";
    auto want = [&](int s) { return find(dump_steps.begin(), dump_steps.end(), s) != dump_steps.end(); };
    char name[64];
    if (want(0)) {
        snprintf(name, sizeof name, "/step_%06d.bin", 0);
        dump_state(out_dir + name, Bodies, Time, Number_iteration, Deltat);
    }
    double t0 = omp_get_wtime();
    for (int s = 1; s <= nsteps; s++) {
        Number_iteration++;
        // Code/main.cpp:148-166
        if (Solver == "Euler") {
            Euler_step(Nb_bodies, Bodies, Nb_regions, Regions, Nb_materials, Materials, Nb_contact_laws,
                       Contact_laws, Contacts_Table, Tend, Xmin_period, Xmax_period, Penalty, Xgravity, Ygravity,
                       Time, Deltat, Number_iteration, flags);
        } else {
            Solver_step(Nb_bodies, Bodies, Nb_regions, Regions, Nb_materials, Materials, Nb_contact_laws,
                        Contact_laws, Contacts_Table, Tend, Xmin_period, Xmax_period, Penalty, Xgravity, Ygravity,
                        Time, Deltat, Number_iteration, Next_contact_update, Contact_update_period, Target_error,
                        Inv_Target_error, Control_parameter, Accepted_ratio, Max_mass_scaling,
                        Control_parameter_mass_scaling, Error_factor_mass_scaling, Decrease_factor_mass_scaling,
                        g_solver.neval, g_solver.max_error, g_solver.mean_error, flags, g_solver.total_mass,
                        g_solver.max_mass, Solver);
            g_solver.stepped = true;
        }
        // Monitoring, Code/main.cpp:172-176 (the harness's neval, max_error
        // and mean_error start at 0 where the driver's are uninitialised)
        if (flags[8] == 0) {
            Monitoring(Nb_bodies, Bodies, Time, Deltat, Number_iteration, g_solver.neval, g_solver.max_error,
                       g_solver.mean_error, flags, monitoring, Nb_monitored, Monitored);
            g_monitor_row = monitoring.back();
            monitoring.clear();
        }
        // contact-update cadence, Code/main.cpp:191-196
        if (Time > Next_contact_update - 1.e-6 * Deltat) {
            Next_contact_update = Next_contact_update + Contact_update_period;
            Update_proximity(Nb_bodies, Bodies, Xmin_period, Xmax_period, flags);
            Update_status(Nb_bodies, Bodies, Nb_deactivated, Deactivated);
        }
        if (want(s)) {
            snprintf(name, sizeof name, "/step_%06d.bin", s);
            dump_state(out_dir + name, Bodies, Time, Number_iteration, Deltat);
        }
    }
    double t1 = omp_get_wtime();
    cout.rdbuf(old);
    long long nn = 0;
    for (int i = 0; i < Nb_bodies; i++) nn += Bodies[i].nb_nodes;
    printf("{\"nodes\": %lld, \"steps\": %d, \"seconds\": %.6f, \"threads\": %d}\n", nn, nsteps, t1 - t0,
           omp_get_max_threads());
    return 0;
}
