// Host-side model: the reference's model files parsed into flat per-body
// arrays (SoA), plus a binary container for models too large for ASCII.
//
// Format semantics follow the reference loaders: Load_static
// (Code/IO.h:598-1862) and Load_dynamic (Code/IO.h:1871-2343).  Material
// parameters are converted exactly as Code/IO.h:678 / :687 do, the Gauss
// points of a body are cut into ceil(G/R) contiguous regions
// (Code/IO.h:1743-1755), the four-node segment stencils are built as
// Body::Initialize_segments does (Code/Body.h:403-605).
#pragma once
#include <cstdint>
#include <istream>
#include <string>
#include <vector>

namespace mel {

enum BcType { BC_NONE = 0, BC_DRIVEN = 1, BC_SOFT = 2, BC_ORIENTED = 3, BC_FOLLOWING = 4 };
enum Periodicity { PER_CLOSED = 0, PER_SIMPLE = 1, PER_PERIODIC = 2 };
enum MaterialType { MAT_ELASTIC_LINEAR = 0, MAT_NEOHOOKEAN = 1 };
enum LawType {
    LAW_FRICTIONLESS = 0, LAW_COHESION = 1, LAW_MOHR_COULOMB = 2, LAW_DAMPED_MOHR_COULOMB = 3,
    LAW_TWO_SLOPES = 4, LAW_BONDED_MOHR_COULOMB = 5, LAW_CZM_LINEAR = 6, LAW_CZM_FATIGUE = 7
};
enum Status { ST_ACTIVE = 0, ST_INACTIVE = 1, ST_NOGRAVITY = 2 };
// time-integration schemes of the SOLVER section (Code/IO.h:821-890)
enum Scheme { SCH_EULER = 0, SCH_ADAPTIVE_EULER = 1, SCH_MASS_SCALING = 2, SCH_ADAPTIVE_MASS_SCALING = 3 };
int scheme_of(const std::string& name);     // -1: not a scheme this library runs

// flag slots of the DYNAMIC file, Code/IO.h:1912-1933
enum Flag {
    F_KILL_VELOCITY = 0, F_MONITOR_ENERGY, F_MONITOR_BOUNDARIES, F_INITIALIZE_CZM, F_UPDATE_MASS,
    F_UPDATE_STIFFNESS, F_UPDATE_DAMPING, F_NO_LOG, F_NO_MONITORING, F_NO_SELF_CONTACT, F_RESET_WORK,
    F_COUNT
};

struct History {
    std::vector<double> t, v;
};

struct BorderBC {
    int type = BC_NONE;
    History hist;             // Dirichlet (velocity) or Neumann (pressure) history
    double stiffness = 0, damping = 0;
};

struct Border {
    int periodicity = PER_CLOSED;
    bool linear = true;       // interpolant "Linear" (Spline is not supported)
    int before = 0, after = 0;
    std::vector<int> nodes;   // a closed border repeats its first node at the end
    BorderBC xbc, ybc;
    // four-node stencils (Body::Initialize_segments)
    std::vector<int> node0, node1, node2, node3, shift0, shift1, shift2, shift3;
    std::vector<double> initial_length;
    // state written by Write_dynamic (Code/IO.h:2530-2630): current tributary
    // lengths, contact and BC pressures per slot
    std::vector<double> length_out, xcp_out, ycp_out, xbp_out, ybp_out;
};

struct Material {
    std::string name;
    int type = MAT_NEOHOOKEAN;
    double raw[6] = {0, 0, 0, 0, 0, 0};     // rho alpha beta E nu c as written
    double p[6] = {0, 0, 0, 0, 0, 0};       // rho alpha beta mu lambda|kappa c
};

struct ContactLaw {
    std::string m1, m2;
    int type = LAW_FRICTIONLESS;
    bool fixed_length = false;               // "Fixed" vs "Evolutive"
    std::vector<double> p;
};

struct ContactElement {                      // Contact_element (Code/Contact_Element.h:11-44)
    int borderS = 0, border_nodeS = 0, nodeS = 0, bodyM = 0, borderM = 0, shiftM = 0, segmentM = 0;
    int nodeM[4] = {0, 0, 0, 0}, shiftMk[4] = {0, 0, 0, 0};
    double shapeM[4] = {0, 0, 0, 0};
    double gapn = 0, vgapn = 0, gapt = 0, vgapt = 0, xsi = 0, xnorm = 0, ynorm = 0, xtan = 0, ytan = 0;
    double effective_mass = 0, length = 0, fx = 0, fy = 0;
    int nb_internal = 0;
    std::vector<double> internal;
};

// absorbing layers (ALID section of a body, Code/IO.h:1395-1485): bounds
// xmin xmax ymin ymax, range, alpha in/out, beta in/out, exponent, and the
// velocity histories of the x and y directions (empty: that direction off)
struct Alid {
    bool on = false;                          // Body::flag_alid
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0, range = 0;
    double alphain = 0, alphaout = 0, betain = 0, betaout = 0, exponent = 0;
    History xh, yh;
};

struct Body {
    bool rigid = false;
    int material = 0;
    int periodicity = PER_SIMPLE;
    int status = ST_ACTIVE;
    double nodal_distance = 0, detection_distance = 0, contact_distance = 0;
    double density = 0;
    int n_nodes = 0;
    std::vector<double> X, Y, domain, mx, my, imx, imy;
    std::vector<Border> borders;
    // deformable
    int n_gauss = 0, n_regions = 1;
    std::vector<double> gx, gy, gw, gjac;
    std::vector<int> gptr, gidx;
    std::vector<double> gphi, gphix, gphiy;
    std::vector<int> dptr, didx;
    std::vector<double> dphi, dphix, dphiy;
    std::vector<int> cells;                  // 3 per triangle
    std::vector<int> mat_dof;                // 2 per matrix entry (upper triangle)
    std::vector<double> mat_stiffness, mat_damping;
    Alid alid;
    std::vector<double> alid_alpha, alid_beta;   // per node (Node::alid_alpha, alid_beta, Code/IO.h:1548-1584)
    // rigid
    double xc = 0, yc = 0, mass = 0, inertia = 0, inverse_mass = 0, inverse_inertia = 0;
    // kinematic state as loaded (deformable: per node; rigid: body level)
    std::vector<double> ux, uy, vx, vy, ax, ay, upx, upy, vpx, vpy, apx, apy;
    double rigid_cur[3] = {0, 0, 0}, rigid_disp[3] = {0, 0, 0}, rigid_vel[3] = {0, 0, 0};
    double works[7] = {0, 0, 0, 0, 0, 0, 0};
    // force fields written by Write_dynamic (Code/IO.h:2455-2500): deformable,
    // 16 per node in the file's column order; rigid, the contact, body,
    // Dirichlet, Neumann, damping and total resultants (x, y, moment each)
    std::vector<double> forces_out;
    double rigid_forces_out[18] = {0};
    std::vector<int> neighbours;             // 3 per neighbour
    std::vector<ContactElement> contacts;
};

struct Model {
    std::string name;
    std::vector<Material> materials;
    std::vector<ContactLaw> laws;
    std::vector<int> contacts_table;          // n_materials^2, -1 = no law
    std::string solver = "Euler";
    int scheme = SCH_EULER;
    double tini = 0, dt = 0, tend = 0;
    // error control of the adaptive schemes (Code/Solver.h:98-486): target
    // error, control exponent, accepted ratio; mass scaling: maximum factor,
    // control exponent, error factor, decrease factor
    double target_error = 0, inv_target_error = 0, control_parameter = 0, accepted_ratio = 0;
    double max_mass_scaling = 0, control_parameter_mass_scaling = 0, error_factor_mass_scaling = 0,
           decrease_factor_mass_scaling = 0;
    double save_period = 0, print_period = 0, contact_update_period = 0;
    double xmin_period = 0, xmax_period = 0, gravity_x = 0, gravity_y = 0;
    int flags[F_COUNT] = {0};
    double time = 0;
    int number_save = 0, number_iteration = 0, number_print = 0;
    double next_save = 0, next_print = 0, next_contact_update = 0;
    std::vector<Body> bodies;
    // DEACTIVATION rules of Code/IO.h:947-968: the spy line of Read_spy_line
    // (Code/IO.h:11-590) followed by the comparison (0 below, 1 equal, 2
    // above) and the threshold
    std::vector<std::vector<double>> deactivated;
    // MONITORING quantities of Code/IO.h:933-945 (Read_spy_line entries)
    std::vector<std::vector<double>> monitored;

    int law_index(int mat_s, int mat_m) const {
        return contacts_table[(size_t)mat_s * materials.size() + mat_m];
    }
    size_t total_nodes() const {
        size_t n = 0;
        for (auto& b : bodies) n += b.n_nodes;
        return n;
    }
};

// Read_spy_line (Code/IO.h:11-590): one spy / monitor / deactivation
// quantity from the stream into out.  False for a keyword outside its
// vocabulary (the reference then leaves out as it was).
bool read_spy_line(std::istream& in, std::vector<double>& out);

// Update_status reads rule entries by position (Code/Proximity.h:550-1007);
// a rule whose entries it would read out of range, or whose body-level
// force branch would index nodes the body does not have, is refused.
bool check_deactivation(const Model& m, std::string& err);

// n DEACTIVATION rules from `in` into m.deactivated (Code/IO.h:947-968).
bool read_deactivation_rules(std::istream& in, int n, Model& m, std::string& err);

// Monitoring (Code/Monitoring.h:8-559) reads the MONITORING entries by
// position too; the same checks for what it reads.
bool check_monitored(const Model& m, std::string& err);

// Parse STATIC_CONTROL.asc, STATIC_DATA.asc and DYNAMIC_<save>.asc from dir.
bool load_model_files(const std::string& dir, int save_index, Model& m, std::string& err);

// Read the binary record container written by the preprocessor package.
bool load_model_binary(const std::string& path, Model& m, std::string& err);

// Segment stencils and initial lengths of every border (Code/Body.h:403-605).
void initialize_segments(Model& m, Body& b);

// Append Monitoring rows (each: its value count, then the values) to
// `path` as Write_monitoring does (Code/IO.h:2644-2666): ten significant
// digits, each value followed by a space, one row per line.
bool write_monitoring_text(const std::string& path, const std::vector<double>& rows, std::string& err);

// Write DYNAMIC_nnnnn-format text of the current state (Code/IO.h:2351-2636 layout,
// 17 significant digits).
bool write_dynamic_text(const std::string& path, const Model& m, std::string& err);

// What Write_graphic draws of one body (Code/Graphic.h:21-946): positions,
// node fields (x then y per vector field), the 19 nodal stress fields of
// Body::Compute_nodal_stresses, the body's works.
struct GraphicBody {
    bool active = true;
    std::vector<double> x0, y0, x, y;
    std::vector<std::vector<double>> vec;    // displacement, velocity, acceleration, force, internal, contact,
                                             // body, dirichlet, neumann, damping: 2 * n values each
    std::vector<double> stress;              // 19 fields of n values
    double works[7] = {0, 0, 0, 0, 0, 0, 0};
};

constexpr int GRAPHIC_FIELDS = 41;           // To_Plot slots 00-40 of Code/Graphic.h

// VTK legacy POLYDATA of every active body in the layout of Write_graphic
// (Code/Graphic.h:21-946): positions wrapped into the period, cells that
// cross the periodic seam drawn on both sides, the point fields whose
// To_Plot flag is set (fields this library does not model - ALID, damage,
// error estimates, temperature, mass-scaling epsilon - written as zeros).
bool write_graphic_text(const std::string& path, const Model& m, const std::vector<GraphicBody>& g, int iteration,
                        int number_save, int number_print, double time, const int* to_plot, std::string& err);

}  // namespace mel
