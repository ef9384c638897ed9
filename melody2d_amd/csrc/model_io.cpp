// Model-file parsing (ASCII and binary) into mel::Model.  See model_io.hpp.
#include "model_io.hpp"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>

namespace mel {

namespace {

bool starts(const std::string& line, const char* kw) {
    size_t n = std::strlen(kw);
    return line.size() >= n && line.compare(0, n, kw) == 0;
}

int material_type_of(const std::string& t) {
    if (t == "ElasticLinear") return MAT_ELASTIC_LINEAR;
    if (t == "NeoHookean") return MAT_NEOHOOKEAN;
    return -1;
}

// Number of parameters each law reads (Code/IO.h:721-793).
int law_nparams(int type) {
    switch (type) {
        case LAW_FRICTIONLESS: return 1;
        case LAW_COHESION: return 4;
        case LAW_MOHR_COULOMB: return 5;
        case LAW_DAMPED_MOHR_COULOMB: return 6;
        case LAW_TWO_SLOPES: return 6;
        case LAW_CZM_LINEAR: return 6;
        case LAW_CZM_FATIGUE: return 9;
        case LAW_BONDED_MOHR_COULOMB: return 11;
    }
    return 0;
}

int law_type_of(const std::string& t) {
    static const char* names[] = {"Frictionless", "Cohesion", "MohrCoulomb", "DampedMohrCoulomb",
                                  "TwoSlopesMohrCoulomb", "BondedMohrCoulomb", "CZMlinear", "CZMfatigue"};
    for (int i = 0; i < 8; i++)
        if (t == names[i]) return i;
    return -1;
}

int periodicity_of(const std::string& p) {
    if (p == "Periodic") return PER_PERIODIC;
    if (p == "Simple") return PER_SIMPLE;
    return PER_CLOSED;
}

int status_of(const std::string& s) {
    if (s == "inactive") return ST_INACTIVE;
    if (s == "nogravity") return ST_NOGRAVITY;
    return ST_ACTIVE;
}

// mu, lambda | kappa conversion of Code/IO.h:678 and Code/IO.h:687
void convert_material(Material& mat) {
    const double* r = mat.raw;
    mat.p[0] = r[0];
    mat.p[1] = r[1];
    mat.p[2] = r[2];
    mat.p[3] = r[3] / (2. * (1. + r[4]));
    if (mat.type == MAT_ELASTIC_LINEAR)
        mat.p[4] = r[3] * r[4] / ((1. + r[4]) * (1. - 2. * r[4]));
    else
        mat.p[4] = r[3] / (3. * (1. - 2. * r[4]));
    mat.p[5] = r[5];
}

void build_table(Model& m) {
    size_t nm = m.materials.size();
    m.contacts_table.assign(nm * nm, -1);
    // later laws overwrite earlier ones, both orders (Code/IO.h:797-813)
    for (size_t i = 0; i < m.laws.size(); i++) {
        int nn = 0, pp = 0;
        for (size_t n = 0; n < nm; n++)
            if (m.materials[n].name == m.laws[i].m1) nn = (int)n;
        for (size_t p = 0; p < nm; p++)
            if (m.materials[p].name == m.laws[i].m2) pp = (int)p;
        m.contacts_table[nn * nm + pp] = (int)i;
        m.contacts_table[pp * nm + nn] = (int)i;
    }
}

void read_bc(std::istream& in, BorderBC& bc) {
    std::string direction, type1, type2, token;
    in >> direction >> type1;
    if (type1 == "None") {
        bc.type = BC_NONE;
        std::getline(in, token);
        return;
    }
    in >> type2;
    if (type1 == "Dirichlet") {
        if (type2 == "Soft") {
            bc.type = BC_SOFT;
            in >> bc.stiffness >> bc.damping;
        } else if (type2 == "Driven") {
            bc.type = BC_DRIVEN;
        }
    } else if (type1 == "Neumann") {
        bc.type = type2 == "Following" ? BC_FOLLOWING : BC_ORIENTED;
    }
    std::getline(in, token);
    int n = 0;
    in >> n;
    std::getline(in, token);
    for (int j = 0; j < n; j++) {
        double t, v;
        in >> t >> v;
        std::getline(in, token);
        bc.hist.t.push_back(t);
        bc.hist.v.push_back(v);
    }
}

void finish_body_state(Model& m, Body& b) {
    size_t n = b.n_nodes;
    auto z = [&](std::vector<double>& v) { if (v.size() != n) v.assign(n, 0.0); };
    z(b.ux); z(b.uy); z(b.vx); z(b.vy); z(b.ax); z(b.ay);
    z(b.upx); z(b.upy); z(b.vpx); z(b.vpy); z(b.apx); z(b.apy);
    (void)m;
}

// a body's borders can be given their stencils: every border has at least
// three entries naming nodes of the body, and before / after name borders of
// the body (initialize_segments indexes with all of them)
bool check_borders(const Body& b, int bi, std::string& err) {
    const int nbr = (int)b.borders.size();
    for (int k = 0; k < nbr; k++) {
        const Border& B = b.borders[k];
        const std::string where = "body " + std::to_string(bi) + " border " + std::to_string(k);
        if (B.nodes.size() < 3) { err = where + ": fewer than three nodes"; return false; }
        if (B.before < 0 || B.before >= nbr || B.after < 0 || B.after >= nbr) {
            err = where + ": before/after border out of range";
            return false;
        }
        if (B.periodicity < PER_CLOSED || B.periodicity > PER_PERIODIC) { err = where + ": periodicity"; return false; }
        for (int v : B.nodes)
            if (v < 0 || v >= b.n_nodes) { err = where + ": node index out of range"; return false; }
    }
    return true;
}

// CSR pointer array of `rows` rows over `n` entries: rows + 1 values, from
// 0 up to n, never decreasing
bool csr_ok(const std::vector<int>& ptr, size_t rows, size_t n) {
    if (ptr.size() != rows + 1 || ptr[0] != 0 || (size_t)ptr[rows] != n) return false;
    for (size_t i = 0; i < rows; i++)
        if (ptr[i + 1] < ptr[i]) return false;
    return true;
}

bool indices_ok(const std::vector<int>& idx, int n) {
    for (int v : idx)
        if (v < 0 || v >= n) return false;
    return true;
}

// every index the library uses on the device, checked once the model is read
bool validate_model(const Model& m, std::string& err) {
    const int nm = (int)m.materials.size();
    if ((int)m.contacts_table.size() != nm * nm) { err = "material table size"; return false; }
    for (size_t bi = 0; bi < m.bodies.size(); bi++) {
        const Body& b = m.bodies[bi];
        const std::string where = "body " + std::to_string(bi) + ": ";
        const size_t n = (size_t)b.n_nodes;
        if (b.n_nodes < 0 || b.n_gauss < 0) { err = where + "negative size"; return false; }
        if (b.material < 0 || b.material >= nm) { err = where + "material index out of range"; return false; }
        if (b.status < ST_ACTIVE || b.status > ST_NOGRAVITY) { err = where + "status"; return false; }
        if (b.X.size() != n || b.Y.size() != n || b.mx.size() != n || b.my.size() != n || b.imx.size() != n ||
            b.imy.size() != n || b.domain.size() != n) {
            err = where + "node array sizes";
            return false;
        }
        if (b.borders.empty()) { err = where + "no border"; return false; }
        if (!check_borders(b, (int)bi, err)) return false;
        if (b.cells.size() % 3 != 0 || !indices_ok(b.cells, b.n_nodes)) { err = where + "cell indices"; return false; }
        if (b.rigid) continue;
        const size_t G = (size_t)b.n_gauss;
        if (b.n_regions < 1) { err = where + "region count"; return false; }
        if (b.gx.size() != G || b.gy.size() != G || b.gw.size() != G || b.gjac.size() != G) {
            err = where + "Gauss point array sizes";
            return false;
        }
        if (!csr_ok(b.gptr, G, b.gidx.size()) || b.gphi.size() != b.gidx.size() ||
            b.gphix.size() != b.gidx.size() || b.gphiy.size() != b.gidx.size() || !indices_ok(b.gidx, b.n_nodes)) {
            err = where + "Gauss point shape functions (pointers or node indices)";
            return false;
        }
        if (!csr_ok(b.dptr, n, b.didx.size()) || b.dphi.size() != b.didx.size() ||
            b.dphix.size() != b.didx.size() || b.dphiy.size() != b.didx.size() || !indices_ok(b.didx, b.n_nodes)) {
            err = where + "DOF_TO_DISP rows (pointers or node indices)";
            return false;
        }
        if (b.mat_dof.size() != 2 * b.mat_stiffness.size() || !indices_ok(b.mat_dof, 2 * b.n_nodes)) {
            err = where + "matrix entries";
            return false;
        }
    }
    const int NB = (int)m.bodies.size();
    if (!check_deactivation(m, err)) return false;
    if (!check_monitored(m, err)) return false;
    for (size_t bi = 0; bi < m.bodies.size(); bi++) {
        const Body& b = m.bodies[bi];
        for (size_t k = 0; k + 2 < b.neighbours.size(); k += 3) {
            const int nb = b.neighbours[k], nbr = b.neighbours[k + 1];
            if (nb < 0 || nb >= NB || nbr < 0 || nbr >= (int)m.bodies[nb].borders.size()) {
                err = "body " + std::to_string(bi) + ": neighbour out of range";
                return false;
            }
        }
        for (const ContactElement& c : b.contacts) {
            const std::string where = "body " + std::to_string(bi) + ": contact element ";
            if (c.borderS < 0 || c.borderS >= (int)b.borders.size() || c.border_nodeS < 0 ||
                c.border_nodeS >= (int)b.borders[c.borderS].nodes.size() || c.nodeS < 0 || c.nodeS >= b.n_nodes) {
                err = where + "slave out of range";
                return false;
            }
            if (c.bodyM < 0 || c.bodyM >= NB) { err = where + "master body out of range"; return false; }
            const Body& M = m.bodies[c.bodyM];
            if (c.borderM < 0 || c.borderM >= (int)M.borders.size() || c.segmentM < 0 ||
                c.segmentM + 1 >= (int)M.borders[c.borderM].nodes.size()) {
                err = where + "master segment out of range";
                return false;
            }
        }
    }
    return true;
}

}  // namespace

// Read_spy_line (Code/IO.h:11-590).  Every quantity is a keyword, mostly a
// component keyword, then its indices; the entries are type, component,
// indices.  Energy and Work read the index before testing the component;
// the Polar form of MassScaling yields type 11 as the reference writes it.
bool read_spy_line(std::istream& in, std::vector<double>& out) {
    auto index_of = [](const std::string& k, std::initializer_list<const char*> names) {
        int i = 0;
        for (const char* n : names) {
            if (k == n) return i;
            i++;
        }
        return -1;
    };
    std::string kw, sub;
    double b = 0, bo = 0, n = 0;
    in >> kw;
    if (kw == "Position") {
        in >> sub;
        const int c = index_of(sub, {"X", "Y", "R", "Length", "LengthInContact", "Area", "AllX", "AllY"});
        if (c < 0) return false;
        if (c <= 2) { in >> b >> n; out = {0., (double)c, b, n}; }
        else if (c <= 5) { in >> b; out = {0., (double)c, b}; }
        else out = {0., (double)c};
        return true;
    }
    const int kin = index_of(kw, {"Displacement", "Velocity", "Acceleration"});
    if (kin >= 0) {
        in >> sub;
        const int c = index_of(sub, {"X", "Y", "Mag", "R"});
        if (c < 0) return false;
        in >> b >> n;
        out = {(double)(kin + 1), (double)c, b, n};
        return true;
    }
    if (kw == "Force") {
        in >> sub;
        const int c = index_of(sub, {"X", "Y", "Mag", "R"});
        if (c < 0) return false;
        in >> sub;
        const int k = index_of(sub, {"Total", "Internal", "Contact", "Body", "Dirichlet", "Neumann", "Damping", "Alid"});
        if (k < 0) return false;
        in >> b >> n;
        out = {4., (double)c, (double)k, b, n};
        return true;
    }
    if (kw == "Jacobian") {
        in >> b >> n;
        out = {5., b, n};
        return true;
    }
    if (kw == "Stress") {
        in >> sub;
        const int c = index_of(sub, {"XX", "YY", "XY", "ZZ", "Tresca", "VM", "Major", "Intermediate", "Minor",
                                     "Spherical"});
        if (c < 0) return false;
        in >> b >> n;
        out = {6., (double)c, b, n};
        return true;
    }
    if (kw == "Contact") {
        in >> sub;
        const int c = index_of(sub, {"Gapn", "Gapt", "Xsi", "Xnorm", "Ynorm", "Damage", "Fx", "Fy", "Length", "Master",
                                     "Px", "Py", "Slave", "Xslave", "Yslave"});
        if (c < 0) return false;
        in >> b >> bo >> n;
        out = {7., (double)c, b, bo, n};
        return true;
    }
    if (kw == "Damage") {
        in >> b;
        out = {8., b};
        return true;
    }
    if (kw == "Energy" || kw == "Work") {
        in >> sub >> b;
        const int c = kw == "Energy" ? index_of(sub, {"Kinetic", "Deformation"})
                                     : index_of(sub, {"Internal", "Contact", "Body", "Dirichlet", "Neumann", "Damping",
                                                      "Alid"});
        if (c < 0) return false;
        out = {kw == "Energy" ? 9. : 10., (double)c, b};
        return true;
    }
    if (kw == "Error") {
        in >> sub;
        const int c = index_of(sub, {"X", "Y", "Mag", "Norm", "Max"});
        if (c < 0) return false;
        in >> b >> n;
        out = {11., (double)c, b, n};
        return true;
    }
    if (kw == "MassScaling") {
        in >> sub;
        if (sub == "Polar") {
            double x0, x1, y0, y1;
            in >> sub >> n >> x0 >> x1 >> y0 >> y1;
            if (sub != "Contact_Normal") return false;
            out = {11., 0., n, x0, x1, y0, y1};
            return true;
        }
        const int c = index_of(sub, {"X", "Y", "Moy", "dX", "dY", "dMoy", "Xmass", "Ymass", "Mass"});
        if (c < 0) return false;
        in >> b >> n;
        out = {12., (double)c, b, n};
        return true;
    }
    return false;
}

// n DEACTIVATION rules (Code/IO.h:947-968): a spy line, the comparison
// (below, equal, above) and the threshold, one rule per line
bool read_deactivation_rules(std::istream& in, int n, Model& m, std::string& err) {
    std::string token;
    std::vector<double> rule;
    for (int i = 0; i < n; i++) {
        if (!read_spy_line(in, rule)) {
            err = "DEACTIVATION rule " + std::to_string(i) + ": unknown quantity keyword";
            return false;
        }
        std::string cmp;
        double threshold = 0.;
        in >> cmp >> threshold;
        if (cmp == "below") rule.push_back(0.);
        else if (cmp == "equal") rule.push_back(1.);
        else if (cmp == "above") rule.push_back(2.);
        else { err = "DEACTIVATION rule " + std::to_string(i) + ": unknown comparison " + cmp; return false; }
        rule.push_back(threshold);
        m.deactivated.push_back(rule);
        std::getline(in, token);
    }
    return true;
}

// Monitoring (Code/Monitoring.h:8-559) per entry type: 0-3 [1] component,
// [2] body, [3] node (< 0: the body); 4 [1] component, [2] force kind,
// [3] body, [4] node - its body-level branch tests the type of body [3] and
// loops over the node count of body [2] (the force kind) reading the nodes
// of body [3]; 5 [1] body, [2] node; 6 [1] component, [2] body, [3] node;
// 7 [1] component, [2] body, [3] border, [4] border node; 8 [1] body;
// 9 and 10 [1] component, [2] body (< 0: all or none).  Errors (11) and
// mass scaling (12) have no branch and add nothing.
bool check_monitored(const Model& m, std::string& err) {
    const int NB = (int)m.bodies.size();
    for (size_t i = 0; i < m.monitored.size(); i++) {
        const std::vector<double>& r = m.monitored[i];
        const std::string where = "MONITORING entry " + std::to_string(i) + ": ";
        auto body_ok = [&](double v) { return v >= 0. && v < NB; };
        auto node_ok = [&](double b, double n) { return n < m.bodies[(int)b].n_nodes; };
        const int type = (int)r[0];
        if (type >= 0 && type <= 3) {
            if (r.size() != 4) { err = where + "the reference reads entries this quantity does not have"; return false; }
            if (!body_ok(r[2])) { err = where + "body out of range"; return false; }
            if (r[3] >= 0. && !node_ok(r[2], r[3])) { err = where + "node out of range"; return false; }
        } else if (type == 4) {
            if (!body_ok(r[3])) { err = where + "body out of range"; return false; }
            const Body& B = m.bodies[(int)r[3]];
            if (r[4] >= 0.) {
                if (!node_ok(r[3], r[4])) { err = where + "node out of range"; return false; }
            } else if (!B.rigid) {
                if (!body_ok(r[2]) || m.bodies[(int)r[2]].n_nodes > B.n_nodes) {
                    err = where + "the reference's body-level force branch loops over the nodes of body " +
                          std::to_string((int)r[2]) + " (the force kind) while reading body " +
                          std::to_string((int)r[3]) + ", which has fewer or does not exist";
                    return false;
                }
            }
        } else if (type == 5) {
            if (!body_ok(r[1]) || r[2] < 0. || !node_ok(r[1], r[2])) { err = where + "node out of range"; return false; }
        } else if (type == 6) {
            if (!body_ok(r[2]) || r[3] < 0. || !node_ok(r[2], r[3])) { err = where + "node out of range"; return false; }
        } else if (type == 7 || type == 8) {
            if (!body_ok(type == 7 ? r[2] : r[1])) { err = where + "body out of range"; return false; }
        } else if (type == 9 || type == 10) {
            if (r[2] >= 0. && !body_ok(r[2])) { err = where + "body out of range"; return false; }
        }
    }
    if (!m.monitored.empty() && m.flags[F_MONITOR_BOUNDARIES]) {
        err = "MONITOR_BOUNDARIES is not supported: its sums start from uninitialised values (Code/Monitoring.h:540-549)";
        return false;
    }
    return true;
}

// Update_status (Code/Proximity.h:550-1007) acts on types 0-4 and 8 and
// reads, by position: type 0-3 [1] component, [2] body, [3] node (< 0: the
// body), [4] comparison, [5] threshold; type 4 [1] component, [2] force
// kind, [3] body, [4] node, [5] comparison, [6] threshold; type 8 [1] body,
// [2] comparison, [3] threshold.  Its body-level force branch tests the
// type of body [2] (the force kind) and loops over that body's node count
// while reading the nodes of body [3]; both must exist for the loop to stay
// in range.  Position components 3-7 carry fewer entries than the branch
// reads.  Other types are ignored by the reference and kept as read.
bool check_deactivation(const Model& m, std::string& err) {
    const int NB = (int)m.bodies.size();
    for (size_t i = 0; i < m.deactivated.size(); i++) {
        const std::vector<double>& r = m.deactivated[i];
        const std::string where = "DEACTIVATION rule " + std::to_string(i) + ": ";
        auto body_ok = [&](double v) { return v >= 0. && v < NB; };
        const int type = (int)r[0];
        if (type >= 0 && type <= 3) {
            if (r.size() != 6 || (type == 0 && r[1] > 2.)) {
                err = where + "the reference reads entries this quantity does not have";
                return false;
            }
            if (!body_ok(r[2])) { err = where + "body out of range"; return false; }
            if (r[3] >= 0. && r[3] >= m.bodies[(int)r[2]].n_nodes) { err = where + "node out of range"; return false; }
        } else if (type == 4) {
            if (r.size() != 7) { err = where + "the reference reads entries this quantity does not have"; return false; }
            if (!body_ok(r[3])) { err = where + "body out of range"; return false; }
            const Body& B = m.bodies[(int)r[3]];
            if (r[4] >= 0.) {
                if (r[4] >= B.n_nodes) { err = where + "node out of range"; return false; }
                if (B.rigid) {
                    err = where + "forces of a rigid body's nodes are not supported (Body::Sum_up_forces copies the "
                                  "body's forces to them, Code/Body.h:1999-2012; use the body-level form)";
                    return false;
                }
            } else {
                if (!body_ok(r[2])) {
                    err = where + "the reference's body-level force branch tests body " + std::to_string((int)r[2]) +
                          " (the force kind), which does not exist";
                    return false;
                }
                const Body& K = m.bodies[(int)r[2]];
                if (!K.rigid && (B.rigid || K.n_nodes > B.n_nodes)) {
                    err = where + "the reference's body-level force branch reads " + std::to_string(K.n_nodes) +
                          " nodes of a body that has " + std::to_string(B.n_nodes) + " deformable nodes";
                    return false;
                }
                if (K.rigid && !B.rigid) {
                    err = where + "the reference's body-level force branch reads the rigid-body forces of a deformable "
                                  "body, which it never sets";
                    return false;
                }
            }
        } else if (type == 8) {
            if (r.size() != 4) { err = where + "the reference reads entries this quantity does not have"; return false; }
            if (!body_ok(r[1])) { err = where + "body out of range"; return false; }
        }
    }
    return true;
}


void initialize_segments(Model& m, Body& body) {
    double period = m.xmax_period - m.xmin_period;
    int nbr = (int)body.borders.size();
    for (int i = 0; i < nbr; i++) {
        Border& B = body.borders[i];
        int n = (int)B.nodes.size();
        B.node0.assign(n, 0); B.node1.assign(n, 0); B.node2.assign(n, 0); B.node3.assign(n, 0);
        B.shift0.assign(n, 0); B.shift1.assign(n, 0); B.shift2.assign(n, 0); B.shift3.assign(n, 0);
        B.initial_length.assign(n, 0.0);
        if (n < 3) continue;
        const Border& bb = body.borders[B.before];
        const Border& ba = body.borders[B.after];
        bool per = body.borders[i].periodicity == PER_PERIODIC;
        // first border node
        B.node0[0] = bb.nodes[bb.nodes.size() - 2];
        B.node1[0] = B.nodes[0];
        B.node2[0] = B.nodes[1];
        B.node3[0] = B.nodes[2];
        if (per) B.shift0[0] = (i == 0) ? -1 : 1;
        for (int j = 1; j < n - 2; j++) {
            B.node0[j] = B.nodes[j - 1];
            B.node1[j] = B.nodes[j];
            B.node2[j] = B.nodes[j + 1];
            B.node3[j] = B.nodes[j + 2];
        }
        // second to last
        B.node0[n - 2] = B.nodes[n - 3];
        B.node1[n - 2] = B.nodes[n - 2];
        B.node2[n - 2] = B.nodes[n - 1];
        B.node3[n - 2] = ba.nodes[1];
        if (per) {
            int s = (i == 0) ? 1 : -1;
            B.shift2[n - 2] = s;
            B.shift3[n - 2] = s;
        }
        // last
        B.node0[n - 1] = B.nodes[n - 2];
        B.node1[n - 1] = B.nodes[n - 1];
        B.node2[n - 1] = ba.nodes[1];
        B.node3[n - 1] = ba.nodes[2];
        if (per) {
            int s = (i == 0) ? 1 : -1;
            B.shift1[n - 1] = s;
            B.shift2[n - 1] = s;
            B.shift3[n - 1] = s;
        }
        for (int j = 0; j < n; j++) {
            double x0 = body.X[B.node0[j]] + B.shift0[j] * period, y0 = body.Y[B.node0[j]];
            double x1 = body.X[B.node1[j]] + B.shift1[j] * period, y1 = body.Y[B.node1[j]];
            double x2 = body.X[B.node2[j]] + B.shift2[j] * period, y2 = body.Y[B.node2[j]];
            B.initial_length[j] = 0.5 * std::sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0)) +
                                  0.5 * std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
        }
    }
}

// Node::alid_alpha, alid_beta of a deformable body (Code/IO.h:1548-1584):
// graded through layers of width `range` inside the bounds, the x layers
// taking the larger value where they overlap the y layers
void alid_coefficients(Body& b) {
    b.alid_alpha.assign(b.n_nodes, 0.);
    b.alid_beta.assign(b.n_nodes, 0.);
    if (!b.alid.on) return;
    const Alid& A = b.alid;
    for (int i = 0; i < b.n_nodes; i++) {
        const double xini = b.X[i], yini = b.Y[i];
        double al = 0., be = 0., ka, kb, temp;
        if (yini < A.ymin + A.range) {
            ka = (A.alphaout - A.alphain) / std::pow(-A.range, A.exponent);
            kb = (A.betaout - A.betain) / std::pow(-A.range, A.exponent);
            al = A.alphain + ka * std::pow(yini - (A.ymin + A.range), A.exponent);
            be = A.betain + kb * std::pow(yini - (A.ymin + A.range), A.exponent);
        } else if (yini > A.ymax - A.range) {
            ka = (A.alphaout - A.alphain) / std::pow(A.range, A.exponent);
            kb = (A.betaout - A.betain) / std::pow(A.range, A.exponent);
            al = A.alphain + ka * std::pow(yini - (A.ymax - A.range), A.exponent);
            be = A.betain + kb * std::pow(yini - (A.ymax - A.range), A.exponent);
        }
        if (xini < A.xmin + A.range) {
            ka = (A.alphaout - A.alphain) / std::pow(-A.range, A.exponent);
            kb = (A.betaout - A.betain) / std::pow(-A.range, A.exponent);
            temp = A.alphain + ka * std::pow(xini - (A.xmin + A.range), A.exponent);
            if (temp > al) al = temp;
            temp = A.betain + kb * std::pow(xini - (A.xmin + A.range), A.exponent);
            if (temp > be) be = temp;
        } else if (xini > A.xmax - A.range) {
            ka = (A.alphaout - A.alphain) / std::pow(A.range, A.exponent);
            kb = (A.betaout - A.betain) / std::pow(A.range, A.exponent);
            temp = A.alphain + ka * std::pow(xini - (A.xmax - A.range), A.exponent);
            if (temp > al) al = temp;
            temp = A.betain + kb * std::pow(xini - (A.xmax - A.range), A.exponent);
            if (temp > be) be = temp;
        }
        b.alid_alpha[i] = al;
        b.alid_beta[i] = be;
    }
}

namespace {
bool load_model_files_impl(const std::string& dir, int save_index, Model& m, std::string& err) {
    m = Model();
    std::string line, token;
    int index_body = 0;
    // ---------------- STATIC_CONTROL.asc (Code/IO.h:646-1487) ----------------
    {
        std::ifstream in(dir + "/STATIC_CONTROL.asc");
        if (!in) { err = "cannot open " + dir + "/STATIC_CONTROL.asc"; return false; }
        while (std::getline(in, line)) {
            if (starts(line, "SIMULATION_NAME")) {
                in >> m.name;
                std::getline(in, token);
            }
            if (starts(line, "MATERIALS")) {
                int nm = 0;
                in >> nm;
                std::getline(in, token);
                std::getline(in, token);
                for (int i = 0; i < nm; i++) {
                    Material mat;
                    std::string type;
                    in >> mat.name;
                    std::getline(in, token);
                    in >> type;
                    std::getline(in, token);
                    mat.type = material_type_of(type);
                    if (mat.type < 0) { err = "unsupported material type " + type; return false; }
                    for (int k = 0; k < 6; k++) in >> mat.raw[k];
                    std::getline(in, token);
                    convert_material(mat);
                    m.materials.push_back(mat);
                    std::getline(in, token);
                }
            }
            if (starts(line, "CONTACT_LAWS")) {
                int nl = 0;
                in >> nl;
                std::getline(in, token);
                std::getline(in, token);
                for (int i = 0; i < nl; i++) {
                    ContactLaw law;
                    std::string type, lev;
                    in >> law.m1; std::getline(in, token);
                    in >> law.m2; std::getline(in, token);
                    in >> type; std::getline(in, token);
                    in >> lev; std::getline(in, token);
                    law.type = law_type_of(type);
                    if (law.type < 0) { err = "unsupported contact law " + type; return false; }
                    law.fixed_length = (lev == "Fixed");
                    law.p.resize(law_nparams(law.type));
                    for (auto& x : law.p) in >> x;
                    std::getline(in, token);
                    m.laws.push_back(law);
                    std::getline(in, token);
                }
            }
            if (starts(line, "SOLVER")) {
                // Code/IO.h:821-890: every scheme reads its times and error
                // control; the mass-scaling schemes one more line
                in >> m.solver;
                std::getline(in, token);
                m.scheme = scheme_of(m.solver);
                if (m.scheme < 0) {
                    err = "unsupported solver " + m.solver +
                          (m.solver == "Euler_2" ? " (the reference reads no target error for it, Code/IO.h:839-847, "
                                                   "yet its time-step update divides by one)"
                                                 : "");
                    return false;
                }
                in >> m.tini >> m.dt >> m.tend;
                std::getline(in, token);
                in >> m.target_error >> m.control_parameter >> m.accepted_ratio;
                std::getline(in, token);
                m.inv_target_error = 1 / m.target_error;
                if (m.scheme == SCH_MASS_SCALING || m.scheme == SCH_ADAPTIVE_MASS_SCALING) {
                    in >> m.max_mass_scaling >> m.control_parameter_mass_scaling >> m.error_factor_mass_scaling >>
                        m.decrease_factor_mass_scaling;
                    std::getline(in, token);
                }
                in >> m.save_period >> m.print_period >> m.contact_update_period;
                std::getline(in, token);
            }
            if (starts(line, "PERIODIC_BOUNDARIES")) {
                in >> m.xmin_period >> m.xmax_period;
                std::getline(in, token);
            }
            if (starts(line, "PENALTY_PARAMETER")) {
                double pen;
                in >> pen;
                std::getline(in, token);
            }
            if (starts(line, "GRAVITY")) {
                in >> m.gravity_x >> m.gravity_y;
                std::getline(in, token);
            }
            if (starts(line, "PLOTTING")) {
                std::getline(in, token);
                std::getline(in, token);
            }
            if (starts(line, "NUMBER_BODIES")) {
                int nb;
                in >> nb;
                std::getline(in, token);
            }
            if (starts(line, "SPIES")) {
                int n = 0;
                in >> n;
                std::getline(in, token);
                if (n != 0) { err = "spies entries are not supported"; return false; }
            }
            if (starts(line, "MONITORING")) {
                // Code/IO.h:933-945
                int n = 0;
                in >> n;
                std::getline(in, token);
                std::vector<double> q;
                for (int i = 0; i < n; i++) {
                    if (!read_spy_line(in, q)) {
                        err = "MONITORING entry " + std::to_string(i) + ": unknown quantity keyword";
                        return false;
                    }
                    m.monitored.push_back(q);
                    std::getline(in, token);
                }
            }
            if (starts(line, "DEACTIVATION")) {
                int n = 0;
                in >> n;
                std::getline(in, token);
                if (!read_deactivation_rules(in, n, m, err)) return false;
            }
            if (starts(line, "GRAPHIC")) {
                for (int i = 0; i < 41; i++) {
                    int j;
                    in >> j >> token;
                    std::getline(in, token);
                }
            }
            bool def = starts(line, "DEFORMABLE"), rig = starts(line, "RIGID");
            if (def || rig) {
                Body b;
                b.rigid = rig;
                std::string mat, per;
                in >> index_body; std::getline(in, token);
                in >> mat; std::getline(in, token);
                in >> per; std::getline(in, token);
                in >> b.nodal_distance >> b.detection_distance >> b.contact_distance;
                std::getline(in, token);
                b.periodicity = periodicity_of(per);
                b.material = -1;
                for (size_t i = 0; i < m.materials.size(); i++)
                    if (m.materials[i].name == mat) { b.material = (int)i; b.density = m.materials[i].p[0]; break; }
                if (b.material < 0) { err = "unknown material " + mat; return false; }
                if (index_body != (int)m.bodies.size()) { err = "bodies must be listed in index order"; return false; }
                m.bodies.push_back(b);
            }
            if (starts(line, "BORDERS")) {
                int nbr = 0;
                in >> nbr;
                std::getline(in, token);
                std::getline(in, token);
                Body& b = m.bodies.at(index_body);
                b.borders.resize(nbr);
                for (int i = 0; i < nbr; i++) {
                    std::string per, interp;
                    int nn;
                    in >> per >> interp;
                    std::getline(in, token);
                    in >> nn;
                    std::getline(in, token);
                    b.borders[i].periodicity = periodicity_of(per);
                    b.borders[i].linear = interp != "Spline";
                    if (!b.borders[i].linear) {
                        // Closest_2_segment's Spline branch (Code/Tools.h:82-105) never sets
                        // flag_detect, which its callers read uninitialised
                        err = "Spline interpolant is not supported (the reference's Spline projection leaves its "
                              "contact flag unset)";
                        return false;
                    }
                    b.borders[i].nodes.resize(nn);
                    read_bc(in, b.borders[i].xbc);
                    read_bc(in, b.borders[i].ybc);
                    std::getline(in, token);
                }
            }
            if (starts(line, "ALID")) {
                // Code/IO.h:1395-1485: twelve values, then the x history,
                // a separator line when both are given, the y history
                Body& b = m.bodies.at(index_body);
                Alid& A = b.alid;
                int nx = 0, ny = 0;
                in >> nx >> ny >> A.xmin >> A.xmax >> A.ymin >> A.ymax >> A.range >> A.alphain >> A.alphaout >>
                    A.betain >> A.betaout >> A.exponent;
                if (!in || nx < 0 || ny < 0 || nx > (1 << 20) || ny > (1 << 20)) { err = "ALID header"; return false; }
                auto pairs = [&](History& h, int n) {
                    for (int j = 0; j < n; j++) {
                        double t, v;
                        in >> t >> v;
                        std::getline(in, token);
                        h.t.push_back(t);
                        h.v.push_back(v);
                    }
                };
                pairs(A.xh, nx);
                if (nx > 0 && ny > 0) std::getline(in, token);
                pairs(A.yh, ny);
                A.on = nx > 0 || ny > 0;
                if (!in) { err = "ALID histories"; return false; }
            }
        }
    }
    // ---------------- STATIC_DATA.asc (Code/IO.h:1491-1852) ----------------
    {
        std::ifstream in(dir + "/STATIC_DATA.asc");
        if (!in) { err = "cannot open " + dir + "/STATIC_DATA.asc"; return false; }
        bool rigid = false;
        while (std::getline(in, line)) {
            if (starts(line, "DEFORMABLE")) {
                rigid = false;
                in >> index_body;
                std::getline(in, token);
            }
            if (starts(line, "RIGID")) {
                rigid = true;
                in >> index_body;
                std::getline(in, token);
                Body& b = m.bodies.at(index_body);
                in >> b.xc >> b.yc;
                std::getline(in, token);
                in >> b.mass >> b.inertia >> b.inverse_mass >> b.inverse_inertia;
                std::getline(in, token);
            }
            if (starts(line, "NODES")) {
                Body& b = m.bodies.at(index_body);
                int n;
                in >> n;
                std::getline(in, token);
                std::getline(in, token);
                if (!in || n < 0 || n > (1 << 28)) { err = "node count"; return false; }
                b.n_nodes = n;
                b.X.resize(n); b.Y.resize(n); b.domain.assign(n, 0); b.mx.assign(n, 0); b.my.assign(n, 0);
                b.imx.assign(n, 0); b.imy.assign(n, 0);
                for (int i = 0; i < n; i++) {
                    double j;
                    if (!rigid) in >> j >> b.X[i] >> b.Y[i] >> b.domain[i] >> b.mx[i] >> b.my[i] >> b.imx[i] >> b.imy[i];
                    else in >> j >> b.X[i] >> b.Y[i];
                    std::getline(in, token);
                }
                if (!rigid) alid_coefficients(b);
            }
            if (starts(line, "BORDERS")) {
                Body& b = m.bodies.at(index_body);
                int nbr;
                in >> nbr;
                std::getline(in, token);
                std::getline(in, token);
                if (nbr != (int)b.borders.size()) { err = "border count differs between control and data files"; return false; }
                for (int i = 0; i < nbr; i++) {
                    int nn, before, after;
                    std::getline(in, token);
                    in >> nn >> before >> after;
                    std::getline(in, token);
                    Border& B = b.borders[i];
                    if (!in || nn < 0 || nn > (1 << 28)) { err = "border node count"; return false; }
                    B.nodes.resize(nn);
                    for (int j = 0; j < nn; j++) {
                        in >> B.nodes[j];
                        std::getline(in, token);
                    }
                    B.before = before;
                    B.after = after;
                    std::getline(in, token);
                }
                if (!check_borders(b, index_body, err)) return false;
                initialize_segments(m, b);
            }
            if (starts(line, "INTEGRATION")) {
                Body& b = m.bodies.at(index_body);
                std::string integ;
                std::getline(in, integ);
                if (!starts(integ, "Gauss")) { err = "only Gauss integration is supported"; return false; }
                int G, R;
                in >> G >> R;
                std::getline(in, token);
                std::getline(in, token);
                b.n_gauss = G;
                b.n_regions = R;
                b.gx.resize(G); b.gy.resize(G); b.gw.resize(G); b.gjac.resize(G);
                b.gptr.assign(G + 1, 0);
                for (int g = 0; g < G; g++) {
                    std::getline(in, token);
                    in >> b.gx[g] >> b.gy[g] >> b.gw[g] >> b.gjac[g];
                    std::getline(in, token);
                    int k;
                    in >> k;
                    std::getline(in, token);
                    for (int j = 0; j < k; j++) {
                        int node, n1, n2;
                        double f, fx, fy;
                        in >> node >> n1 >> n2 >> f >> fx >> fy;
                        std::getline(in, token);
                        b.gidx.push_back(node);
                        b.gphi.push_back(f);
                        b.gphix.push_back(fx);
                        b.gphiy.push_back(fy);
                    }
                    b.gptr[g + 1] = b.gptr[g] + k;
                    std::getline(in, token);
                }
            }
            if (starts(line, "TRIANGULATION")) {
                Body& b = m.bodies.at(index_body);
                int nc;
                in >> nc;
                std::getline(in, token);
                b.cells.resize(3 * nc);
                for (int i = 0; i < nc; i++) {
                    int n0;
                    in >> n0 >> b.cells[3 * i] >> b.cells[3 * i + 1] >> b.cells[3 * i + 2];
                    std::getline(in, token);
                }
                std::getline(in, token);
            }
            if (starts(line, "DOF_TO_DISP")) {
                Body& b = m.bodies.at(index_body);
                b.dptr.assign(b.n_nodes + 1, 0);
                for (int i = 0; i < b.n_nodes; i++) {
                    std::getline(in, token);
                    std::getline(in, token);
                    int k;
                    in >> k;
                    std::getline(in, token);
                    for (int j = 0; j < k; j++) {
                        int nn;
                        double ss, sx, sy;
                        in >> nn >> ss >> sx >> sy;
                        std::getline(in, token);
                        b.didx.push_back(nn);
                        b.dphi.push_back(ss);
                        b.dphix.push_back(sx);
                        b.dphiy.push_back(sy);
                    }
                    b.dptr[i + 1] = b.dptr[i] + k;
                }
            }
            if (starts(line, "MATRICES")) {
                Body& b = m.bodies.at(index_body);
                int nmat;
                in >> nmat;
                std::getline(in, token);
                std::getline(in, token);
                for (int i = 0; i < nmat; i++) {
                    int d0, d1;
                    double s, d;
                    in >> d0 >> d1 >> s >> d;
                    std::getline(in, token);
                    b.mat_dof.push_back(d0);
                    b.mat_dof.push_back(d1);
                    b.mat_stiffness.push_back(s);
                    b.mat_damping.push_back(d);
                }
            }
        }
    }
    build_table(m);
    // ---------------- DYNAMIC_nnnnn.asc (Code/IO.h:1871-2343) ----------------
    {
        char name[64];
        std::snprintf(name, sizeof name, "/DYNAMIC_%05d.asc", save_index);
        std::ifstream in(dir + name);
        if (!in) { err = "cannot open " + dir + name; return false; }
        for (auto& b : m.bodies) finish_body_state(m, b);
        bool rigid = false;
        static const char* flag_names[F_COUNT] = {"KILL_VELOCITY", "MONITOR_ENERGY", "MONITOR_BOUNDARIES",
                                                  "INITIALIZE_CZM", "UPDATE_MASS_MATRIX", "UPDATE_STIFFNESS_MATRIX",
                                                  "UPDATE_DAMPING_MATRIX", "NO_LOG", "NO_MONITORING",
                                                  "NO_SELF_CONTACT", "RESET_WORK"};
        while (std::getline(in, line)) {
            for (int f = 0; f < F_COUNT; f++)
                if (starts(line, flag_names[f])) m.flags[f] = 1;
            if (starts(line, "SOLVER")) {
                // the DYNAMIC file names the scheme that runs (Load_dynamic,
                // Code/IO.h:1936-1940); its parameters come from the static file
                std::string solver;
                in >> solver;
                std::getline(in, token);
                const int sch = scheme_of(solver);
                if (sch < 0) { err = "unsupported solver " + solver + " in " + name; return false; }
                if (sch != m.scheme && (sch == SCH_MASS_SCALING || sch == SCH_ADAPTIVE_MASS_SCALING) &&
                    m.scheme != SCH_MASS_SCALING && m.scheme != SCH_ADAPTIVE_MASS_SCALING) {
                    err = "the DYNAMIC file asks for " + solver + ", whose mass-scaling parameters the static file "
                          "(solver " + m.solver + ") does not give";
                    return false;
                }
                m.solver = solver;
                m.scheme = sch;
                in >> m.number_save >> m.number_iteration >> m.time >> m.number_print;
                std::getline(in, token);
                in >> m.dt;
                std::getline(in, token);
                in >> m.next_save >> m.next_print >> m.next_contact_update;
                std::getline(in, token);
            }
            bool def = starts(line, "DEFORMABLE"), rig = starts(line, "RIGID");
            if (def || rig) {
                std::string st;
                in >> st;
                std::getline(in, token);
                in >> index_body;
                std::getline(in, token);
                m.bodies.at(index_body).status = status_of(st);
                rigid = rig;
            }
            if (starts(line, "KINEMATICS")) {
                Body& b = m.bodies.at(index_body);
                if (!rigid) {
                    int n;
                    in >> n;
                    std::getline(in, token);
                    for (int i = 0; i < n && i < b.n_nodes; i++) {
                        double j, xc, yc;
                        in >> j >> xc >> yc >> b.ux[i] >> b.uy[i] >> b.vx[i] >> b.vy[i] >> b.ax[i] >> b.ay[i] >>
                            b.upx[i] >> b.upy[i] >> b.vpx[i] >> b.vpy[i] >> b.apx[i] >> b.apy[i];
                        std::getline(in, token);
                        if (m.flags[F_KILL_VELOCITY]) { b.vx[i] = b.vy[i] = b.vpx[i] = b.vpy[i] = 0.0; }
                    }
                } else {
                    in >> b.rigid_cur[0] >> b.rigid_cur[1] >> b.rigid_cur[2];
                    std::getline(in, token);
                    in >> b.rigid_disp[0] >> b.rigid_disp[1] >> b.rigid_disp[2];
                    std::getline(in, token);
                    in >> b.rigid_vel[0] >> b.rigid_vel[1] >> b.rigid_vel[2];
                    std::getline(in, token);
                    if (m.flags[F_KILL_VELOCITY]) b.rigid_vel[0] = b.rigid_vel[1] = b.rigid_vel[2] = 0.0;
                }
            }
            if (starts(line, "WORKS")) {
                Body& b = m.bodies.at(index_body);
                double w[7];
                for (int k = 0; k < 7; k++) in >> w[k];
                std::getline(in, token);
                for (int k = 0; k < 7; k++) b.works[k] = m.flags[F_RESET_WORK] ? 0.0 : w[k];
            }
            if (starts(line, "NEIGHBOURS")) {
                Body& b = m.bodies.at(index_body);
                int n;
                in >> n;
                std::getline(in, token);
                b.neighbours.resize(3 * n);
                for (int i = 0; i < n; i++) {
                    in >> b.neighbours[3 * i] >> b.neighbours[3 * i + 1] >> b.neighbours[3 * i + 2];
                    std::getline(in, token);
                }
            }
            if (starts(line, "BORDERS")) {
                Body& b = m.bodies.at(index_body);
                for (size_t i = 0; i < b.borders.size(); i++) {
                    std::getline(in, token);
                    std::getline(in, token);
                    int n;
                    in >> n;
                    std::getline(in, token);
                    for (int j = 0; j < n; j++) std::getline(in, token);
                }
            }
            if (starts(line, "CONTACTS")) {
                Body& b = m.bodies.at(index_body);
                int n;
                in >> n;
                std::getline(in, token);
                b.contacts.assign(n, ContactElement());
                for (int i = 0; i < n; i++) {
                    ContactElement& c = b.contacts[i];
                    int ni;
                    double N[4];
                    int nd[4], sh[4];
                    in >> c.borderS >> c.border_nodeS >> c.nodeS >> c.bodyM >> c.borderM >> c.shiftM >> c.segmentM >>
                        c.gapn >> c.gapt >> c.xsi;
                    in >> c.xnorm >> c.ynorm >> c.xtan >> c.ytan;
                    in >> c.length >> c.fx >> c.fy;
                    for (int k = 0; k < 4; k++) in >> nd[k] >> N[k] >> sh[k];
                    in >> c.effective_mass >> ni;
                    // Code/IO.h:2270-2281 stores all four master nodes into slot 0;
                    // Update_contacts rebuilds them from segmentM before use.
                    c.nodeM[0] = nd[3]; c.shapeM[0] = N[3]; c.shiftMk[0] = sh[3];
                    if (!in || ni < 0 || ni > 64) { err = "contact element internal variables"; return false; }
                    c.nb_internal = ni;
                    if (ni == 0) c.internal.assign(1, 0.0);
                    for (int k = 0; k < ni; k++) {
                        double v;
                        in >> v;
                        c.internal.push_back(v);
                    }
                }
            }
        }
    }
    return true;
}
}  // namespace

bool load_model_files(const std::string& dir, int save_index, Model& m, std::string& err) {
    try {
        if (!load_model_files_impl(dir, save_index, m, err)) return false;
    } catch (std::exception& e) {
        err = std::string("malformed model files: ") + e.what();
        return false;
    }
    return validate_model(m, err);
}

// ---------------------------------------------------------------------------
// binary container
// ---------------------------------------------------------------------------

namespace {
struct Rec {
    int kind = 0;
    std::vector<double> d;
    std::vector<int> i;
};
}  // namespace

bool load_model_binary(const std::string& path, Model& m, std::string& err) {
    m = Model();
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) { err = "cannot open " + path; return false; }
    // every length is checked against the bytes left before anything is
    // allocated, so a damaged container is an error, not a huge allocation
    const uint64_t size = (uint64_t)in.tellg();
    in.seekg(0);
    auto left = [&]() { return size - (uint64_t)in.tellg(); };
    std::map<std::string, Rec> recs;
    while (true) {
        uint32_t l;
        if (!in.read((char*)&l, 4)) break;
        if ((uint64_t)l + 9 > left()) { err = "malformed record header in " + path; return false; }
        std::string name(l, '\0');
        in.read(&name[0], l);
        uint8_t kind;
        uint64_t n;
        in.read((char*)&kind, 1);
        in.read((char*)&n, 8);
        if (!in || kind > 1 || n > left() / (kind == 0 ? 8 : 4)) {
            err = "malformed record " + name.substr(0, 64) + " in " + path;
            return false;
        }
        Rec& r = recs[name];
        r.kind = kind;
        if (kind == 0) { r.d.resize(n); in.read((char*)r.d.data(), 8 * n); }
        else { r.i.resize(n); in.read((char*)r.i.data(), 4 * n); }
        if (!in) { err = "truncated record " + name; return false; }
    }
    auto D = [&](const std::string& k) -> std::vector<double>& {
        auto it = recs.find(k);
        if (it == recs.end()) throw std::runtime_error("missing record " + k);
        return it->second.d;
    };
    auto I = [&](const std::string& k) -> std::vector<int>& {
        auto it = recs.find(k);
        if (it == recs.end()) throw std::runtime_error("missing record " + k);
        return it->second.i;
    };
    auto need = [&](bool ok, const std::string& what) {
        if (!ok) throw std::runtime_error("malformed container " + path + ": " + what);
    };
    try {
        auto& h = D("model.header");
        need(h.size() >= 17, "model.header");
        m.tini = h[0]; m.dt = h[1]; m.tend = h[2]; m.save_period = h[3]; m.print_period = h[4];
        m.contact_update_period = h[5]; m.xmin_period = h[6]; m.xmax_period = h[7];
        m.gravity_x = h[8]; m.gravity_y = h[9]; m.time = h[10]; m.number_save = (int)h[11];
        m.number_iteration = (int)h[12]; m.number_print = (int)h[13]; m.next_save = h[14];
        m.next_print = h[15]; m.next_contact_update = h[16];
        auto& fl = I("model.flags");
        need(fl.size() >= (size_t)F_COUNT, "model.flags");
        for (int f = 0; f < F_COUNT; f++) m.flags[f] = fl[f];
        // optional: the scheme and its error control (absent: Euler)
        if (recs.count("model.monitoring")) {
            // the MONITORING entries' text, one character per entry
            auto& tx = I("model.monitoring");
            std::string text;
            for (int ch : tx) {
                need(ch > 0 && ch < 128, "model.monitoring text");
                text.push_back((char)ch);
            }
            std::istringstream rs(text);
            int n = 0;
            rs >> n;
            std::string rest;
            std::getline(rs, rest);
            need(n >= 0 && n < (1 << 20), "model.monitoring count");
            std::vector<double> q;
            for (int i = 0; i < n; i++) {
                need(read_spy_line(rs, q), "model.monitoring entry");
                m.monitored.push_back(q);
                std::getline(rs, rest);
            }
        }
        if (recs.count("model.deactivation")) {
            // the rules' text, one character per entry, read as the ASCII
            // reader reads the DEACTIVATION section
            auto& tx = I("model.deactivation");
            std::string text;
            for (int ch : tx) {
                need(ch > 0 && ch < 128, "model.deactivation text");
                text.push_back((char)ch);
            }
            std::istringstream rs(text);
            int n = 0;
            rs >> n;
            std::string rest;
            std::getline(rs, rest);
            need(n >= 0 && n < (1 << 20), "model.deactivation count");
            if (!read_deactivation_rules(rs, n, m, err)) return false;
        }
        if (recs.count("model.solver")) {
            auto& sv = D("model.solver");
            need(sv.size() == 8, "model.solver size");
            const int sch = (int)sv[0];
            need(sch >= SCH_EULER && sch <= SCH_ADAPTIVE_MASS_SCALING && sv[0] == (double)sch, "model.solver scheme");
            static const char* names[] = {"Euler", "Adaptive_Euler", "Mass_Scaling", "Adaptive_Mass_Scaling"};
            m.scheme = sch;
            m.solver = names[sch];
            m.target_error = sv[1]; m.control_parameter = sv[2]; m.accepted_ratio = sv[3];
            m.inv_target_error = 1 / m.target_error;
            m.max_mass_scaling = sv[4]; m.control_parameter_mass_scaling = sv[5];
            m.error_factor_mass_scaling = sv[6]; m.decrease_factor_mass_scaling = sv[7];
        }
        auto& mt = I("model.material_type");
        auto& mr = D("model.material_raw");
        need(mr.size() == 6 * mt.size(), "model.material_raw size");
        for (size_t k = 0; k < mt.size(); k++) {
            need(mt[k] == MAT_ELASTIC_LINEAR || mt[k] == MAT_NEOHOOKEAN, "material type");
            Material mat;
            mat.name = "Material" + std::to_string(k);
            mat.type = mt[k];
            for (int q = 0; q < 6; q++) mat.raw[q] = mr[6 * k + q];
            convert_material(mat);
            m.materials.push_back(mat);
        }
        auto& lm = I("model.law_materials");
        auto& lt = I("model.law_type");
        auto& lf = I("model.law_fixed");
        auto& lp = I("model.law_ptr");
        auto& lv = D("model.law_params");
        need(lm.size() == 2 * lt.size() && lf.size() == lt.size(), "law records sizes");
        need(csr_ok(lp, lt.size(), lv.size()), "model.law_ptr");
        for (size_t k = 0; k < lt.size(); k++) {
            need(lt[k] >= LAW_FRICTIONLESS && lt[k] <= LAW_CZM_FATIGUE, "law type");
            ContactLaw law;
            law.m1 = m.materials.at(lm[2 * k]).name;
            law.m2 = m.materials.at(lm[2 * k + 1]).name;
            law.type = lt[k];
            law.fixed_length = lf[k] != 0;
            law.p.assign(lv.begin() + lp[k], lv.begin() + lp[k + 1]);
            if ((int)law.p.size() != law_nparams(law.type)) { err = "law parameter count"; return false; }
            m.laws.push_back(law);
        }
        build_table(m);
        auto& nbr_rec = I("model.n_bodies");
        need(nbr_rec.size() == 1 && nbr_rec[0] >= 0 && (size_t)nbr_rec[0] <= recs.size(), "model.n_bodies");
        int nb = nbr_rec[0];
        m.bodies.resize(nb);
        for (int bi = 0; bi < nb; bi++) {
            std::string p = "b" + std::to_string(bi) + ".";
            Body& b = m.bodies[bi];
            auto& info = I(p + "info");
            need(info.size() >= 7, p + "info");
            need(info[1] >= 0 && info[1] < (int)m.materials.size(), p + "info material");
            need(info[4] >= 0 && info[5] >= 0 && info[6] >= 1, p + "info sizes");
            b.rigid = info[0] != 0;
            b.material = info[1];
            b.periodicity = info[2];
            b.status = info[3];
            b.n_nodes = info[4];
            b.n_gauss = info[5];
            b.n_regions = info[6];
            b.density = m.materials.at(b.material).p[0];
            auto& dist = D(p + "dist");
            need(dist.size() >= 3, p + "dist");
            b.nodal_distance = dist[0]; b.detection_distance = dist[1]; b.contact_distance = dist[2];
            b.X = D(p + "X");
            b.Y = D(p + "Y");
            auto& bptr = I(p + "border_ptr");
            auto& bnod = I(p + "border_nodes");
            auto& meta = I(p + "border_meta");
            auto& bcp = D(p + "border_bcparams");
            auto& hptr = I(p + "border_hist_ptr");
            auto& ht = D(p + "hist_t");
            auto& hv = D(p + "hist_v");
            const size_t nbr_z = bptr.empty() ? 0 : bptr.size() - 1;
            need(nbr_z >= 1 && csr_ok(bptr, nbr_z, bnod.size()), p + "border_ptr");
            need(meta.size() == 5 * nbr_z && bcp.size() == 4 * nbr_z, p + "border_meta / border_bcparams size");
            need(hptr.size() == 2 * nbr_z + 1 && ht.size() == hv.size() && csr_ok(hptr, 2 * nbr_z, ht.size()),
                 p + "border_hist_ptr");
            for (size_t k = 0; k < nbr_z; k++)
                need(meta[5 * k + 3] >= BC_NONE && meta[5 * k + 3] <= BC_FOLLOWING && meta[5 * k + 4] >= BC_NONE &&
                         meta[5 * k + 4] <= BC_FOLLOWING, p + "border BC type");
            need(b.X.size() == (size_t)b.n_nodes && b.Y.size() == (size_t)b.n_nodes, p + "X / Y size");
            int nbr = (int)nbr_z;
            b.borders.resize(nbr);
            for (int k = 0; k < nbr; k++) {
                Border& B = b.borders[k];
                B.nodes.assign(bnod.begin() + bptr[k], bnod.begin() + bptr[k + 1]);
                B.periodicity = meta[5 * k];
                B.before = meta[5 * k + 1];
                B.after = meta[5 * k + 2];
                B.xbc.type = meta[5 * k + 3];
                B.ybc.type = meta[5 * k + 4];
                B.xbc.stiffness = bcp[4 * k]; B.xbc.damping = bcp[4 * k + 1];
                B.ybc.stiffness = bcp[4 * k + 2]; B.ybc.damping = bcp[4 * k + 3];
                B.xbc.hist.t.assign(ht.begin() + hptr[2 * k], ht.begin() + hptr[2 * k + 1]);
                B.xbc.hist.v.assign(hv.begin() + hptr[2 * k], hv.begin() + hptr[2 * k + 1]);
                B.ybc.hist.t.assign(ht.begin() + hptr[2 * k + 1], ht.begin() + hptr[2 * k + 2]);
                B.ybc.hist.v.assign(hv.begin() + hptr[2 * k + 1], hv.begin() + hptr[2 * k + 2]);
            }
            {
                std::string berr;
                need(check_borders(b, bi, berr), berr);
            }
            initialize_segments(m, b);
            b.cells = I(p + "cells");
            if (!b.rigid) {
                b.domain = D(p + "domain");
                b.mx = D(p + "mx"); b.my = D(p + "my");
                need(b.mx.size() == (size_t)b.n_nodes && b.my.size() == (size_t)b.n_nodes, p + "mx / my size");
                b.imx.resize(b.n_nodes); b.imy.resize(b.n_nodes);
                for (int i = 0; i < b.n_nodes; i++) { b.imx[i] = 1.0 / b.mx[i]; b.imy[i] = 1.0 / b.my[i]; }
                auto& gq = D(p + "gauss");
                need(gq.size() == 4 * (size_t)b.n_gauss, p + "gauss size");
                b.gx.resize(b.n_gauss); b.gy.resize(b.n_gauss); b.gw.resize(b.n_gauss); b.gjac.resize(b.n_gauss);
                for (int g = 0; g < b.n_gauss; g++) {
                    b.gx[g] = gq[4 * g]; b.gy[g] = gq[4 * g + 1]; b.gw[g] = gq[4 * g + 2]; b.gjac[g] = gq[4 * g + 3];
                }
                b.gptr = I(p + "gptr"); b.gidx = I(p + "gidx");
                b.gphi = D(p + "gphi"); b.gphix = D(p + "gphix"); b.gphiy = D(p + "gphiy");
                b.dptr = I(p + "dptr"); b.didx = I(p + "didx");
                b.dphi = D(p + "dphi"); b.dphix = D(p + "dphix"); b.dphiy = D(p + "dphiy");
                finish_body_state(m, b);
                auto& kin = D(p + "kin");
                need(kin.size() == 12 * (size_t)b.n_nodes, p + "kin size");
                std::vector<double>* f[12] = {&b.ux, &b.uy, &b.vx, &b.vy, &b.ax, &b.ay,
                                              &b.upx, &b.upy, &b.vpx, &b.vpy, &b.apx, &b.apy};
                for (int q = 0; q < 12; q++)
                    for (int i = 0; i < b.n_nodes; i++) (*f[q])[i] = kin[(size_t)q * b.n_nodes + i];
            } else {
                auto& r = D(p + "rigid");
                need(r.size() >= 6, p + "rigid");
                b.xc = r[0]; b.yc = r[1]; b.mass = r[2]; b.inertia = r[3]; b.inverse_mass = r[4]; b.inverse_inertia = r[5];
                auto& rk = D(p + "rkin");
                need(rk.size() >= 9, p + "rkin");
                for (int q = 0; q < 3; q++) { b.rigid_cur[q] = rk[q]; b.rigid_disp[q] = rk[3 + q]; b.rigid_vel[q] = rk[6 + q]; }
                b.domain.assign(b.n_nodes, 0.0);
                b.mx.assign(b.n_nodes, 0.0); b.my.assign(b.n_nodes, 0.0);
                b.imx.assign(b.n_nodes, 0.0); b.imy.assign(b.n_nodes, 0.0);
                finish_body_state(m, b);
            }
        }
    } catch (std::exception& e) {
        err = e.what();
        return false;
    }
    return validate_model(m, err);
}

// ---------------------------------------------------------------------------
// DYNAMIC writer (layout of Code/IO.h:2351-2636, full precision)
// ---------------------------------------------------------------------------

bool write_monitoring_text(const std::string& path, const std::vector<double>& rows, std::string& err) {
    std::ofstream f(path, std::ios::app);
    if (!f) { err = "cannot open " + path; return false; }
    f << std::setprecision(10);
    for (size_t i = 0; i < rows.size();) {
        const size_t nv = (size_t)rows[i];
        if (i + 1 + nv > rows.size()) { err = "malformed monitoring rows"; return false; }
        for (size_t k = 0; k < nv; k++) f << rows[i + 1 + k] << ' ';
        f << '\n';
        i += 1 + nv;
    }
    f.flush();
    if (!f) { err = "cannot write " + path; return false; }
    return true;
}

bool write_dynamic_text(const std::string& path, const Model& m, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) { err = "cannot write " + path; return false; }
    auto g = [&](double x) { std::fprintf(f, "%.17g ", x); };
    std::fprintf(f, " \nSIMULATION_NAME\n%s\n \nSOLVER\n%s\n%d %d %.17g %d \n%.17g\n%.17g %.17g %.17g \n",
                 m.name.c_str(), m.solver.c_str(), m.number_save, m.number_iteration, m.time, m.number_print,
                 m.dt, m.next_save, m.next_print, m.next_contact_update);
    static const char* flag_names[F_COUNT] = {"KILL_VELOCITY", "MONITOR_ENERGY", "MONITOR_BOUNDARIES",
                                              "INITIALIZE_CZM", "UPDATE_MASS_MATRIX", "UPDATE_STIFFNESS_MATRIX",
                                              "UPDATE_DAMPING_MATRIX", "NO_LOG", "NO_MONITORING",
                                              "NO_SELF_CONTACT", "RESET_WORK"};
    for (int k = 0; k < F_COUNT; k++)
        if (m.flags[k] && k != F_INITIALIZE_CZM && k != F_KILL_VELOCITY) std::fprintf(f, "%s\n", flag_names[k]);
    std::fprintf(f, " \n");
    for (size_t bi = 0; bi < m.bodies.size(); bi++) {
        const Body& b = m.bodies[bi];
        const char* st = b.status == ST_INACTIVE ? "inactive" : b.status == ST_NOGRAVITY ? "nogravity" : "active";
        std::fprintf(f, " \n%s\n%s\n%zu\n \nKINEMATICS\n", b.rigid ? "RIGID" : "DEFORMABLE", st, bi);
        if (!b.rigid) {
            std::fprintf(f, "%d\n", b.n_nodes);
            for (int i = 0; i < b.n_nodes; i++) {
                std::fprintf(f, "%d ", i);
                g(b.X[i] + b.ux[i]); g(b.Y[i] + b.uy[i]);
                g(b.ux[i]); g(b.uy[i]); g(b.vx[i]); g(b.vy[i]); g(b.ax[i]); g(b.ay[i]);
                g(b.upx[i]); g(b.upy[i]); g(b.vpx[i]); g(b.vpy[i]); g(b.apx[i]); g(b.apy[i]);
                std::fprintf(f, "\n");
            }
        } else {
            for (int q = 0; q < 3; q++) g(b.rigid_cur[q]);
            std::fprintf(f, "\n");
            for (int q = 0; q < 3; q++) g(b.rigid_disp[q]);
            std::fprintf(f, "\n");
            for (int q = 0; q < 3; q++) g(b.rigid_vel[q]);
            std::fprintf(f, "\n");
        }
        // FORCES (Code/IO.h:2455-2500)
        std::fprintf(f, " \nFORCES\n");
        if (!b.rigid) {
            for (int i = 0; i < b.n_nodes; i++) {
                std::fprintf(f, "%d ", i);
                for (int k = 0; k < 16; k++)
                    g(b.forces_out.size() == (size_t)16 * b.n_nodes ? b.forces_out[16 * (size_t)i + k] : 0.);
                std::fprintf(f, "\n");
            }
        } else {
            for (int r = 0; r < 6; r++) {
                for (int q = 0; q < 3; q++) g(b.rigid_forces_out[3 * r + q]);
                std::fprintf(f, "\n");
            }
        }
        std::fprintf(f, " \nWORKS\n");
        for (int k = 0; k < 7; k++) g(b.works[k]);
        std::fprintf(f, "\n \nNEIGHBOURS\n%zu\n", b.neighbours.size() / 3);
        for (size_t k = 0; k < b.neighbours.size() / 3; k++)
            std::fprintf(f, "%d %d %d\n", b.neighbours[3 * k], b.neighbours[3 * k + 1], b.neighbours[3 * k + 2]);
        // BORDERS: lengths and segment stencils (Code/IO.h:2530-2552)
        std::fprintf(f, " \nBORDERS\n \n");
        for (size_t j = 0; j < b.borders.size(); j++) {
            const Border& B = b.borders[j];
            std::fprintf(f, "%zu\n%zu\n", j, B.nodes.size());
            for (size_t k = 0; k < B.nodes.size(); k++) {
                g(B.length_out.size() == B.nodes.size() ? B.length_out[k] : B.initial_length[k]);
                std::fprintf(f, "%d %d %d %d %d %d %d %d\n", B.node0[k], B.shift0[k], B.node1[k], B.shift1[k],
                             B.node2[k], B.shift2[k], B.node3[k], B.shift3[k]);
            }
            std::fprintf(f, " \n");
        }
        std::fprintf(f, " \nCONTACTS\n%zu\n", b.contacts.size());
        for (const ContactElement& c : b.contacts) {
            std::fprintf(f, "%d %d %d %d %d %d %d ", c.borderS, c.border_nodeS, c.nodeS, c.bodyM, c.borderM, c.shiftM,
                         c.segmentM);
            g(c.gapn); g(c.gapt); g(c.xsi); g(c.xnorm); g(c.ynorm); g(c.xtan); g(c.ytan);
            g(c.length); g(c.fx); g(c.fy);
            for (int k = 0; k < 4; k++) { std::fprintf(f, "%d ", c.nodeM[k]); g(c.shapeM[k]); std::fprintf(f, "%d ", c.shiftMk[k]); }
            g(c.effective_mass);
            std::fprintf(f, "%d ", c.nb_internal);
            for (int k = 0; k < c.nb_internal; k++) g(c.internal[k]);
            std::fprintf(f, "\n");
        }
        // CONTACT_PRESSURES and BC_PRESSURES per slot (Code/IO.h:2600-2630)
        const char* sec[2] = {"CONTACT_PRESSURES", "BC_PRESSURES"};
        for (int q = 0; q < 2; q++) {
            std::fprintf(f, " \n%s\n \n", sec[q]);
            for (size_t j = 0; j < b.borders.size(); j++) {
                const Border& B = b.borders[j];
                const std::vector<double>& px = q == 0 ? B.xcp_out : B.xbp_out;
                const std::vector<double>& py = q == 0 ? B.ycp_out : B.ybp_out;
                std::fprintf(f, "%zu\n%zu\n", j, B.nodes.size());
                for (size_t k = 0; k < B.nodes.size(); k++) {
                    g(px.size() == B.nodes.size() ? px[k] : 0.);
                    g(py.size() == B.nodes.size() ? py[k] : 0.);
                    std::fprintf(f, "\n");
                }
                std::fprintf(f, " \n");
            }
        }
        std::fprintf(f, " \n");
    }
    std::fclose(f);
    return true;
}

// ---------------------------------------------------------------------------
// Graphic writer (layout of Write_graphic, Code/Graphic.h:21-946)
// ---------------------------------------------------------------------------
bool write_graphic_text(const std::string& path, const Model& m, const std::vector<GraphicBody>& g, int iteration,
                        int number_save, int number_print, double time, const int* to_plot, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) { err = "cannot write " + path; return false; }
    const double xmin = m.xmin_period, xmax = m.xmax_period, period = xmax - xmin;
    // points (body, node, x, y) and triangles of every active body
    struct Pt { int b, i; double x, y; };
    std::vector<Pt> pts;
    std::vector<std::array<int, 3>> tris;
    for (size_t bi = 0; bi < m.bodies.size(); bi++) {
        const Body& b = m.bodies[bi];
        const GraphicBody& G = g[bi];
        if (!G.active) continue;
        const int base = (int)pts.size();
        double ymin = 1.e6, ymax = -1.e6;
        for (int i = 0; i < b.n_nodes; i++) {
            const double x = G.x[i] - period * std::floor((G.x[i] - xmin) / period);
            pts.push_back({(int)bi, i, x, G.y[i]});
            ymin = std::min(ymin, G.y[i]);
            ymax = std::max(ymax, G.y[i]);
        }
        // a cell with vertices within `dist` of both ends of the period
        // crosses the seam: it is drawn at each end, its vertices from the
        // other end copied across by one period
        const double dist = b.n_nodes > 0 ? 6 * std::sqrt((period * (ymax - ymin)) / b.n_nodes) : 0.;
        for (size_t c = 0; c + 2 < b.cells.size(); c += 3) {
            const int v[3] = {b.cells[c], b.cells[c + 1], b.cells[c + 2]};
            bool lo = false, hi = false;
            for (int k = 0; k < 3; k++) {
                lo |= pts[base + v[k]].x <= xmin + dist;
                hi |= pts[base + v[k]].x >= xmax - dist;
            }
            if (!(lo && hi)) {
                tris.push_back({base + v[0], base + v[1], base + v[2]});
                continue;
            }
            for (int side = 0; side < 2; side++) {
                std::array<int, 3> t;
                for (int k = 0; k < 3; k++) {
                    const Pt& p = pts[base + v[k]];
                    const bool far = side == 0 ? p.x >= xmax - dist : p.x <= xmin + dist;
                    if (!far) { t[k] = base + v[k]; continue; }
                    t[k] = (int)pts.size();
                    pts.push_back({p.b, p.i, p.x + (side == 0 ? -period : period), p.y});
                }
                tris.push_back(t);
            }
        }
    }
    auto clamp = [](double v) { return std::fabs(v) < 1.e-16 ? 0. : v; };
    std::fprintf(f, "# vtk DataFile Version 2.0\nMELODY 2D graphic file ; Iteration %d ; Save %d ; Print %d ; Time %.12g\n"
                    "ASCII\n\nDATASET POLYDATA\n\n", iteration, number_save, number_print, time);
    std::fprintf(f, "POINTS %zu float\n", pts.size());
    for (const Pt& p : pts) std::fprintf(f, "%.12g %.12g 0\n", clamp(p.x), clamp(p.y));
    std::fprintf(f, "\nPOLYGONS %zu %zu\n", tris.size(), 4 * tris.size());
    for (const auto& t : tris) std::fprintf(f, "3 %d %d %d\n", t[0], t[1], t[2]);
    std::fprintf(f, "\nPOINT_DATA %zu\n", pts.size());
    static const char* names[GRAPHIC_FIELDS] = {
        "00_Body_index", "01_Initial_position", "02_Current_position", "03_Displacement", "04_Velocity",
        "05_Acceleration", "06_Force", "07_Internal_force", "08_Contact_force", "09_Body_force", "10_Dirichlet_force",
        "11_Neumann_force", "12_Damping_force", "13_Alid_force", "14_Jacobian", "15_Cauchy_XX_stress",
        "16_Cauchy_YY_stress", "17_Cauchy_XY_stress", "18_Cauchy_ZZ_stress", "19_Tresca_stress",
        "20_Von_Mises_stress", "21_Major_principal_stress", "22_Intermediate_principal_stress",
        "23_Minor_principal_stress", "24_Spherical_stress", "25_Green-Lagrange_XX_Strain",
        "26_Green-Lagrange_YY_Strain", "27_Green-Lagrange_XY_Strain", "28_Green-Lagrange_Norm", "29_Body_Damage",
        "30_Normalized_displacement_error", "31_Displacement_error", "32_Internal_work", "33_Contact_work",
        "34_Body_work", "35_Dirichlet_work", "36_Neumann_work", "37_Damping_work", "38_Alid_work", "39_Temperature",
        "40-Epsilon_Mass_Scaling"};
    // fields of the 19 stress fields behind slots 14-28 (mel_stress_field order)
    static const int stress_of[15] = {3, 8, 9, 10, 11, 12, 13, 0, 1, 2, 14, 15, 16, 17, 18};
    for (int s = 0; s < GRAPHIC_FIELDS; s++) {
        if (!to_plot[s]) continue;
        const bool two = (s >= 1 && s <= 13) || s == 31;
        std::fprintf(f, "SCALARS %s float %d\nLOOKUP_TABLE default\n", names[s], two ? 2 : 1);
        for (const Pt& p : pts) {
            const GraphicBody& G = g[p.b];
            const size_t n = G.x.size();
            double a = 0., b = 0.;
            if (s == 0) { std::fprintf(f, "%d\n", p.b); continue; }
            if (s == 1) { a = G.x0[p.i]; b = G.y0[p.i]; }
            else if (s == 2) { a = G.x[p.i]; b = G.y[p.i]; }
            else if (s >= 3 && s <= 12) { const std::vector<double>& v = G.vec[s - 3]; a = v[p.i]; b = v[n + p.i]; }
            else if (s >= 14 && s <= 28) { a = G.stress.empty() ? (s == 14 ? 1. : 0.) : G.stress[(size_t)stress_of[s - 14] * n + p.i]; }
            else if (s >= 32 && s <= 37) a = G.works[s - 32];   // internal contact body dirichlet neumann damping
            if (two) std::fprintf(f, "%.6g %.6g\n", clamp(a), clamp(b));
            else if (s == 14 || s >= 32) std::fprintf(f, "%.6g\n", a);
            else std::fprintf(f, "%.6g\n", clamp(a));
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return true;
}

int scheme_of(const std::string& name) {
    if (name == "Euler") return SCH_EULER;
    if (name == "Adaptive_Euler") return SCH_ADAPTIVE_EULER;
    if (name == "Mass_Scaling") return SCH_MASS_SCALING;
    if (name == "Adaptive_Mass_Scaling") return SCH_ADAPTIVE_MASS_SCALING;
    return -1;
}

}  // namespace mel
