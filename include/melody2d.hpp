/*
 * melody2d.hpp - C++ façade over melody2d.h with the reference's call
 * structure: a Model loaded from the reference's model files (Load_static,
 * Load_dynamic and the initialisation block of Code/main.cpp:83-126), its
 * Bodies, and a Solver whose Euler_step advances the model by one step
 * (Code/Solver.h:10-88) with the contact-update cadence of
 * Code/main.cpp:191-196.  Header-only; it adds no state of its own beyond
 * the context handle, and every call goes through the C-ABI.
 *
 *   melody2d::Model model = melody2d::Model::load(".", number_save);
 *   melody2d::Solver solver;
 *   while (model.time() < tend) solver.Euler_step(model);
 *   model.write_dynamic("DYNAMIC_00001.asc");
 */
#ifndef MELODY2D_HPP
#define MELODY2D_HPP

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "melody2d.h"

namespace melody2d {

class Error : public std::runtime_error {
public:
    Error(mel_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
    mel_status status;
};

inline void check(mel_status st, const mel_ctx* ctx) {
    if (st != MEL_OK) {
        const char* m = mel_last_error(ctx);
        throw Error(st, m ? m : "melody2d error");
    }
}

class Model;

// one body of a loaded model: node fields, contact elements, rigid state
class Body {
public:
    Body(mel_ctx* ctx, int index, int n_nodes) : ctx_(ctx), index_(index), n_nodes_(n_nodes) {}
    int index() const { return index_; }
    int nb_nodes() const { return n_nodes_; }
    // a node field of every node of the body (mel_node_field order)
    std::vector<double> field(mel_node_field f) const {
        std::vector<double> v((size_t)n_nodes_);
        check(mel_get_node_field(ctx_, index_, (int)f, v.data(), v.size()), ctx_);
        return v;
    }
    void set_field(mel_node_field f, const std::vector<double>& v) {
        check(mel_set_node_field(ctx_, index_, (int)f, v.data(), v.size()), ctx_);
    }
    // contact elements with this body as slave (16 ints, 18 doubles each)
    void contacts(std::vector<int32_t>& ids, std::vector<double>& state) const {
        size_t n = 0;
        check(mel_get_contacts(ctx_, index_, nullptr, nullptr, 0, &n), ctx_);
        ids.resize(16 * n);
        state.resize(18 * n);
        check(mel_get_contacts(ctx_, index_, ids.data(), state.data(), n, &n), ctx_);
    }
    // rigid-body record (26 values) and the 7 works
    void state(std::vector<double>& rigid26, std::vector<double>& works7) const {
        rigid26.assign(26, 0.);
        works7.assign(7, 0.);
        check(mel_get_body_state(ctx_, index_, rigid26.data(), works7.data()), ctx_);
    }
    // Body::status (0 active, 1 inactive, 2 nogravity) and Body::damage
    int status(double* damage = nullptr) const {
        int st = 0;
        check(mel_get_body_status(ctx_, index_, &st, damage), ctx_);
        return st;
    }

private:
    mel_ctx* ctx_;
    int index_, n_nodes_;
};

// a loaded model on one HIP device; owns its context
class Model {
public:
    static Model load(const std::string& dir, int number_save, mel_precision precision = MEL_FP64, int device = 0,
                      const mel_options* opts = nullptr) {
        mel_ctx* c = nullptr;
        check(mel_load_ex(dir.c_str(), number_save, precision, device, opts, &c), nullptr);
        return Model(c);
    }
    static Model load_binary(const std::string& path, mel_precision precision = MEL_FP64, int device = 0,
                             const mel_options* opts = nullptr) {
        mel_ctx* c = nullptr;
        check(mel_load_binary_ex(path.c_str(), precision, device, opts, &c), nullptr);
        return Model(c);
    }
    Model(Model&& o) noexcept : ctx_(o.ctx_), bodies_(std::move(o.bodies_)) { o.ctx_ = nullptr; }
    Model& operator=(Model&& o) noexcept {
        std::swap(ctx_, o.ctx_);
        std::swap(bodies_, o.bodies_);
        return *this;
    }
    Model(const Model&) = delete;
    Model& operator=(const Model&) = delete;
    ~Model() { mel_destroy(ctx_); }

    mel_info_t info() const {
        mel_info_t i;
        check(mel_get_info(ctx_, &i), ctx_);
        return i;
    }
    double time() const { return info().time; }
    double Deltat() const { return info().dt; }
    int Number_iteration() const { return info().iteration; }
    int Nb_bodies() const { return (int)bodies_.size(); }
    Body& body(int i) { return bodies_.at((size_t)i); }
    // time, works summed over bodies, kinetic energies (Monitoring's energy block)
    void scalars(double& time, double works7[7], double energy3[3]) const {
        check(mel_get_scalars(ctx_, &time, works7, energy3), ctx_);
    }
    // Write_dynamic (Code/IO.h:2351-2636)
    void write_dynamic(const std::string& path) { check(mel_write_dynamic(ctx_, path.c_str()), ctx_); }
    // Monitoring rows after every step from now on (Code/Monitoring.h:8-559);
    // false when the model sets NO_MONITORING
    bool start_monitoring() { return mel_set_monitoring(ctx_, 1) == MEL_OK; }
    // Write_monitoring (Code/IO.h:2644-2666): the rows since the last write, appended
    void write_monitoring(const std::string& path) { check(mel_write_monitoring(ctx_, path.c_str()), ctx_); }
    void synchronize() { check(mel_synchronize(ctx_), ctx_); }
    mel_ctx* handle() { return ctx_; }

private:
    explicit Model(mel_ctx* c) : ctx_(c) {
        mel_info_t i;
        check(mel_get_info(ctx_, &i), ctx_);
        for (int b = 0; b < i.n_bodies; b++) {
            int n = 0;
            check(mel_get_body_sizes(ctx_, b, &n, nullptr, nullptr, nullptr), ctx_);
            bodies_.emplace_back(ctx_, b, n);
        }
    }
    mel_ctx* ctx_;
    std::vector<Body> bodies_;
};

// the explicit solver: Euler_step and the contact updates at their cadence
class Solver {
public:
    // one Euler_step (Code/Solver.h:10-88), then Update_proximity when the
    // contact-update period has elapsed (Code/main.cpp:191-196)
    void Euler_step(Model& m) { check(mel_euler_steps(m.handle(), 1), m.handle()); }
    // one iteration of the stepping loop with the model's scheme: Euler_step,
    // or Solver_step (Code/Solver.h:98-486) for Adaptive_Euler, Mass_Scaling
    // and Adaptive_Mass_Scaling (Code/main.cpp:148-166), then the cadence
    void Solver_step(Model& m) { check(mel_solver_steps(m.handle(), 1), m.handle()); }
    void Solver_steps(Model& m, int n) { check(mel_solver_steps(m.handle(), n), m.handle()); }
    mel_solver_info_t info(Model& m) {
        mel_solver_info_t i;
        check(mel_get_solver_info(m.handle(), &i), m.handle());
        return i;
    }
    // n of them, queued on the device at once (a run replays captured graphs)
    void Euler_steps(Model& m, int n) { check(mel_euler_steps(m.handle(), n), m.handle()); }
    // Update_proximity (Code/Proximity.h:48-473) now
    void Update_proximity(Model& m) { check(mel_update_proximity(m.handle()), m.handle()); }
    // one force evaluation without integration (what: MEL_EVAL_* bits)
    void eval_forces(Model& m, unsigned what = MEL_EVAL_ALL) {
        check(mel_eval_forces(m.handle(), what), m.handle());
    }
};

}  // namespace melody2d

#endif
