/*
 * melody2d.h - C-ABI of the MI355X (gfx950) implementation of the MELODY 2D
 * explicit dynamics step.
 *
 * The reference (GMollon/MELODY2D) has no plugin or FFI layer: its step sits
 * behind C++ free functions over vector<Body>& called from Code/main.cpp.
 * Each entry point below names the reference function or call sequence it
 * replaces.  No C++ or HIP types cross this boundary: plain pointers, sizes
 * and status codes.  Device memory is owned by the context; host buffers are
 * owned by the caller.  A context is not re-entrant; distinct contexts may be
 * used from different threads.
 */
#ifndef MELODY2D_H
#define MELODY2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mel_ctx mel_ctx;

typedef enum {
    MEL_OK = 0,
    MEL_E_ARG = 1,          /* bad argument (index out of range, buffer too small) */
    MEL_E_IO = 2,           /* model file missing / unreadable */
    MEL_E_PARSE = 3,        /* model file content not understood */
    MEL_E_OOM = 4,          /* device allocation failed */
    MEL_E_HIP = 5,          /* HIP runtime error */
    MEL_E_NAN = 6,          /* non-finite kinematic state (Code/main.cpp:227-237 stops the run) */
    MEL_E_CAPACITY = 7,     /* a fixed-capacity device buffer overflowed */
    MEL_E_UNSUPPORTED = 8,  /* model uses a feature this library does not implement */
    MEL_E_TIMEOUT = 9       /* a bounded cross-workgroup wait of the step ran out of time */
} mel_status;

typedef enum {
    MEL_FP64 = 0,           /* plain FP64 drop-in */
    MEL_FP32_MIXED = 1      /* FP32 shape-function products / stresses / laws, FP64 state */
} mel_precision;

/* node fields, in the order of the dump records used by the tests */
typedef enum {
    MEL_X_CURRENT = 0, MEL_Y_CURRENT, MEL_X_DISP, MEL_Y_DISP, MEL_X_VEL, MEL_Y_VEL, MEL_X_ACC, MEL_Y_ACC,
    MEL_X_DISP_PARAM, MEL_Y_DISP_PARAM, MEL_X_VEL_PARAM, MEL_Y_VEL_PARAM, MEL_X_ACC_PARAM, MEL_Y_ACC_PARAM,
    MEL_X_INTERNAL_FORCE, MEL_Y_INTERNAL_FORCE, MEL_X_CONTACT_FORCE, MEL_Y_CONTACT_FORCE,
    MEL_X_BODY_FORCE, MEL_Y_BODY_FORCE, MEL_X_DIRICHLET_FORCE, MEL_Y_DIRICHLET_FORCE,
    MEL_X_NEUMANN_FORCE, MEL_Y_NEUMANN_FORCE, MEL_X_DAMPING_FORCE, MEL_Y_DAMPING_FORCE,
    MEL_X_FORCE, MEL_Y_FORCE, MEL_X_MASS, MEL_Y_MASS,
    MEL_X_ALID_FORCE, MEL_Y_ALID_FORCE,     /* Node::x_alid_force (absorbing layers, Code/Body.h:1163-1296) */
    MEL_NODE_FIELD_COUNT
} mel_node_field;

typedef struct {
    int n_bodies, n_nodes, n_gauss, n_borders, n_border_nodes, n_rigid;
    int n_contacts, contact_capacity;
    long long nnz_gauss;    /* Gauss-point support entries */
    long long nnz_d2d;      /* DOF_TO_DISP entries */
    double time, dt;
    int iteration;
    int precision;
} mel_info_t;

/* Load_static (Code/IO.h:598) + Load_dynamic (Code/IO.h:1871) on the files
 * STATIC_CONTROL.asc, STATIC_DATA.asc, DYNAMIC_<save_index>.asc of `dir`,
 * then the initialisation block of Code/main.cpp:103-126 (borders, BCs,
 * Update_proximity, Initialize_CZM, Update_material masses, Update_proximity,
 * Initialize_CZM) on HIP device `hip_device`. */
mel_status mel_load(const char* dir, int save_index, mel_precision precision, int hip_device, mel_ctx** out);

/* Same as mel_load for a model held in the binary record container written by
 * melody2d_amd.preprocessor (for models too large for the ASCII files). */
mel_status mel_load_binary(const char* path, mel_precision precision, int hip_device, mel_ctx** out);

/* Load-time choices of the step's launch configuration (no reference
 * counterpart: they change how the step is laid out on the GPU, never its
 * results beyond rounding order).  -1 leaves a field to the library's default
 * (or to the MELODY2D_* environment switch of the same meaning). */
typedef struct {
    int block_nodes;    /* owned nodes per node block of a body split over blocks, 1..256 (default 256) */
    int fused_step;     /* 1: the step in two launches where every body fits one block; 0: the launch sequence */
    int graph_replay;   /* 1: runs of steps between contact updates replayed from captured HIP graphs */
    int wide_window;    /* 1: the wide register window (30 entries per lane); 0: never; -1: by support sizes */
} mel_options;

void mel_default_options(mel_options* opts);
/* mel_load / mel_load_binary with options (NULL: defaults) */
mel_status mel_load_ex(const char* dir, int save_index, mel_precision precision, int hip_device,
                       const mel_options* opts, mel_ctx** out);
mel_status mel_load_binary_ex(const char* path, mel_precision precision, int hip_device, const mel_options* opts,
                              mel_ctx** out);
/* the configuration a context runs with (fields resolved, no -1 left except
 * block_nodes when it was not set) */
mel_status mel_get_options(const mel_ctx* ctx, mel_options* opts);

/* Autotuner: loads the model of `binary_path` once per candidate launch
 * configuration (node-block size, two-launch step or launch sequence, graph
 * replay, register window), times `steps` steps of each after a warm-up and
 * returns the fastest in *best (and its time per step, and how many
 * configurations were timed).  No reference counterpart (the reference has
 * no launch configuration). */
mel_status mel_autotune(const char* binary_path, mel_precision precision, int hip_device, int steps,
                        mel_options* best, double* best_ms_per_step, int* n_tried);

/* Update_proximity (Code/Proximity.h:48-473): neighbour lists, contact
 * candidates with history transfer, self-contact unless NO_MONITORING.
 * (The DEACTIVATION rules run only at the stepping loop's contact updates,
 * as Code/main.cpp:191-196 calls Update_status.) */
mel_status mel_update_proximity(mel_ctx* ctx);

/* nsteps x [Euler_step (Code/Solver.h:10-88) + the contact-update cadence of
 * Code/main.cpp:191-196: Update_proximity, then Update_status
 * (Code/Proximity.h:550-1007) when the model has DEACTIVATION rules].
 * Asynchronous with respect to the host except at contact updates; returns
 * MEL_E_NAN / MEL_E_CAPACITY when detected. */
mel_status mel_euler_steps(mel_ctx* ctx, int nsteps);

/* wait for queued device work; reports deferred errors */
mel_status mel_synchronize(mel_ctx* ctx);

/* The model's time-integration scheme (SOLVER section, Code/IO.h:821-890;
 * the DYNAMIC file names the one that runs, Code/IO.h:1936-1940). */
typedef enum {
    MEL_SCHEME_EULER = 0, MEL_SCHEME_ADAPTIVE_EULER = 1, MEL_SCHEME_MASS_SCALING = 2,
    MEL_SCHEME_ADAPTIVE_MASS_SCALING = 3
} mel_scheme;

/* nsteps iterations of the stepping loop of Code/main.cpp:141-238 with the
 * model's scheme: Euler_step (as mel_euler_steps), or Solver_step
 * (Code/Solver.h:98-486: two half steps per attempt, the error estimate,
 * rejected attempts restored and retried with a smaller time step,
 * mass-scaling factors), each followed by the contact-update cadence of
 * Code/main.cpp:191-196 with the updated time step.  Every attempt of an
 * adaptive scheme waits for its error scalars (a few doubles) on the host.
 * MEL_E_NAN when a body's total error is NaN (Code/main.cpp:227-237), after
 * that step.  The "Euler_2" scheme is refused at load (MEL_E_PARSE): the
 * reference reads no target error for it, yet divides by one. */
mel_status mel_solver_steps(mel_ctx* ctx, int nsteps);

typedef struct {
    int scheme;             /* mel_scheme */
    double dt;              /* current time step (Deltat) */
    int neval;              /* attempts of the last accepted step */
    long long attempts;     /* attempts since load */
    double max_error, mean_error;   /* Solver_step's max_error, mean_error of the last attempt */
    double total_mass, max_mass;    /* mass-scaling schemes: total scaled mass, largest scaled/initial ratio */
} mel_solver_info_t;
mel_status mel_get_solver_info(mel_ctx* ctx, mel_solver_info_t* info);

/* Body::total_error, max_error, node_for_max_error, mass_mass_scaling,
 * factor_mass_scaling (Code/Body.h:2524-2585) after the last attempt;
 * MEL_E_ARG for an Euler model. */
mel_status mel_get_body_error(mel_ctx* ctx, int body, double* out5);

/* Node::x_factor_mass_scaling, y_factor_mass_scaling of a body's n nodes
 * (Code/Node.h:274-320; all 1 unless a mass-scaling scheme changed them). */
mel_status mel_get_mass_scaling(mel_ctx* ctx, int body, double* fx, double* fy, size_t n);

/* node field of one body (Node members, Code/Node.h:10-99) */
mel_status mel_get_node_field(mel_ctx* ctx, int body, int field, double* out, size_t n);
mel_status mel_set_node_field(mel_ctx* ctx, int body, int field, const double* in, size_t n);

/* contact elements of slave body `body` in the reference's order: 16 ints
 * (borderS border_nodeS nodeS bodyM borderM shiftM segmentM nodeM0..3
 * shiftM0..3 nb_internal) and 18 doubles (shapeM0..3 gapn vgapn gapt vgapt
 * xsi xnorm ynorm xtan ytan effective_mass length fx fy internal0) each,
 * node ids local to their body (Contact_element, Code/Contact_Element.h:11-44) */
mel_status mel_get_contacts(mel_ctx* ctx, int body, int32_t* ids, double* state, size_t cap, size_t* count);

/* Body::Compute_nodal_stresses (Code/Body.h:2310-2519), the stress output
 * of Write_graphic (Code/Graphic.h:117-122): stresses, strains and principal
 * and equivalent stresses at every node from the derivatives of its
 * DOF_TO_DISP row and the current displacement parameters (rigid bodies:
 * zeros, jacobian 1).  mel_compute_nodal_stresses evaluates them for the
 * current state; mel_get_nodal_stress reads one field of one body. */
typedef enum {
    MEL_SIGMA_I = 0, MEL_SIGMA_II, MEL_SIGMA_III, MEL_JACOBIAN, MEL_SXX, MEL_SYY, MEL_SXY, MEL_SZZ,
    MEL_SIGMA_XX, MEL_SIGMA_YY, MEL_SIGMA_XY, MEL_SIGMA_ZZ, MEL_SIGMA_TRESCA, MEL_SIGMA_VM, MEL_SIGMA_SPH,
    MEL_EXX, MEL_EYY, MEL_EXY, MEL_NORM_E,
    MEL_STRESS_FIELD_COUNT
} mel_stress_field;

mel_status mel_compute_nodal_stresses(mel_ctx* ctx);
mel_status mel_get_nodal_stress(mel_ctx* ctx, int body, int field, double* out, size_t n);

/* Write_graphic (Code/Graphic.h:21-946): a VTK legacy POLYDATA file of the
 * current state with the point fields whose flag is set in to_plot[0..40]
 * (the reference's To_Plot slots 00_Body_index .. 40-Epsilon_Mass_Scaling;
 * NULL: slots 00-12 and 14-28); nodal stresses are evaluated for it.
 * number_print labels the file as the reference's print counter does. */
mel_status mel_write_graphic(mel_ctx* ctx, const char* path, const int* to_plot, int number_print);

/* sizes of a body: nodes, Gauss points, borders, and whether it is rigid
 * (Body::nb_nodes, nb_gauss, nb_borders, type; Code/Body.h:24-60) */
mel_status mel_get_body_sizes(mel_ctx* ctx, int body, int* n_nodes, int* n_gauss, int* n_borders, int* rigid);

/* the 3 internal variables of each contact element of body (in the order of
 * mel_get_contacts; Contact_element::internal, Code/Contact_Element.h:44:
 * damage, and for BondedMohrCoulomb and CZMfatigue the last normal and
 * tangential tractions; zeros where the law keeps fewer) */
mel_status mel_get_contact_internals(mel_ctx* ctx, int body, double* out3, size_t cap, size_t* count);

/* neighbour triples (body2 border2 shift) of body (Code/Proximity.h:160) */
mel_status mel_get_neighbours(mel_ctx* ctx, int body, int32_t* out, size_t cap, size_t* count);

/* Monitoring (Code/Monitoring.h:8-559): with on = 1 a row is formed on the
 * device after every step of mel_euler_steps / mel_solver_steps (the
 * reference forms one per iteration unless NO_MONITORING, Code/main.cpp:
 * 172-176): iteration, time, time step, neval, max_error, mean_error (0
 * under Euler_step, where the reference leaves them uninitialised), one
 * value per MONITORING entry that yields one (Read_spy_line types 0-10),
 * then with MONITOR_ENERGY the boundary work, elastic energy (0: the laws
 * set their energy to 0, Code/Model.h:28, 73) and kinetic energy.  Rows
 * collect in host memory until read.  MEL_E_UNSUPPORTED when the model sets
 * NO_MONITORING. */
mel_status mel_set_monitoring(mel_ctx* ctx, int on);

/* The rows formed since the last read, each as its value count followed by
 * the values, and clears them; out = NULL only reports the count of
 * doubles. */
mel_status mel_get_monitoring(mel_ctx* ctx, double* out, size_t cap, size_t* count);

/* Write_monitoring (Code/IO.h:2644-2666): appends the rows formed since the
 * last read to path (MONITORING.asc in the reference's run directory), ten
 * significant digits, and clears them. */
mel_status mel_write_monitoring(mel_ctx* ctx, const char* path);

/* Body::status as a mel status code (0 active, 1 inactive, 2 nogravity;
 * Code/Body.h:46, set to inactive by Update_status, Code/Proximity.h:550-1007)
 * and Body::damage (Code/Body.h:110, formed by Body::Update_damage,
 * Code/Body.h:2667-2700, for the DEACTIVATION rules that read it); either
 * pointer may be null */
mel_status mel_get_body_status(mel_ctx* ctx, int body, int* status, double* damage);

/* 26 body-level values (rigid kinematics and resultants; mass, inertia) and
 * the 7 work accumulators internal contact body dirichlet neumann damping alid
 * (Code/Body.h:112-118, accumulated as in Code/Body.h:2058-2081) */
mel_status mel_get_body_state(mel_ctx* ctx, int body, double* rigid26, double* works7);

/* energy scalars of one body, energy3 = {kinetic energy from the physical
 * velocities and nodal masses, and boundary work of the last step, sum over
 * border slots of (px vx + py vy) * length * dt: the MONITOR_ENERGY block of
 * Code/Monitoring.h:508-537; kinetic energy from the velocity parameters
 * (rigid bodies: from the body's velocity, mass and inertia), monitor type 9
 * of Code/Monitoring.h:395-412}, in the order {kinetic, kinetic from
 * parameters, boundary work} */
mel_status mel_get_body_energy(mel_ctx* ctx, int body, double* energy3);

/* whole-model scalars for the energy balance (Code/Monitoring.h:508-537):
 * the time, the 7 works summed over bodies (as mel_get_body_state) and the 3
 * energy scalars summed over bodies (as mel_get_body_energy); any pointer
 * may be NULL.  Elastic energy is not returned: the reference never
 * computes it (Code/Model.h:28, Code/Model.h:72). */
mel_status mel_get_scalars(mel_ctx* ctx, double* time, double* works7, double* energy3);

/* In-memory checkpoint of the dynamic state at full precision: node
 * kinematics, parameters and force fields, rigid-body records and works,
 * BC values, border lengths, neighbour lists, contact elements, time and
 * the contact-update schedule (what Write_dynamic / Load_dynamic,
 * Code/IO.h:1871-2636, carry between runs, without the ASCII rounding).
 * mel_load_state accepts a state saved from a context of the same model,
 * in either precision; it is checked before anything is written
 * (MEL_E_PARSE when it does not fit the model, MEL_E_CAPACITY when its
 * contact elements exceed this context's capacity). */
mel_status mel_get_state_size(mel_ctx* ctx, size_t* bytes);
mel_status mel_save_state(mel_ctx* ctx, void* buf, size_t cap, size_t* written);
mel_status mel_load_state(mel_ctx* ctx, const void* buf, size_t size);

/* what mel_eval_forces evaluates */
typedef enum {
    MEL_EVAL_INTERNAL = 1,  /* Body::Update_internal_forces (Code/Body.h:937-1049) */
    MEL_EVAL_CONTACT = 2,   /* Update_contacts + Update_contact_forces + sends (Code/Body.h:1341-1967) */
    MEL_EVAL_EXTERNAL = 4,  /* Update_bc_forces, Update_body_forces, Update_damping_forces (Code/Body.h:669-1154) */
    MEL_EVAL_ALL = 7
} mel_eval_what;

/* One force evaluation at the current state without integration: the force
 * stage of Euler_step (Code/Solver.h:32-67) and Sum_up_forces
 * (Code/Node.h:212-225), for the parts named in `what` (the others read 0).
 * The node force fields (MEL_X_INTERNAL_FORCE ... MEL_Y_FORCE) of deformable
 * bodies then hold the result; contact elements are evaluated on a copy of
 * their state, so the step's state is left as it was.  Rigid bodies'
 * resultants are not re-formed. */
mel_status mel_eval_forces(mel_ctx* ctx, unsigned what);

/* mel_eval_forces without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The contact
 * elements' state is saved in a device scratch array of the library (at
 * most 65536 elements, 14 MiB, held by every GPU the library runs on until
 * the process ends) instead of a buffer the context allocates on first
 * use; larger element counts return MEL_E_CAPACITY.  All contexts on one
 * GPU share that scratch: calls that overlap on one GPU race. */
mel_status mel_eval_forces_noalloc(mel_ctx* ctx, unsigned what);

/* per border-node slot: 0 tributary length (Code/Body.h:614-661),
 * 1 / 2 x / y contact pressure (Update_contact_pressures, Code/Proximity.h:1014-1069) */
mel_status mel_get_border_field(mel_ctx* ctx, int body, int border, int field, double* out, size_t n);

mel_status mel_get_info(mel_ctx* ctx, mel_info_t* info);

/* state in the DYNAMIC_nnnnn.asc layout (Write_dynamic, Code/IO.h:2351-2636),
 * 17 significant digits */
mel_status mel_write_dynamic(mel_ctx* ctx, const char* path);

/* the HIP stream (hipStream_t) all of the context's kernels run on, for
 * event timing by the caller */
mel_status mel_get_stream(mel_ctx* ctx, void** stream);

/* Optional HIP-event instrumentation of kernel launches of the step (and of
 * each contact update as a whole).  `on` = 0 off, -1 every kernel, otherwise a
 * bit mask of kernel ids.  Setting it resets the accumulators;
 * mel_get_kernel_timing waits for the stream and returns the summed elapsed
 * time and launch count of kernel `kernel` (0 <= kernel < mel_kernel_count()).
 * No reference counterpart (the reference has no timers); used by the
 * benchmark's roofline figures. */
mel_status mel_set_kernel_timing(mel_ctx* ctx, int on);
mel_status mel_get_kernel_timing(mel_ctx* ctx, int kernel, double* total_ms, long long* launches);
int mel_kernel_count(void);
const char* mel_kernel_name(int kernel);

void mel_destroy(mel_ctx* ctx);
const char* mel_last_error(const mel_ctx* ctx);
const char* mel_version(void);

/* device memory the context holds, in bytes (its footprint for admission by
 * the job server below); no reference counterpart */
mel_status mel_get_device_bytes(const mel_ctx* ctx, size_t* bytes);

/* runs of steps replayed from captured HIP graphs (on = 1, the default of
 * the two-launch step) or launched step by step (on = 0: the replays of
 * several contexts overlap less on the device than their launches do); no
 * reference counterpart */
mel_status mel_set_graph_replay(mel_ctx* ctx, int on);

/* ---- job server: many independent jobs on one GPU at once ----
 * The reference runs one model per process (Code/main.cpp, with the
 * Euler_step loop of Code/main.cpp:141-238 until Tend); a batch of models is
 * a series of such processes.  The server replaces that series: each job is
 * one binary model container (as for mel_load_binary) run for a step count,
 * in its own context on its own stream, so the jobs' work overlaps on the
 * device and each job's state is exactly what it would be alone.  Jobs are
 * admitted in arrival order while their device footprint fits the budget
 * (estimated from the container's size until the job is loaded); running
 * jobs are kept `steps_per_slice`-step slices ahead of the device.  One
 * thread drives a server (loading included); distinct servers are
 * independent. */
typedef struct mel_server mel_server;

typedef enum {
    MEL_JOB_WAITING = 0,    /* submitted, not yet arrived or not yet admitted */
    MEL_JOB_RUNNING = 1,
    MEL_JOB_DONE = 2,       /* all steps ran; the context is held if the job was kept */
    MEL_JOB_FAILED = 3,     /* see status and mel_server_job_error */
    MEL_JOB_RELEASED = 4    /* done, and its context given back by mel_server_release */
} mel_job_state;

typedef struct {
    int state;              /* mel_job_state */
    int status;             /* mel_status of a failed job */
    long long steps, steps_done;
    double arrival;         /* seconds from the first mel_server_run, as submitted */
    double admitted;        /* when its loading began; -1 when not yet */
    double finished;        /* when its last step was seen complete; -1 when not yet */
    size_t device_bytes;    /* footprint once loaded */
    int n_nodes;
} mel_job_info_t;

mel_status mel_server_create(int hip_device, size_t hbm_budget, long long steps_per_slice, mel_server** out);
/* queue a job that arrives `arrival_s` seconds after the first
 * mel_server_run; keep != 0 holds its context once done (read its state
 * through mel_server_job_context, then mel_server_release) */
mel_status mel_server_submit(mel_server* s, const char* binary_path, mel_precision precision, long long nsteps,
                             double arrival_s, int keep, int* job);
/* drive the loop until every submitted job is done or failed, or until
 * `timeout_s` seconds have passed; may be called again (more submissions in
 * between are fine) */
mel_status mel_server_run(mel_server* s, double timeout_s, int* n_unfinished);
mel_status mel_server_job(const mel_server* s, int job, mel_job_info_t* info);
mel_status mel_server_job_context(mel_server* s, int job, mel_ctx** ctx);
mel_status mel_server_release(mel_server* s, int job);
mel_status mel_server_usage(const mel_server* s, size_t* used, size_t* budget);
/* graph replay (mel_set_graph_replay) of the jobs loaded from now on;
 * default on */
mel_status mel_server_set_graph_replay(mel_server* s, int on);
const char* mel_server_job_error(const mel_server* s, int job);
void mel_server_destroy(mel_server* s);

#ifdef __cplusplus
}
#endif
#endif /* MELODY2D_H */
