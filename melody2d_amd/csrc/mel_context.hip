// Host orchestration and C-ABI: builds the flattened device model from a
// parsed mel::Model, runs the initialisation block of Code/main.cpp:103-126
// and the Euler_step sequence of Code/Solver.h:10-88 as kernel launches on
// one HIP stream.
#include <algorithm>
#include <cmath>
#include <functional>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/melody2d.h"
#include "k_block.hpp"
#include "k_fused.hpp"
#include "k_adaptive.hpp"
#include "k_status.hpp"
#include "k_rigid_lists.hpp"
#include "k_step.hpp"
#include "mel_device.hpp"
#include "model_io.hpp"

using namespace mel;

namespace {

struct HipError {
    std::string msg;
};

#define HIP_OK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

struct CapacityError {
    std::string msg;
};

struct NanError {
    std::string msg;
};

struct TimeoutError {
    std::string msg;
};

inline int blocks(long long n, int b = BLOCK) { return (int)std::max<long long>(1, (n + b - 1) / b); }

#define TIMED(ctx, id, st, launch)                 \
    do {                                           \
        if ((ctx).timing && ((ctx).timing_mask >> (id) & 1u)) { \
            int ev_ = (ctx).ev_begin(id, st);      \
            launch;                                \
            (ctx).ev_end(ev_, st);                 \
        } else {                                   \
            launch;                                \
        }                                          \
    } while (0)

}  // namespace

// ---------------------------------------------------------------------------
// Device scratch of the no-allocation force evaluation
// (mel_eval_forces_noalloc): the contact elements' state saved around the
// evaluation, at most SCRATCH_MAX_CONTACTS elements, in a file-scope device
// array that every GPU the library runs on holds until the process ends.
// Layout (fixed capacity, so offsets do not depend on the count): the 19
// double fields of ContactArrays, SCRATCH_MAX_CONTACTS each, then its 18 int
// fields.  The array's host-side name is not a device address: host code
// uses the per-device address from hipGetSymbolAddress (scratch_address).
// ---------------------------------------------------------------------------
constexpr int SCRATCH_MAX_CONTACTS = 65536;
constexpr int SCRATCH_DBL_FIELDS = 21, SCRATCH_INT_FIELDS = 18;
constexpr size_t SCRATCH_BYTES =
    (size_t)SCRATCH_MAX_CONTACTS * (SCRATCH_DBL_FIELDS * sizeof(double) + SCRATCH_INT_FIELDS * sizeof(int));
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

__device__ inline void contact_fields(const ContactArrays& A, double** dp, int** ip) {
    double* d[SCRATCH_DBL_FIELDS] = {A.N0, A.N1, A.N2, A.N3, A.gapn, A.vgapn, A.gapt, A.vgapt, A.xsi, A.xn,
                                     A.yn, A.xt, A.yt, A.emass, A.len, A.fx, A.fy, A.int0, A.wacc, A.int1, A.int2};
    int* i[SCRATCH_INT_FIELDS] = {A.bodyS, A.borderS, A.bnodeS, A.nodeS, A.bodyM, A.borderM, A.shiftM, A.segM,
                                  A.nbint, A.fresh, A.nM0, A.nM1, A.nM2, A.nM3, A.sM0, A.sM1, A.sM2, A.sM3};
    for (int k = 0; k < SCRATCH_DBL_FIELDS; k++) dp[k] = d[k];
    for (int k = 0; k < SCRATCH_INT_FIELDS; k++) ip[k] = i[k];
}

// element e of A to (to_buf) or from the scratch buffer `buf`, for e < nc
__global__ void k_contacts_scratch(ContactArrays A, unsigned char* buf, int nc, int to_buf) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc) return;
    double* dp[SCRATCH_DBL_FIELDS];
    int* ip[SCRATCH_INT_FIELDS];
    contact_fields(A, dp, ip);
    double* bd = reinterpret_cast<double*>(buf);
    int* bi = reinterpret_cast<int*>(buf + (size_t)SCRATCH_DBL_FIELDS * SCRATCH_MAX_CONTACTS * sizeof(double));
    for (int k = 0; k < SCRATCH_DBL_FIELDS; k++) {
        double* q = bd + (size_t)k * SCRATCH_MAX_CONTACTS + e;
        if (to_buf) *q = dp[k][e]; else dp[k][e] = *q;
    }
    for (int k = 0; k < SCRATCH_INT_FIELDS; k++) {
        int* q = bi + (size_t)k * SCRATCH_MAX_CONTACTS + e;
        if (to_buf) *q = ip[k][e]; else ip[k][e] = *q;
    }
    // the restored element's info is recomputed from its restored indices
    if (!to_buf) A.info[e] = make_int4(0, 0, 0, 0);
}

// the device address of `scratch` on `device` (current), taken once per
// device when the first context is created on it
void scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return;
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
}

void* scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < SCRATCH_MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

// kernels timed individually by the optional event instrumentation
enum KernelId {
    KT_CONTACTS = 0, KT_GATHER, KT_DAMPING, KT_BLOCK_UPDATE, KT_FINISH, KT_PROXIMITY, KT_STEP_FORCES,
    KT_STEP_FINISH, KT_COUNT
};
static const char* kKernelNames[KT_COUNT] = {
    "contacts_bc", "gather_rigid", "damping", "block_force_update", "update_finish", "proximity_update",
    "step_forces", "step_finish"};

struct mel_ctx {
    Model model;
    mel_options opts{-1, -1, -1, -1};   // load options (-1: library default, or the environment's)
    // event instrumentation (off by default)
    bool timing = false;
    unsigned timing_mask = 0;     // bit k set: kernel k is instrumented
    struct EvPair { int id; hipEvent_t a, b; };
    std::vector<EvPair> pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    double kt_ms[KT_COUNT] = {0};
    long long kt_n[KT_COUNT] = {0};
    int ev_begin(int id, hipStream_t st) {
        std::pair<hipEvent_t, hipEvent_t> p;
        if (!pool.empty()) { p = pool.back(); pool.pop_back(); }
        else { HIP_OK(hipEventCreate(&p.first)); HIP_OK(hipEventCreate(&p.second)); }
        pending.push_back({id, p.first, p.second});
        HIP_OK(hipEventRecord(p.first, st));
        return (int)pending.size() - 1;
    }
    void ev_end(int k, hipStream_t st) { HIP_OK(hipEventRecord(pending[k].b, st)); }
    void harvest() {
        if (pending.empty()) return;
        HIP_OK(hipStreamSynchronize(aux));
        HIP_OK(hipStreamSynchronize(stream));
        for (auto& e : pending) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, e.a, e.b));
            kt_ms[e.id] += ms;
            kt_n[e.id]++;
            pool.push_back({e.a, e.b});
        }
        pending.clear();
    }
    DevModel d{};
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;            // contact pipeline, concurrent with internal forces
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int device = 0;
    int precision = MEL_FP64;
    std::vector<void*> allocs;
    size_t dev_bytes = 0;        // device memory held by allocs
    std::string err;
    // host mirrors
    std::vector<int> h_node_off, h_border_off, h_gp_off, h_rigid_ids;
    std::vector<int> h_br_slot0, h_br_nslots;
    long long nnz_g = 0, nnz_d = 0;
    int n_contacts = 0;          // *d.c_count, read back when an API call needs it exactly
    // element count and status flags of the last proximity update, copied
    // back asynchronously into pinned memory: the count sizes the contact
    // workgroups of later steps (their kernels read the exact count), the
    // flags are checked once the copy has landed
    int* h_async = nullptr;      // [0] element count, [1..2] status flags
    hipEvent_t ev_async = nullptr;
    bool async_pending = false;
    int nc_hint = 0;
    int row_cap = 0;
    int max_row_chunks = 0;
    int* d_rigid_ids = nullptr;
    double time = 0.0;
    double next_contact_update = 0.0;
    int iteration = 0;
    ContactArrays bufA{}, bufB{};
    BlockArrays blocks_dev{};
    size_t block_lds = 0;
    // internal-force kernel with the wide register window (some Gauss point
    // has more support entries than the default window holds)
    bool wide = false;
    // step in two launches (every deformable body in one node block), else
    // the launch sequence on two streams
    bool two_launch = false;
    // captured runs of steps (two-launch form): run length, contact
    // workgroups, current contact buffer -> executable graph
    struct GraphKey {
        int n, n_cwg;
        const void* cbuf;
        bool operator<(const GraphKey& o) const {
            return n != o.n ? n < o.n : n_cwg != o.n_cwg ? n_cwg < o.n_cwg : cbuf < o.cbuf;
        }
    };
    std::map<GraphKey, hipGraphExec_t> graphs;
    bool use_graphs = false;
    size_t finish_lds = 0;
    int finish_slots = 0;      // workgroups of launch B resident at once on the device
    int* d_kin_all = nullptr;     // every deformable node (kinematics on request)
    int n_kin_all = 0;
    bool outputs_stale = false;   // interior kinematics / force fields lag the last steps
    std::vector<int> h_blk_body;  // body of each block (work sums)
    bool work_law = false;        // some contact law does work (DampedMohrCoulomb)
    // a proximity update ran after the last step: the elements of that step
    // (whose forces the step's contact pressures come from) are the previous
    // buffer
    bool prox_after_step = false;
    double* d_energy = nullptr;   // per body: kinetic, kinetic from parameters, boundary work
    ContactArrays eval_save{};    // contact elements saved around mel_eval_forces
    // nodal stresses (allocated by the first mel_compute_nodal_stresses):
    // DOF_TO_DISP derivatives aligned with d_idx, per-body material
    // parameters, and the MEL_STRESS_FIELD_COUNT output fields
    double *st_dphix = nullptr, *st_dphiy = nullptr, *st_bpar = nullptr, *st_out = nullptr;
    int eval_save_cap = 0;
    std::vector<int> h_bn_list;   // border nodes finished by each block (BlockArrays::bn_list)
    // adaptive schemes (Solver_step): per-node and per-body arrays, the
    // Store snapshot and its copy tables (one per contact buffer that can be
    // current, stores then restores), the attempt's scalars in pinned memory
    AdArrays ad{};
    CopyDesc* d_copy[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    const void* copy_key[2] = {nullptr, nullptr};
    int n_copy = 0;
    double* snap_rigid = nullptr;
    double* h_ad = nullptr;
    hipEvent_t ev_ad = nullptr;
    int neval = 0;                // attempts of the last accepted step
    long long attempts = 0;       // attempts since load
    bool nan_stop = false;        // some body's total error was NaN
    // Monitoring rows (mel_set_monitoring): device rows since the last
    // drain, the MONITORING entries, the rows drained to the host (each: its
    // value count, then the values)
    bool mon_on = false;
    double* d_mon_rows = nullptr;
    int* d_mon_n = nullptr;
    double* d_mon_iter = nullptr;
    double* d_mon_items = nullptr;
    int mon_stride = 0, mon_cap = 0, mon_pending = 0;
    std::vector<double> h_mon;

    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (hipMalloc(&p, bytes) != hipSuccess) throw HipError{"hipMalloc of " + std::to_string(bytes) + " bytes failed"};
        allocs.push_back(p);
        dev_bytes += bytes;
        HIP_OK(hipMemsetAsync(p, 0, bytes, stream));
        return (T*)p;
    }
    template <typename T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        return p;
    }
    template <typename T>
    void download(std::vector<T>& v, const T* p, size_t n) {
        v.resize(n);
        if (n) HIP_OK(hipMemcpyAsync(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    ~mel_ctx() {
        if (aux) (void)hipStreamSynchronize(aux);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& e : pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
        for (auto& e : pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_async) (void)hipEventDestroy(ev_async);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        if (h_async) (void)hipHostFree(h_async);
        if (h_ad) (void)hipHostFree(h_ad);
        if (ev_ad) (void)hipEventDestroy(ev_ad);
        if (ev_join) (void)hipEventDestroy(ev_join);
        for (void* p : allocs) (void)hipFree(p);
        if (aux) (void)hipStreamDestroy(aux);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void alloc_contacts(mel_ctx& c, ContactArrays& A, int cap) {
    int** ints[] = {&A.bodyS, &A.borderS, &A.bnodeS, &A.nodeS, &A.bodyM, &A.borderM, &A.shiftM, &A.segM, &A.nbint,
                    &A.fresh, &A.nM0, &A.nM1, &A.nM2, &A.nM3, &A.sM0, &A.sM1, &A.sM2, &A.sM3};
    for (int** p : ints) *p = c.alloc<int>(cap);
    double** dbl[] = {&A.N0, &A.N1, &A.N2, &A.N3, &A.gapn, &A.vgapn, &A.gapt, &A.vgapt, &A.xsi, &A.xn, &A.yn,
                      &A.xt, &A.yt, &A.emass, &A.len, &A.fx, &A.fy, &A.int0, &A.wacc, &A.int1, &A.int2};
    for (double** p : dbl) *p = c.alloc<double>(cap);
    A.info = c.alloc<int4>(cap);
    A.sinfo = c.alloc<int4>(cap);
}

// Body::Update_material masses (Code/Body.h:2751-2783) and rigid mass and
// inertia (Code/Body.h:2877-2905), in the reference's summation order.
void compute_masses(Model& m, std::vector<double>& body_mass, std::vector<double>& body_inv_mass,
                    std::vector<double>& body_inertia, std::vector<double>& body_inv_inertia) {
    size_t NB = m.bodies.size();
    body_mass.assign(NB, 0.); body_inv_mass.assign(NB, 0.); body_inertia.assign(NB, 0.); body_inv_inertia.assign(NB, 0.);
    for (size_t bi = 0; bi < NB; bi++) {
        Body& b = m.bodies[bi];
        double rho = m.materials[b.material].p[0];
        if (!b.rigid) {
            if (m.flags[F_UPDATE_MASS]) {
                std::vector<double> mx(b.n_nodes, 0.), my(b.n_nodes, 0.);
                for (int g = 0; g < b.n_gauss; g++)
                    for (int j = b.gptr[g]; j < b.gptr[g + 1]; j++)
                        for (int k = b.gptr[g]; k < b.gptr[g + 1]; k++) {
                            double dm = b.gw[g] * b.gjac[g] * b.gphi[j] * b.gphi[k];
                            mx[b.gidx[j]] += dm;
                            my[b.gidx[j]] += dm;
                        }
                double mass = 0., inv = 0.;
                for (int i = 0; i < b.n_nodes; i++) {
                    mx[i] *= rho;
                    my[i] *= rho;
                    b.imx[i] = 1. / mx[i];
                    b.imy[i] = 1. / my[i];
                    mass += std::sqrt(mx[i] * mx[i] + my[i] * my[i]);
                    inv = 1 / mass;
                }
                b.mx = mx; b.my = my;
                body_mass[bi] = mass; body_inv_mass[bi] = inv;
            }
            // Rayleigh damping rebuilt from the stored stiffness (Code/Body.h:2862-2875)
            if (m.flags[F_UPDATE_DAMPING]) {
                double alpha = m.materials[b.material].p[1], beta = m.materials[b.material].p[2];
                for (size_t i = 0; i < b.mat_damping.size(); i++) {
                    if (b.mat_dof[2 * i] == b.mat_dof[2 * i + 1])
                        b.mat_damping[i] = alpha * b.mat_stiffness[i] + beta * b.mx[b.mat_dof[2 * i] / 2];
                    else
                        b.mat_damping[i] = alpha * b.mat_stiffness[i];
                }
            }
        } else {
            double mass = 0., inertia = 0.;
            auto surf = [](double a0, double b0, double a1, double b1, double a2, double b2) {
                return 0.5 * std::abs((a1 - a0) * (b2 - b0) - (a2 - a0) * (b1 - b0));
            };
            for (size_t c = 0; c < b.cells.size() / 3; c++) {
                int n0 = b.cells[3 * c], n1 = b.cells[3 * c + 1], n2 = b.cells[3 * c + 2];
                double x0 = b.X[n0], y0 = b.Y[n0], x1 = b.X[n1], y1 = b.Y[n1], x2 = b.X[n2], y2 = b.Y[n2];
                double xc = (x0 + x1 + x2) / 3., yc = (y0 + y1 + y2) / 3.;
                double sf = surf(x0, y0, x1, y1, x2, y2);
                mass += sf;
                inertia += sf * ((xc - b.xc) * (xc - b.xc) + (yc - b.yc) * (yc - b.yc));
                double px[3] = {x0, x1, x2}, py[3] = {y0, y1, y2};
                double ti = 0.;
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
            body_mass[bi] = mass * rho;
            body_inertia[bi] = inertia * rho;
            body_inv_mass[bi] = 1. / body_mass[bi];
            body_inv_inertia[bi] = 1. / body_inertia[bi];
        }
    }
}


void raise_flags(const int* f) {
    if (f[0] & 16) throw TimeoutError{"a cross-workgroup wait of the step ran out of its time limit"};
    if (f[0]) {
        std::string what;
        if (f[0] & 1) what += " point-force bucket";
        if (f[0] & 2) what += " neighbour list";
        if (f[0] & 4) what += " history index";
        if (f[0] & 8) what += " contact elements";
        if (f[0] & 32) what += " rigid-body element list";
        if (f[0] & 64) what += " inactive bodies' contact sends";
        throw CapacityError{"device capacity exceeded:" + what};
    }
    if (f[1]) throw NanError{"non-finite displacement parameters (the run is no longer meaningful)"};
}

void check_flags(mel_ctx& c) {
    int f[2];
    HIP_OK(hipMemcpyAsync(f, c.d.status_flags, sizeof f, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    raise_flags(f);
}

// takes in the asynchronous copy of the last proximity update's count and
// flags once it has landed (wait: block until it has)
void poll_async(mel_ctx& c, bool wait) {
    if (!c.async_pending) return;
    if (wait) HIP_OK(hipEventSynchronize(c.ev_async));
    else if (hipEventQuery(c.ev_async) != hipSuccess) return;
    c.async_pending = false;
    c.nc_hint = c.h_async[0];
    raise_flags(c.h_async + 1);
}

// exact element count (waits for the queued work)
int contact_count(mel_ctx& c) {
    HIP_OK(hipMemcpyAsync(&c.n_contacts, c.d.c_count, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    c.nc_hint = c.n_contacts;
    return c.n_contacts;
}

// workgroups of BLOCK_THREADS that the contact kernels start with: the
// last known count with room to grow (the kernels loop over the exact one)
int contact_workgroups(const mel_ctx& c) {
    long long n = (long long)c.nc_hint + c.nc_hint / 4 + 64 + c.blocks_dev.n_bc_bn + (long long)c.d.n_rbc * BLOCK_THREADS;
    return (int)std::max<long long>(1, (n + BLOCK_THREADS - 1) / BLOCK_THREADS);
}


// dynamic LDS of the block kernel: displacement parameters of the touched
// nodes (FP64 pairs in both forms), then the region of every contribution
// slot (reference order only)
size_t block_lds_size(const BlockArrays& B, bool f32) {
    (void)f32;
    size_t u = 16 * (size_t)B.max_touch;
    u = (u + 15) & ~size_t(15);
    return u + (B.any_multi_region ? 2 * ((size_t)CHUNK_ENTRIES + 8) : 0);
}

// Node blocks of the fused internal-force kernel: whole bodies when they fit,
// otherwise groups of nodes in Morton order; chunked Gauss-point streams and
// the transposed (node -> entry position) index per chunk.
void build_blocks(mel_ctx& c, const std::vector<int>& g_ptr, const std::vector<int>& g_idx,
                  const std::vector<double>& g_phix, const std::vector<double>& g_phiy,
                  const std::vector<double>& g_c0, const std::vector<int>& t_ptr, const std::vector<int>& t_gp,
                  const std::vector<int>& bn_node, const std::vector<int>& node_off, const std::vector<int>& gp_off,
                  const std::vector<int>& perreg, const std::vector<int>& d_ptr, const std::vector<int>& d_idx,
                  const std::vector<int>& bt_ptr, const std::vector<int>& bt_bn) {
    Model& m = c.model;
    const int ECAP = CHUNK_ENTRIES, CHCAP = BLOCK_THREADS;   // slots and Gauss points per chunk
    int N = node_off.back();
    std::vector<int> blk_body, blk_own_off, blk_nown, blk_touch_off, blk_ntouch, blk_chunk_off, blk_nchunk;
    std::vector<int> blk_single, own_nodes, touch_nodes;
    std::vector<int4> ch_meta;
    std::vector<int2> ch_ptr_off;
    std::vector<double> bgp_c0, be_fxy, blk_work;
    std::vector<int> bgp_region;
    std::vector<int> tptr, gl_cnt;
    std::vector<uint32_t> be_pk;
    std::vector<int> local(N, -1);
    int max_touch = 1, max_own = 1;
    // MELODY2D_BLOCK_NODES lowers the nodes per block (exercises the split
    // path on small models); default and maximum BLOCK_MAX_OWN
    int own_cap = BLOCK_MAX_OWN;
    if (const char* e = std::getenv("MELODY2D_BLOCK_NODES")) {
        int v = std::atoi(e);
        if (v >= 1 && v < BLOCK_MAX_OWN) own_cap = v;
    }
    if (c.opts.block_nodes >= 1) own_cap = std::min(c.opts.block_nodes, BLOCK_MAX_OWN);
    // MELODY2D_EXACT_SUMS=1: internal forces summed in the reference's order
    // (bit-identical where pow rounds alike); default: balanced order.  The
    // FP32-mixed form, which rounds differently anyway, always takes the
    // balanced order (its kernels assume single-region blocks).
    const char* ex = std::getenv("MELODY2D_EXACT_SUMS");
    const bool exact = ex && std::atoi(ex) != 0 && c.precision != MEL_FP32_MIXED;
    for (size_t b = 0; b < m.bodies.size(); b++) {
        const Body& Bd = m.bodies[b];
        if (Bd.rigid) continue;
        int n = Bd.n_nodes, o = node_off[b];
        std::vector<std::vector<int>> groups;
        if (n <= own_cap) {
            std::vector<int> g(n);
            for (int i = 0; i < n; i++) g[i] = o + i;
            groups.push_back(g);
        } else {
            double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
            for (int i = 0; i < n; i++) {
                xmin = std::min(xmin, Bd.X[i]); xmax = std::max(xmax, Bd.X[i]);
                ymin = std::min(ymin, Bd.Y[i]); ymax = std::max(ymax, Bd.Y[i]);
            }
            double sc = std::max(xmax - xmin, ymax - ymin);
            if (sc <= 0) sc = 1;
            std::vector<std::pair<uint64_t, int>> key(n);
            for (int i = 0; i < n; i++) {
                uint32_t qx = (uint32_t)std::min(65535.0, (Bd.X[i] - xmin) / sc * 65535.0);
                uint32_t qy = (uint32_t)std::min(65535.0, (Bd.Y[i] - ymin) / sc * 65535.0);
                uint64_t z = 0;
                for (int bit = 0; bit < 16; bit++) {
                    z |= (uint64_t)((qx >> bit) & 1u) << (2 * bit);
                    z |= (uint64_t)((qy >> bit) & 1u) << (2 * bit + 1);
                }
                key[i] = {z, i};
            }
            std::sort(key.begin(), key.end());
            for (int s0 = 0; s0 < n; s0 += own_cap) {
                std::vector<int> g;
                for (int i = s0; i < std::min(n, s0 + own_cap); i++) g.push_back(o + key[i].second);
                std::sort(g.begin(), g.end());
                groups.push_back(g);
            }
        }
        bool first = true;
        for (auto& owned : groups) {
            int blk = (int)blk_body.size();
            blk_body.push_back((int)b);
            blk_own_off.push_back((int)own_nodes.size());
            blk_nown.push_back((int)owned.size());
            max_own = std::max(max_own, (int)owned.size());
            std::vector<int> gps;
            for (int j : owned)
                for (int e = t_ptr[j]; e < t_ptr[j + 1]; e++) gps.push_back(t_gp[e]);
            std::sort(gps.begin(), gps.end());
            gps.erase(std::unique(gps.begin(), gps.end()), gps.end());
            std::vector<int> touched = owned;
            for (size_t i = 0; i < owned.size(); i++) local[owned[i]] = (int)i;
            for (int g : gps)
                for (int k = g_ptr[g]; k < g_ptr[g + 1]; k++) {
                    int j = g_idx[k];
                    if (local[j] < 0) { local[j] = (int)touched.size(); touched.push_back(j); }
                }
            if (touched.size() > 65535) throw CapacityError{"node block touches more than 65535 nodes"};
            max_touch = std::max(max_touch, (int)touched.size());
            blk_touch_off.push_back((int)touch_nodes.size());
            blk_ntouch.push_back((int)touched.size());
            own_nodes.insert(own_nodes.end(), owned.begin(), owned.end());
            touch_nodes.insert(touch_nodes.end(), touched.begin(), touched.end());
            blk_chunk_off.push_back((int)ch_meta.size());
            {
                // regions only shape the rounding order; the balanced order
                // sums all of a node's entries in one accumulator
                int r0 = gps.empty() ? 0 : (gps.front() - gp_off[b]) / perreg[b];
                int r1 = gps.empty() ? 0 : (gps.back() - gp_off[b]) / perreg[b];
                blk_single.push_back(!exact || r0 == r1 ? 1 : 0);
            }
            int nown = (int)owned.size();
            for (int g : gps)
                if (g_ptr[g + 1] - g_ptr[g] > ECAP) throw CapacityError{"Gauss point support larger than the chunk capacity"};
            std::vector<std::vector<int>> chunks;
            if (exact) {
                // contiguous Gauss-point ranges: every node sums its entries
                // in Gauss-point order, region by region, as the reference does
                size_t q = 0;
                while (q < gps.size()) {
                    std::vector<int> ch;
                    int ne = 0;
                    while (q < gps.size() && (int)ch.size() < CHCAP) {
                        int cnt = g_ptr[gps[q] + 1] - g_ptr[gps[q]];
                        if (ne + cnt > ECAP) break;
                        ne += cnt;
                        ch.push_back(gps[q++]);
                    }
                    chunks.push_back(ch);
                }
            } else {
                // waves of 64 Gauss points of nearly equal support size, so
                // that a lane pads little to its wave's largest support: the
                // Gauss points by decreasing support size (equal sizes in
                // round-robin order over the chunk count, which spreads a
                // node's entries over the chunks and balances the per-node
                // gather), cut into waves; a chunk pairs the heaviest
                // remaining wave with the lightest one when both fit.  Sum
                // order: chunk by chunk, Gauss-point order inside a chunk.
                // a chunk's capacity is in contribution slots, which only
                // the entries of owned nodes take (the others go to the sink):
                // a block of a split body streams the whole support of its
                // Gauss points but fills slots with about half of it
                std::vector<int> own_k(gps.size(), 0);   // per position in gps
                long long tot = 0;
                for (size_t a = 0; a < gps.size(); a++) {
                    for (int k = g_ptr[gps[a]]; k < g_ptr[gps[a] + 1]; k++) own_k[a] += local[g_idx[k]] < nown;
                    tot += own_k[a];
                }
                const int nrr = (int)std::max<long long>({1, (tot + ECAP * 15 / 16 - 1) / (ECAP * 15 / 16),
                                                          ((long long)gps.size() + CHCAP - 1) / CHCAP});
                std::vector<int> ord(gps.size());
                for (size_t a = 0; a < gps.size(); a++) ord[a] = (int)a;
                std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
                    const int kx = g_ptr[gps[x] + 1] - g_ptr[gps[x]], ky = g_ptr[gps[y] + 1] - g_ptr[gps[y]];
                    if (kx != ky) return kx > ky;
                    return x % nrr < y % nrr;
                });
                std::vector<std::vector<int>> waves;
                std::vector<int> wsum;
                for (size_t a = 0; a < ord.size();) {
                    // at most 64 lanes and ECAP entries per wave
                    std::vector<int> wv;
                    int s = 0;
                    while (a < ord.size() && wv.size() < 64) {
                        const int cnt = own_k[ord[a]];
                        if (s + cnt > ECAP) break;
                        wv.push_back(gps[ord[a++]]);
                        s += cnt;
                    }
                    waves.push_back(wv);
                    wsum.push_back(s);
                }
                int i = 0, j = (int)waves.size() - 1;
                while (i <= j) {
                    std::vector<int> ch = waves[i];
                    if (i < j && wsum[i] + wsum[j] <= ECAP) {
                        ch.insert(ch.end(), waves[j].begin(), waves[j].end());
                        j--;
                    }
                    i++;
                    // Gauss-point order inside the chunk
                    std::sort(ch.begin(), ch.end());
                    chunks.push_back(ch);
                }
                // a block's time grows with its chunk count: when the pairing
                // needs more chunks than dealing the Gauss points round-robin
                // over the chunks, deal them (more padding, fewer chunks)
                for (int nch = nrr; nch < (int)chunks.size(); nch++) {
                    std::vector<std::vector<int>> rr(nch);
                    std::vector<int> ne(nch, 0);
                    bool ok = true;
                    for (size_t a = 0; a < gps.size() && ok; a++) {
                        const int k = (int)(a % nch), g = gps[a];
                        ne[k] += own_k[a];
                        rr[k].push_back(g);
                        ok = ne[k] <= ECAP && (int)rr[k].size() <= CHCAP;
                    }
                    if (ok) {
                        chunks = rr;
                        break;
                    }
                }
            }
            for (const std::vector<int>& chunk : chunks) {
                // lanes: the chunk's Gauss points by decreasing support size
                // (stable), 64 per wave; a wave's entries are stored
                // entry-major (entry i of its n_w lanes contiguous), each
                // lane padded to the wave's largest support K_w with zero
                // derivatives and the sink slot
                const int ngp = (int)chunk.size();
                std::vector<int> order(ngp);
                for (int i = 0; i < ngp; i++) order[i] = i;
                std::stable_sort(order.begin(), order.end(), [&](int a, int b2) {
                    return g_ptr[chunk[a] + 1] - g_ptr[chunk[a]] > g_ptr[chunk[b2] + 1] - g_ptr[chunk[b2]];
                });
                // destination slot of every entry: per owned node, its
                // entries in Gauss-point order
                std::vector<int> first_entry(ngp + 1, 0);
                for (int i = 0; i < ngp; i++)
                    first_entry[i + 1] = first_entry[i] + g_ptr[chunk[i] + 1] - g_ptr[chunk[i]];
                const int ne = first_entry[ngp];
                std::vector<int> cnt(nown, 0), dst(ne, CHUNK_ENTRIES);
                for (int i = 0; i < ngp; i++)
                    for (int k = g_ptr[chunk[i]]; k < g_ptr[chunk[i] + 1]; k++) {
                        int li = local[g_idx[k]];
                        if (li < nown) cnt[li]++;
                    }
                int acc = 0;
                tptr.push_back(0);
                std::vector<int> next(nown);
                for (int i = 0; i < nown; i++) { next[i] = acc; acc += cnt[i]; tptr.push_back(acc); }
                if (acc > CHUNK_ENTRIES) throw CapacityError{"chunk has more contribution slots than CHUNK_ENTRIES"};
                for (int i = 0; i < ngp; i++)
                    for (int k = g_ptr[chunk[i]]; k < g_ptr[chunk[i] + 1]; k++) {
                        int li = local[g_idx[k]];
                        if (li < nown) dst[first_entry[i] + k - g_ptr[chunk[i]]] = next[li]++;
                    }
                int Kw[2] = {0, 0}, nw[2] = {0, 0};
                for (int l = 0; l < ngp; l++) {
                    int gi = order[l];
                    nw[l / 64]++;
                    Kw[l / 64] = std::max(Kw[l / 64], g_ptr[chunk[gi] + 1] - g_ptr[chunk[gi]]);
                }
                if (Kw[0] > 255 || Kw[1] > 255) throw CapacityError{"Gauss point support larger than 255 entries"};
                if (be_pk.size() > (size_t)INT32_MAX) throw CapacityError{"support entries exceed the chunk index range"};
                ch_meta.push_back(make_int4((int)bgp_c0.size(), ngp, (int)be_pk.size(),
                                            Kw[0] | (Kw[1] << 8) | (nw[0] << 16) | (nw[1] << 24)));
                ch_ptr_off.push_back(make_int2(ne, (int)tptr.size() - nown - 1));
                for (int l = 0; l < ngp; l++) {
                    int g = chunk[order[l]];
                    bgp_c0.push_back(g_c0[g]);
                    bgp_region.push_back((g - gp_off[b]) / perreg[b]);
                    gl_cnt.push_back(g_ptr[g + 1] - g_ptr[g]);
                }
                for (int w = 0; w < 2; w++) {
                    const size_t base = be_pk.size();
                    be_pk.resize(base + (size_t)Kw[w] * nw[w], (uint32_t)CHUNK_ENTRIES << 16);
                    be_fxy.resize(2 * (base + (size_t)Kw[w] * nw[w]), 0.);
                    for (int lw = 0; lw < nw[w]; lw++) {
                        int gi = order[64 * w + lw], g = chunk[gi];
                        for (int k = g_ptr[g]; k < g_ptr[g + 1]; k++) {
                            int i = k - g_ptr[g];
                            size_t e = base + (size_t)i * nw[w] + lw;
                            int li = local[g_idx[k]];
                            be_pk[e] = (uint32_t)li | ((uint32_t)dst[first_entry[gi] + i] << 16);
                            be_fxy[2 * e] = g_phix[k];
                            be_fxy[2 * e + 1] = g_phiy[k];
                        }
                    }
                }
            }
            blk_nchunk.push_back((int)ch_meta.size() - blk_chunk_off.back());
            if (blk_nchunk.back() > MAX_BLOCK_CHUNKS)
                throw CapacityError{"node block streams more than " + std::to_string(MAX_BLOCK_CHUNKS) +
                                    " chunks (set MELODY2D_BLOCK_NODES lower)"};
            for (int j : touched) local[j] = -1;
            std::vector<double> w(14, 0.);
            if (first)
                for (int k = 0; k < 7; k++)
                    if (k != W_CONTACT) w[2 * k] = Bd.works[k];
            first = false;
            blk_work.insert(blk_work.end(), w.begin(), w.end());
            c.h_blk_body.push_back((int)b);
            (void)blk;
        }
    }
    // border nodes finished by each block (its own, when the block holds a
    // whole body) and those of split bodies (finished by separate launches)
    std::vector<int> node_bn(N, -1), kin_all, blk_bn_off, blk_bn_n, bn_list, split_bn, split_kin;
    for (size_t q = 0; q < bn_node.size(); q++) node_bn[bn_node[q]] = (int)q;
    {
        std::vector<int> nblk_of_body(m.bodies.size(), 0);
        for (int b : blk_body) nblk_of_body[b]++;
        for (size_t k = 0; k < blk_body.size(); k++) {
            blk_bn_off.push_back((int)bn_list.size());
            int n0 = (int)bn_list.size();
            if (nblk_of_body[blk_body[k]] == 1)
                for (int o = 0; o < blk_nown[k]; o++) {
                    int bn = node_bn[own_nodes[blk_own_off[k] + o]];
                    if (bn >= 0) bn_list.push_back(bn);
                }
            blk_bn_n.push_back((int)bn_list.size() - n0);
        }
        for (size_t q = 0; q < bn_node.size(); q++) {
            int b = c.d.node_body_host[bn_node[q]];
            if (!m.bodies[b].rigid && nblk_of_body[b] > 1) {
                split_bn.push_back((int)q);
                split_kin.push_back(bn_node[q]);
            }
        }
    }
    for (size_t b = 0; b < m.bodies.size(); b++)
        if (!m.bodies[b].rigid)
            for (int i = 0; i < m.bodies[b].n_nodes; i++) kin_all.push_back(node_off[b] + i);
    // the one-launch step's finish (node_block_finish_lds): owned nodes
    // contiguous, spreading entries of owned nodes name the block's border
    // nodes, kinematics rows of its border nodes name owned nodes, and the
    // staged point forces + parameters fit the contribution buffer
    std::vector<uint16_t> bt_lbn(std::max<size_t>(bt_bn.size(), 1), 0);
    bool lds_ok = split_bn.empty() && split_kin.empty();
    {
        std::vector<int> bn_local(bn_node.size(), -1);
        for (size_t k = 0; k < blk_body.size() && lds_ok; k++) {
            const int* own = own_nodes.data() + blk_own_off[k];
            const int nown = blk_nown[k], n0 = own[0];
            const int* bl = bn_list.data() + blk_bn_off[k];
            const int nbn = blk_bn_n[k];
            if (6 * (size_t)(nbn + nown) * sizeof(double) > sizeof(double2) * ((size_t)CHUNK_ENTRIES + 1) || nbn > 65535) {
                lds_ok = false;
                break;
            }
            for (int o = 0; o < nown; o++)
                if (own[o] != n0 + o) lds_ok = false;
            for (int i = 0; i < nbn; i++) bn_local[bl[i]] = i;
            for (int o = 0; o < nown && lds_ok; o++)
                for (int e = bt_ptr[n0 + o]; e < bt_ptr[n0 + o + 1]; e++) {
                    if (bn_local[bt_bn[e]] < 0) { lds_ok = false; break; }
                    bt_lbn[e] = (uint16_t)bn_local[bt_bn[e]];
                }
            for (int i = 0; i < nbn && lds_ok; i++) {
                const int node = bn_node[bl[i]];
                for (int q = d_ptr[node]; q < d_ptr[node + 1]; q++)
                    if (d_idx[q] < n0 || d_idx[q] >= n0 + nown) { lds_ok = false; break; }
            }
            for (int i = 0; i < nbn; i++) bn_local[bl[i]] = -1;
        }
    }
    // block headers (what a block reads before its first loads, in one go)
    std::vector<int4> blk_hdr;
    std::vector<double2> blk_par;
    for (size_t k = 0; k < blk_body.size(); k++) {
        const Body& Bd = m.bodies[blk_body[k]];
        const int* own = own_nodes.data() + blk_own_off[k];
        const int* touch = touch_nodes.data() + blk_touch_off[k];
        const int nown = blk_nown[k], ntouch = blk_ntouch[k], own0 = nown > 0 ? own[0] : 0;
        bool own_contig = true, touch_contig = true;
        for (int o = 0; o < nown; o++) own_contig = own_contig && own[o] == own0 + o;
        for (int i = 0; i < ntouch; i++) touch_contig = touch_contig && touch[i] == own0 + i;
        const int flags = (blk_single[k] ? BLK_SINGLE : 0) | (Bd.status == 1 ? BLK_INACTIVE : 0) |
                          (Bd.status == 2 ? BLK_NO_GRAVITY : 0) |
                          (m.materials[Bd.material].type != 0 ? BLK_NEO_HOOKEAN : 0) |
                          (own_contig ? BLK_OWN_CONTIGUOUS : 0);
        blk_hdr.push_back(make_int4(own0, nown, ntouch, touch_contig ? -1 : blk_touch_off[k]));
        blk_hdr.push_back(make_int4(blk_chunk_off[k], blk_nchunk[k], flags, blk_own_off[k]));
        const bool any = blk_nchunk[k] > 0;
        blk_hdr.push_back(any ? ch_meta[blk_chunk_off[k]] : make_int4(0, 0, 0, 0));
        const int2 po = any ? ch_ptr_off[blk_chunk_off[k]] : make_int2(0, 0);
        blk_hdr.push_back(make_int4(po.x, po.y, blk_bn_off[k], blk_bn_n[k]));
        blk_par.push_back(make_double2(m.materials[Bd.material].p[3], m.materials[Bd.material].p[4]));
    }
    BlockArrays& B = c.blocks_dev;
    B.blk_hdr = c.upload(blk_hdr);
    B.blk_par = c.upload(blk_par);
    B.bt_lbn = c.upload(bt_lbn);
    B.lds_finish_ok = lds_ok ? 1 : 0;
    B.n_blocks = (int)blk_body.size();
    B.max_touch = max_touch;
    B.max_own = max_own;
    B.blk_body = c.upload(blk_body); B.blk_own_off = c.upload(blk_own_off); B.blk_nown = c.upload(blk_nown);
    B.blk_touch_off = c.upload(blk_touch_off); B.blk_ntouch = c.upload(blk_ntouch);
    B.blk_chunk_off = c.upload(blk_chunk_off); B.blk_nchunk = c.upload(blk_nchunk);
    B.blk_single_region = c.upload(blk_single);
    B.any_multi_region = 0;
    for (int v : blk_single) B.any_multi_region |= v ? 0 : 1;
    B.own_nodes = c.upload(own_nodes); B.touch_nodes = c.upload(touch_nodes);
    B.ch_meta = c.upload(ch_meta); B.ch_ptr_off = c.upload(ch_ptr_off);
    B.bgp_c0 = c.upload(bgp_c0); B.bgp_region = c.upload(bgp_region);
    // slack so that a lane reading a full register window past the end of
    // the stream stays inside the allocation
    be_pk.resize(be_pk.size() + 64 * KREG_ENTRIES, (uint32_t)CHUNK_ENTRIES << 16);
    be_fxy.resize(2 * be_pk.size(), 0.);
    B.gl_cnt = c.upload(gl_cnt); B.be_pk = c.upload(be_pk); B.tptr = c.upload(tptr);
    if (c.precision == MEL_FP32_MIXED) {
        std::vector<float> f32(be_fxy.begin(), be_fxy.end());
        B.be_fxy32 = reinterpret_cast<float2*>(c.upload(f32));
        B.be_fxy = nullptr;
    } else {
        B.be_fxy = reinterpret_cast<double2*>(c.upload(be_fxy));
        B.be_fxy32 = nullptr;
    }
    B.blk_work = c.upload(blk_work);
    B.blk_bn_off = c.upload(blk_bn_off); B.blk_bn_n = c.upload(blk_bn_n); B.bn_list = c.upload(bn_list);
    c.h_bn_list = bn_list;
    B.split_bn = c.upload(split_bn); B.n_split_bn = (int)split_bn.size();
    B.split_kin = c.upload(split_kin); B.n_split_kin = (int)split_kin.size();
    c.d_kin_all = c.upload(kin_all);
    c.n_kin_all = (int)kin_all.size();
    c.block_lds = block_lds_size(B, c.precision == MEL_FP32_MIXED);
    if (c.block_lds + 16 * ((size_t)CHUNK_ENTRIES + 1) + 2048 > 160 * 1024)
        throw CapacityError{"node block needs more LDS than a workgroup has"};
    const bool f32 = c.precision == MEL_FP32_MIXED;
    // two launches per step unless a body is split over several node blocks
    // (its blocks read each other's displacement parameters);
    // MELODY2D_FUSED_STEP=0 asks for the launch sequence
    const char* fe = getenv("MELODY2D_FUSED_STEP");
    const bool fused_off = c.opts.fused_step >= 0 ? c.opts.fused_step == 0 : (fe && fe[0] == '0');
    // (the adaptive schemes run their half steps as the launch sequence, and
    // so do models with absorbing layers, deactivation rules or inactive
    // bodies)
    bool any_inactive = false;
    for (const Body& b : c.model.bodies) any_inactive = any_inactive || b.status == ST_INACTIVE;
    c.two_launch = B.lds_finish_ok && !fused_off && c.model.scheme == SCH_EULER && !c.d.alid_on &&
                   c.d.n_rules == 0 && !any_inactive;
    // the launch sequence takes the wide-window kernel when Gauss points
    // with more support entries than the default window are common (RPIM
    // supports of 30 nodes; a few long MLS supports read their tail from
    // memory at less cost than the wide window's register pressure);
    // MELODY2D_WIDE_LANES=1 asks for the launch sequence with that kernel
    const int window = f32 ? KREG_ENTRIES_F32 : KREG_ENTRIES;
    long long n_gp = 0, n_long = 0;
    for (size_t g = 0; g + 1 < g_ptr.size(); g++) {
        n_gp++;
        n_long += g_ptr[g + 1] - g_ptr[g] > window;
    }
    const char* we = getenv("MELODY2D_WIDE_LANES");
    const bool force_wide = c.opts.wide_window >= 0 ? c.opts.wide_window == 1 : (we && we[0] == '1');
    if (force_wide) c.two_launch = false;
    c.wide = force_wide || (c.opts.wide_window != 0 && !c.two_launch && 10 * n_long >= n_gp && n_long > 0);
    HIP_OK(hipFuncSetAttribute(f32 ? (c.wide ? (const void*)k_block_force_update32_wide : (const void*)k_block_force_update32)
                                   : (c.wide ? (const void*)k_block_force_update_wide : (const void*)k_block_force_update),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.block_lds));
    // runs of steps replayed from captured graphs; MELODY2D_GRAPHS=0 launches
    // every step
    const char* ge = getenv("MELODY2D_GRAPHS");
    const bool graphs_off = c.opts.graph_replay >= 0 ? c.opts.graph_replay == 0 : (ge && ge[0] == '0');
    c.use_graphs = c.two_launch && !graphs_off;
    if (c.two_launch) {
        HIP_OK(hipFuncSetAttribute(f32 ? (const void*)k_step_forces32 : (const void*)k_step_forces,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.block_lds));
        int max_bn = 0;
        for (int v : blk_bn_n) max_bn = std::max(max_bn, v);
        c.finish_lds = step_finish_lds(max_bn, max_own);
        HIP_OK(hipFuncSetAttribute((const void*)k_step_finish, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)c.finish_lds));
        int per_cu = 0, cus = 0;
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_step_finish, BLOCK_THREADS,
                                                            c.finish_lds));
        HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
        c.finish_slots = per_cu * cus;
    }
}

// interior kinematics and force fields for output, after steps have run
void refresh_outputs(mel_ctx& c) {
    if (!c.outputs_stale) return;
    if (c.n_kin_all > 0) k_kinematics_list<<<blocks(c.n_kin_all), BLOCK, 0, c.stream>>>(c.d, c.d_kin_all, c.n_kin_all);
    k_node_outputs<<<blocks(c.d.n_nodes), BLOCK, 0, c.stream>>>(c.d, 1);
    HIP_OK(hipGetLastError());
    c.outputs_stale = false;
}

// contact work held by the elements (Code/Body.h:1847-1848) added to their
// bodies' works, in a fixed order; before each proximity update (which
// replaces the elements) and before works are read
void flush_contact_work(mel_ctx& c) {
    if (!c.work_law) return;
    k_flush_contact_work<<<c.d.n_bodies, BLOCK, 0, c.stream>>>(c.d);
    k_clear_contact_work<<<blocks(std::max(1, c.d.c_cap)), BLOCK, 0, c.stream>>>(c.d);
    HIP_OK(hipGetLastError());
}

// Update_proximity (Code/Proximity.h:48-473) without host round trips: the
// kernels size their loops from counts on the device
void update_proximity(mel_ctx& c) {
    DevModel& d = c.d;
    hipStream_t s = c.stream;
    flush_contact_work(c);
    const int persistent = 4 * 256;     // workgroups of the grid-stride loops
    c.prox_after_step = true;
    // (the chunk-box pass also clears the cell list's counters)
    if (d.n_seg_chunks > 0) k_prox_chunk_boxes<<<blocks((long long)d.n_seg_chunks * SEG_CHUNK), BLOCK, 0, s>>>(d);
    else k_cells_clear<<<std::min(blocks(d.cell_n), 1024), BLOCK, 0, s>>>(d);
    k_prox_aabb<<<blocks((long long)d.n_borders * 64), BLOCK, 0, s>>>(d);
    // broad phase: the cell list, then the neighbour lists from it
    const int cblocks = (d.cell_n + 1023) / 1024;
    k_cells_assign<<<blocks(d.n_borders), BLOCK, 0, s>>>(d);
    k_cells_scan_blocks<<<cblocks, 1024, 0, s>>>(d);
    k_cells_scan_top<<<1, 1024, 0, s>>>(d, cblocks);
    k_cells_scatter<<<blocks(d.n_borders), BLOCK, 0, s>>>(d);
    k_prox_neighbours<<<d.n_bodies, NB_THREADS, 0, s>>>(d);
    k_prox_chunk_offsets<<<1, 1024, 0, s>>>(d);
    k_prox_chunk_bodies<<<d.n_bodies, BLOCK, 0, s>>>(d, c.max_row_chunks);
    // previous elements become the history source
    std::swap(c.bufA, c.bufB);
    d.C = c.bufA;
    d.Cold = c.bufB;
    // one counting workgroup per row chunk up to 16 per CU: the hardware
    // hands the next chunk to whichever workgroup slot frees first
    const int n_count = std::min(16 * 256, std::max(1, c.max_row_chunks));
    k_prox_count<<<n_count + 64, ROW_CHUNK, 0, s>>>(d, n_count);
    k_prox_elem_offsets<<<1, 1024, 0, s>>>(d);
    k_prox_write<<<persistent, BLOCK, 0, s>>>(d);
    k_prox_history<<<persistent, BLOCK, 0, s>>>(d);
    k_prox_chain<<<persistent, BLOCK, 0, s>>>(d);
    HIP_OK(build_rigid_lists(d, s, false));     // the chunk-offset and chain passes collected the elements
    HIP_OK(hipMemcpyAsync(c.h_async, d.c_count, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(c.h_async + 1, d.status_flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipEventRecord(c.ev_async, s));
    c.async_pending = true;
}

// Initialize_CZM (Code/Proximity.h:481-541); it clears the INITIALIZE_CZM
// flag (Code/Proximity.h:486), so the second call of Code/main.cpp:124 is a no-op
void initialize_czm(mel_ctx& c) {
    c.model.flags[F_INITIALIZE_CZM] = 0;
    if (contact_count(c) == 0) return;
    k_project_all<<<blocks(c.n_contacts), BLOCK, 0, c.stream>>>(c.d);
    k_init_czm<<<blocks(c.n_contacts), BLOCK, 0, c.stream>>>(c.d);
}

// the contact update of the stepping loop (Code/main.cpp:191-196):
// Update_proximity, then Update_status when the model has deactivation rules
void contact_update(mel_ctx& c) {
    update_proximity(c);
    if (c.d.n_rules > 0) k_update_status<<<1, STATUS_THREADS, 0, c.stream>>>(c.d, c.blocks_dev);
}

// Monitoring rows written on the device since the last drain appended to the
// host's rows (Monitoring keeps them in memory until Write_monitoring)
void mon_drain(mel_ctx& c) {
    if (!c.mon_on || c.mon_pending == 0) return;
    int n = 0;
    HIP_OK(hipMemcpyAsync(&n, c.d_mon_n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    std::vector<double> rows((size_t)n * c.mon_stride);
    if (n > 0)
        HIP_OK(hipMemcpyAsync(rows.data(), c.d_mon_rows, rows.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    const int zero = 0;
    HIP_OK(hipMemcpyAsync(c.d_mon_n, &zero, sizeof(int), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    for (int r = 0; r < n; r++) {
        const double* q = rows.data() + (size_t)r * c.mon_stride;
        const int nv = (int)q[0];
        c.h_mon.push_back(nv);
        c.h_mon.insert(c.h_mon.end(), q + 1, q + 1 + nv);
    }
    c.mon_pending = 0;
}

// Monitoring (Code/Monitoring.h:8-559) after a step (Code/main.cpp:172-176):
// the energies it reads first (physical velocities of every node for
// MONITOR_ENERGY), then the row
void launch_monitor(mel_ctx& c, double neval, double max_error, double mean_error) {
    if (!c.mon_on) return;
    if (c.mon_pending >= c.mon_cap) mon_drain(c);
    const Model& m = c.model;
    bool energies = m.flags[F_MONITOR_ENERGY] == 1;
    for (const std::vector<double>& q : m.monitored) energies = energies || (int)q[0] == 9;
    if (energies) {
        if (m.flags[F_MONITOR_ENERGY] == 1 && c.n_kin_all > 0)
            k_kinematics_list<<<blocks(c.n_kin_all), BLOCK, 0, c.stream>>>(c.d, c.d_kin_all, c.n_kin_all);
        k_body_energy<<<c.d.n_bodies, BLOCK, 0, c.stream>>>(c.d, c.d_energy);
    }
    MonitorArgs a{};
    a.rows = c.d_mon_rows;
    a.n_rows = c.d_mon_n;
    a.stride = c.mon_stride;
    a.cap = c.mon_cap;
    a.items = c.d_mon_items;
    a.n_items = (int)m.monitored.size();
    a.energy = m.flags[F_MONITOR_ENERGY] == 1;
    a.body_energy = c.d_energy;
    a.stress = c.st_out;
    a.iter = c.d_mon_iter;
    a.neval = neval;
    a.max_error = max_error;
    a.mean_error = mean_error;
    k_monitor<<<1, STATUS_THREADS, 0, c.stream>>>(c.d, c.blocks_dev, a);
    HIP_OK(hipGetLastError());
    c.mon_pending++;
}

// the launches of one step in two launches (the device advances its time)
void launch_step(mel_ctx& c, int n_cwg) {
    DevModel& d = c.d;
    const BlockArrays& B = c.blocks_dev;
    hipStream_t s = c.stream;
    // Update_damping_forces reads the velocity parameters of the start of
    // the step, so it runs before the step kernels update them
    if (d.n_matrix > 0) TIMED(c, KT_DAMPING, s, (k_damping<<<blocks(d.n_nodes), BLOCK, 0, s>>>(d)));
    StepLaunch S{n_cwg};
    const int ga = step_forces_grid(d, B, S);
    TIMED(c, KT_STEP_FORCES, s,
          (c.precision == MEL_FP32_MIXED ? k_step_forces32<<<ga, BLOCK_THREADS, c.block_lds, s>>>(d, B, S)
                                         : k_step_forces<<<ga, BLOCK_THREADS, c.block_lds, s>>>(d, B, S)));
    const int n_rwg = step_rigid_node_wgs(d.n_rigid_nodes, c.finish_slots - d.n_rigid - B.n_blocks - 1);
    TIMED(c, KT_STEP_FINISH, s,
          (k_step_finish<<<step_finish_grid(d, B, n_rwg), BLOCK_THREADS, c.finish_lds, s>>>(d, B, n_rwg)));
}

// a run of steps between contact updates: replayed from a captured graph
// (one per run length, contact workgroup count and current contact
// buffer), or launched step by step while kernels are being timed
void run_steps(mel_ctx& c, int n) {
    poll_async(c, false);
    const int n_cwg = contact_workgroups(c);
    if (n > 0) c.prox_after_step = false;
    if (!c.use_graphs || c.timing || n < 2) {
        for (int k = 0; k < n; k++) launch_step(c, n_cwg);
        return;
    }
    const mel_ctx::GraphKey key{n, n_cwg, (const void*)c.d.C.bodyS};
    auto it = c.graphs.find(key);
    if (it == c.graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        HIP_OK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < n; k++) launch_step(c, n_cwg);
        HIP_OK(hipStreamEndCapture(c.stream, &g));
        hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_OK(err);
        it = c.graphs.emplace(key, e).first;
    }
    HIP_OK(hipGraphLaunch(it->second, c.stream));
}

// Update_internal_forces (Code/Body.h:937-1049) over every node block
void launch_internal_forces(mel_ctx& c, hipStream_t s) {
    const DevModel& d = c.d;
    const BlockArrays& B = c.blocks_dev;
    if (B.n_blocks == 0) return;
    if (c.precision == MEL_FP32_MIXED) {
        if (c.wide) k_block_force_update32_wide<<<B.n_blocks, BLOCK_THREADS, c.block_lds, s>>>(d, B);
        else k_block_force_update32<<<B.n_blocks, BLOCK_THREADS, c.block_lds, s>>>(d, B);
    } else {
        if (c.wide) k_block_force_update_wide<<<B.n_blocks, BLOCK_THREADS, c.block_lds, s>>>(d, B);
        else k_block_force_update<<<B.n_blocks, BLOCK_THREADS, c.block_lds, s>>>(d, B);
    }
}

// Euler_step (Code/Solver.h:10-88)
void euler_step(mel_ctx& c) {
    DevModel& d = c.d;
    const BlockArrays& B = c.blocks_dev;
    hipStream_t s = c.stream, a = c.aux;
    poll_async(c, false);
    c.prox_after_step = false;
    const int n_cwg = contact_workgroups(c);
    const double t_next = c.time + d.dt;
    if (c.two_launch) {
        launch_step(c, n_cwg);
        c.time = t_next;
        c.outputs_stale = true;
        return;
    }
    // the contact pipeline runs on the second stream, concurrently with the
    // internal forces, which do not depend on it; its launches are a few
    // workgroups, which fit beside the internal-force blocks
    HIP_OK(hipEventRecord(c.ev_fork, s));
    HIP_OK(hipStreamWaitEvent(a, c.ev_fork, 0));
    // Initialize_contact_forces, Update_contacts, Update_contact_forces
    // (records into buckets) | Update_bc_forces
    TIMED(c, KT_CONTACTS, a, (k_step_contacts<<<n_cwg, BLOCK_THREADS, 0, a>>>(d, B.bc_bn, B.n_bc_bn)));
    // Send_contact_forces of the inactive bodies
    if (d.n_rules > 0) k_ghost_sends<<<64, BLOCK, 0, a>>>(d);
    // Update_damping_forces, Update_alid_forces(Time)
    if (d.n_matrix > 0) TIMED(c, KT_DAMPING, a, (k_damping<<<blocks(d.n_nodes), BLOCK, 0, a>>>(d)));
    if (d.n_alid_nodes > 0) k_alid<<<blocks(d.n_alid_nodes), BLOCK, 0, a>>>(d, c.time);
    // rigid resultants, Newton, Euler | sends to the border nodes of split bodies
    if (d.n_rigid + B.n_split_bn > 0)
        TIMED(c, KT_GATHER, a, (k_step_gather<<<d.n_rigid + blocks(B.n_split_bn, BLOCK_THREADS), BLOCK_THREADS, 0, a>>>(d, B.split_bn, B.n_split_bn)));
    HIP_OK(hipEventRecord(c.ev_join, a));
    // Update_internal_forces
    if (B.n_blocks > 0) TIMED(c, KT_BLOCK_UPDATE, s, launch_internal_forces(c, s));
    HIP_OK(hipStreamWaitEvent(s, c.ev_join, 0));
    // sends, Sum_up_forces, Apply_Newton, Apply_Euler, Update_borders,
    // Update_kinematics of border nodes (interior ones are rebuilt on request)
    // | rigid nodes | Update_bc(t + dt) | time = t + dt (the host's copy of the
    // time is the one the device takes)
    const int extra = std::max(1, (d.n_rigid_nodes + d.n_borders + BLOCK_THREADS - 1) / BLOCK_THREADS);
    TIMED(c, KT_FINISH, s, (k_update_finish<<<B.n_blocks + extra, BLOCK_THREADS, 0, s>>>(d, B, t_next)));
    if (B.n_split_kin > 0)
        k_kinematics_list<<<blocks(B.n_split_kin), BLOCK, 0, s>>>(d, B.split_kin, B.n_split_kin);
    // the absorbing layers read every node's physical velocity of the end of the step
    if (d.n_alid_nodes > 0) k_kinematics_list<<<blocks(d.n_alid_nodes), BLOCK, 0, s>>>(d, d.alid_nodes, d.n_alid_nodes);
    c.time = t_next;
    c.outputs_stale = true;
}

// ---------------------------------------------------------------------------
// Solver_step (Code/Solver.h:98-486): set-up of the adaptive schemes
// ---------------------------------------------------------------------------
// Per-node inverse distance estimators (Load_static, Code/IO.h:1620, 1819),
// unit mass-scaling factors (Code/Node.h:195-199), the deformable bodies'
// driven directions, and the tables of the Store / Restore copies: what
// Body::Store keeps (Code/Body.h:2593-2617) and the step changes - node
// kinematics, parameters and mass-scaling factors, border lengths, BC
// values and BC pressures, contact elements, works - into a snapshot of the
// same layout.
void setup_adaptive(mel_ctx& c) {
    const Model& m = c.model;
    DevModel& d = c.d;
    AdArrays& A = c.ad;
    const size_t N = d.n_nodes, NB = d.n_bodies;
    A.uptx = c.alloc<double>(N); A.upty = c.alloc<double>(N);
    A.xerr = c.alloc<double>(N); A.yerr = c.alloc<double>(N);
    A.enorm = c.alloc<double>(N); A.msm = c.alloc<double>(N);
    std::vector<double> ide(N, 0.), one(N, 1.), nd(NB, 0.), body(NB * AD_BODY_FIELDS, 0.);
    std::vector<int> drv(NB, 0);
    for (size_t b = 0; b < NB; b++) {
        const Body& B = m.bodies[b];
        nd[b] = B.nodal_distance;
        body[b * AD_BODY_FIELDS + AB_FACTOR] = 1.;
        for (const Border& br : B.borders) {
            if (br.xbc.type == BC_DRIVEN) drv[b] |= 1;
            if (br.ybc.type == BC_DRIVEN) drv[b] |= 2;
        }
        for (int i = 0; i < B.n_nodes; i++) {
            const size_t g = c.h_node_off[b] + i;
            if (B.rigid) ide[g] = 1. / B.nodal_distance;
            else ide[g] = 1. / (B.domain[i] * std::sqrt(3.1415927 / 4. / (B.dptr[i + 1] - B.dptr[i])));
        }
    }
    A.ide = c.upload(ide);
    A.fsx = c.upload(one); A.fsy = c.upload(one); A.fsn = c.upload(one);
    A.body = c.upload(body);
    A.body_nd = c.upload(nd);
    A.body_drv = c.upload(drv);
    A.out = c.alloc<double>(AD_OUT_FIELDS);
    A.kin_nodes = c.d_kin_all;
    A.n_kin = c.n_kin_all;
    A.mass_scaling = m.scheme == SCH_MASS_SCALING || m.scheme == SCH_ADAPTIVE_MASS_SCALING;
    A.target_error = m.target_error;
    A.max_ms = m.max_mass_scaling;
    A.ctrl_ms = m.control_parameter_mass_scaling;
    A.err_factor_ms = m.error_factor_mass_scaling;
    A.decrease_ms = m.decrease_factor_mass_scaling;
    // the rigid bodies' driven DOFs of the initialisation block's
    // Update_bc(Time) (Code/main.cpp:109-110): the first half of the first
    // attempt moves driven bodies with them (the Euler step forms them
    // afresh at t + dt every step and never reads these)
    k_ad_bc<<<blocks(d.n_borders + d.n_rigid), BLOCK, 0, c.stream>>>(d, c.time);
    HIP_OK(hipGetLastError());
    HIP_OK(hipHostMalloc((void**)&c.h_ad, AD_OUT_FIELDS * sizeof(double), hipHostMallocDefault));
    HIP_OK(hipEventCreateWithFlags(&c.ev_ad, hipEventDisableTiming));
    // the snapshot: fixed arrays, then the contact elements at capacity
    struct Arr { void* p; unsigned elem; size_t n; bool per_contact; };
    std::vector<Arr> fixed;
    for (double* p : {d.xc, d.yc, d.ux, d.uy, d.vx, d.vy, d.ax, d.ay, d.upx, d.upy, d.vpx, d.vpy, d.apx, d.apy, A.fsx,
                      A.fsy})
        fixed.push_back({p, 8, N, false});
    for (double* p : {d.sl_len, d.sl_xbp, d.sl_ybp}) fixed.push_back({p, 8, (size_t)d.n_slots, false});
    for (double* p : {d.br_xval, d.br_yval, d.br_xvel, d.br_yvel}) fixed.push_back({p, 8, (size_t)d.n_borders, false});
    if (c.blocks_dev.n_blocks > 0) fixed.push_back({c.blocks_dev.blk_work, 8, 14 * (size_t)c.blocks_dev.n_blocks, false});
    auto contact_arrays = [&](const ContactArrays& C) {
        std::vector<Arr> v;
        for (int* p : {C.bodyS, C.borderS, C.bnodeS, C.nodeS, C.bodyM, C.borderM, C.shiftM, C.segM, C.nbint, C.fresh,
                       C.nM0, C.nM1, C.nM2, C.nM3, C.sM0, C.sM1, C.sM2, C.sM3})
            v.push_back({p, 4, 0, true});
        for (double* p : {C.N0, C.N1, C.N2, C.N3, C.gapn, C.vgapn, C.gapt, C.vgapt, C.xsi, C.xn, C.yn, C.xt, C.yt,
                          C.emass, C.len, C.fx, C.fy, C.int0, C.wacc, C.int1, C.int2})
            v.push_back({p, 8, 0, true});
        v.push_back({C.info, 16, 0, true});
        v.push_back({C.sinfo, 16, 0, true});
        return v;
    };
    size_t snap_bytes = 0;
    std::vector<size_t> off;
    auto layout = [&](const std::vector<Arr>& v) {
        for (const Arr& a : v) {
            off.push_back(snap_bytes);
            const size_t n = a.per_contact ? (size_t)d.c_cap : a.n;
            snap_bytes += (n * a.elem + 255) / 256 * 256;
        }
    };
    const std::vector<Arr> ca = contact_arrays(c.bufA), cb = contact_arrays(c.bufB);
    layout(fixed);
    layout(ca);
    unsigned char* snap = c.alloc<unsigned char>(snap_bytes);
    c.snap_rigid = c.alloc<double>(NB * RIGID_FIELDS);
    const ContactArrays* bufs[2] = {&c.bufA, &c.bufB};
    for (int k = 0; k < 2; k++) {
        std::vector<Arr> all = fixed;
        const std::vector<Arr>& cc = k == 0 ? ca : cb;
        all.insert(all.end(), cc.begin(), cc.end());
        std::vector<CopyDesc> st(all.size()), rs(all.size());
        for (size_t q = 0; q < all.size(); q++) {
            const unsigned long long live = (unsigned long long)all[q].p, copy = (unsigned long long)(snap + off[q]);
            st[q] = {live, copy, all[q].n, all[q].elem, all[q].per_contact ? 1u : 0u};
            rs[q] = {copy, live, all[q].n, all[q].elem, all[q].per_contact ? 1u : 0u};
        }
        c.d_copy[k][0] = c.upload(st);
        c.d_copy[k][1] = c.upload(rs);
        c.copy_key[k] = bufs[k]->bodyS;
        c.n_copy = (int)all.size();
    }
    HIP_OK(hipStreamSynchronize(c.stream));
}

// The contact elements read from the DYNAMIC file (Load_dynamic,
// Code/IO.h:2222-2287) as the current elements, so that the first
// Update_proximity of the initialisation block transfers their history
// (Code/Proximity.h:271-319), as the reference does on a restart.
void upload_loaded_contacts(mel_ctx& c) {
    const Model& m = c.model;
    const DevModel& d = c.d;
    std::vector<int> iv[18];
    std::vector<double> dv[20];
    for (size_t b = 0; b < m.bodies.size(); b++)
        for (const ContactElement& E : m.bodies[b].contacts) {
            // node ids are local to their body in the file, global on the device
            const int oS = c.h_node_off[b], oM = c.h_node_off.at(E.bodyM);
            const int ints[18] = {(int)b, E.borderS, E.border_nodeS, E.nodeS + oS, E.bodyM, E.borderM, E.shiftM,
                                  E.segmentM, E.nb_internal, E.bodyM == (int)b ? 2 : 1, E.nodeM[0] + oM,
                                  E.nodeM[1] + oM, E.nodeM[2] + oM, E.nodeM[3] + oM, E.shiftMk[0], E.shiftMk[1],
                                  E.shiftMk[2], E.shiftMk[3]};
            auto internal = [&](size_t k) { return k < E.internal.size() ? E.internal[k] : 0.; };
            const double dbl[20] = {E.shapeM[0], E.shapeM[1], E.shapeM[2], E.shapeM[3], E.gapn, E.vgapn, E.gapt,
                                    E.vgapt, E.xsi, E.xnorm, E.ynorm, E.xtan, E.ytan, E.effective_mass, E.length,
                                    E.fx, E.fy, internal(0), internal(1), internal(2)};
            for (int k = 0; k < 18; k++) iv[k].push_back(ints[k]);
            for (int k = 0; k < 20; k++) dv[k].push_back(dbl[k]);
        }
    const int n = (int)iv[0].size();
    if (n > d.c_cap)
        throw CapacityError{"the loaded state holds " + std::to_string(n) + " contact elements, above the capacity " +
                            std::to_string(d.c_cap)};
    if (n > 0) {
        const ContactArrays& A = c.bufA;
        int* ip[18] = {A.bodyS, A.borderS, A.bnodeS, A.nodeS, A.bodyM, A.borderM, A.shiftM, A.segM, A.nbint,
                       A.fresh, A.nM0, A.nM1, A.nM2, A.nM3, A.sM0, A.sM1, A.sM2, A.sM3};
        double* dp[20] = {A.N0, A.N1, A.N2, A.N3, A.gapn, A.vgapn, A.gapt, A.vgapt, A.xsi, A.xn, A.yn, A.xt, A.yt,
                          A.emass, A.len, A.fx, A.fy, A.int0, A.int1, A.int2};
        for (int k = 0; k < 18; k++)
            HIP_OK(hipMemcpyAsync(ip[k], iv[k].data(), n * sizeof(int), hipMemcpyHostToDevice, c.stream));
        for (int k = 0; k < 20; k++)
            HIP_OK(hipMemcpyAsync(dp[k], dv[k].data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream));
        // computed from the indices by the first step that reads them
        HIP_OK(hipMemsetAsync(A.info, 0, n * sizeof(int4), c.stream));
    }
    HIP_OK(hipMemcpyAsync(d.c_count, &n, sizeof(int), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
}

// flatten the parsed model into device arrays
void build(mel_ctx& c) {
    Model& m = c.model;
    DevModel& d = c.d;
    int NB = (int)m.bodies.size();
    d.n_bodies = NB;
    d.n_materials = (int)m.materials.size();
    d.n_laws = (int)m.laws.size();
    d.period = m.xmax_period - m.xmin_period;
    d.xmin_period = m.xmin_period;
    d.gravity_x = m.gravity_x;
    d.gravity_y = m.gravity_y;
    d.dt = m.dt;
    d.self_contact = m.flags[F_NO_MONITORING] ? 0 : 1;
    d.alid_on = 0;
    for (const Body& B : m.bodies)
        if (!B.rigid && B.alid.on) d.alid_on = 1;
    d.n_rules = (int)m.deactivated.size();
    d.law_f32 = c.precision == MEL_FP32_MIXED ? 1 : 0;
    for (auto& b : m.bodies)
        if (b.borders.empty()) throw CapacityError{"every body needs at least one border"};
    for (auto& b : m.bodies)
        if (m.flags[F_UPDATE_STIFFNESS] && !b.mat_dof.empty())
            throw CapacityError{"UPDATE_STIFFNESS_MATRIX is not supported"};

    std::vector<double> body_mass, body_inv_mass, body_inertia, body_inv_inertia;
    compute_masses(m, body_mass, body_inv_mass, body_inertia, body_inv_inertia);

    // bodies
    std::vector<int> rigid(NB), material(NB), periodicity(NB), status(NB), node_off(NB + 1, 0), nn(NB), gp_off(NB + 1, 0),
        ng(NB), nreg(NB), perreg(NB), border_off(NB + 1, 0), nbr(NB);
    std::vector<double> detect(NB), contact(NB);
    for (int b = 0; b < NB; b++) {
        Body& B = m.bodies[b];
        rigid[b] = B.rigid;
        material[b] = B.material;
        periodicity[b] = B.periodicity;
        status[b] = B.status;
        nn[b] = B.n_nodes;
        node_off[b + 1] = node_off[b] + B.n_nodes;
        ng[b] = B.rigid ? 0 : B.n_gauss;
        gp_off[b + 1] = gp_off[b] + ng[b];
        nreg[b] = std::max(1, B.n_regions);
        perreg[b] = ng[b] > 0 ? (int)std::ceil((double)ng[b] / (double)nreg[b]) : 1;
        nbr[b] = (int)B.borders.size();
        border_off[b + 1] = border_off[b] + nbr[b];
        detect[b] = B.detection_distance;
        contact[b] = B.contact_distance;
        if (B.rigid) c.h_rigid_ids.push_back(b);
    }
    int N = node_off[NB], G = gp_off[NB], NBR = border_off[NB];
    d.n_nodes = N;
    d.n_gauss = G;
    d.n_borders = NBR;
    d.n_rigid = (int)c.h_rigid_ids.size();
    c.h_node_off = node_off;
    c.h_gp_off = gp_off;
    c.h_border_off = border_off;

    // nodes
    std::vector<int> node_body(N);
    std::vector<double> X(N), Y(N), ux(N), uy(N), vx(N), vy(N), ax(N), ay(N), upx(N), upy(N), vpx(N), vpy(N), apx(N),
        apy(N), mx(N), my(N), imx(N), imy(N), xc(N), yc(N);
    for (int b = 0; b < NB; b++) {
        Body& B = m.bodies[b];
        for (int i = 0; i < B.n_nodes; i++) {
            int g = node_off[b] + i;
            node_body[g] = b;
            X[g] = B.X[i]; Y[g] = B.Y[i];
            ux[g] = B.ux[i]; uy[g] = B.uy[i]; vx[g] = B.vx[i]; vy[g] = B.vy[i]; ax[g] = B.ax[i]; ay[g] = B.ay[i];
            upx[g] = B.upx[i]; upy[g] = B.upy[i]; vpx[g] = B.vpx[i]; vpy[g] = B.vpy[i];
            apx[g] = B.apx[i]; apy[g] = B.apy[i];
            if (!B.rigid) {
                mx[g] = B.mx[i]; my[g] = B.my[i]; imx[g] = B.imx[i]; imy[g] = B.imy[i];
                xc[g] = B.X[i] + B.ux[i]; yc[g] = B.Y[i] + B.uy[i];
            }
        }
    }
    // Gauss points and the transposed gather
    std::vector<int> g_ptr(G + 1, 0), g_idx, g_body(G);
    std::vector<double> g_phix, g_phiy, g_c0(G);
    for (int b = 0; b < NB; b++) {
        Body& B = m.bodies[b];
        if (B.rigid) continue;
        for (int g = 0; g < B.n_gauss; g++) {
            int gg = gp_off[b] + g;
            g_body[gg] = b;
            g_c0[gg] = B.gw[g] * B.gjac[g];
            for (int k = B.gptr[g]; k < B.gptr[g + 1]; k++) {
                g_idx.push_back(node_off[b] + B.gidx[k]);
                g_phix.push_back(B.gphix[k]);
                g_phiy.push_back(B.gphiy[k]);
            }
            g_ptr[gg + 1] = (int)g_idx.size();
        }
    }
    c.nnz_g = (long long)g_idx.size();
    std::vector<int> t_ptr(N + 1, 0), t_gp(g_idx.size());
    std::vector<double> t_phix(g_idx.size()), t_phiy(g_idx.size());
    for (int j : g_idx) t_ptr[j + 1]++;
    for (int i = 0; i < N; i++) t_ptr[i + 1] += t_ptr[i];
    {
        std::vector<int> cur(t_ptr.begin(), t_ptr.end() - 1);
        for (int g = 0; g < G; g++)
            for (int k = g_ptr[g]; k < g_ptr[g + 1]; k++) {
                int q = cur[g_idx[k]]++;
                t_gp[q] = g;
                t_phix[q] = g_phix[k];
                t_phiy[q] = g_phiy[k];
            }
    }
    // DOF_TO_DISP
    std::vector<int> d_ptr(N + 1, 0), d_idx;
    std::vector<double> d_phi;
    for (int b = 0; b < NB; b++) {
        Body& B = m.bodies[b];
        for (int i = 0; i < B.n_nodes; i++) {
            int g = node_off[b] + i;
            if (!B.rigid)
                for (int k = B.dptr[i]; k < B.dptr[i + 1]; k++) {
                    d_idx.push_back(node_off[b] + B.didx[k]);
                    d_phi.push_back(B.dphi[k]);
                }
            d_ptr[g + 1] = (int)d_idx.size();
        }
    }
    c.nnz_d = (long long)d_idx.size();
    // damping matrices
    std::vector<int> m_dof;
    std::vector<double> m_damp, m_stiff;
    for (int b = 0; b < NB; b++) {
        Body& B = m.bodies[b];
        for (size_t i = 0; i < B.mat_damping.size(); i++) {
            m_dof.push_back(2 * node_off[b] + B.mat_dof[2 * i]);
            m_dof.push_back(2 * node_off[b] + B.mat_dof[2 * i + 1]);
            m_damp.push_back(B.mat_damping[i]);
            m_stiff.push_back(B.mat_stiffness[i]);
        }
    }
    int NM = (int)m_damp.size();
    std::vector<int> mt_ptr(N + 1, 0), mt_entry;
    {
        std::vector<std::vector<int>> lists(NM > 0 ? N : 0);
        for (int i = 0; i < NM; i++) {
            int dof0 = m_dof[2 * i], dof1 = m_dof[2 * i + 1];
            lists[dof0 / 2].push_back(2 * i);
            if (dof0 != dof1) lists[dof1 / 2].push_back(2 * i + 1);
        }
        for (int j = 0; j < (NM > 0 ? N : 0); j++) {
            mt_ptr[j + 1] = mt_ptr[j] + (int)lists[j].size();
            mt_entry.insert(mt_entry.end(), lists[j].begin(), lists[j].end());
        }
    }
    d.n_matrix = NM;
    // borders and slots
    std::vector<int> br_body(NBR), br_slot0(NBR), br_nslots(NBR), br_before(NBR), br_after(NBR), br_per(NBR),
        br_xt(NBR), br_yt(NBR), br_hist_ptr(2 * NBR + 1, 0);
    std::vector<double> br_xparam(2 * NBR), br_yparam(2 * NBR), hist_t, hist_v;
    std::vector<int> sl_border, sl_j, sl_node, sl_n0, sl_n1, sl_n2, sl_n3, sl_s0, sl_s1, sl_s2, sl_s3;
    std::vector<double> sl_len0;
    for (int b = 0; b < NB; b++) {
        Body& B = m.bodies[b];
        for (int r = 0; r < nbr[b]; r++) {
            Border& R = B.borders[r];
            int gb = border_off[b] + r;
            br_body[gb] = b;
            br_slot0[gb] = (int)sl_node.size();
            br_nslots[gb] = (int)R.nodes.size();
            br_before[gb] = R.before;
            br_after[gb] = R.after;
            br_per[gb] = R.periodicity;
            br_xt[gb] = R.xbc.type;
            br_yt[gb] = R.ybc.type;
            br_xparam[2 * gb] = R.xbc.stiffness; br_xparam[2 * gb + 1] = R.xbc.damping;
            br_yparam[2 * gb] = R.ybc.stiffness; br_yparam[2 * gb + 1] = R.ybc.damping;
            hist_t.insert(hist_t.end(), R.xbc.hist.t.begin(), R.xbc.hist.t.end());
            hist_v.insert(hist_v.end(), R.xbc.hist.v.begin(), R.xbc.hist.v.end());
            br_hist_ptr[2 * gb + 1] = (int)hist_t.size();
            hist_t.insert(hist_t.end(), R.ybc.hist.t.begin(), R.ybc.hist.t.end());
            hist_v.insert(hist_v.end(), R.ybc.hist.v.begin(), R.ybc.hist.v.end());
            br_hist_ptr[2 * gb + 2] = (int)hist_t.size();
            if ((int)R.nodes.size() < 3) throw CapacityError{"borders need at least 3 node positions"};
            for (size_t j = 0; j < R.nodes.size(); j++) {
                int o = node_off[b];
                sl_border.push_back(gb);
                sl_j.push_back((int)j);
                sl_node.push_back(o + R.nodes[j]);
                sl_n0.push_back(o + R.node0[j]); sl_n1.push_back(o + R.node1[j]);
                sl_n2.push_back(o + R.node2[j]); sl_n3.push_back(o + R.node3[j]);
                sl_s0.push_back(R.shift0[j]); sl_s1.push_back(R.shift1[j]);
                sl_s2.push_back(R.shift2[j]); sl_s3.push_back(R.shift3[j]);
                sl_len0.push_back(R.initial_length[j]);
            }
        }
    }
    int NS = (int)sl_node.size();
    d.n_slots = NS;
    c.h_br_slot0 = br_slot0;
    c.h_br_nslots = br_nslots;
    // distinct border nodes in first-appearance order
    std::vector<int> node_bn(N, -1), bn_node, sl_bn(NS);
    for (int s = 0; s < NS; s++) {
        int nd = sl_node[s];
        if (node_bn[nd] < 0) { node_bn[nd] = (int)bn_node.size(); bn_node.push_back(nd); }
        sl_bn[s] = node_bn[nd];
    }
    int NBN = (int)bn_node.size();
    d.n_bnodes = NBN;
    std::vector<int> bn_slot_ptr(NBN + 1, 0), bn_slots(NS);
    for (int s = 0; s < NS; s++) bn_slot_ptr[sl_bn[s] + 1]++;
    for (int i = 0; i < NBN; i++) bn_slot_ptr[i + 1] += bn_slot_ptr[i];
    {
        std::vector<int> cur(bn_slot_ptr.begin(), bn_slot_ptr.end() - 1);
        for (int s = 0; s < NS; s++) bn_slots[cur[sl_bn[s]]++] = s;
    }
    // transposed DOF_TO_DISP rows of border nodes: node j -> (bn, phi_j(x_bn))
    std::vector<int> bt_ptr(N + 1, 0), bt_bn;
    std::vector<double> bt_phi;
    {
        std::vector<int> cnt(N, 0);
        for (int q = 0; q < NBN; q++) {
            int i = bn_node[q];
            for (int k = d_ptr[i]; k < d_ptr[i + 1]; k++) cnt[d_idx[k]]++;
        }
        for (int j = 0; j < N; j++) bt_ptr[j + 1] = bt_ptr[j] + cnt[j];
        bt_bn.resize(bt_ptr[N]);
        bt_phi.resize(bt_ptr[N]);
        std::vector<int> cur(bt_ptr.begin(), bt_ptr.end() - 1);
        for (int q = 0; q < NBN; q++) {
            int i = bn_node[q];
            for (int k = d_ptr[i]; k < d_ptr[i + 1]; k++) {
                int p = cur[d_idx[k]]++;
                bt_bn[p] = q;
                bt_phi[p] = d_phi[k];
            }
        }
    }
    // materials and laws
    std::vector<int> mat_type(d.n_materials), law_type(d.n_laws), law_fixed(d.n_laws);
    std::vector<double> mat_p(6 * d.n_materials), law_p(12 * std::max(1, d.n_laws), 0.);
    for (int i = 0; i < d.n_materials; i++) {
        mat_type[i] = m.materials[i].type;
        for (int k = 0; k < 6; k++) mat_p[6 * i + k] = m.materials[i].p[k];
    }
    for (int i = 0; i < d.n_laws; i++) {
        law_type[i] = m.laws[i].type;
        if (m.laws[i].type == LAW_DAMPED_MOHR_COULOMB) c.work_law = true;
        law_fixed[i] = m.laws[i].fixed_length;
        for (size_t k = 0; k < m.laws[i].p.size() && k < 12; k++) law_p[12 * i + k] = m.laws[i].p[k];
    }
    // first law naming a material pair, in either order (materials indexed
    // by name once: the batched workloads carry a thousand materials)
    std::vector<int> law_first(d.n_materials * d.n_materials, -1);
    {
        std::map<std::string, std::vector<int>> by_name;
        for (int a = 0; a < d.n_materials; a++) by_name[m.materials[a].name].push_back(a);
        static const std::vector<int> none;
        auto named = [&](const std::string& n) -> const std::vector<int>& {
            auto it = by_name.find(n);
            return it == by_name.end() ? none : it->second;
        };
        for (int i = 0; i < d.n_laws; i++) {
            const ContactLaw& L = m.laws[i];
            for (int a : named(L.m1))
                for (int bm : named(L.m2)) {
                    int& x = law_first[a * d.n_materials + bm];
                    if (x < 0) x = i;
                    int& y = law_first[bm * d.n_materials + a];
                    if (y < 0) y = i;
                }
        }
    }
    // rigid / per-body state
    std::vector<double> body_rigid((size_t)NB * RIGID_FIELDS, 0.);
    for (int b = 0; b < NB; b++) {
        Body& B = m.bodies[b];
        double* R = &body_rigid[(size_t)b * RIGID_FIELDS];
        for (int q = 0; q < 7; q++) R[R_WORK + q] = B.works[q];
        if (!B.rigid) continue;
        for (int q = 0; q < 3; q++) { R[R_DISP + q] = B.rigid_disp[q]; R[R_VEL + q] = B.rigid_vel[q]; }
        R[R_XC] = B.xc;
        R[R_YC] = B.yc;
        // Update_current_positions at load (Code/IO.h:2050)
        R[R_CUR] = B.xc + R[R_DISP];
        R[R_CUR + 1] = B.yc + R[R_DISP + 1];
        R[R_CUR + 2] = 0. + R[R_DISP + 2];
    }
    // works of deformable bodies accumulate per node block (the loaded values
    // are put into the body's first block by build_blocks); the contact work
    // stays in the body slot
    for (int b = 0; b < NB; b++) {
        if (m.bodies[b].rigid) continue;
        for (int q = 0; q < 7; q++)
            if (q != W_CONTACT) body_rigid[(size_t)b * RIGID_FIELDS + R_WORK + q] = 0.;
    }

    // ---- upload ----
    d.body_rigid_flag = c.upload(rigid);
    d.body_material = c.upload(material);
    d.body_periodicity = c.upload(periodicity);
    d.body_status = c.upload(status);
    d.node_off = c.upload(node_off);
    d.body_nnodes = c.upload(nn);
    d.gp_off = c.upload(gp_off);
    d.body_ngauss = c.upload(ng);
    d.body_nregions = c.upload(nreg);
    d.body_per_region = c.upload(perreg);
    d.border_off = c.upload(border_off);
    d.body_nborders = c.upload(nbr);
    d.body_detect = c.upload(detect);
    d.body_contact = c.upload(contact);
    // masses are uploaded in two stages by the init sequence
    std::vector<double> mass0(NB, 0.), inv0(NB, 0.), inertia0(NB, 0.), invi0(NB, 0.);
    for (int b = 0; b < NB; b++)
        if (m.bodies[b].rigid) {
            mass0[b] = m.bodies[b].mass; inv0[b] = m.bodies[b].inverse_mass;
            inertia0[b] = m.bodies[b].inertia; invi0[b] = m.bodies[b].inverse_inertia;
        }
    d.body_mass = c.upload(mass0);
    d.body_inv_mass = c.upload(inv0);
    d.body_inertia = c.upload(inertia0);
    d.body_inv_inertia = c.upload(invi0);
    d.body_rigid = c.upload(body_rigid);
    d.mat_type = c.upload(mat_type);
    d.mat_p = c.upload(mat_p);
    d.law_table = c.upload(m.contacts_table);
    d.law_first = c.upload(law_first);
    d.law_type = c.upload(law_type);
    d.law_fixed = c.upload(law_fixed);
    d.law_p = c.upload(law_p);
    d.node_body = c.upload(node_body);
    d.X = c.upload(X); d.Y = c.upload(Y);
    d.xc = c.upload(xc); d.yc = c.upload(yc);
    d.ux = c.upload(ux); d.uy = c.upload(uy); d.vx = c.upload(vx); d.vy = c.upload(vy);
    d.ax = c.upload(ax); d.ay = c.upload(ay);
    d.upx = c.upload(upx); d.upy = c.upload(upy); d.vpx = c.upload(vpx); d.vpy = c.upload(vpy);
    d.apx = c.upload(apx); d.apy = c.upload(apy);
    d.mx = c.upload(mx); d.my = c.upload(my); d.imx = c.upload(imx); d.imy = c.upload(imy);
    double** forces[] = {&d.fix, &d.fiy, &d.fcx, &d.fcy, &d.fbx, &d.fby, &d.fdx, &d.fdy, &d.fnx, &d.fny, &d.fkx,
                         &d.fky, &d.fx, &d.fy};
    for (double** p : forces) *p = c.alloc<double>(N);
    d.fax = c.alloc<double>(N);
    d.fay = c.alloc<double>(N);
    c.d.node_body_host = node_body.data();
    build_blocks(c, g_ptr, g_idx, g_phix, g_phiy, g_c0, t_ptr, t_gp, bn_node, node_off, gp_off, perreg, d_ptr, d_idx,
                 bt_ptr, bt_bn);
    {
        // what the finish of a node block reads first, per listed border
        // node and per slot (static model data gathered so that one load
        // replaces a chain of index loads)
        const std::vector<int>& bl = c.h_bn_list;
        std::vector<int4> fr0(std::max<size_t>(bl.size(), 1)), fr1(std::max<size_t>(bl.size(), 1));
        std::vector<double2> frX(std::max<size_t>(bl.size(), 1));
        for (size_t i = 0; i < bl.size(); i++) {
            const int q = bl[i], node = bn_node[q];
            const int a0 = bn_slot_ptr[q], a1 = bn_slot_ptr[q + 1];
            fr0[i] = make_int4(q, node, d_ptr[node], d_ptr[node + 1]);
            fr1[i] = make_int4(a0, a1, a1 > a0 ? bn_slots[a0] : 0, a1 > a0 + 1 ? bn_slots[a0 + 1] : 0);
            frX[i] = make_double2(X[node], Y[node]);
        }
        std::vector<int4> sr0(std::max(NS, 1)), sr1(std::max(NS, 1));
        for (int s2 = 0; s2 < NS; s2++) {
            const int br = sl_border[s2];
            sr0[s2] = make_int4(sl_n0[s2], sl_n1[s2], sl_n2[s2], sl_j[s2]);
            sr1[s2] = make_int4(sl_s0[s2], sl_s1[s2], sl_s2[s2], br_nslots[br] | (br_per[br] == 1 ? 1 << 30 : 0));
        }
        BlockArrays& B = c.blocks_dev;
        B.fin_rec0 = c.upload(fr0); B.fin_rec1 = c.upload(fr1); B.fin_recX = c.upload(frX);
        B.sl_rec0 = c.upload(sr0); B.sl_rec1 = c.upload(sr1);
    }
    {
        // deformable border nodes with a slot on a border that carries a BC
        std::vector<int> bc_bn;
        for (int q = 0; q < NBN; q++) {
            if (m.bodies[node_body[bn_node[q]]].rigid) continue;
            bool any = false;
            for (int k = bn_slot_ptr[q]; k < bn_slot_ptr[q + 1]; k++) {
                int br = sl_border[bn_slots[k]];
                any |= br_xt[br] != 0 || br_yt[br] != 0;
            }
            if (any) bc_bn.push_back(q);
        }
        c.blocks_dev.bc_bn = c.upload(bc_bn);
        c.blocks_dev.n_bc_bn = (int)bc_bn.size();
    }
    c.d.node_body_host = nullptr;
    d.d_ptr = c.upload(d_ptr);
    d.d_idx = c.upload(d_idx);
    d.d_phi = c.upload(d_phi);
    d.m_dof = c.upload(m_dof);
    d.m_damp = c.upload(m_damp);
    d.mt_ptr = c.upload(mt_ptr);
    d.mt_entry = c.upload(mt_entry);
    d.br_body = c.upload(br_body);
    d.br_slot0 = c.upload(br_slot0);
    d.br_nslots = c.upload(br_nslots);
    d.br_before = c.upload(br_before);
    d.br_after = c.upload(br_after);
    d.br_periodicity = c.upload(br_per);
    d.br_xtype = c.upload(br_xt);
    d.br_ytype = c.upload(br_yt);
    {
        // borders of rigid bodies whose BCs give forces: their terms are
        // formed by the force launch, run by run
        auto force = [](int t) { return t == 2 || t == 3 || t == 4; };
        std::vector<int> rbc, br_rbc(std::max(NBR, 1), -1);
        for (int br = 0; br < NBR; br++)
            if (rigid[br_body[br]] && (force(br_xt[br]) || force(br_yt[br]))) {
                br_rbc[br] = (int)rbc.size();
                rbc.push_back(br);
            }
        d.n_rbc = (int)rbc.size();
        d.rbc_border = c.upload(rbc.empty() ? std::vector<int>{0} : rbc);
        d.br_rbc = c.upload(br_rbc);
        d.rbc_part = c.alloc<double>((size_t)std::max(1, d.n_rbc) * BLOCK_THREADS * 6);
    }
    {
        // Update_bc changes the values of types 2, 3 and 4 only (type 0 stays
        // zero, type 1 drives the DOFs directly)
        std::vector<int> timed;
        auto moves = [](int t) { return t == 2 || t == 3 || t == 4; };
        for (size_t br = 0; br < br_xt.size(); br++)
            if (moves(br_xt[br]) || moves(br_yt[br])) timed.push_back((int)br);
        d.n_bc_timed = (int)timed.size();
        d.bc_timed = c.upload(timed);
    }
    d.br_xparam = c.upload(br_xparam);
    d.br_yparam = c.upload(br_yparam);
    {
        // absorbing layers: per-node coefficients, the bodies' velocity
        // histories after the borders' ones, the matrix stiffness, the nodes
        std::vector<int> body_alid(4 * (size_t)NB, 0), alid_nodes;
        std::vector<double> aa, ab;
        if (d.alid_on) { aa.assign(N, 0.); ab.assign(N, 0.); }
        for (int b = 0; b < NB; b++) {
            const Body& B = m.bodies[b];
            if (B.rigid || !B.alid.on) continue;
            const History* hs[2] = {&B.alid.xh, &B.alid.yh};
            for (int k = 0; k < 2; k++) {
                body_alid[4 * b + 2 * k] = (int)hist_t.size();
                hist_t.insert(hist_t.end(), hs[k]->t.begin(), hs[k]->t.end());
                hist_v.insert(hist_v.end(), hs[k]->v.begin(), hs[k]->v.end());
                body_alid[4 * b + 2 * k + 1] = (int)hist_t.size();
            }
            for (int i = 0; i < B.n_nodes; i++) {
                aa[node_off[b] + i] = B.alid_alpha[i];
                ab[node_off[b] + i] = B.alid_beta[i];
                alid_nodes.push_back(node_off[b] + i);
            }
        }
        d.body_alid = c.upload(body_alid);
        d.alid_a = c.upload(aa);
        d.alid_b = c.upload(ab);
        d.alid_nodes = c.upload(alid_nodes);
        d.n_alid_nodes = (int)alid_nodes.size();
        d.m_stiff = c.upload(d.alid_on ? m_stiff : std::vector<double>());
    }
    d.br_hist_ptr = c.upload(br_hist_ptr);
    d.hist_t = c.upload(hist_t);
    d.hist_v = c.upload(hist_v);
    d.br_xval = c.alloc<double>(NBR); d.br_yval = c.alloc<double>(NBR);
    d.br_xvel = c.alloc<double>(NBR); d.br_yvel = c.alloc<double>(NBR);
    d.sl_border = c.upload(sl_border); d.sl_j = c.upload(sl_j); d.sl_node = c.upload(sl_node);
    d.sl_n0 = c.upload(sl_n0); d.sl_n1 = c.upload(sl_n1); d.sl_n2 = c.upload(sl_n2); d.sl_n3 = c.upload(sl_n3);
    d.sl_s0 = c.upload(sl_s0); d.sl_s1 = c.upload(sl_s1); d.sl_s2 = c.upload(sl_s2); d.sl_s3 = c.upload(sl_s3);
    d.sl_bn = c.upload(sl_bn);
    d.sl_len = c.upload(sl_len0);
    d.sl_len0 = c.upload(sl_len0);
    d.sl_xcp = c.alloc<double>(NS);
    d.sl_ycp = c.alloc<double>(NS);
    d.sl_xbp = c.alloc<double>(NS);
    d.sl_ybp = c.alloc<double>(NS);
    c.d_energy = c.alloc<double>(3 * (size_t)std::max(1, NB));
    d.bn_node = c.upload(bn_node);
    d.bn_slot_ptr = c.upload(bn_slot_ptr);
    d.bn_slots = c.upload(bn_slots);
    double** bnp[] = {&d.bn_pcx, &d.bn_pcy, &d.bn_pdx, &d.bn_pdy, &d.bn_pnx, &d.bn_pny};
    for (double** p : bnp) *p = c.alloc<double>(NBN);
    d.bt_ptr = c.upload(bt_ptr);
    d.bt_bn = c.upload(bt_bn);
    d.bt_phi = c.upload(bt_phi);
    d.bk_count = c.alloc<int>(NBN);
    d.bk_cursor = c.alloc<int>(NBN);
    d.bk_rec = c.alloc<int>((size_t)NBN * BUCKET_CAP);
    d.bk_val = c.alloc<double2>((size_t)NBN * BUCKET_CAP);
    // contacts
    d.c_cap = 4 * NS + 1024;
    d.c_count = c.alloc<int>(1);
    d.c_count_old = c.alloc<int>(1);
    alloc_contacts(c, c.bufA, d.c_cap);
    alloc_contacts(c, c.bufB, d.c_cap);
    d.C = c.bufA;
    d.Cold = c.bufB;
    upload_loaded_contacts(c);
    // deactivation rules: the ghost records of inactive bodies take bucket
    // record ids past the elements' (at most two master records per
    // element of the bodies they stop)
    d.ghost_cap = d.n_rules > 0 ? 2 * d.c_cap : 0;
    d.rec_bn = c.alloc<int>(3 * (size_t)d.c_cap + d.ghost_cap);
    HIP_OK(hipMemsetAsync(d.rec_bn, 0xff, (3 * (size_t)d.c_cap + d.ghost_cap) * sizeof(int), c.stream));
    d.rec_fx = c.alloc<double>(3 * (size_t)d.c_cap + d.ghost_cap);
    d.rec_fy = c.alloc<double>(3 * (size_t)d.c_cap + d.ghost_cap);
    d.rig_rec = c.alloc<double>(6 * (size_t)d.c_cap);
    if (d.n_rules > 0 || !m.monitored.empty()) {
        // Body::damage, formed by Update_status and Monitoring (Body::Update_damage)
        d.body_damage = c.alloc<double>(NB);
        d.body_flip = c.alloc<int>(NB);
        d.slot_damage = c.alloc<double>(NS);
    }
    if (d.n_rules > 0) {
        std::vector<double> rules((size_t)RULE_STRIDE * d.n_rules, 0.);
        for (int i = 0; i < d.n_rules; i++)
            for (size_t k = 0; k < m.deactivated[i].size() && k < (size_t)RULE_STRIDE; k++)
                rules[(size_t)RULE_STRIDE * i + k] = m.deactivated[i][k];
        d.rules = c.upload(rules);
        d.gh_count = c.alloc<int>(2);
        d.gh_bn = c.alloc<int>(d.ghost_cap);
        d.gh_fx = c.alloc<double>(d.ghost_cap);
        d.gh_fy = c.alloc<double>(d.ghost_cap);
        d.gh_rb = c.alloc<int>(d.ghost_cap);
        d.gh_rf = c.alloc<double>(3 * (size_t)d.ghost_cap);
        d.bn_acx = c.alloc<double>(d.n_bnodes);
        d.bn_acy = c.alloc<double>(d.n_bnodes);
    }
    // proximity scratch
    d.pb_xmin = c.alloc<double>(NBR); d.pb_xmax = c.alloc<double>(NBR);
    d.pb_ymin = c.alloc<double>(NBR); d.pb_ymax = c.alloc<double>(NBR);
    d.pb_shmin = c.alloc<int>(NBR); d.pb_shmax = c.alloc<int>(NBR);
    {
        std::vector<int> chunk0(NBR + 1, 0), sc_border;
        for (int br = 0; br < NBR; br++) {
            int nch = (br_nslots[br] + SEG_CHUNK - 1) / SEG_CHUNK;
            chunk0[br + 1] = chunk0[br] + nch;
            for (int q = 0; q < nch; q++) sc_border.push_back(br);
        }
        d.n_seg_chunks = chunk0[NBR];
        d.br_chunk0 = c.upload(chunk0);
        d.sc_border = c.upload(sc_border);
        size_t nc = (size_t)d.n_seg_chunks;
        d.sc_nxmin = c.alloc<double>(nc); d.sc_nxmax = c.alloc<double>(nc);
        d.sc_nymin = c.alloc<double>(nc); d.sc_nymax = c.alloc<double>(nc);
        d.sc_box = c.alloc<double4>(nc);
        d.sc_box2 = c.alloc<double4>((nc + SC_GROUP - 1) / SC_GROUP);
    }
    d.seg_ns = c.alloc<int4>(NS);
    d.seg_xy = c.alloc<double4>(NS);
    // cell list of the broad phase: cells 1.25 x the largest border AABB
    // (grown by the detection distance) among the borders up to 8 x the
    // median size; larger borders (walls) are listed apart
    {
        std::vector<double> ext;
        double ylo = 1e300, yhi = -1e300;
        for (const Body& B : m.bodies) {
            for (const Border& R : B.borders) {
                double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
                for (int n : R.nodes) {
                    double x = B.X[n], y = B.Y[n];
                    if (!B.rigid && !B.ux.empty()) { x += B.ux[n]; y += B.uy[n]; }
                    if (B.rigid) { x += B.rigid_disp[0]; y += B.rigid_disp[1]; }
                    x0 = std::min(x0, x); x1 = std::max(x1, x); y0 = std::min(y0, y); y1 = std::max(y1, y);
                }
                if (R.nodes.empty()) continue;
                const double dd = B.detection_distance;
                ext.push_back(std::max(x1 - x0, y1 - y0) + 2 * dd);
                ylo = std::min(ylo, y0 - dd);
                yhi = std::max(yhi, y1 + dd);
            }
        }
        for (const Body& B : m.bodies)
            if (B.borders.size() > 255)
                throw CapacityError{"a body with more than 255 borders is not supported"};
        double cell = 1.0;
        if (!ext.empty()) {
            std::vector<double> e = ext;
            std::nth_element(e.begin(), e.begin() + e.size() / 2, e.end());
            const double med = e[e.size() / 2];
            cell = 0.;
            for (double x : ext)
                if (x <= 8 * med) cell = std::max(cell, x);
            cell = cell > 0. ? 1.25 * cell : 1.0;
        } else {
            ylo = 0.; yhi = 1.;
        }
        const double P = d.period;
        d.cell_nx = (int)std::max(1.0, std::min(4096.0, std::floor(P / cell)));
        d.cell_dx = P / d.cell_nx;
        d.cell_dy = cell;
        const double margin = 8 * cell;
        d.cell_y0 = ylo - margin;
        double rows = std::ceil((yhi - ylo + 2 * margin) / d.cell_dy);
        const double max_cells = 1 << 22;
        if (rows * d.cell_nx > max_cells) {
            d.cell_dy *= rows * d.cell_nx / max_cells;
            rows = std::ceil((yhi - ylo + 2 * margin) / d.cell_dy);
        }
        d.cell_ny = (int)std::max(1.0, rows);
        d.cell_n = d.cell_nx * d.cell_ny;
        d.cell_count = c.alloc<int>(d.cell_n);
        d.cell_fill = c.alloc<int>(d.cell_n);
        d.cell_start = c.alloc<int>(d.cell_n);
        d.cell_bsum = c.alloc<int>((d.cell_n + 1023) / 1024 + 1);
        d.cell_boff = c.alloc<int>((d.cell_n + 1023) / 1024 + 1);
        d.cell_items = c.alloc<int>(NBR);
        d.border_cell = c.alloc<int>(NBR);
        d.big_cap = std::max(64, NBR);
        d.big_count = c.alloc<int>(1);
        d.big_list = c.alloc<int>(d.big_cap);
    }
    d.nb_cap = 256;
    d.nb_count = c.alloc<int>(NB);
    d.nb_list = c.alloc<int>((size_t)NB * d.nb_cap * 3);
    d.nb_gm = c.alloc<int>((size_t)NB * d.nb_cap);
    {
        // narrow-phase row chunks: at most nb_cap neighbours per body
        long long ch = 0;
        for (int b = 0; b < NB; b++) {
            long long ns = 0;
            for (auto& br : m.bodies[b].borders) ns += (long long)br.nodes.size();
            long long rows = (long long)d.nb_cap * ns + ns * (long long)m.bodies[b].borders.size();
            ch += (rows + ROW_CHUNK - 1) / ROW_CHUNK;
        }
        if (ch > INT32_MAX / 2) throw CapacityError{"narrow-phase row chunks exceed the index range"};
        c.max_row_chunks = (int)ch;
    }
    d.body_chunks = c.alloc<int>(NB);
    d.chunk_off = c.alloc<int>(NB + 1);
    d.chunk_cnt = c.alloc<int>(c.max_row_chunks);
    d.chunk_body = c.alloc<int>(c.max_row_chunks);
    d.body_rowrec = c.alloc<int4>(NB);
    d.body_gs0 = c.alloc<int2>(NB);
    d.slot_xy = c.alloc<double2>(NS);
    d.chunk_elem_off = c.alloc<int>((size_t)c.max_row_chunks + 1);
    d.hit_rec = c.alloc<int4>(d.c_cap);
    d.hit_cnt = c.alloc<int>(d.c_cap);
    d.pb_segbox = c.alloc<double4>(NBR);
    d.hit_n = c.alloc<int>(1);
    d.slot_old_list = c.alloc<int>((size_t)NS * 16);
    d.slot_old_cursor = c.alloc<int>(NS);
    d.status_flags = c.alloc<int>(4);
    d.step_sync = c.alloc<unsigned>(4);
    d.rigid_dv = c.alloc<double>(8 * (size_t)std::max(1, d.n_rigid));
    d.time = c.alloc<double>(1);
    HIP_OK(hipMemcpyAsync(d.time, &m.time, sizeof(double), hipMemcpyHostToDevice, c.stream));
    c.d_rigid_ids = c.upload(c.h_rigid_ids);
    d.rigid_ids = c.d_rigid_ids;
    {
        std::vector<int> rindex(NB, -1);
        for (size_t r = 0; r < c.h_rigid_ids.size(); r++) rindex[c.h_rigid_ids[r]] = (int)r;
        d.body_rindex = c.upload(rindex);
        d.rm_cnt = c.alloc<int>(std::max(1, d.n_rigid));
        d.rm_list = c.alloc<int>((size_t)std::max(1, d.n_rigid) * RIGID_LIST_CAP);
        d.rs_range = c.alloc<int>(2 * (size_t)std::max(1, d.n_rigid));
    }
    {
        std::vector<int> rn;
        for (int b : c.h_rigid_ids)
            for (int i = 0; i < m.bodies[b].n_nodes; i++) rn.push_back(node_off[b] + i);
        d.n_rigid_nodes = (int)rn.size();
        d.rigid_node_ids = c.upload(rn);
    }
    HIP_OK(hipStreamSynchronize(c.stream));

    c.time = m.time;
    c.iteration = m.number_iteration;
    c.next_contact_update = m.next_contact_update;

    // rigid node positions, Update_current_positions at load (Code/IO.h:2050)
    if (d.n_rigid_nodes > 0) k_rigid_nodes<<<blocks(d.n_rigid_nodes), BLOCK, 0, c.stream>>>(d, -1);
    // ---- initialisation block, Code/main.cpp:103-126 ----
    k_border_length<<<blocks(NS), BLOCK, 0, c.stream>>>(d);          // Update_borders
    k_bc_eval<<<blocks(NBR), BLOCK, 0, c.stream>>>(d, 0);            // Update_bc(Time)
    update_proximity(c);
    if (m.flags[F_INITIALIZE_CZM]) initialize_czm(c);
    // Update_material: masses of deformable bodies, rigid mass and inertia
    {
        std::vector<double> mx2(N), my2(N), imx2(N), imy2(N);
        for (int b = 0; b < NB; b++) {
            Body& B = m.bodies[b];
            for (int i = 0; i < B.n_nodes; i++) {
                int g = node_off[b] + i;
                if (!B.rigid) { mx2[g] = B.mx[i]; my2[g] = B.my[i]; imx2[g] = B.imx[i]; imy2[g] = B.imy[i]; }
            }
        }
        HIP_OK(hipMemcpyAsync(d.mx, mx2.data(), N * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(d.my, my2.data(), N * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(d.imx, imx2.data(), N * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(d.imy, imy2.data(), N * sizeof(double), hipMemcpyHostToDevice, c.stream));
        std::vector<double> bm(NB), bim(NB), bi(NB), bii(NB);
        for (int b = 0; b < NB; b++) {
            bool upd = m.bodies[b].rigid || m.flags[F_UPDATE_MASS];
            bm[b] = upd ? body_mass[b] : mass0[b];
            bim[b] = upd ? body_inv_mass[b] : inv0[b];
            bi[b] = upd ? body_inertia[b] : inertia0[b];
            bii[b] = upd ? body_inv_inertia[b] : invi0[b];
        }
        HIP_OK(hipMemcpyAsync(d.body_mass, bm.data(), NB * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(d.body_inv_mass, bim.data(), NB * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(d.body_inertia, bi.data(), NB * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(d.body_inv_inertia, bii.data(), NB * sizeof(double), hipMemcpyHostToDevice, c.stream));
    }
    update_proximity(c);
    if (m.flags[F_INITIALIZE_CZM]) initialize_czm(c);
    HIP_OK(hipStreamSynchronize(c.stream));
    poll_async(c, true);
    contact_count(c);
    if (m.scheme != SCH_EULER) setup_adaptive(c);
}

thread_local std::string g_load_err;

mel_status create(mel_ctx** out, int device, int precision, const std::function<bool(Model&, std::string&)>& loader,
                  const mel_options* opts = nullptr) {
    if (!out) return MEL_E_ARG;
    *out = nullptr;
    if (precision != MEL_FP64 && precision != MEL_FP32_MIXED) { g_load_err = "unknown precision"; return MEL_E_ARG; }
    if (opts && (opts->block_nodes > BLOCK_MAX_OWN || opts->block_nodes == 0)) {
        g_load_err = "block_nodes must be 1.." + std::to_string(BLOCK_MAX_OWN) + " or -1";
        return MEL_E_ARG;
    }
    std::unique_ptr<mel_ctx> c(new mel_ctx());
    if (opts) c->opts = *opts;
    c->device = device;
    c->precision = precision;
    std::string err;
    if (!loader(c->model, err)) {
        g_load_err = err;
        return err.rfind("cannot open", 0) == 0 ? MEL_E_IO : MEL_E_PARSE;
    }
    try {
        HIP_OK(hipSetDevice(device));
        try {
            scratch_setup(device);
        } catch (std::runtime_error& e) {
            throw HipError{e.what()};
        }
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&c->ev_async, hipEventDisableTiming));
        HIP_OK(hipHostMalloc((void**)&c->h_async, 8 * sizeof(int), hipHostMallocDefault));
        build(*c);
    } catch (HipError& e) {
        g_load_err = e.msg;
        return e.msg.find("hipMalloc") != std::string::npos ? MEL_E_OOM : MEL_E_HIP;
    } catch (CapacityError& e) {
        g_load_err = e.msg;
        return e.msg.find("capacity") != std::string::npos ? MEL_E_CAPACITY : MEL_E_UNSUPPORTED;
    } catch (NanError& e) {
        g_load_err = e.msg;
        return MEL_E_NAN;
    } catch (TimeoutError& e) {
        g_load_err = e.msg;
        return MEL_E_TIMEOUT;
    }
    *out = c.release();
    return MEL_OK;
}

template <typename F>
mel_status guarded(mel_ctx* c, F&& f) {
    if (!c) return MEL_E_ARG;
    try {
        HIP_OK(hipSetDevice(c->device));
        return f();
    } catch (HipError& e) {
        c->err = e.msg;
        return MEL_E_HIP;
    } catch (CapacityError& e) {
        c->err = e.msg;
        return MEL_E_CAPACITY;
    } catch (NanError& e) {
        c->err = e.msg;
        return MEL_E_NAN;
    } catch (TimeoutError& e) {
        c->err = e.msg;
        return MEL_E_TIMEOUT;
    }
}

// the 7 works of every body (Code/Body.h:112-118): rigid bodies and the
// contact work from the body records, the other works of deformable bodies
// as compensated (Neumaier) sums of their node blocks' accumulators
std::vector<double> all_works(mel_ctx& c) {
    const int NB = c.d.n_bodies;
    flush_contact_work(c);
    std::vector<double> R, w, out(7 * (size_t)NB);
    c.download(R, c.d.body_rigid, (size_t)NB * RIGID_FIELDS);
    if (c.blocks_dev.n_blocks > 0) c.download(w, c.blocks_dev.blk_work, 14 * (size_t)c.blocks_dev.n_blocks);
    for (int b = 0; b < NB; b++)
        for (int q = 0; q < 7; q++) out[7 * b + q] = R[(size_t)b * RIGID_FIELDS + R_WORK + q];
    std::vector<double> sum(7 * (size_t)NB, 0.), comp(7 * (size_t)NB, 0.);
    std::vector<char> has(NB, 0);
    for (size_t k = 0; k < c.h_blk_body.size(); k++) {
        const int b = c.h_blk_body[k];
        has[b] = 1;
        for (int q = 0; q < 7; q++) {
            if (q == W_CONTACT) continue;
            for (int h = 0; h < 2; h++) {
                double& S = sum[7 * b + q];
                double x = w[14 * k + 2 * q + h], t = S + x;
                comp[7 * b + q] += std::fabs(S) >= std::fabs(x) ? (S - t) + x : (x - t) + S;
                S = t;
            }
        }
    }
    for (int b = 0; b < NB; b++)
        if (has[b] && !c.model.bodies[b].rigid)
            for (int q = 0; q < 7; q++)
                if (q != W_CONTACT) out[7 * b + q] = sum[7 * b + q] + comp[7 * b + q];
    return out;
}

// per-body energy scalars (k_body_energy), 3 per body
std::vector<double> all_energy(mel_ctx& c) {
    refresh_outputs(c);
    k_body_energy<<<c.d.n_bodies, BLOCK, 0, c.stream>>>(c.d, c.d_energy);
    HIP_OK(hipGetLastError());
    std::vector<double> e;
    c.download(e, c.d_energy, 3 * (size_t)c.d.n_bodies);
    return e;
}

// the current contact elements copied to `dst` (the first nc of each array)
void copy_contacts(mel_ctx& c, const ContactArrays& dst, const ContactArrays& src, int nc) {
    if (nc <= 0) return;
    int* const* di[] = {&dst.bodyS, &dst.borderS, &dst.bnodeS, &dst.nodeS, &dst.bodyM, &dst.borderM, &dst.shiftM,
                        &dst.segM, &dst.nbint, &dst.fresh, &dst.nM0, &dst.nM1, &dst.nM2, &dst.nM3, &dst.sM0,
                        &dst.sM1, &dst.sM2, &dst.sM3};
    int* const* si[] = {&src.bodyS, &src.borderS, &src.bnodeS, &src.nodeS, &src.bodyM, &src.borderM, &src.shiftM,
                        &src.segM, &src.nbint, &src.fresh, &src.nM0, &src.nM1, &src.nM2, &src.nM3, &src.sM0,
                        &src.sM1, &src.sM2, &src.sM3};
    for (size_t k = 0; k < sizeof(di) / sizeof(di[0]); k++)
        HIP_OK(hipMemcpyAsync(*di[k], *si[k], nc * sizeof(int), hipMemcpyDeviceToDevice, c.stream));
    double* const* dd[] = {&dst.N0, &dst.N1, &dst.N2, &dst.N3, &dst.gapn, &dst.vgapn, &dst.gapt, &dst.vgapt,
                           &dst.xsi, &dst.xn, &dst.yn, &dst.xt, &dst.yt, &dst.emass, &dst.len, &dst.fx, &dst.fy,
                           &dst.int0, &dst.wacc, &dst.int1, &dst.int2};
    double* const* sd[] = {&src.N0, &src.N1, &src.N2, &src.N3, &src.gapn, &src.vgapn, &src.gapt, &src.vgapt,
                           &src.xsi, &src.xn, &src.yn, &src.xt, &src.yt, &src.emass, &src.len, &src.fx, &src.fy,
                           &src.int0, &src.wacc, &src.int1, &src.int2};
    for (size_t k = 0; k < sizeof(dd) / sizeof(dd[0]); k++)
        HIP_OK(hipMemcpyAsync(*dd[k], *sd[k], nc * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dst.info, src.info, nc * sizeof(int4), hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dst.sinfo, src.sinfo, nc * sizeof(int4), hipMemcpyDeviceToDevice, c.stream));
}

// One force evaluation at the current state, without integration
// (mel_eval_forces).  The contact elements are re-projected and their laws
// applied on a copy of their state, which is put back afterwards, so the
// evaluation leaves the step's state as it found it.
// save_buf: the device scratch (no allocation), or nullptr to save the
// contact elements into a buffer the context allocates on first use
void eval_forces(mel_ctx& c, unsigned what, unsigned char* save_buf = nullptr) {
    DevModel& d = c.d;
    hipStream_t s = c.stream;
    const bool internal = what & MEL_EVAL_INTERNAL, contact = what & MEL_EVAL_CONTACT,
               external = what & MEL_EVAL_EXTERNAL;
    refresh_outputs(c);
    const int nc = contact ? contact_count(c) : 0;
    if (nc > 0 && save_buf) {
        if (nc > SCRATCH_MAX_CONTACTS)
            throw CapacityError{"mel_eval_forces_noalloc: " + std::to_string(nc) + " contact elements exceed the " +
                                std::to_string(SCRATCH_MAX_CONTACTS) + " its device scratch holds"};
        k_contacts_scratch<<<blocks(nc), BLOCK, 0, s>>>(d.C, save_buf, nc, 1);
    } else if (nc > 0) {
        if (c.eval_save_cap < d.c_cap) {
            alloc_contacts(c, c.eval_save, d.c_cap);
            c.eval_save_cap = d.c_cap;
        }
        copy_contacts(c, c.eval_save, d.C, nc);
    }
    const BlockArrays& B = c.blocks_dev;
    const int n_items = nc + (external ? B.n_bc_bn : 0);
    if (n_items > 0) k_eval_items<<<blocks(n_items), BLOCK, 0, s>>>(d, B.bc_bn, B.n_bc_bn, contact, external);
    if (d.n_bnodes > 0) {
        k_bucket_sums<<<blocks(d.n_bnodes), BLOCK, 0, s>>>(d);
        if (!external) k_clear_point_forces<<<blocks(d.n_bnodes), BLOCK, 0, s>>>(d, 0, 1);
    }
    if (external && d.n_matrix > 0) k_damping<<<blocks(d.n_nodes), BLOCK, 0, s>>>(d);
    if (external && d.n_alid_nodes > 0) k_alid<<<blocks(d.n_alid_nodes), BLOCK, 0, s>>>(d, c.time);
    if (!external || d.n_matrix == 0) k_clear_node_forces<<<blocks(d.n_nodes), BLOCK, 0, s>>>(d, 0, 1);
    if (internal) launch_internal_forces(c, s);
    else k_clear_node_forces<<<blocks(d.n_nodes), BLOCK, 0, s>>>(d, 1, 0);
    k_node_outputs<<<blocks(d.n_nodes), BLOCK, 0, s>>>(d, external ? 1 : 0);
    HIP_OK(hipGetLastError());
    if (nc > 0 && save_buf) k_contacts_scratch<<<blocks(nc), BLOCK, 0, s>>>(d.C, save_buf, nc, 0);
    else if (nc > 0) copy_contacts(c, d.C, c.eval_save, nc);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
    c.outputs_stale = false;
}

// ---------------------------------------------------------------------------
// Solver_step (Code/Solver.h:98-486)
// ---------------------------------------------------------------------------
// Body::Store (dir 0) or Body::Restore (dir 1) of every body: the copy table
// of the current contact buffer, then the rigid records' restored fields
void store_restore(mel_ctx& c, int dir) {
    const int k = c.d.C.bodyS == c.copy_key[0] ? 0 : 1;
    hipStream_t s = c.stream;
    k_ad_copy<<<dim3(64, c.n_copy), BLOCK, 0, s>>>(c.d_copy[k][dir], c.n_copy, c.d.c_count);
    const int nr = 16 * c.d.n_bodies;
    if (dir == 0) k_ad_rigid_copy<<<blocks(nr), BLOCK, 0, s>>>(c.d.body_rigid, c.snap_rigid, c.d.n_bodies);
    else k_ad_rigid_copy<<<blocks(nr), BLOCK, 0, s>>>(c.snap_rigid, c.d.body_rigid, c.d.n_bodies);
    HIP_OK(hipGetLastError());
}

// one half of an attempt of step dt from time t (Code/Solver.h:155-214,
// 218-279): Update_bc(t + dt/2) first in the second half; contacts, BC
// forces and damping with the laws' time step dt/2 and the rigid bodies'
// half on the second stream beside the internal forces; then the half's end
void solver_half(mel_ctx& c, int half, double dt, double t) {
    DevModel& d = c.d;
    const BlockArrays& B = c.blocks_dev;
    hipStream_t s = c.stream, a = c.aux;
    const int n_cwg = contact_workgroups(c);
    DevModel dh = d;
    dh.dt = dt * 0.5;
    if (half == 2) k_ad_bc<<<blocks(d.n_borders + d.n_rigid), BLOCK, 0, s>>>(d, t + dt * 0.5);
    HIP_OK(hipEventRecord(c.ev_fork, s));
    HIP_OK(hipStreamWaitEvent(a, c.ev_fork, 0));
    TIMED(c, KT_CONTACTS, a, (k_step_contacts<<<n_cwg, BLOCK_THREADS, 0, a>>>(dh, B.bc_bn, B.n_bc_bn)));
    if (d.n_rules > 0) k_ghost_sends<<<64, BLOCK, 0, a>>>(d);
    if (d.n_matrix > 0) TIMED(c, KT_DAMPING, a, (k_damping<<<blocks(d.n_nodes), BLOCK, 0, a>>>(d)));
    // Update_alid_forces(Time) in the first half, (Time + dt/2) in the second
    if (d.n_alid_nodes > 0) k_alid<<<blocks(d.n_alid_nodes), BLOCK, 0, a>>>(d, half == 1 ? t : t + dt * 0.5);
    if (d.n_rigid + B.n_split_bn > 0)
        TIMED(c, KT_GATHER, a, (k_ad_gather<<<d.n_rigid + blocks(B.n_split_bn, BLOCK_THREADS), BLOCK_THREADS, 0, a>>>(
                                   d, c.ad, B.split_bn, B.n_split_bn, half, dt, t + dt)));
    HIP_OK(hipEventRecord(c.ev_join, a));
    if (B.n_blocks > 0) TIMED(c, KT_BLOCK_UPDATE, s, launch_internal_forces(c, s));
    HIP_OK(hipStreamWaitEvent(s, c.ev_join, 0));
    const int extra = std::max(1, (d.n_rigid_nodes + d.n_borders + BLOCK_THREADS - 1) / BLOCK_THREADS);
    TIMED(c, KT_FINISH, s, (k_ad_finish<<<B.n_blocks + extra, BLOCK_THREADS, 0, s>>>(d, B, c.ad, half, dt, t + dt)));
    if (B.n_split_kin > 0)
        k_kinematics_list<<<blocks(B.n_split_kin), BLOCK, 0, s>>>(d, B.split_kin, B.n_split_kin);
    // Update_kinematics of the first half's end, for the absorbing layers of
    // the second (the second half's end rebuilds every node's)
    if (half == 1 && d.n_alid_nodes > 0)
        k_kinematics_list<<<blocks(d.n_alid_nodes), BLOCK, 0, s>>>(d, d.alid_nodes, d.n_alid_nodes);
    HIP_OK(hipGetLastError());
}

// One Solver_step: Store, then attempts of the current time step until one
// is accepted; after every attempt the device reduces the error estimate
// (and the scaled masses) and the host decides and updates the time step
// as Code/Solver.h:328-484 does.
void solver_step(mel_ctx& c) {
    const Model& m = c.model;
    DevModel& d = c.d;
    poll_async(c, false);
    c.prox_after_step = false;
    store_restore(c, 0);
    c.neval = 0;
    for (;;) {
        c.neval++;
        c.attempts++;
        const double dt = d.dt;
        solver_half(c, 1, dt, c.time);
        solver_half(c, 2, dt, c.time);
        if (c.ad.n_kin > 0) k_ad_error<<<blocks(c.ad.n_kin), BLOCK, 0, c.stream>>>(d, c.ad);
        k_ad_body_reduce<<<d.n_bodies, BLOCK, 0, c.stream>>>(d, c.ad);
        k_ad_global<<<1, 64, 0, c.stream>>>(d, c.ad);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(c.h_ad, c.ad.out, AD_OUT_FIELDS * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIP_OK(hipEventRecord(c.ev_ad, c.stream));
        HIP_OK(hipEventSynchronize(c.ev_ad));
        const double max_error = c.h_ad[AO_MAX_ERROR], max_mass = c.h_ad[AO_MAX_MASS];
        bool fail;
        double correction1;
        if (m.scheme == SCH_ADAPTIVE_EULER) {
            fail = max_error > m.target_error * m.accepted_ratio;
            correction1 = max_error / m.target_error;
        } else if (m.scheme == SCH_MASS_SCALING) {
            fail = max_error > m.target_error * m.accepted_ratio;
            correction1 = max_error * m.inv_target_error;
        } else {
            fail = max_mass > m.accepted_ratio + 0.05;
            correction1 = max_mass;
        }
        const double correction2 = 1. / m.accepted_ratio;
        if (fail) store_restore(c, 1);
        else c.time += dt;
        double ndt;
        if (m.scheme == SCH_ADAPTIVE_MASS_SCALING && !fail) ndt = dt / std::pow(correction2, m.control_parameter);
        else if (correction1 > correction2) ndt = dt / std::pow(correction1, m.control_parameter);
        else ndt = dt / std::pow(correction2, m.control_parameter);
        if (ndt > m.contact_update_period) ndt = m.contact_update_period;
        d.dt = ndt;
        if (!fail) {
            c.nan_stop = c.h_ad[AO_NAN] != 0.;
            break;
        }
    }
    k_set_time<<<1, 64, 0, c.stream>>>(d.time, c.time);
    HIP_OK(hipGetLastError());
    c.outputs_stale = true;
}

// ---------------------------------------------------------------------------
// In-memory checkpoint of the dynamic state (mel_save_state / mel_load_state)
// ---------------------------------------------------------------------------
constexpr uint64_t STATE_MAGIC = 0x3245544154534c4dull;   // "MLSTATE2"

struct StateArray {
    void* p;
    size_t bytes;
};

// the device arrays that carry the state between steps, in blob order;
// contact element arrays hold the first nc elements
std::vector<StateArray> state_arrays(mel_ctx& c, int nc) {
    DevModel& d = c.d;
    std::vector<StateArray> v;
    const size_t N = d.n_nodes, NB = d.n_bodies, NS = d.n_slots, NBR = d.n_borders, NBN = d.n_bnodes;
    double* nf[] = {d.xc, d.yc, d.ux, d.uy, d.vx, d.vy, d.ax, d.ay, d.upx, d.upy, d.vpx, d.vpy, d.apx, d.apy,
                    d.fix, d.fiy, d.fcx, d.fcy, d.fbx, d.fby, d.fdx, d.fdy, d.fnx, d.fny, d.fkx, d.fky,
                    d.fx, d.fy, d.mx, d.my, d.imx, d.imy, d.fax, d.fay};
    for (double* p : nf) v.push_back({p, N * sizeof(double)});
    v.push_back({d.body_rigid, NB * RIGID_FIELDS * sizeof(double)});
    v.push_back({d.body_mass, NB * sizeof(double)});
    v.push_back({d.body_inv_mass, NB * sizeof(double)});
    double* bf[] = {d.br_xval, d.br_yval, d.br_xvel, d.br_yvel};
    for (double* p : bf) v.push_back({p, NBR * sizeof(double)});
    double* sf[] = {d.sl_len, d.sl_xbp, d.sl_ybp};
    for (double* p : sf) v.push_back({p, NS * sizeof(double)});
    v.push_back({d.slot_xy, NS * sizeof(double2)});
    double* bnf[] = {d.bn_pcx, d.bn_pcy, d.bn_pdx, d.bn_pdy, d.bn_pnx, d.bn_pny};
    for (double* p : bnf) v.push_back({p, NBN * sizeof(double)});
    v.push_back({d.nb_count, NB * sizeof(int)});
    v.push_back({d.nb_list, NB * (size_t)d.nb_cap * 3 * sizeof(int)});
    if (d.n_rigid > 0) v.push_back({d.rigid_dv, (size_t)d.n_rigid * 8 * sizeof(double)});
    if (c.blocks_dev.n_blocks > 0) v.push_back({c.blocks_dev.blk_work, 14 * (size_t)c.blocks_dev.n_blocks * sizeof(double)});
    v.push_back({d.time, sizeof(double)});
    if (d.n_rules > 0) {
        // statuses, damage, the inactive bodies' ghost sends and what their
        // border nodes received, the node blocks' flags
        v.push_back({d.body_status, NB * sizeof(int)});
        v.push_back({d.body_damage, NB * sizeof(double)});
        v.push_back({d.gh_count, 2 * sizeof(int)});
        v.push_back({d.gh_bn, (size_t)d.ghost_cap * sizeof(int)});
        v.push_back({d.gh_fx, (size_t)d.ghost_cap * sizeof(double)});
        v.push_back({d.gh_fy, (size_t)d.ghost_cap * sizeof(double)});
        v.push_back({d.gh_rb, (size_t)d.ghost_cap * sizeof(int)});
        v.push_back({d.gh_rf, 3 * (size_t)d.ghost_cap * sizeof(double)});
        v.push_back({d.bn_acx, NBN * sizeof(double)});
        v.push_back({d.bn_acy, NBN * sizeof(double)});
        if (c.blocks_dev.n_blocks > 0) v.push_back({c.blocks_dev.blk_hdr, 4 * (size_t)c.blocks_dev.n_blocks * sizeof(int4)});
    }
    if (c.model.scheme != SCH_EULER) {
        // mass-scaling factors and the bodies' adaptive scalars
        v.push_back({c.ad.fsx, N * sizeof(double)});
        v.push_back({c.ad.fsy, N * sizeof(double)});
        v.push_back({c.ad.body, NB * AD_BODY_FIELDS * sizeof(double)});
    }
    const ContactArrays& A = d.C;
    int* ci[] = {A.bodyS, A.borderS, A.bnodeS, A.nodeS, A.bodyM, A.borderM, A.shiftM, A.segM, A.nbint, A.fresh,
                 A.nM0, A.nM1, A.nM2, A.nM3, A.sM0, A.sM1, A.sM2, A.sM3};
    for (int* p : ci) v.push_back({p, (size_t)nc * sizeof(int)});
    double* cd[] = {A.N0, A.N1, A.N2, A.N3, A.gapn, A.vgapn, A.gapt, A.vgapt, A.xsi, A.xn, A.yn, A.xt, A.yt,
                    A.emass, A.len, A.fx, A.fy, A.int0, A.wacc, A.int1, A.int2};
    for (double* p : cd) v.push_back({p, (size_t)nc * sizeof(double)});
    v.push_back({A.info, (size_t)nc * sizeof(int4)});
    v.push_back({A.sinfo, (size_t)nc * sizeof(int4)});
    return v;
}

struct StateHeader {
    uint64_t magic;
    uint64_t n_arrays;
    int64_t n_nodes, n_bodies, n_slots, n_contacts, iteration, reserved;
    double time, next_contact_update, dt;
};

size_t state_size(mel_ctx& c, int nc) {
    size_t n = sizeof(StateHeader);
    for (const StateArray& a : state_arrays(c, nc)) n += sizeof(uint64_t) + a.bytes;
    return n;
}

double* node_field_ptr(mel_ctx& c, int field) {
    DevModel& d = c.d;
    double* f[] = {d.xc, d.yc, d.ux, d.uy, d.vx, d.vy, d.ax, d.ay, d.upx, d.upy, d.vpx, d.vpy, d.apx, d.apy,
                   d.fix, d.fiy, d.fcx, d.fcy, d.fbx, d.fby, d.fdx, d.fdy, d.fnx, d.fny, d.fkx, d.fky,
                   d.fx, d.fy, d.mx, d.my, d.fax, d.fay};
    static_assert(sizeof(f) / sizeof(f[0]) == MEL_NODE_FIELD_COUNT, "one array per node field");
    if (field < 0 || field >= MEL_NODE_FIELD_COUNT) return nullptr;
    return f[field];
}
}  // namespace

extern "C" {

const char* mel_version(void) { return "melody2d-mi355x 0.1"; }

mel_status mel_load(const char* dir, int save_index, mel_precision precision, int hip_device, mel_ctx** out) {
    return mel_load_ex(dir, save_index, precision, hip_device, nullptr, out);
}

mel_status mel_load_binary(const char* path, mel_precision precision, int hip_device, mel_ctx** out) {
    return mel_load_binary_ex(path, precision, hip_device, nullptr, out);
}

void mel_default_options(mel_options* o) {
    if (o) *o = mel_options{-1, -1, -1, -1};
}

mel_status mel_load_ex(const char* dir, int save_index, mel_precision precision, int hip_device,
                       const mel_options* opts, mel_ctx** out) {
    if (!dir) return MEL_E_ARG;
    std::string d(dir);
    return create(out, hip_device, precision,
                  [&](Model& m, std::string& err) { return load_model_files(d, save_index, m, err); }, opts);
}

mel_status mel_load_binary_ex(const char* path, mel_precision precision, int hip_device, const mel_options* opts,
                              mel_ctx** out) {
    if (!path) return MEL_E_ARG;
    std::string p(path);
    return create(out, hip_device, precision,
                  [&](Model& m, std::string& err) { return load_model_binary(p, m, err); }, opts);
}

mel_status mel_get_options(const mel_ctx* c, mel_options* o) {
    if (!c || !o) return MEL_E_ARG;
    o->block_nodes = c->opts.block_nodes;
    o->fused_step = c->two_launch ? 1 : 0;
    o->graph_replay = c->use_graphs ? 1 : 0;
    o->wide_window = c->wide ? 1 : 0;
    return MEL_OK;
}

mel_status mel_autotune(const char* path, mel_precision precision, int hip_device, int steps, mel_options* best,
                        double* best_ms, int* n_tried) {
    if (!path || !best || steps < 1) return MEL_E_ARG;
    Model base;
    std::string err;
    if (!load_model_binary(path, base, err)) {
        g_load_err = err;
        return err.rfind("cannot open", 0) == 0 ? MEL_E_IO : MEL_E_PARSE;
    }
    int max_nodes = 0;
    for (auto& b : base.bodies)
        if (!b.rigid) max_nodes = std::max(max_nodes, b.n_nodes);
    // node-block sizes only matter for bodies split over blocks; a body that
    // fits one block is never split on purpose (that forces the launch
    // sequence)
    std::vector<int> blocks = {-1};
    if (max_nodes > BLOCK_MAX_OWN) blocks = {BLOCK_MAX_OWN, 192, 128};
    std::vector<mel_options> cand;
    for (int bn : blocks) {
        cand.push_back(mel_options{bn, 1, 1, -1});
        cand.push_back(mel_options{bn, 1, 0, -1});
        cand.push_back(mel_options{bn, 0, 0, 0});
        cand.push_back(mel_options{bn, 0, 0, 1});
    }
    std::vector<mel_options> seen;
    double t_best = 1e300;
    mel_status first_err = MEL_OK;
    int tried = 0;
    for (const mel_options& o : cand) {
        mel_ctx* c = nullptr;
        mel_status st = create(&c, hip_device, precision,
                               [&](Model& m, std::string&) { m = base; return true; }, &o);
        if (st != MEL_OK) {
            if (first_err == MEL_OK) first_err = st;
            continue;
        }
        mel_options r;
        mel_get_options(c, &r);
        bool dup = false;
        for (auto& q : seen)
            dup |= q.block_nodes == r.block_nodes && q.fused_step == r.fused_step &&
                   q.graph_replay == r.graph_replay && q.wide_window == r.wide_window;
        if (dup) { mel_destroy(c); continue; }
        seen.push_back(r);
        // warm-up: first graph captures and first contact update out of the timing
        st = mel_euler_steps(c, std::min(steps, 20));
        if (st == MEL_OK) st = mel_synchronize(c);
        auto t0 = std::chrono::steady_clock::now();
        if (st == MEL_OK) st = mel_euler_steps(c, steps);
        if (st == MEL_OK) st = mel_synchronize(c);
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
        if (st != MEL_OK) g_load_err = c->err;
        mel_destroy(c);
        if (st != MEL_OK) {
            if (first_err == MEL_OK) first_err = st;
            continue;
        }
        tried++;
        if (ms < t_best) {
            t_best = ms;
            *best = o;
            best->fused_step = r.fused_step;
            best->graph_replay = r.graph_replay;
            best->wide_window = r.wide_window;
        }
    }
    if (n_tried) *n_tried = tried;
    if (!tried) return first_err == MEL_OK ? MEL_E_UNSUPPORTED : first_err;
    if (best_ms) *best_ms = t_best;
    return MEL_OK;
}

mel_status mel_update_proximity(mel_ctx* c) {
    return guarded(c, [&]() { update_proximity(*c); return MEL_OK; });
}

mel_status mel_euler_steps(mel_ctx* c, int nsteps) {
    if (nsteps < 0) return MEL_E_ARG;
    return guarded(c, [&]() {
        if (c->two_launch) {
            // runs of steps up to and including the one after which the
            // contact update falls due (the host keeps the same time sums as
            // the device)
            int k = 0;
            while (k < nsteps) {
                int n = 0;
                double t = c->time;
                bool due = false;
                while (k + n < nsteps) {
                    t = t + c->d.dt;
                    n++;
                    if (t > c->next_contact_update - 1.e-6 * c->d.dt) { due = true; break; }
                    if (c->mon_on) break;      // a Monitoring row after every step
                }
                run_steps(*c, n);
                launch_monitor(*c, 0., 0., 0.);
                c->time = t;
                c->iteration += n;
                c->outputs_stale = true;
                k += n;
                if (due) {
                    // contact-update cadence, Code/main.cpp:191-196
                    c->next_contact_update = c->next_contact_update + c->model.contact_update_period;
                    if (c->timing && (c->timing_mask >> KT_PROXIMITY & 1u)) {
                        int ev = c->ev_begin(KT_PROXIMITY, c->stream);
                        contact_update(*c);
                        c->ev_end(ev, c->stream);
                    } else {
                        contact_update(*c);
                    }
                }
            }
            HIP_OK(hipGetLastError());
            return MEL_OK;
        }
        for (int k = 0; k < nsteps; k++) {
            c->iteration++;
            euler_step(*c);
            launch_monitor(*c, 0., 0., 0.);
            // contact-update cadence, Code/main.cpp:191-196
            if (c->time > c->next_contact_update - 1.e-6 * c->d.dt) {
                c->next_contact_update = c->next_contact_update + c->model.contact_update_period;
                if (c->timing && (c->timing_mask >> KT_PROXIMITY & 1u)) {
                    int ev = c->ev_begin(KT_PROXIMITY, c->stream);
                    contact_update(*c);
                    c->ev_end(ev, c->stream);
                } else {
                    contact_update(*c);
                }
            }
        }
        HIP_OK(hipGetLastError());
        return MEL_OK;
    });
}

mel_status mel_solver_steps(mel_ctx* c, int nsteps) {
    if (nsteps < 0) return MEL_E_ARG;
    if (c && c->model.scheme == SCH_EULER) return mel_euler_steps(c, nsteps);
    return guarded(c, [&]() {
        for (int k = 0; k < nsteps; k++) {
            c->iteration++;
            solver_step(*c);
            if (c->mon_on) {
                HIP_OK(hipStreamSynchronize(c->stream));
                launch_monitor(*c, c->neval, c->h_ad[AO_MAX_ERROR], c->h_ad[AO_MEAN_ERROR]);
            }
            // contact-update cadence with the new time step, Code/main.cpp:191-196
            if (c->time > c->next_contact_update - 1.e-6 * c->d.dt) {
                c->next_contact_update = c->next_contact_update + c->model.contact_update_period;
                contact_update(*c);
            }
            if (c->nan_stop) {
                c->err = "a body's total error is NaN (the reference stops the run, Code/main.cpp:227-237)";
                return MEL_E_NAN;
            }
        }
        HIP_OK(hipGetLastError());
        return MEL_OK;
    });
}

mel_status mel_get_solver_info(mel_ctx* c, mel_solver_info_t* info) {
    if (!c || !info) return MEL_E_ARG;
    return guarded(c, [&]() {
        const Model& m = c->model;
        info->scheme = m.scheme;
        info->dt = c->d.dt;
        info->neval = c->neval;
        info->attempts = c->attempts;
        info->max_error = info->mean_error = info->total_mass = info->max_mass = 0.;
        if (m.scheme != SCH_EULER && c->h_ad) {
            HIP_OK(hipStreamSynchronize(c->stream));
            info->max_error = c->h_ad[AO_MAX_ERROR];
            info->mean_error = c->h_ad[AO_MEAN_ERROR];
            info->total_mass = c->h_ad[AO_TOTAL_MASS];
            info->max_mass = c->h_ad[AO_MAX_MASS];
        }
        return MEL_OK;
    });
}

mel_status mel_get_body_error(mel_ctx* c, int body, double* out5) {
    if (!c || !out5 || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    if (c->model.scheme == SCH_EULER) {
        c->err = "mel_get_body_error: the model's scheme is Euler (no error estimate)";
        return MEL_E_ARG;
    }
    return guarded(c, [&]() {
        std::vector<double> v;
        c->download(v, c->ad.body + (size_t)body * AD_BODY_FIELDS, AD_BODY_FIELDS);
        out5[0] = v[AB_TOTAL]; out5[1] = v[AB_MAX]; out5[2] = v[AB_NODE];
        out5[3] = v[AB_MASS_MS]; out5[4] = v[AB_FACTOR];
        return MEL_OK;
    });
}

mel_status mel_get_mass_scaling(mel_ctx* c, int body, double* fx, double* fy, size_t n) {
    if (!c || !fx || !fy || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    if ((size_t)c->model.bodies[body].n_nodes != n) return MEL_E_ARG;
    return guarded(c, [&]() {
        const size_t o = c->h_node_off[body];
        if (c->model.scheme == SCH_EULER) {
            for (size_t i = 0; i < n; i++) fx[i] = fy[i] = 1.;
            return MEL_OK;
        }
        HIP_OK(hipMemcpyAsync(fx, c->ad.fsx + o, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(fy, c->ad.fsy + o, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        return MEL_OK;
    });
}

mel_status mel_synchronize(mel_ctx* c) {
    return guarded(c, [&]() {
        HIP_OK(hipStreamSynchronize(c->aux));
        HIP_OK(hipStreamSynchronize(c->stream));
        poll_async(*c, true);
        check_flags(*c);
        return MEL_OK;
    });
}

mel_status mel_compute_nodal_stresses(mel_ctx* c) {
    return guarded(c, [&]() {
        const Model& m = c->model;
        if (!c->st_out) {
            std::vector<double> fx, fy, par;
            for (const Body& b : m.bodies) {
                const Material& mat = m.materials[b.material];
                par.push_back(b.rigid ? -1. : (mat.type == MAT_ELASTIC_LINEAR ? 0. : 1.));
                par.push_back(mat.p[3]);
                par.push_back(mat.p[4]);
                // the same entry order as the device's DOF_TO_DISP rows
                if (b.rigid) continue;
                for (int i = 0; i < b.n_nodes; i++)
                    for (int k = b.dptr[i]; k < b.dptr[i + 1]; k++) {
                        fx.push_back(b.dphix[k]);
                        fy.push_back(b.dphiy[k]);
                    }
            }
            if ((long long)fx.size() != c->nnz_d)
                throw CapacityError{"nodal stresses: DOF_TO_DISP rows differ from the device's"};
            c->st_dphix = c->upload(fx);
            c->st_dphiy = c->upload(fy);
            c->st_bpar = c->upload(par);
            c->st_out = c->alloc<double>((size_t)MEL_STRESS_FIELD_COUNT * c->d.n_nodes);
        }
        k_nodal_stresses<<<blocks(c->d.n_nodes), BLOCK, 0, c->stream>>>(c->d, c->st_dphix, c->st_dphiy, c->st_bpar,
                                                                      c->st_out);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(c->stream));
        return MEL_OK;
    });
}

mel_status mel_get_nodal_stress(mel_ctx* c, int body, int field, double* out, size_t n) {
    if (!c || !out || body < 0 || body >= c->d.n_bodies || field < 0 || field >= MEL_STRESS_FIELD_COUNT)
        return MEL_E_ARG;
    const size_t nn = c->model.bodies[body].n_nodes;
    if (n < nn || !c->st_out) return MEL_E_ARG;
    return guarded(c, [&]() {
        HIP_OK(hipMemcpyAsync(out, c->st_out + (size_t)field * c->d.n_nodes + c->h_node_off[body], nn * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        return MEL_OK;
    });
}

mel_status mel_write_graphic(mel_ctx* c, const char* path, const int* to_plot, int number_print) {
    if (!c || !path) return MEL_E_ARG;
    mel_status st = mel_compute_nodal_stresses(c);
    if (st != MEL_OK) return st;
    int plot[GRAPHIC_FIELDS];
    for (int k = 0; k < GRAPHIC_FIELDS; k++) plot[k] = to_plot ? to_plot[k] : ((k <= 12) || (k >= 14 && k <= 28));
    const Model& m = c->model;
    std::vector<GraphicBody> g(m.bodies.size());
    static const int vec_fields[10][2] = {
        {MEL_X_DISP, MEL_Y_DISP}, {MEL_X_VEL, MEL_Y_VEL}, {MEL_X_ACC, MEL_Y_ACC}, {MEL_X_FORCE, MEL_Y_FORCE},
        {MEL_X_INTERNAL_FORCE, MEL_Y_INTERNAL_FORCE}, {MEL_X_CONTACT_FORCE, MEL_Y_CONTACT_FORCE},
        {MEL_X_BODY_FORCE, MEL_Y_BODY_FORCE}, {MEL_X_DIRICHLET_FORCE, MEL_Y_DIRICHLET_FORCE},
        {MEL_X_NEUMANN_FORCE, MEL_Y_NEUMANN_FORCE}, {MEL_X_DAMPING_FORCE, MEL_Y_DAMPING_FORCE}};
    for (size_t bi = 0; bi < m.bodies.size(); bi++) {
        const Body& b = m.bodies[bi];
        GraphicBody& G = g[bi];
        const size_t n = b.n_nodes;
        G.active = b.status != ST_INACTIVE;
        G.x0 = b.X;
        G.y0 = b.Y;
        G.x.resize(n);
        G.y.resize(n);
        auto get = [&](int field, double* out) { return mel_get_node_field(c, (int)bi, field, out, n); };
        if ((st = get(MEL_X_CURRENT, G.x.data())) != MEL_OK || (st = get(MEL_Y_CURRENT, G.y.data())) != MEL_OK)
            return st;
        G.vec.assign(10, std::vector<double>(2 * n));
        for (int v = 0; v < 10; v++)
            if ((st = get(vec_fields[v][0], G.vec[v].data())) != MEL_OK ||
                (st = get(vec_fields[v][1], G.vec[v].data() + n)) != MEL_OK)
                return st;
        G.stress.resize((size_t)MEL_STRESS_FIELD_COUNT * n);
        for (int f = 0; f < MEL_STRESS_FIELD_COUNT; f++)
            if ((st = mel_get_nodal_stress(c, (int)bi, f, G.stress.data() + (size_t)f * n, n)) != MEL_OK) return st;
        double rigid26[26];
        if ((st = mel_get_body_state(c, (int)bi, rigid26, G.works)) != MEL_OK) return st;
    }
    mel_info_t info;
    if ((st = mel_get_info(c, &info)) != MEL_OK) return st;
    std::string err;
    if (!write_graphic_text(path, m, g, info.iteration, m.number_save, number_print, info.time, plot, err)) {
        c->err = err;
        return MEL_E_IO;
    }
    return MEL_OK;
}

mel_status mel_get_body_sizes(mel_ctx* c, int body, int* n_nodes, int* n_gauss, int* n_borders, int* rigid) {
    if (!c || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    const Body& b = c->model.bodies[body];
    if (n_nodes) *n_nodes = b.n_nodes;
    if (n_gauss) *n_gauss = b.n_gauss;
    if (n_borders) *n_borders = (int)b.borders.size();
    if (rigid) *rigid = b.rigid ? 1 : 0;
    return MEL_OK;
}

mel_status mel_get_node_field(mel_ctx* c, int body, int field, double* out, size_t n) {
    if (!c || !out || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    double* p = node_field_ptr(*c, field);
    size_t nn = c->model.bodies[body].n_nodes;
    if (!p || n < nn) return MEL_E_ARG;
    return guarded(c, [&]() {
        refresh_outputs(*c);
        HIP_OK(hipMemcpyAsync(out, p + c->h_node_off[body], nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        return MEL_OK;
    });
}

mel_status mel_set_node_field(mel_ctx* c, int body, int field, const double* in, size_t n) {
    if (!c || !in || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    double* p = node_field_ptr(*c, field);
    size_t nn = c->model.bodies[body].n_nodes;
    if (!p || n < nn) return MEL_E_ARG;
    return guarded(c, [&]() {
        refresh_outputs(*c);
        HIP_OK(hipMemcpyAsync(p + c->h_node_off[body], in, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        return MEL_OK;
    });
}

mel_status mel_get_contacts(mel_ctx* c, int body, int32_t* ids, double* state, size_t cap, size_t* count) {
    if (!c || !count || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    return guarded(c, [&]() {
        int n = contact_count(*c);
        ContactArrays& A = c->bufA;
        std::vector<int> I[18];
        int* ip[18] = {A.bodyS, A.borderS, A.bnodeS, A.nodeS, A.bodyM, A.borderM, A.shiftM, A.segM, A.nM0, A.nM1,
                       A.nM2, A.nM3, A.sM0, A.sM1, A.sM2, A.sM3, A.nbint, A.fresh};
        for (int k = 0; k < 18; k++) c->download(I[k], ip[k], n);
        std::vector<double> D[18];
        double* dp[18] = {A.N0, A.N1, A.N2, A.N3, A.gapn, A.vgapn, A.gapt, A.vgapt, A.xsi, A.xn, A.yn, A.xt, A.yt,
                          A.emass, A.len, A.fx, A.fy, A.int0};
        for (int k = 0; k < 18; k++) c->download(D[k], dp[k], n);
        size_t w = 0;
        for (int e = 0; e < n; e++) {
            if (I[0][e] != body) continue;
            if (ids && state && w < cap) {
                int oS = c->h_node_off[body], oM = c->h_node_off[I[4][e]];
                int32_t* o = ids + 16 * w;
                o[0] = I[1][e]; o[1] = I[2][e]; o[2] = I[3][e] - oS; o[3] = I[4][e]; o[4] = I[5][e]; o[5] = I[6][e];
                o[6] = I[7][e];
                for (int k = 0; k < 4; k++) { o[7 + k] = I[8 + k][e] - oM; o[11 + k] = I[12 + k][e]; }
                o[15] = I[16][e];
                for (int k = 0; k < 18; k++) state[18 * w + k] = D[k][e];
            }
            w++;
        }
        *count = w;
        return w > cap && ids ? MEL_E_ARG : MEL_OK;
    });
}

mel_status mel_get_contact_internals(mel_ctx* c, int body, double* out, size_t cap, size_t* count) {
    if (!c || !count || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    return guarded(c, [&]() {
        int n = contact_count(*c);
        const ContactArrays& A = c->bufA;
        std::vector<int> bs;
        std::vector<double> v[3];
        c->download(bs, A.bodyS, n);
        c->download(v[0], A.int0, n);
        c->download(v[1], A.int1, n);
        c->download(v[2], A.int2, n);
        size_t w = 0;
        for (int e = 0; e < n; e++) {
            if (bs[e] != body) continue;
            if (out && w < cap)
                for (int k = 0; k < 3; k++) out[3 * w + k] = v[k][e];
            w++;
        }
        *count = w;
        return w > cap && out ? MEL_E_ARG : MEL_OK;
    });
}

mel_status mel_get_neighbours(mel_ctx* c, int body, int32_t* out, size_t cap, size_t* count) {
    if (!c || !count || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    return guarded(c, [&]() {
        std::vector<int> nb, lst;
        c->download(nb, c->d.nb_count + body, 1);
        c->download(lst, c->d.nb_list + (size_t)body * c->d.nb_cap * 3, 3 * (size_t)nb[0]);
        *count = nb[0];
        if (out) {
            if ((size_t)nb[0] > cap) return MEL_E_ARG;
            std::copy(lst.begin(), lst.end(), out);
        }
        return MEL_OK;
    });
}

mel_status mel_set_monitoring(mel_ctx* c, int on) {
    if (!c) return MEL_E_ARG;
    return guarded(c, [&]() {
        if (!on) {
            mon_drain(*c);
            c->mon_on = false;
            return MEL_OK;
        }
        if (c->model.flags[F_NO_MONITORING]) {
            c->err = "the model sets NO_MONITORING (Code/main.cpp:172 then skips Monitoring)";
            return MEL_E_UNSUPPORTED;
        }
        if (!c->d_mon_rows) {
            const Model& m = c->model;
            // iteration, time, time step, neval, max_error, mean_error, one
            // value per entry at most, three energies
            c->mon_stride = 1 + 6 + (int)m.monitored.size() + 3;
            c->mon_cap = 1024;
            c->d_mon_rows = c->alloc<double>((size_t)c->mon_stride * c->mon_cap);
            c->d_mon_n = c->alloc<int>(1);
            c->d_mon_iter = c->alloc<double>(1);
            std::vector<double> items((size_t)RULE_STRIDE * std::max<size_t>(1, m.monitored.size()), 0.);
            for (size_t i = 0; i < m.monitored.size(); i++)
                for (size_t k = 0; k < m.monitored[i].size() && k < (size_t)RULE_STRIDE; k++)
                    items[RULE_STRIDE * i + k] = m.monitored[i][k];
            c->d_mon_items = c->upload(items);
        }
        if (!c->mon_on) {
            const double it = c->iteration;
            HIP_OK(hipMemcpyAsync(c->d_mon_iter, &it, sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIP_OK(hipStreamSynchronize(c->stream));
        }
        c->mon_on = true;
        return MEL_OK;
    });
}

mel_status mel_get_monitoring(mel_ctx* c, double* out, size_t cap, size_t* count) {
    if (!c || !count) return MEL_E_ARG;
    return guarded(c, [&]() {
        mon_drain(*c);
        *count = c->h_mon.size();
        if (!out) return MEL_OK;
        if (cap < c->h_mon.size()) {
            c->err = "mel_get_monitoring: " + std::to_string(c->h_mon.size()) + " values, room for " + std::to_string(cap);
            return MEL_E_ARG;
        }
        std::copy(c->h_mon.begin(), c->h_mon.end(), out);
        c->h_mon.clear();
        return MEL_OK;
    });
}

mel_status mel_write_monitoring(mel_ctx* c, const char* path) {
    if (!c || !path) return MEL_E_ARG;
    return guarded(c, [&]() {
        mon_drain(*c);
        std::string err;
        if (!write_monitoring_text(path, c->h_mon, err)) { c->err = err; return MEL_E_IO; }
        c->h_mon.clear();
        return MEL_OK;
    });
}

mel_status mel_get_body_status(mel_ctx* c, int body, int* status, double* damage) {
    if (!c || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    return guarded(c, [&]() {
        std::vector<int> st;
        c->download(st, c->d.body_status + body, 1);
        if (status) *status = st[0];
        if (damage) {
            *damage = 0.;
            if (c->d.body_damage) {
                std::vector<double> dm;
                c->download(dm, c->d.body_damage + body, 1);
                *damage = dm[0];
            }
        }
        return MEL_OK;
    });
}

mel_status mel_get_body_state(mel_ctx* c, int body, double* rigid26, double* works7) {
    if (!c || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    return guarded(c, [&]() {
        std::vector<double> R, bm, bi;
        c->download(R, c->d.body_rigid + (size_t)body * RIGID_FIELDS, RIGID_FIELDS);
        c->download(bm, c->d.body_mass + body, 1);
        c->download(bi, c->d.body_inertia + body, 1);
        if (rigid26) {
            double v[26] = {R[R_CUR], R[R_CUR + 1], R[R_CUR + 2], R[R_DISP], R[R_DISP + 1], R[R_DISP + 2],
                            R[R_VEL], R[R_VEL + 1], R[R_VEL + 2], R[R_ACC], R[R_ACC + 1], R[R_ACC + 2],
                            R[R_FCON], R[R_FCON + 1], R[R_FCON + 2], R[R_FTOT], R[R_FTOT + 1], R[R_FTOT + 2],
                            bm[0], bi[0], R[R_FDIR], R[R_FDIR + 1], R[R_FDIR + 2], R[R_FNEU], R[R_FNEU + 1],
                            R[R_FNEU + 2]};
            std::memcpy(rigid26, v, sizeof v);
        }
        if (works7) {
            const std::vector<double> w = all_works(*c);
            for (int q = 0; q < 7; q++) works7[q] = w[7 * (size_t)body + q];
        }
        return MEL_OK;
    });
}

mel_status mel_get_body_energy(mel_ctx* c, int body, double* energy3) {
    if (!c || !energy3 || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    return guarded(c, [&]() {
        const std::vector<double> e = all_energy(*c);
        for (int k = 0; k < 3; k++) energy3[k] = e[3 * (size_t)body + k];
        return MEL_OK;
    });
}

mel_status mel_get_scalars(mel_ctx* c, double* time, double* works7, double* energy3) {
    if (!c) return MEL_E_ARG;
    return guarded(c, [&]() {
        if (time) *time = c->time;
        if (works7) {
            const std::vector<double> w = all_works(*c);
            for (int q = 0; q < 7; q++) {
                works7[q] = 0.;
                for (int b = 0; b < c->d.n_bodies; b++) works7[q] += w[7 * (size_t)b + q];
            }
        }
        if (energy3) {
            const std::vector<double> e = all_energy(*c);
            for (int k = 0; k < 3; k++) {
                energy3[k] = 0.;
                for (int b = 0; b < c->d.n_bodies; b++) energy3[k] += e[3 * (size_t)b + k];
            }
        }
        return MEL_OK;
    });
}

mel_status mel_eval_forces(mel_ctx* c, unsigned what) {
    if (!c || (what & ~(unsigned)MEL_EVAL_ALL)) return MEL_E_ARG;
    return guarded(c, [&]() {
        HIP_OK(hipStreamSynchronize(c->stream));
        poll_async(*c, true);
        check_flags(*c);
        eval_forces(*c, what);
        return MEL_OK;
    });
}

mel_status mel_get_state_size(mel_ctx* c, size_t* bytes) {
    if (!c || !bytes) return MEL_E_ARG;
    return guarded(c, [&]() {
        HIP_OK(hipStreamSynchronize(c->stream));
        poll_async(*c, true);
        *bytes = state_size(*c, contact_count(*c));
        return MEL_OK;
    });
}

mel_status mel_save_state(mel_ctx* c, void* buf, size_t cap, size_t* written) {
    if (!c || !buf) return MEL_E_ARG;
    return guarded(c, [&]() {
        HIP_OK(hipStreamSynchronize(c->stream));
        poll_async(*c, true);
        check_flags(*c);
        flush_contact_work(*c);
        refresh_outputs(*c);
        const int nc = contact_count(*c);
        const size_t need = state_size(*c, nc);
        if (cap < need) {
            c->err = "mel_save_state: buffer of " + std::to_string(cap) + " bytes, the state needs " + std::to_string(need);
            return MEL_E_ARG;
        }
        const std::vector<StateArray> arrs = state_arrays(*c, nc);
        StateHeader h{STATE_MAGIC, arrs.size(), c->d.n_nodes, c->d.n_bodies, c->d.n_slots, nc, c->iteration, 0,
                      c->time, c->next_contact_update, c->d.dt};
        unsigned char* o = static_cast<unsigned char*>(buf);
        std::memcpy(o, &h, sizeof h);
        size_t pos = sizeof h;
        for (const StateArray& a : arrs) {
            uint64_t nb = a.bytes;
            std::memcpy(o + pos, &nb, sizeof nb);
            pos += sizeof nb;
            if (a.bytes) HIP_OK(hipMemcpyAsync(o + pos, a.p, a.bytes, hipMemcpyDeviceToHost, c->stream));
            pos += a.bytes;
        }
        HIP_OK(hipStreamSynchronize(c->stream));
        if (written) *written = pos;
        return MEL_OK;
    });
}

mel_status mel_load_state(mel_ctx* c, const void* buf, size_t size) {
    if (!c || !buf) return MEL_E_ARG;
    return guarded(c, [&]() {
        StateHeader h;
        if (size < sizeof h) { c->err = "mel_load_state: truncated state"; return MEL_E_PARSE; }
        std::memcpy(&h, buf, sizeof h);
        if (h.magic != STATE_MAGIC || h.n_nodes != c->d.n_nodes || h.n_bodies != c->d.n_bodies ||
            h.n_slots != c->d.n_slots) {
            c->err = "mel_load_state: the state is not one of this model";
            return MEL_E_PARSE;
        }
        if (h.n_contacts < 0 || h.n_contacts > c->d.c_cap) {
            c->err = "mel_load_state: " + std::to_string(h.n_contacts) + " contact elements exceed the capacity " +
                     std::to_string(c->d.c_cap);
            return MEL_E_CAPACITY;
        }
        const int nc = (int)h.n_contacts;
        HIP_OK(hipStreamSynchronize(c->stream));
        poll_async(*c, true);
        const std::vector<StateArray> arrs = state_arrays(*c, nc);
        if (h.n_arrays != arrs.size()) { c->err = "mel_load_state: the state is not one of this model"; return MEL_E_PARSE; }
        // check every record before writing any
        const unsigned char* in = static_cast<const unsigned char*>(buf);
        size_t pos = sizeof h;
        for (const StateArray& a : arrs) {
            uint64_t nb;
            if (pos + sizeof nb > size) { c->err = "mel_load_state: truncated state"; return MEL_E_PARSE; }
            std::memcpy(&nb, in + pos, sizeof nb);
            if (nb != a.bytes || pos + sizeof nb + nb > size) {
                c->err = "mel_load_state: a record does not match this model";
                return MEL_E_PARSE;
            }
            pos += sizeof nb + nb;
        }
        pos = sizeof h;
        for (const StateArray& a : arrs) {
            pos += sizeof(uint64_t);
            if (a.bytes) HIP_OK(hipMemcpyAsync(a.p, in + pos, a.bytes, hipMemcpyHostToDevice, c->stream));
            pos += a.bytes;
        }
        HIP_OK(hipMemcpyAsync(c->d.c_count, &nc, sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIP_OK(build_rigid_lists(c->d, c->stream, true));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->n_contacts = nc;
        c->nc_hint = nc;
        c->time = h.time;
        c->next_contact_update = h.next_contact_update;
        c->iteration = (int)h.iteration;
        c->d.dt = h.dt;
        c->outputs_stale = false;
        if (c->mon_on) {
            // rows carry the restored iteration numbers
            mon_drain(*c);
            const double it = c->iteration;
            HIP_OK(hipMemcpyAsync(c->d_mon_iter, &it, sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIP_OK(hipStreamSynchronize(c->stream));
        }
        return MEL_OK;
    });
}

mel_status mel_eval_forces_noalloc(mel_ctx* c, unsigned what) {
    if (!c || (what & ~(unsigned)MEL_EVAL_ALL)) return MEL_E_ARG;
    return guarded(c, [&]() {
        unsigned char* buf = static_cast<unsigned char*>(scratch_address(c->device));
        if (!buf) {
            c->err = "mel_eval_forces_noalloc: no device scratch on device " + std::to_string(c->device);
            return MEL_E_HIP;
        }
        HIP_OK(hipStreamSynchronize(c->stream));
        poll_async(*c, true);
        check_flags(*c);
        eval_forces(*c, what, buf);
        return MEL_OK;
    });
}

mel_status mel_get_border_field(mel_ctx* c, int body, int border, int field, double* out, size_t n) {
    if (!c || !out || body < 0 || body >= c->d.n_bodies) return MEL_E_ARG;
    if (border < 0 || border >= (int)c->model.bodies[body].borders.size() || field < 0 || field > 2) return MEL_E_ARG;
    int gb = c->h_border_off[body] + border;
    size_t ns = c->h_br_nslots[gb];
    if (n < ns) return MEL_E_ARG;
    return guarded(c, [&]() {
        if (field == 0) {
            HIP_OK(hipMemcpyAsync(out, c->d.sl_len + c->h_br_slot0[gb], ns * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIP_OK(hipStreamSynchronize(c->stream));
            return MEL_OK;
        }
        // Update_contact_pressures (Code/Proximity.h:1014-1069): evaluated on
        // request, in element order, from the elements of the last step: the
        // current ones, or the previous buffer when a proximity update has
        // replaced them since (the reference forms the pressures at the end
        // of Euler_step, before the contact update, Code/Solver.h:86-87)
        int nc = contact_count(*c);
        ContactArrays& A = c->prox_after_step ? c->bufB : c->bufA;
        if (c->prox_after_step) {
            HIP_OK(hipMemcpyAsync(&nc, c->d.c_count_old, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIP_OK(hipStreamSynchronize(c->stream));
        }
        std::vector<int> bS, brS, jS, bM, brM, seg;
        std::vector<double> fx, fy, len, N1, N2;
        c->download(bS, A.bodyS, nc); c->download(brS, A.borderS, nc); c->download(jS, A.bnodeS, nc);
        c->download(bM, A.bodyM, nc); c->download(brM, A.borderM, nc); c->download(seg, A.segM, nc);
        c->download(fx, A.fx, nc); c->download(fy, A.fy, nc); c->download(len, A.len, nc);
        c->download(N1, A.N1, nc); c->download(N2, A.N2, nc);
        std::vector<double> p(ns, 0.);
        for (int e = 0; e < nc; e++) {
            double f = field == 1 ? fx[e] : fy[e];
            double il = 1. / len[e];
            if (bS[e] == body && brS[e] == border) p[jS[e]] += f * il;
            if (bM[e] == body && brM[e] == border) {
                p[seg[e]] -= f * il * N1[e];
                p[seg[e] + 1] -= f * il * N2[e];
            }
        }
        std::copy(p.begin(), p.end(), out);
        return MEL_OK;
    });
}

mel_status mel_get_info(mel_ctx* c, mel_info_t* info) {
    if (!c || !info) return MEL_E_ARG;
    mel_status st = guarded(c, [&]() { contact_count(*c); return MEL_OK; });
    if (st != MEL_OK) return st;
    info->n_bodies = c->d.n_bodies;
    info->n_nodes = c->d.n_nodes;
    info->n_gauss = c->d.n_gauss;
    info->n_borders = c->d.n_borders;
    info->n_border_nodes = c->d.n_bnodes;
    info->n_rigid = c->d.n_rigid;
    info->n_contacts = c->n_contacts;
    info->contact_capacity = c->d.c_cap;
    info->nnz_gauss = c->nnz_g;
    info->nnz_d2d = c->nnz_d;
    info->time = c->time;
    info->dt = c->d.dt;
    info->iteration = c->iteration;
    info->precision = c->precision;
    return MEL_OK;
}

mel_status mel_write_dynamic(mel_ctx* c, const char* path) {
    if (!c || !path) return MEL_E_ARG;
    return guarded(c, [&]() {
        refresh_outputs(*c);
        Model& m = c->model;
        m.time = c->time;
        m.number_iteration = c->iteration;
        m.next_contact_update = c->next_contact_update;
        m.dt = c->d.dt;     // the adaptive schemes' current time step (Write_dynamic, Code/IO.h:2396)
        if (c->d.n_rules > 0) {
            // statuses the deactivation rules changed (Write_dynamic writes them, Code/IO.h:2429, 2481)
            std::vector<int> st;
            c->download(st, c->d.body_status, c->d.n_bodies);
            for (int b = 0; b < c->d.n_bodies; b++) m.bodies[b].status = st[b];
        }
        for (int b = 0; b < c->d.n_bodies; b++) {
            Body& B = m.bodies[b];
            size_t n = B.n_nodes;
            std::vector<double>* f[12] = {&B.ux, &B.uy, &B.vx, &B.vy, &B.ax, &B.ay, &B.upx, &B.upy, &B.vpx, &B.vpy, &B.apx, &B.apy};
            int fld[12] = {MEL_X_DISP, MEL_Y_DISP, MEL_X_VEL, MEL_Y_VEL, MEL_X_ACC, MEL_Y_ACC, MEL_X_DISP_PARAM,
                           MEL_Y_DISP_PARAM, MEL_X_VEL_PARAM, MEL_Y_VEL_PARAM, MEL_X_ACC_PARAM, MEL_Y_ACC_PARAM};
            for (int k = 0; k < 12; k++) c->download(*f[k], node_field_ptr(*c, fld[k]) + c->h_node_off[b], n);
            double r[26], w[7];
            mel_get_body_state(c, b, r, w);
            if (!B.rigid) {
                // internal, contact, self-contact (part of contact here), body,
                // Dirichlet, Neumann, damping, total, x then y each
                const int ff[16] = {MEL_X_INTERNAL_FORCE, MEL_Y_INTERNAL_FORCE, MEL_X_CONTACT_FORCE, MEL_Y_CONTACT_FORCE,
                                    -1, -1, MEL_X_BODY_FORCE, MEL_Y_BODY_FORCE, MEL_X_DIRICHLET_FORCE,
                                    MEL_Y_DIRICHLET_FORCE, MEL_X_NEUMANN_FORCE, MEL_Y_NEUMANN_FORCE,
                                    MEL_X_DAMPING_FORCE, MEL_Y_DAMPING_FORCE, MEL_X_FORCE, MEL_Y_FORCE};
                B.forces_out.assign(16 * n, 0.);
                std::vector<double> col;
                for (int k = 0; k < 16; k++) {
                    if (ff[k] < 0) continue;
                    c->download(col, node_field_ptr(*c, ff[k]) + c->h_node_off[b], n);
                    for (size_t i = 0; i < n; i++) B.forces_out[16 * i + k] = col[i];
                }
            } else {
                std::vector<double> R;
                c->download(R, c->d.body_rigid + (size_t)b * RIGID_FIELDS, RIGID_FIELDS);
                const int slot[6] = {R_FCON, R_FBODY, R_FDIR, R_FNEU, R_FDAMP, R_FTOT};
                for (int q = 0; q < 6; q++)
                    for (int k = 0; k < 3; k++) B.rigid_forces_out[3 * q + k] = R[slot[q] + k];
            }
            for (size_t j = 0; j < B.borders.size(); j++) {
                Border& Br = B.borders[j];
                const size_t ns = Br.nodes.size();
                const int gb = c->h_border_off[b] + (int)j;
                Br.length_out.resize(ns); Br.xcp_out.resize(ns); Br.ycp_out.resize(ns);
                mel_get_border_field(c, b, (int)j, 0, Br.length_out.data(), ns);
                mel_get_border_field(c, b, (int)j, 1, Br.xcp_out.data(), ns);
                mel_get_border_field(c, b, (int)j, 2, Br.ycp_out.data(), ns);
                c->download(Br.xbp_out, c->d.sl_xbp + c->h_br_slot0[gb], ns);
                c->download(Br.ybp_out, c->d.sl_ybp + c->h_br_slot0[gb], ns);
            }
            for (int q = 0; q < 3; q++) { B.rigid_cur[q] = r[q]; B.rigid_disp[q] = r[3 + q]; B.rigid_vel[q] = r[6 + q]; }
            for (int q = 0; q < 7; q++) B.works[q] = w[q];
            size_t cnt = 0;
            mel_get_neighbours(c, b, nullptr, 0, &cnt);
            B.neighbours.assign(3 * cnt, 0);
            mel_get_neighbours(c, b, B.neighbours.data(), cnt, &cnt);
            mel_get_contacts(c, b, nullptr, nullptr, 0, &cnt);
            std::vector<int32_t> ids(16 * cnt);
            std::vector<double> st(18 * cnt);
            mel_get_contacts(c, b, ids.data(), st.data(), cnt, &cnt);
            std::vector<double> iv(3 * cnt);
            mel_get_contact_internals(c, b, iv.data(), cnt, &cnt);
            B.contacts.assign(cnt, ContactElement());
            for (size_t e = 0; e < cnt; e++) {
                ContactElement& E = B.contacts[e];
                int32_t* o = &ids[16 * e];
                double* s = &st[18 * e];
                E.borderS = o[0]; E.border_nodeS = o[1]; E.nodeS = o[2]; E.bodyM = o[3]; E.borderM = o[4];
                E.shiftM = o[5]; E.segmentM = o[6];
                for (int k = 0; k < 4; k++) { E.nodeM[k] = o[7 + k]; E.shiftMk[k] = o[11 + k]; E.shapeM[k] = s[k]; }
                E.nb_internal = o[15];
                E.gapn = s[4]; E.vgapn = s[5]; E.gapt = s[6]; E.vgapt = s[7]; E.xsi = s[8]; E.xnorm = s[9];
                E.ynorm = s[10]; E.xtan = s[11]; E.ytan = s[12]; E.effective_mass = s[13]; E.length = s[14];
                E.fx = s[15]; E.fy = s[16];
                E.internal.assign(std::max(1, E.nb_internal), 0.);
                for (int k = 0; k < std::min(3, (int)E.internal.size()); k++) E.internal[k] = iv[3 * e + k];
            }
        }
        std::string err;
        if (!write_dynamic_text(path, m, err)) { c->err = err; return MEL_E_IO; }
        return MEL_OK;
    });
}

mel_status mel_set_graph_replay(mel_ctx* c, int on) {
    if (!c) return MEL_E_ARG;
    c->use_graphs = on != 0 && c->two_launch;
    return MEL_OK;
}

mel_status mel_get_device_bytes(const mel_ctx* c, size_t* bytes) {
    if (!c || !bytes) return MEL_E_ARG;
    *bytes = c->dev_bytes;
    return MEL_OK;
}

mel_status mel_get_stream(mel_ctx* c, void** stream) {
    if (!c || !stream) return MEL_E_ARG;
    *stream = (void*)c->stream;
    return MEL_OK;
}

mel_status mel_set_kernel_timing(mel_ctx* c, int on) {
    return guarded(c, [&]() {
        c->harvest();
        for (int k = 0; k < KT_COUNT; k++) { c->kt_ms[k] = 0.; c->kt_n[k] = 0; }
        c->timing = on != 0;
        c->timing_mask = on == -1 ? 0xffffffffu : (unsigned)on;
        return MEL_OK;
    });
}

mel_status mel_get_kernel_timing(mel_ctx* c, int kernel, double* total_ms, long long* launches) {
    if (!c || kernel < 0 || kernel >= KT_COUNT) return MEL_E_ARG;
    return guarded(c, [&]() {
        c->harvest();
        if (total_ms) *total_ms = c->kt_ms[kernel];
        if (launches) *launches = c->kt_n[kernel];
        return MEL_OK;
    });
}

int mel_kernel_count(void) { return KT_COUNT; }

const char* mel_kernel_name(int kernel) { return kernel >= 0 && kernel < KT_COUNT ? kKernelNames[kernel] : ""; }

void mel_destroy(mel_ctx* c) { delete c; }

const char* mel_last_error(const mel_ctx* c) { return c ? c->err.c_str() : g_load_err.c_str(); }

}  // extern "C"
