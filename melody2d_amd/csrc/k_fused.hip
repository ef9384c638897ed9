// The explicit step (Euler_step, Code/Solver.h:10-88) in two launches, for
// models whose deformable bodies each fit one node block (the build checks
// the preconditions of node_block_finish_lds).
//
// Launch A, workgroups by index range, all of BLOCK_THREADS threads:
//   [0, n_cwg)         contact elements (Update_contacts +
//                      Update_contact_forces, point-force records into the
//                      border nodes' buckets and the rigid bodies' records)
//                      and BC point forces of deformable border nodes
//                      (Update_bc_forces), then the BC terms of rigid
//                      borders run by run (rigid_bc_item); the first also
//                      forms the rigid bodies' driven DOFs at the end of the
//                      step
//   next n_blocks      one node block each: internal forces
//                      (Update_internal_forces), which depend on nothing else
//                      of the step
// No workgroup of launch A waits for another.
// Launch B:
//   [0, n_rigid)       one rigid body each: resultants of its contact records
//                      and of the BC terms launch A formed, Sum_up_forces,
//                      Apply_Newton, Apply_Euler with the driven DOFs launch
//                      A formed; signal [1]
//   next n_blocks      one node block each: bucket sums (the contact sends),
//                      Sum_up_forces + Apply_Newton + Apply_Euler with works,
//                      Update_borders and the border-node kinematics, with
//                      the block's point forces and new parameters staged in
//                      LDS
//   next               one rigid-body node per thread: wait [1]; its new
//                      position and velocity (Update_kinematics of rigid
//                      bodies), which only the next step reads
//   last               wait [1] (the rigid bodies read the BC values and the
//                      time); Update_bc(t + dt), time = t + dt
//
// A workgroup only ever waits for workgroups of lower index, which the
// hardware dispatches first, so every wait ends; a wait is also bounded in
// time: one that runs out sets bit 16 of the status flags (reported to the
// host as an error) instead of hanging.  The counter (DevModel::step_sync[1])
// is zero when launch B starts: the first thread of launch A clears it, when
// every waiter of the previous launch B has passed.  Publication: every
// thread fences its writes at device scope, the workgroup meets at a
// barrier, then one thread increments the counter.  A waiting workgroup
// polls with one thread, meets at a barrier, then every thread fences
// (acquire, device scope) before it reads what the producers wrote.
#include <hip/hip_runtime.h>
#include "mel_device.hpp"
#include "k_block.hpp"
#include "k_block_dev.hpp"
#include "k_step_dev.hpp"
#include "k_fused.hpp"

namespace mel {

namespace {

constexpr long long WAIT_LIMIT_TICKS = 20000000;   // 0.2 s of the 100 MHz constant clock

__device__ inline void wg_signal(unsigned* ctr) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void wg_wait(const DevModel& m, unsigned* ctr, unsigned target) {
    if (threadIdx.x == 0 && target > 0) {
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (wall_clock64() - t0 > WAIT_LIMIT_TICKS) {
                atomicOr(&m.status_flags[0], 16);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <class PR>
__device__ __forceinline__ void step_forces(DevModel& m, const BlockArrays& B, const StepLaunch& S,
                                            typename PR::slot* s_cxy) {
    const int t = threadIdx.x;
    int w = blockIdx.x;
    if (w >= B.n_blocks) {
        w -= B.n_blocks;
        // launch B of the previous step has ended: its counter starts again at zero
        if (w == 0 && t == 0) __hip_atomic_store(&m.step_sync[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the rigid bodies' driven DOFs at the end of the step (time and BC
        // tables only), ready for launch B
        if (w == 0)
            for (int k = t; k < m.n_rigid; k += BLOCK_THREADS) {
                double dv[8];
                rigid_dv_end_of_step(m, k, dv);
                for (int q = 0; q < 8; q++) m.rigid_dv[8 * k + q] = dv[q];
            }
        const int nc = *m.c_count;
        const int n_items = nc + B.n_bc_bn + m.n_rbc * BLOCK_THREADS;
        for (int i = w * BLOCK_THREADS + t; i < n_items; i += S.n_cwg * BLOCK_THREADS) {
            if (i < nc + B.n_bc_bn) contact_or_bc_item(m, i, nc, B.bc_bn, B.n_bc_bn);
            else rigid_bc_item(m, i - nc - B.n_bc_bn);
        }
        return;
    }
    double fi[2][NODES_PER_THREAD];
    block_force_update<PR>(m, B, w, s_cxy, fi);
}

}  // namespace

__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_step_forces(DevModel m, BlockArrays B, StepLaunch S) {
    __shared__ __align__(16) blockdev::Fp64::slot s_cxy[CHUNK_ENTRIES + 1];
    step_forces<blockdev::Fp64>(m, B, S, s_cxy);
}

__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_step_forces32(DevModel m, BlockArrays B, StepLaunch S) {
    __shared__ __align__(16) blockdev::Fp32::slot s_cxy[2 * (CHUNK_ENTRIES + 1)];
    step_forces<blockdev::Fp32>(m, B, S, s_cxy);
}

// two waves per SIMD (four workgroups per CU) keep a 1000-body model's
// blocks resident at once
__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_step_finish(DevModel m, BlockArrays B, int n_rwg) {
    extern __shared__ __align__(16) double s_fin[];
    unsigned* sync = m.step_sync;
    const int t = threadIdx.x;
    int w = blockIdx.x;
    if (w < m.n_rigid) {
        rigid_body_wg(m, w, s_fin, m.rigid_dv);
        wg_signal(&sync[1]);
        return;
    }
    w -= m.n_rigid;
    if (w < B.n_blocks) {
        double fi[2][NODES_PER_THREAD];
        const int4 h0 = B.blk_hdr[4 * w];
        const int n0 = h0.x, nown = h0.y;
#pragma unroll
        for (int q = 0; q < NODES_PER_THREAD; q++) {
            const int o = t + q * BLOCK_THREADS;
            fi[0][q] = o < nown ? m.fix[n0 + o] : 0.;
            fi[1][q] = o < nown ? m.fiy[n0 + o] : 0.;
        }
        node_block_finish_lds(m, B, w, fi, s_fin);
        return;
    }
    w -= B.n_blocks;
    if (w < n_rwg) {
        // the nodes q = w * BLOCK_THREADS + t + k * n_rwg * BLOCK_THREADS;
        // the first RIGID_NODES_PREFETCH of them are loaded before the wait
        // for their bodies
        const int stride = n_rwg * BLOCK_THREADS;
        RigidNodePre p[RIGID_NODES_PREFETCH];
#pragma unroll
        for (int u = 0; u < RIGID_NODES_PREFETCH; u++) {
            const int q = w * BLOCK_THREADS + t + u * stride;
            p[u] = RigidNodePre{};
            if (q < m.n_rigid_nodes) p[u] = rigid_node_pre(m, m.rigid_node_ids[q]);
        }
        wg_wait(m, &sync[1], m.n_rigid);
        // the rotation's cosine and sine once per body and thread
        int cb = -1;
        double c = 0., sn = 0.;
        auto place = [&](const RigidNodePre& r) {
            if (!r.on) return;
            const double* R = m.body_rigid + r.b * RIGID_FIELDS;
            if (r.b != cb) {
                const double dr = R[R_DISP + 2] - 0.;
                c = cos(dr);
                sn = sin(dr);
                cb = r.b;
            }
            rigid_node_at(m, r.i, R, c, sn, 1, r.X, r.Y);
        };
#pragma unroll
        for (int u = 0; u < RIGID_NODES_PREFETCH; u++) place(p[u]);
        for (int q = w * BLOCK_THREADS + t + RIGID_NODES_PREFETCH * stride; q < m.n_rigid_nodes; q += stride)
            place(rigid_node_pre(m, m.rigid_node_ids[q]));
        return;
    }
    // last workgroup: the time of the end of the step from the device's copy
    // of the time (the same sum the host forms), so that a launch does not
    // carry it and a captured run of steps can be replayed
    wg_wait(m, &sync[1], m.n_rigid);
    const double t_next = *m.time + m.dt;
    __syncthreads();           // every thread has read the time before it is replaced
    for (int k = t; k < m.n_bc_timed; k += BLOCK_THREADS) bc_eval_at(m, m.bc_timed[k], t_next);
    if (t == 0) *m.time = t_next;
}

}  // namespace mel
