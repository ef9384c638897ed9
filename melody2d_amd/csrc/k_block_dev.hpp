// Internal forces of one node block (device code of k_block.hip and the
// fused step).
//
// Replaces, per block of owned nodes (a whole grain when it fits):
//   Body::Update_internal_forces  Code/Body.h:937-1049 (+ Code/Model.h:19-73)
//   Node::Sum_up_forces           Code/Node.h:212-225
//   Node::Apply_Newton            Code/Node.h:227-231
//   Body::Apply_Euler (works)     Code/Body.h:2051-2068, Node::Apply_Euler Code/Node.h:239-245
//
// Data path per block: the displacement parameters of the touched nodes are
// staged in LDS once.  Chunk by chunk, every Gauss point is a lane that holds
// its own support entries (dphi/dx, dphi/dy, node index, destination slot)
// in registers, loaded with coalesced 16-byte loads from the wave's
// entry-major stream while the previous chunk is summed.  The lane gathers
// all its displacements from LDS at once, forms F and the stress terms
// (Code/Model.h:19-73) summing in the reference's entry order, and writes
// each entry's contribution c0 * (dphi . P) to the entry's slot in
// owned-node order.  Then each thread sums, for its owned nodes, their
// contiguous slots - a node's entries in Gauss-point order - per region when
// the reference order is requested.  A block is 128 threads and needs < 40 KB
// of LDS, so four blocks share a CU and a 1000-grain model is resident at
// once.  HBM traffic per entry: 16 (dphi/dx, dphi/dy) + 4 (index and slot),
// times the padding of each wave to its largest support.
#pragma once
#include "mel_device.hpp"
#include "k_block.hpp"

namespace mel {

namespace blockdev {

static_assert(BLOCK_THREADS == 128, "chunk descriptors describe two waves of lanes");
// native vector types (LLVM vectors) keep in-flight loads in registers
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// stress terms of Code/Model.h:19-73 for one Gauss point from its gradient
// sums (Code/Body.h:975-1014): linear material (a, b, c) = (E11, 2 E12, E22),
// neo-Hookean (a, b, c, d) = (F11, F12, F21, F22).  Written so that a support
// entry contributes c0 * (phix * P.x + phiy * P.y), c0 * (phix * P.z + phiy * P.w).
template <int MAT>
__device__ inline double4 stress_terms(double Mu, double K2, double a, double b, double c, double d) {
    if (MAT == 0) {
        double E11 = a, E12 = 0.5 * b, E22 = c;
        double J = 1. + E11 + E22;
        double S33 = K2 * (J - 1.);
        double M2 = 2 * Mu;
        double S11 = S33 + M2 * E11, S22 = S33 + M2 * E22, S12 = M2 * E12;
        return make_double4(S11, S12, S12, S22);
    }
    double F11 = a, F12 = b, F21 = c, F22 = d;
    double F11F12 = F11 * F11 + F12 * F12;
    double F21F22 = F21 * F21 + F22 * F22;
    double J = F11 * F22 - F12 * F21;
    double invJ = 1. / J;
    double K = K2 * (J - 1.);
    double MuJ53 = Mu * pow(J, -1.666666666667);
    double trB3 = 0.333333333333 * (F11F12 + F21F22 + 1.);
    double s11 = K + MuJ53 * (F11F12 - trB3);
    double s22 = K + MuJ53 * (F21F22 - trB3);
    double s12 = MuJ53 * (F11 * F21 + F12 * F22);
    double S11 = (F12 * (F12 * s22 - F22 * s12) - F22 * (F12 * s12 - F22 * s11)) * invJ;
    double S22 = (F11 * (F11 * s22 - F21 * s12) - F21 * (F11 * s12 - F21 * s11)) * invJ;
    double S12 = (F21 * (F12 * s12 - F22 * s11) - F11 * (F12 * s22 - F22 * s12)) * invJ;
    return make_double4(F11 * S11 + F12 * S12, F12 * S22 + F11 * S12, F21 * S11 + F22 * S12, F22 * S22 + F21 * S12);
}

// one support entry added in the reference's entry order; a padding entry
// has zero derivatives and adds +-0, which leaves every sum unchanged (none
// of them is ever -0)
template <int MAT>
__device__ inline void grad_add(double& a, double& b, double& c, double& d, f64x2 f, f64x2 w) {
    const double u = w.x, v = w.y;
    if (MAT == 0) {
        a += f.x * u;
        b += f.x * v + f.y * u;
        c += f.y * v;
    } else {
        a += f.x * u;
        b += f.y * u;
        c += f.x * v;
        d += f.y * v;
    }
}

__device__ inline double2 contribution(double c0, f64x2 f, const double4& P) {
    return make_double2(c0 * (f.x * P.x + f.y * P.y), c0 * (f.x * P.z + f.y * P.w));
}

// ---- FP32-mixed form --------------------------------------------------
// The gradient is summed as H = F - I from anchored displacements: the
// displacement of the Gauss point's first support node is subtracted from
// every support node's in FP64 before rounding to FP32, so the FP32 terms are
// of the size of the displacement differences across one support (an anchor
// further away, such as a node block's first node, makes them larger than
// their sum by the ratio of the distances and the gradient loses as many
// digits); and
// the stress terms are rewritten in H so that no step takes a small
// difference of O(1) numbers: J - 1 = H11 + H22 + H11 H22 - H12 H21,
// F11^2 + F12^2 - 1 = 2 H11 + H11^2 + H12^2, ... .  The reference's literal
// 0.333333333333 leaves an offset 1 - 3 * 0.333333333333 in F11F12 - trB3,
// which is kept.
template <int MAT>
__device__ inline float4 stress_terms32(float Mu, float K2, float a, float b, float c, float d) {
    if (MAT == 0) {
        float E11 = a, E12 = 0.5f * b, E22 = c;
        float S33 = K2 * (E11 + E22);
        float M2 = 2.f * Mu;
        float S11 = S33 + M2 * E11, S22 = S33 + M2 * E22, S12 = M2 * E12;
        return make_float4(S11, S12, S12, S22);
    }
    const float H11 = a, H12 = b, H21 = c, H22 = d;
    const float jm1 = H11 + H22 + (H11 * H22 - H12 * H21);
    const float invJ = 1.f / (1.f + jm1);
    const float K = K2 * jm1;
    const float MuJ53 = Mu * expf(-1.666666666667f * log1pf(jm1));
    const float A = fmaf(H11, H11, fmaf(H12, H12, 2.f * H11));     // F11^2 + F12^2 - 1
    const float Bq = fmaf(H22, H22, fmaf(H21, H21, 2.f * H22));    // F21^2 + F22^2 - 1
    const float third = 0.333333333333f;
    const float off = (float)(1.0 - 3.0 * 0.333333333333);
    const float tsum = third * (A + Bq);
    const float s11 = K + MuJ53 * ((A - tsum) + off);
    const float s22 = K + MuJ53 * ((Bq - tsum) + off);
    const float s12 = MuJ53 * (H21 + H12 + fmaf(H11, H21, H12 * H22));
    const float F11 = 1.f + H11, F12 = H12, F21 = H21, F22 = 1.f + H22;
    const float S11 = (F12 * (F12 * s22 - F22 * s12) - F22 * (F12 * s12 - F22 * s11)) * invJ;
    const float S22 = (F11 * (F11 * s22 - F21 * s12) - F21 * (F11 * s12 - F21 * s11)) * invJ;
    const float S12 = (F21 * (F12 * s12 - F22 * s11) - F11 * (F12 * s22 - F22 * s12)) * invJ;
    return make_float4(F11 * S11 + F12 * S12, F12 * S22 + F11 * S12, F21 * S11 + F22 * S12, F22 * S22 + F21 * S12);
}

template <int MAT>
__device__ inline void grad_add32(float& a, float& b, float& c, float& d, f32x2 f, f32x2 w) {
    const float u = w.x, v = w.y;
    if (MAT == 0) {
        a = fmaf(f.x, u, a);
        b = fmaf(f.x, v, fmaf(f.y, u, b));
        c = fmaf(f.y, v, c);
    } else {
        a = fmaf(f.x, u, a);
        b = fmaf(f.y, u, b);
        c = fmaf(f.x, v, c);
        d = fmaf(f.y, v, d);
    }
}

__device__ inline float2 contribution32(float c0, f32x2 f, const float4& P) {
    return make_float2(c0 * fmaf(f.x, P.x, f.y * P.y), c0 * fmaf(f.x, P.z, f.y * P.w));
}

// precision policies of the block kernel.  nt_f / nt_pk: the derivative and
// index streams are read with non-temporal loads (read once per step, they
// then do not displace the node data the rest of the step reads from the
// caches; measured per form)
struct Fp64 {
    static constexpr bool f32 = false;
    static constexpr bool dbuf = false;     // one chunk's entries in flight (registers)
    static constexpr bool pipe = false;     // one contribution buffer
    static constexpr bool nt_f = true, nt_pk = false;
    static constexpr int KREG = KREG_ENTRIES;
    static constexpr int GATHER = KREG_ENTRIES / 2;   // displacement gathers in flight per batch
    typedef f64x2 vec;
    typedef double2 slot;
    typedef double2 disp;
    typedef double real;
};
struct Fp32 {
    static constexpr bool f32 = true;
    static constexpr bool dbuf = true;      // two chunks' entries in flight
    static constexpr bool pipe = true;      // two contribution buffers (2 * (CHUNK_ENTRIES + 1) slots)
    static constexpr bool nt_f = true, nt_pk = false;
    static constexpr int KREG = KREG_ENTRIES_F32;
    static constexpr int GATHER = KREG_ENTRIES_F32 / 2;
    typedef f32x2 vec;
    typedef float2 slot;
    typedef double2 disp;     // staged in FP64: each Gauss point anchors its own differences
    typedef float real;
};
// wide register windows, for models whose Gauss points have more support
// entries than the windows above (RPIM supports of 30 nodes): entries past
// the window are read from memory one dependent load after another, which
// is what a window that holds them avoids
struct Fp64W : Fp64 {
    static constexpr int KREG = KREG_ENTRIES_WIDE;
    static constexpr int GATHER = 8;
};
struct Fp32W : Fp32 {
    static constexpr bool dbuf = false;     // one chunk's entries in flight: two wide windows do not fit
    static constexpr bool pipe = false;
    static constexpr int KREG = KREG_ENTRIES_WIDE;
    static constexpr int GATHER = 8;
};

// a value that is the same in every lane of the wave, moved to a scalar
// register: loops and branches on it (a wave's support size, its lane count)
// are then scalar branches instead of per-entry lane masks, and entry
// addresses are a scalar base plus the lane's 32-bit offset
__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// where a lane's entries live in the chunk's stream
struct LaneMap {
    int valid;      // lane holds a Gauss point
    int lane;       // lane index within the chunk
    int K;          // the wave's largest support: entries per lane (wave-uniform)
    int stride;     // lanes in the wave (wave-uniform)
    size_t wbase;   // entry 0 of the wave's lane 0 (wave-uniform)
    unsigned lw;    // lane within the wave
};

// entry i of the lane (wbase + i * stride + lw)
template <class T>
__device__ __forceinline__ const T& lane_entry(const T* s, const LaneMap& L, int i) {
    return (s + (L.wbase + (size_t)i * L.stride))[L.lw];
}

__device__ inline LaneMap lane_map(int4 meta, int t) {
    const int w = wave_uniform(t >> 6);
    const unsigned lw = (unsigned)t & 63u;
    const int mw = wave_uniform(meta.w);
    const int K0 = mw & 255, K1 = (mw >> 8) & 255, n0 = (mw >> 16) & 255, n1 = (mw >> 24) & 255;
    LaneMap L;
    L.stride = w ? n1 : n0;
    L.K = w ? K1 : K0;
    L.valid = (int)lw < L.stride;
    L.lane = 64 * w + (int)lw;
    L.wbase = (size_t)(unsigned)wave_uniform(meta.z) + (w ? (size_t)K0 * n0 : 0);
    L.lw = lw;
    return L;
}

// one lane's chunk data in flight, held in registers until its Gauss-point work
template <class PR>
struct LaneRegs {
    typename PR::vec f[PR::KREG];
    unsigned pk[PR::KREG];
    double c0;      // kept as loaded: a use before the chunk's barrier would wait for every load in flight
    int reg;
    int tp0[NODES_PER_THREAD], tp1[NODES_PER_THREAD];
};

template <class PR>
__device__ inline const typename PR::vec* entry_stream(const BlockArrays& B) {
    if constexpr (PR::f32) return reinterpret_cast<const f32x2*>(B.be_fxy32);
    else return reinterpret_cast<const f64x2*>(B.be_fxy);
}

template <class PR>
__device__ inline void load_lane(const BlockArrays& B, int4 meta, int2 po, int nown, int t, LaneRegs<PR>& R) {
    const LaneMap L = lane_map(meta, t);
    const typename PR::vec* fxy = entry_stream<PR>(B);
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) {
        int o = t + q * BLOCK_THREADS;
        if (o < nown) {
            R.tp0[q] = B.tptr[po.y + o];
            R.tp1[q] = B.tptr[po.y + o + 1];
        }
    }
    if (L.valid) {
        R.c0 = B.bgp_c0[meta.x + L.lane];
        R.reg = B.bgp_region[meta.x + L.lane];
#pragma unroll
        for (int i = 0; i < PR::KREG; i++) {
            if (i < L.K) {
                if constexpr (PR::nt_f) R.f[i] = __builtin_nontemporal_load(&lane_entry(fxy, L, i));
                else R.f[i] = lane_entry(fxy, L, i);
                if constexpr (PR::nt_pk) R.pk[i] = __builtin_nontemporal_load(&lane_entry(B.be_pk, L, i));
                else R.pk[i] = lane_entry(B.be_pk, L, i);
            }
        }
    }
}

// the Gauss-point work of one lane: gradient, stress, contributions to slots
// (PR::KREG entries from registers, in batches of PR::GATHER displacement
// gathers; entries past the window from memory)
template <int MAT, class PR>
__device__ inline void gp_lane(const BlockArrays& B, const LaneMap& L, const LaneRegs<PR>& R, const f64x2* uv,
                               double2* s_cxy, uint16_t* s_creg, bool single, double Mu, double K2) {
    constexpr int KR = PR::KREG, GB = PR::GATHER;
    double a = MAT == 0 ? 0. : 1., b = 0., c = 0., d = MAT == 0 ? 0. : 1.;
    f64x2 u[GB];
#pragma unroll
    for (int j0 = 0; j0 < KR; j0 += GB) {
#pragma unroll
        for (int i = 0; i < GB; i++)
            if (j0 + i < KR && j0 + i < L.K) u[i] = uv[R.pk[j0 + i] & 0xFFFF];
#pragma unroll
        for (int i = 0; i < GB; i++)
            if (j0 + i < KR && j0 + i < L.K) grad_add<MAT>(a, b, c, d, R.f[j0 + i], u[i]);
    }
    // supports longer than the register window: straight from memory
    const f64x2* fxy = reinterpret_cast<const f64x2*>(B.be_fxy);
    for (int i = KR; i < L.K; i++)
        grad_add<MAT>(a, b, c, d, lane_entry(fxy, L, i), uv[lane_entry(B.be_pk, L, i) & 0xFFFF]);
    const double4 P = stress_terms<MAT>(Mu, K2, a, b, c, d);
#pragma unroll
    for (int i = 0; i < KR; i++)
        if (i < L.K) s_cxy[R.pk[i] >> 16] = contribution(R.c0, R.f[i], P);
    for (int i = KR; i < L.K; i++)
        s_cxy[lane_entry(B.be_pk, L, i) >> 16] = contribution(R.c0, lane_entry(fxy, L, i), P);
    if (!single) {
#pragma unroll
        for (int i = 0; i < KR; i++)
            if (i < L.K) s_creg[R.pk[i] >> 16] = (uint16_t)R.reg;
        for (int i = KR; i < L.K; i++)
            s_creg[lane_entry(B.be_pk, L, i) >> 16] = (uint16_t)R.reg;
    }
}

// displacement difference to the Gauss point's anchor, formed in FP64 and
// rounded once
__device__ __forceinline__ f32x2 anchored(f64x2 u, f64x2 a) {
    f32x2 r;
    r.x = (float)(u.x - a.x);
    r.y = (float)(u.y - a.y);
    return r;
}

// the same in the FP32-mixed form (H sums from displacements anchored at the
// Gauss point's first support node)
template <int MAT, class PR>
__device__ inline void gp_lane32(const BlockArrays& B, const LaneMap& L, const LaneRegs<PR>& R, const f64x2* uv,
                                 float2* s_cxy, uint16_t* s_creg, bool single, float Mu, float K2) {
    constexpr int KR = PR::KREG, GB = PR::GATHER;
    float a = 0.f, b = 0.f, c = 0.f, d = 0.f;
    f64x2 u[GB];
    const f64x2 u0 = uv[R.pk[0] & 0xFFFF];
#pragma unroll
    for (int j0 = 0; j0 < KR; j0 += GB) {
#pragma unroll
        for (int i = 0; i < GB; i++)
            if (j0 + i < KR && j0 + i < L.K) u[i] = uv[R.pk[j0 + i] & 0xFFFF];
#pragma unroll
        for (int i = 0; i < GB; i++)
            if (j0 + i < KR && j0 + i < L.K) grad_add32<MAT>(a, b, c, d, R.f[j0 + i], anchored(u[i], u0));
    }
    const f32x2* fxy = reinterpret_cast<const f32x2*>(B.be_fxy32);
    for (int i = KR; i < L.K; i++)
        grad_add32<MAT>(a, b, c, d, lane_entry(fxy, L, i), anchored(uv[lane_entry(B.be_pk, L, i) & 0xFFFF], u0));
    const float4 P = stress_terms32<MAT>(Mu, K2, a, b, c, d);
    const float c0 = (float)R.c0;
#pragma unroll
    for (int i = 0; i < KR; i++)
        if (i < L.K) s_cxy[R.pk[i] >> 16] = contribution32(c0, R.f[i], P);
    for (int i = KR; i < L.K; i++)
        s_cxy[lane_entry(B.be_pk, L, i) >> 16] = contribution32(c0, lane_entry(fxy, L, i), P);
    if (!single) {
#pragma unroll
        for (int i = 0; i < KR; i++)
            if (i < L.K) s_creg[R.pk[i] >> 16] = (uint16_t)R.reg;
        for (int i = KR; i < L.K; i++)
            s_creg[lane_entry(B.be_pk, L, i) >> 16] = (uint16_t)R.reg;
    }
}

}  // namespace blockdev

// Internal forces of the owned nodes of node block blk, written to fix, fiy
// and returned in fi[0][q], fi[1][q] for the thread's owned node
// o = threadIdx.x + q * BLOCK_THREADS.  s_cxy: the caller's LDS array of
// CHUNK_ENTRIES + 1 contribution slots (the last one is the sink), twice
// that when PR::pipe.
template <class PR>
__device__ __forceinline__ void block_force_update(DevModel& m, const BlockArrays& B, int blk,
                                                   typename PR::slot* s_cxy, double (&fi)[2][NODES_PER_THREAD]) {
    using namespace blockdev;
    typedef typename PR::disp disp_t;
    typedef typename PR::slot slot_t;
    extern __shared__ __align__(16) unsigned char s_dyn[];   // (ux, uy) of the touched nodes [, slot regions]
    __shared__ int4 s_meta[MAX_BLOCK_CHUNKS];
    __shared__ int2 s_po[MAX_BLOCK_CHUNKS];
    const int t = threadIdx.x;
    // the block's header: one round trip before the first chunk's loads
    const int4 h0 = B.blk_hdr[4 * blk], h1 = B.blk_hdr[4 * blk + 1], meta0 = B.blk_hdr[4 * blk + 2];
    const int4 h3 = B.blk_hdr[4 * blk + 3];
    const double2 par = B.blk_par[blk];
    const int flags = h1.z;
    if (flags & BLK_INACTIVE) return;   // inactive body: no update (uniform over the block)
    const bool single = (flags & BLK_SINGLE) != 0;
    const int own0 = h0.x, nown = h0.y, ntouch = h0.z;
    const bool own_contig = (flags & BLK_OWN_CONTIGUOUS) != 0;
    const int* own = B.own_nodes + h1.w;
    const int mtype = (flags & BLK_NEO_HOOKEAN) ? 1 : 0;
    const double Mu = par.x, K2 = par.y;
    disp_t* s_uv = reinterpret_cast<disp_t*>(s_dyn);
    uint16_t* s_creg = reinterpret_cast<uint16_t*>(s_dyn + sizeof(disp_t) * B.max_touch);

    const int nch = h1.y;
    const int4* gmeta = B.ch_meta + h1.x;
    const int2* gpo = B.ch_ptr_off + h1.x;
    LaneRegs<PR> R;
    if (nch > 0) load_lane<PR>(B, meta0, make_int2(h3.x, h3.y), nown, t, R);
    // chunk descriptors in LDS: read without waiting on the entry stream
    for (int i = t; i < nch; i += BLOCK_THREADS) {
        s_meta[i] = gmeta[i];
        s_po[i] = gpo[i];
    }
    // displacement parameters of the touched nodes (a contiguous range when
    // they are the owned nodes: no index load first)
    const int* touch = h0.w >= 0 ? B.touch_nodes + h0.w : nullptr;
    for (int i = t; i < ntouch; i += BLOCK_THREADS) {
        int g = touch ? touch[i] : own0 + i;
        s_uv[i] = make_double2(m.upx[g], m.upy[g]);
    }
    double tx[NODES_PER_THREAD], ty[NODES_PER_THREAD], rx[NODES_PER_THREAD], ry[NODES_PER_THREAD];
    int cur[NODES_PER_THREAD];
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) { tx[q] = ty[q] = rx[q] = ry[q] = 0.; cur[q] = -1; }

    // one chunk with its entries in R; R is refilled with chunk nxt once its
    // Gauss-point work is done
    auto chunk = [&](int c, LaneRegs<PR>& R, int nxt) {
        __syncthreads();      // descriptors and displacements staged; previous chunk's slots read
        const int4 meta = s_meta[c];
        const LaneMap L = lane_map(meta, t);
        if (L.valid) {
            if constexpr (PR::f32) {
                const f64x2* uv = reinterpret_cast<const f64x2*>(s_uv);
                if (mtype == 0) gp_lane32<0>(B, L, R, uv, s_cxy, s_creg, single, (float)Mu, (float)K2);
                else gp_lane32<1>(B, L, R, uv, s_cxy, s_creg, single, (float)Mu, (float)K2);
            } else {
                const f64x2* uv = reinterpret_cast<const f64x2*>(s_uv);
                if (mtype == 0) gp_lane<0>(B, L, R, uv, s_cxy, s_creg, single, Mu, K2);
                else gp_lane<1>(B, L, R, uv, s_cxy, s_creg, single, Mu, K2);
            }
        }
        int tp0[NODES_PER_THREAD], tp1[NODES_PER_THREAD];
#pragma unroll
        for (int q = 0; q < NODES_PER_THREAD; q++) { tp0[q] = R.tp0[q]; tp1[q] = R.tp1[q]; }
        if (nxt < nch) load_lane<PR>(B, s_meta[nxt], s_po[nxt], nown, t, R);
        __syncthreads();      // contributions visible
        if (single) {
#pragma unroll
            for (int q = 0; q < NODES_PER_THREAD; q++) {
                int o = t + q * BLOCK_THREADS;
                if (o >= nown) break;
                if (tp1[q] > tp0[q]) cur[q] = 0;
#pragma unroll 4
                for (int p = tp0[q]; p < tp1[q]; p++) {
                    slot_t v = s_cxy[p];
                    rx[q] -= (double)v.x;
                    ry[q] -= (double)v.y;
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < NODES_PER_THREAD; q++) {
                int o = t + q * BLOCK_THREADS;
                if (o >= nown) break;
                for (int p = tp0[q]; p < tp1[q]; p++) {
                    slot_t v = s_cxy[p];
                    int r = s_creg[p];
                    if (r != cur[q]) {
                        if (cur[q] >= 0) { tx[q] += rx[q]; ty[q] += ry[q]; }
                        rx[q] = 0.; ry[q] = 0.; cur[q] = r;
                    }
                    rx[q] -= (double)v.x;
                    ry[q] -= (double)v.y;
                }
            }
        }
    };
    if constexpr (PR::pipe) {
        // two contribution buffers: chunk c's Gauss points write buffer
        // c & 1 while the slots of chunk c - 1 are summed from the other,
        // with one barrier per chunk; each node's slots are still summed
        // chunk by chunk, in their order.  Single-region blocks only (the
        // FP32-mixed form always builds them).
        LaneRegs<PR> R2;
        if (nch > 1) load_lane<PR>(B, gmeta[1], gpo[1], nown, t, R2);
        int ptp0[NODES_PER_THREAD], ptp1[NODES_PER_THREAD];
        auto sum = [&](int c) {
            const slot_t* buf = s_cxy + (c & 1) * (CHUNK_ENTRIES + 1);
#pragma unroll
            for (int q = 0; q < NODES_PER_THREAD; q++) {
                int o = t + q * BLOCK_THREADS;
                if (o >= nown) break;
                if (ptp1[q] > ptp0[q]) cur[q] = 0;
#pragma unroll 4
                for (int p = ptp0[q]; p < ptp1[q]; p++) {
                    slot_t v = buf[p];
                    rx[q] -= (double)v.x;
                    ry[q] -= (double)v.y;
                }
            }
        };
        auto gp = [&](int c, LaneRegs<PR>& Rc) {
            __syncthreads();      // chunk c - 1's contributions visible; buffer c & 1 summed
            const int4 meta = s_meta[c];
            const LaneMap L = lane_map(meta, t);
            slot_t* buf = s_cxy + (c & 1) * (CHUNK_ENTRIES + 1);
            if (L.valid) {
                const f64x2* uv = reinterpret_cast<const f64x2*>(s_uv);
                if (mtype == 0) gp_lane32<0>(B, L, Rc, uv, buf, s_creg, true, (float)Mu, (float)K2);
                else gp_lane32<1>(B, L, Rc, uv, buf, s_creg, true, (float)Mu, (float)K2);
            }
            int tp0[NODES_PER_THREAD], tp1[NODES_PER_THREAD];
#pragma unroll
            for (int q = 0; q < NODES_PER_THREAD; q++) { tp0[q] = Rc.tp0[q]; tp1[q] = Rc.tp1[q]; }
            if (c + 2 < nch) load_lane<PR>(B, s_meta[c + 2], s_po[c + 2], nown, t, Rc);
            if (c > 0) sum(c - 1);
#pragma unroll
            for (int q = 0; q < NODES_PER_THREAD; q++) { ptp0[q] = tp0[q]; ptp1[q] = tp1[q]; }
        };
        for (int c = 0; c < nch; c += 2) {
            gp(c, R);
            if (c + 1 < nch) gp(c + 1, R2);
        }
        if (nch > 0) {
            __syncthreads();      // the last chunk's contributions visible
            sum(nch - 1);
        }
    } else if constexpr (PR::dbuf) {
        // chunk c + 1's entries load during the whole of chunk c
        LaneRegs<PR> R2;
        if (nch > 1) load_lane<PR>(B, gmeta[1], gpo[1], nown, t, R2);
        for (int c = 0; c < nch; c += 2) {
            chunk(c, R, c + 2);
            if (c + 1 < nch) chunk(c + 1, R2, c + 3);
        }
    } else {
        // the next chunk's entries load while this chunk is summed
        for (int c = 0; c < nch; c++) chunk(c, R, c + 1);
    }

    // internal force of the owned nodes (Node::Sum_up_forces region sum)
#pragma unroll
    for (int q = 0; q < NODES_PER_THREAD; q++) {
        int o = t + q * BLOCK_THREADS;
        if (o >= nown) break;
        int j = own_contig ? own0 + o : own[o];
        if (cur[q] >= 0) { tx[q] += rx[q]; ty[q] += ry[q]; }
        m.fix[j] = tx[q];
        m.fiy[j] = ty[q];
        fi[0][q] = tx[q];
        fi[1][q] = ty[q];
    }
}


}  // namespace mel
