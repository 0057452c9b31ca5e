// Per-vertex gather over incident (element, quadrature point) pairs: the hot
// kernel of VBD (one launch per colour, reference kernel_vbd_solve_v3/v4,
// src/GPU/VBD/GPU_VBD_FEM.cu:234-400) and of the explicit integrator's internal
// force pass (reference kernel_fem_eval_force,
// src/GPU/Explicit/GPU_FEM_Explicit.cu:5-64).
//
// Work decomposition (CDNA4, wave64): a group of L consecutive lanes owns one
// vertex.  The vertex's incident elements are split into work items of QL
// quadrature points; lane t handles items t, t+L, ...  A work item loads the S
// node positions of its element once (one 16-byte load each) and reuses them
// for its QL quadrature points.  Each lane accumulates the force and the
// 3x3 Hessian block in registers; the group then reduces with cross-lane
// shuffles (no LDS, no barrier) and its first lane solves the 3x3 system in
// double precision.  L and QL are tile parameters chosen per element type.
//
// Cached form (CACHED = true, one quadrature point per work item): the
// displacement gradient Ju(e, q) = sum_k u_k (x) dN_k(q) is read from a
// per-(element, quadrature point) array instead of being rebuilt from the S
// node positions (S loads and 9 S FMAs per point), and after the solve the
// group adds du (x) dN_r(q) to the records of the vertex's elements.  The
// array is rebuilt from the positions at the start of every iteration
// (k_ju_refresh), after the predictor, constraints and any other position
// change.  Within one colour each element has at most one vertex being
// solved, so each record has a single reader-writer per launch.
//
// This file is compiled once per element type (OE_GATHER_S / OE_GATHER_Q set
// on the command line) so that the template instances build in parallel.
#include "oe_gather.hpp"
#include "oe_host.hpp"
#include "oe_kernels.hpp"

namespace oe {

// Work items of a lane whose gradient records stay in registers between the
// gather and the update in the cached form: enough for vertices with up to 4
// incident elements (ceil(4 Q / L), at most 2); further items are re-read.
template <int Q, int L>
constexpr int kept_items() {
    return (4 * Q + L - 1) / L < 2 ? (4 * Q + L - 1) / L : 2;
}

// (e, q, local node) of work item t of a vertex whose owner records start at own
__device__ __forceinline__ void work_item(const int* __restrict__ own, int t, int QC, int& eid, int& qc, int& r) {
    const int j = t / QC;
    qc = t - j * QC;
    const int pk = gload(own, j);
    eid = pk >> 5;
    r = pk & 31;
}

// dN_r at one point, from the shape-derivative table in device memory
__device__ __forceinline__ void load_dnr(const float* dn, int at, float dnr[3]) {
    dnr[0] = gload(dn, at);
    dnr[1] = gload(dn, at + 1);
    dnr[2] = gload(dn, at + 2);
}

// column-major: Ju[c*3 + k] += du[k] dN_r[c]
__device__ __forceinline__ void rank1_update(float* g, const float Ju[9], const float du[3],
                                             const float* __restrict__ dn) {
    const float d0 = dn[0], d1 = dn[1], d2 = dn[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        g[k] = fmaf(du[k], d0, Ju[k]);
        g[3 + k] = fmaf(du[k], d1, Ju[3 + k]);
        g[6 + k] = fmaf(du[k], d2, Ju[6 + k]);
    }
}

// Gather + local solve of one vertex slot by a group of L lanes.
template <int S, int Q, int L, int QL, bool CACHED>
__device__ __forceinline__ void gather_body(const GatherArgs& a, int grp, int lane) {
    static_assert(!CACHED || QL == 1, "cached gradients take one quadrature point per work item");
    constexpr int QC = Q / QL;
    constexpr int KEEP = CACHED ? kept_items<Q, L>() : 1;
    if (grp >= a.count) return;                 // whole group leaves together
    const int slot = a.first_slot + grp;
    const int vid = gload(a.own.slot_vid, slot);
    if (a.mode != GATHER_EXPLICIT_FORCE && gload(a.ps.mask, vid) == 0) return;
    const int2 nbo = gload(a.own.slot_nbo, slot);
    const int nwork = nbo.x * QC;
    const int* __restrict__ own = a.own.own + nbo.y;
    Acc9 acc;
    acc_zero(acc);
    const float4* __restrict__ P = a.ps.p;
    float keep[KEEP][9];
    if constexpr (CACHED) {
#pragma unroll
        for (int it = 0; it < KEEP; ++it) {
            const int t = lane + it * L;
            if (t < nwork) {
                int eid, q, r;
                work_item(own, t, Q, eid, q, r);
                const size_t g = (static_cast<size_t>(eid) * Q + q) * 9;
                float JXi[9], V;
#pragma unroll
                for (int i = 0; i < 9; ++i) keep[it][i] = gload(a.ju, g + i);
                load_element_record(a.fem, eid, q, JXi, V);
                float dnr[3];
                load_dnr(a.fem.dN, (q * S + r) * 3, dnr);
                nh_contrib(keep[it], JXi, V, dnr, a.mt, acc);
            }
        }
        for (int t = lane + KEEP * L; t < nwork; t += L) {
            int eid, q, r;
            work_item(own, t, Q, eid, q, r);
            const size_t g = (static_cast<size_t>(eid) * Q + q) * 9;
            float Ju[9], JXi[9], V;
#pragma unroll
            for (int i = 0; i < 9; ++i) Ju[i] = gload(a.ju, g + i);
            load_element_record(a.fem, eid, q, JXi, V);
            float dnr[3];
            load_dnr(a.fem.dN, (q * S + r) * 3, dnr);
            nh_contrib(Ju, JXi, V, dnr, a.mt, acc);
        }
    } else {
        for (int t = lane; t < nwork; t += L) {
            int eid, qc, r;
            work_item(own, t, QC, eid, qc, r);
            const size_t tp = static_cast<size_t>(eid) * S;
            float4 h[S];
#pragma unroll
            for (int k = 0; k < S; ++k) h[k] = gload(static_cast<const float4*>(P), gload(a.fem.topo, tp + k));
            float xs[S][3];
#pragma unroll
            for (int k = 0; k < S; ++k) rel_node(h[k], h[0], a.ps.lo_unit, xs[k]);
#pragma unroll
            for (int qq = 0; qq < QL; ++qq) {
                const int q = qc * QL + qq;
                const float* __restrict__ dn = a.fem.dN + q * S * 3;
                float Ju[9], JXi[9], V;
                load_element_record(a.fem, eid, q, JXi, V);
                jx_from_nodes<S>(xs, dn, Ju);
                float dnr[3];
                load_dnr(dn, 3 * r, dnr);
                nh_contrib(Ju, JXi, V, dnr, a.mt, acc);
            }
        }
    }
    group_sum<L>(acc);
    if (a.mode == GATHER_EXPLICIT_FORCE) {
        if (lane != 0) return;
        // f = sum(-V P g) - k H v  (damping uses the vertex's own block)
        float4 f = make_float4(acc.v[0], acc.v[1], acc.v[2], 0.f);
        if (a.damping != 0.f) {
            const float4 v = a.ps.v[vid];
            f.x -= a.damping * (acc.v[3] * v.x + acc.v[4] * v.y + acc.v[5] * v.z);
            f.y -= a.damping * (acc.v[4] * v.x + acc.v[6] * v.y + acc.v[7] * v.z);
            f.z -= a.damping * (acc.v[5] * v.x + acc.v[7] * v.y + acc.v[8] * v.z);
        }
        a.ps.f[vid] = f;
        return;
    }
    if (a.mode == GATHER_VBD_RESIDUAL) {
        if (lane == 0) {
            const float u = a.ps.lo_unit;
            const float4 rv = vbd_residual(acc, dunpack(gload(static_cast<const float4*>(P), vid), u),
                                           dunpack(gload(static_cast<const float4*>(a.ps.last_p), vid), u),
                                           dunpack(gload(a.y, vid), u), a.ps.m[vid], a.damping, a.dt);
            if (a.out3) {
                a.out3[3 * static_cast<size_t>(vid)] = rv.x;
                a.out3[3 * static_cast<size_t>(vid) + 1] = rv.y;
                a.out3[3 * static_cast<size_t>(vid) + 2] = rv.z;
            } else {
                a.p_new[vid] = rv;
            }
        }
        return;
    }
    float du[3] = {0.f, 0.f, 0.f};
    if (lane == 0) {
        const float4 pv = gload(static_cast<const float4*>(P), vid);
        const float un = a.ps.lo_unit;
        const dvec3 p = dunpack(pv, un);
        double dx[3];
        vbd_local_solve(acc, p, dunpack(gload(static_cast<const float4*>(a.ps.last_p), vid), un),
                        dunpack(gload(a.y, vid), un), gload(a.ps.m, vid), a.damping, a.dt, dx);
        const float4 out = dpack(dvec3{p.x + dx[0], p.y + dx[1], p.z + dx[2]}, a.ps.lo_inv);
        gstore(a.mode == GATHER_VBD_JACOBI ? a.p_new : a.ps.p, vid, out);
        // the move of the full value (the cached gradients are built from
        // hi + lo); zero when the solve left the vertex where it was
        du[0] = static_cast<float>(dx[0]);
        du[1] = static_cast<float>(dx[1]);
        du[2] = static_cast<float>(dx[2]);
    }
    if constexpr (CACHED) {
        // broadcast the move from the group's first lane, then update the
        // gradients of every (element, quadrature point) the group read
#pragma unroll
        for (int k = 0; k < 3; ++k) du[k] = __shfl(du[k], 0, L);
        if (du[0] == 0.f && du[1] == 0.f && du[2] == 0.f) return;
#pragma unroll
        for (int it = 0; it < KEEP; ++it) {
            const int t = lane + it * L;
            if (t < nwork) {
                int eid, q, r;
                work_item(own, t, Q, eid, q, r);
                rank1_update(a.ju + (static_cast<size_t>(eid) * Q + q) * 9, keep[it], du, a.fem.dN + (q * S + r) * 3);
            }
        }
        for (int t = lane + KEEP * L; t < nwork; t += L) {
            int eid, q, r;
            work_item(own, t, Q, eid, q, r);
            float* g = a.ju + (static_cast<size_t>(eid) * Q + q) * 9;
            float Ju[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) Ju[i] = g[i];
            rank1_update(g, Ju, du, a.fem.dN + (q * S + r) * 3);
        }
    }
}

// One quadrature point per work item keeps the live state small; ask for at
// least 4 waves per SIMD there so that the scheduler does not trade occupancy
// for early issue of all S node loads.
template <int QL>
constexpr int gather_min_waves() { return QL == 1 ? 4 : 1; }

template <int S, int Q, int L, int QL, bool CACHED>
__global__ __launch_bounds__(256, gather_min_waves<QL>()) void k_vertex_gather(GatherArgs a) {
    const int gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    gather_body<S, Q, L, QL, CACHED>(a, gt / L, gt & (L - 1));
}

// Batched form: one launch covers one colour of many scenes.  Every block
// belongs to one (scene, colour) segment, so the scene's description is read
// once per block with scalar loads.
template <int S, int Q, int L, int QL, bool CACHED>
__global__ __launch_bounds__(256, gather_min_waves<QL>()) void k_batch_gather(BatchGatherArgs b) {
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const BatchSeg seg = find_seg(b.segs, b.seg_first, b.seg_count, blk);
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;
    GatherArgs a;
    a.ps = sd.ps;
    a.fem = sd.fem;
    a.own = sd.own;
    a.mt = sd.mt;
    a.y = sd.y;
    a.p_new = nullptr;
    a.ju = sd.ju;
    a.damping = sd.damping;
    a.dt = b.dt;
    a.first_slot = seg.first;
    a.count = seg.count;
    a.mode = GATHER_VBD_INPLACE;
    a.out3 = nullptr;
    const int local = (blk - seg.block_begin) * 256 + threadIdx.x;
    gather_body<S, Q, L, QL, CACHED>(a, local / L, local & (L - 1));
}

// Refresh of the cached gradients, Ju(e, q) = sum_k u_k (x) dN_k(q): a block
// takes REFRESH_EPB(Q) = 256 / Q consecutive elements, stages their S node
// displacements in LDS once (one 16-byte load per node instead of one per
// (node, quadrature point)), then thread (element, q) builds its record; the
// records of one element are contiguous, so the stores are coalesced.
template <int Q>
constexpr int refresh_epb() { return 256 / Q; }

template <int S, int Q>
__device__ __forceinline__ void ju_refresh_block(const float4* __restrict__ P, float unit, const FemView& f, float* ju,
                                                 int e0, int ne) {
    constexpr int EPB = refresh_epb<Q>();
    __shared__ float4 sx[EPB * S];
    for (int t = threadIdx.x; t < ne * S; t += blockDim.x) sx[t] = P[f.topo[static_cast<size_t>(e0) * S + t]];
    __syncthreads();
    const int le = threadIdx.x / Q;
    const int q = threadIdx.x - le * Q;
    if (le >= ne) return;
    float xs[S][3];
    const float4 ref = sx[le * S];
#pragma unroll
    for (int k = 0; k < S; ++k) rel_node(sx[le * S + k], ref, unit, xs[k]);
    float Ju[9];
    jx_from_nodes<S>(xs, f.dN + q * S * 3, Ju);
    float* g = ju + (static_cast<size_t>(e0 + le) * Q + q) * 9;
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = Ju[i];
}

template <int S, int Q>
__global__ __launch_bounds__(256) void k_ju_refresh(PSView ps, FemView f, float* ju) {
    constexpr int EPB = refresh_epb<Q>();
    const int e0 = xcd_block(blockIdx.x, gridDim.x) * EPB;
    const int ne = min(EPB, f.E - e0);
    ju_refresh_block<S, Q>(ps.p, ps.lo_unit, f, ju, e0, ne);
}

// Batched form: VSeg.count is the scene's element count, each block takes
// REFRESH_EPB elements of one scene.
template <int S, int Q>
__global__ __launch_bounds__(256) void k_batch_ju_refresh(BatchRefreshArgs b) {
    constexpr int EPB = refresh_epb<Q>();
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const VSeg seg = find_seg(b.segs, b.seg_first, b.seg_count, blk);
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;      // uniform over the block
    const int e0 = (blk - seg.block_begin) * EPB;
    const int ne = min(EPB, seg.count - e0);
    ju_refresh_block<S, Q>(sd.ps.p, sd.ps.lo_unit, sd.fem, sd.ju, e0, ne);
}

#if !defined(OE_GATHER_S) || OE_GATHER_S != 27   // 27-node hexahedra have their own task kernels
// Fused element tasks (see FusedTaskArgs).  Phase 1: thread t of the block
// takes the incidences t, t + kFusedThreads, ... of the vertex block.  Fixed
// vertices are not skipped there: phase 2 skips them.
//
// The shape-derivative table is read through an accessor.  Tables of up to
// 3 KB travel by value in the kernel arguments, so the Jx contraction (the
// same point q on every lane) reads them with scalar loads; larger ones stay
// in device memory.  dN_r(q) differs between lanes and always comes from the
// table in device memory.
template <int S, int Q>
struct ShapeGrad {
    float dN[Q * S * 3];
};
template <int S, int Q>
constexpr bool grad_by_value() {
    return Q * S * 3 * sizeof(float) <= 3072;
}
template <int S, int Q>
struct GradByValue {
    const ShapeGrad<S, Q>& g;
    __device__ __forceinline__ float at(int i) const { return g.dN[i]; }
};
struct GradInMemory {
    const float* dN;
    __device__ __forceinline__ float at(int i) const { return gload(dN, i); }
};
// the kernel-argument table: the element's own, or a 1-point stand-in for
// types whose table stays in device memory
template <int S, int Q>
using ShapeGradArg = ShapeGrad<S, (grad_by_value<S, Q>() ? Q : 1)>;

template <int S, class Grad>
__device__ __forceinline__ void jx_from_grad(const float (&xs)[S][3], const Grad& grad, int base, float Jx[9]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) Jx[i] = 0.f;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const float d0 = grad.at(base + 3 * k), d1 = grad.at(base + 3 * k + 1), d2 = grad.at(base + 3 * k + 2);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Jx[0 + c] = fmaf(xs[k][c], d0, Jx[0 + c]);
            Jx[3 + c] = fmaf(xs[k][c], d1, Jx[3 + c]);
            Jx[6 + c] = fmaf(xs[k][c], d2, Jx[6 + c]);
        }
    }
}

// In the throughput form of the fused blocks, shape-derivative tables of at
// most kDnLdsFloats floats (P1, P2, Q1) are copied to LDS, one copy per wave,
// at the start of a block: a lane's dN_r at each point is then an LDS read,
// not a device-memory load that the point loop waits on (P2 batch: 84.9
// against 89.1 us per colour pass).  The latency form keeps the loads: it
// issues the vertex state's loads first, and vector-memory loads complete in
// order, so the copy's LDS write would wait for them (single P2 scene: 1.22
// against 1.19 ms per step).
constexpr int kDnLdsFloats = 256;
template <int S, int Q, bool LAT>
constexpr bool dn_in_lds() { return !LAT && S * Q * 3 <= kDnLdsFloats; }
template <int S, int Q, bool LAT>
constexpr int dn_lds_size() { return dn_in_lds<S, Q, LAT>() ? S * Q * 3 : 1; }

template <int S, int Q, bool DNL>
__device__ __forceinline__ void dnr_at(const FemView& f, const float* dnw, int q, int r, float dnr[3]) {
    if constexpr (DNL) {
        const float* d = dnw + (q * S + r) * 3;
        dnr[0] = d[0];
        dnr[1] = d[1];
        dnr[2] = d[2];
    } else {
        load_dnr(f.dN, (q * S + r) * 3, dnr);
    }
}

// Partial (force, Hessian) of local node r of element eid over its Q points.
// dnw: the wave's LDS copy of the shape-derivative table (DNL only).
template <int S, int Q, bool DNL, class Grad>
__device__ __forceinline__ void elem_partial(const FemView& f, const MaterialView& mt, const Grad& grad, int eid,
                                             int r, const float (&xs)[S][3], const float* dnw, Acc9& acc) {
    acc_zero(acc);
    if (f.affine) {
        // one record per element: JX_inv and |det B|, V_q = |det B| w_q
        float JXi[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) JXi[i] = gload(f.jxv, static_cast<size_t>(eid) * 10 + i);
        const float detB = gload(f.jxv, static_cast<size_t>(eid) * 10 + 9);
        // one point at a time: its 3S table entries fit the scalar registers
#pragma unroll 1
        for (int q = 0; q < Q; ++q) {
            float Ju[9], dnr[3];
            jx_from_grad<S>(xs, grad, q * S * 3, Ju);
            dnr_at<S, Q, DNL>(f, dnw, q, r, dnr);
            nh_contrib(Ju, JXi, detB * gload(f.wq, q), dnr, mt, acc);
        }
    } else {
#pragma unroll 1
        for (int q = 0; q < Q; ++q) {
            float Ju[9], JXi[9], V, dnr[3];
            load_element_record(f, eid, q, JXi, V);
            jx_from_grad<S>(xs, grad, q * S * 3, Ju);
            dnr_at<S, Q, DNL>(f, dnw, q, r, dnr);
            nh_contrib(Ju, JXi, V, dnr, mt, acc);
        }
    }
}

// Phase 1 of a fused block: the partials of the block's incidences into LDS.
template <int S, int Q, bool DNL, class Grad>
__device__ __forceinline__ void fused_phase1(const PSView& ps, const FemView& f, const OwnersView& own,
                                             const MaterialView& mt, const Grad& grad, const int4 blk,
                                             float (*part)[9], float* dnw) {
    const int first_own = blk.z, n_inc = blk.w;
    const int t0 = threadIdx.x;
    // the table copy's loads go first and its LDS writes after the first
    // owners load is issued, so waiting for them overlaps that load; every
    // lane of the wave writes its share before any lane leaves
    constexpr int kDnPerLane = DNL ? (S * Q * 3 + 63) / 64 : 0;
    float dn_stage[kDnPerLane > 0 ? kDnPerLane : 1];
    const int lane = t0 & 63;
    if constexpr (kDnPerLane > 0) {
#pragma unroll
        for (int j = 0; j < kDnPerLane; ++j)
            if (lane + 64 * j < S * Q * 3) dn_stage[j] = gload(f.dN, lane + 64 * j);
    }
    const int pk0 = t0 < n_inc ? gload(own.own, first_own + t0) : 0;
    if constexpr (kDnPerLane > 0) {
#pragma unroll
        for (int j = 0; j < kDnPerLane; ++j)
            if (lane + 64 * j < S * Q * 3) dnw[lane + 64 * j] = dn_stage[j];
        // a wave's LDS operations complete in order; keep the compiler from
        // moving the table reads above the writes
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    if (t0 >= n_inc) return;
#pragma unroll 1
    for (int t = t0; t < n_inc; t += kFusedThreads) {
        const int pk = t == t0 ? pk0 : gload(own.own, first_own + t);
        float4 h[S];
#pragma unroll
        for (int k = 0; k < S; ++k)
            h[k] = gload(static_cast<const float4*>(ps.p), gload(f.topo, static_cast<size_t>(pk >> 5) * S + k));
        float xs[S][3];
#pragma unroll
        for (int k = 0; k < S; ++k) rel_node(h[k], h[0], ps.lo_unit, xs[k]);
        const int eid = pk >> 5, r = pk & 31;
        Acc9 acc;
        elem_partial<S, Q, DNL>(f, mt, grad, eid, r, xs, dnw, acc);
#pragma unroll
        for (int i = 0; i < 9; ++i) part[t][i] = acc.v[i];
    }
}

// The latency form's phase 1: one incidence per thread, and the slot state's
// loads issued by after_owner / after_topo right after the owners-entry and
// node-id loads, so that no wait of the element chain covers them.  A wave past
// the block's incidences returns at once (it holds no slot either).  The rest
// is straight-line: a lane past the incidences works on the last one and
// stores nothing (every vertex block has an incidence), so that no wait has to
// cover the loads of another path, which would make the compiler wait for
// every load in flight, the slot state's included.  (The throughput form keeps
// the loop above: this form's prologue cost it 2%, 86.7 against 84.5 us per P2
// batch pass on one GPU.)
template <int S, int Q, class Grad, class AfterOwner, class AfterTopo>
__device__ __forceinline__ void fused_phase1_lat(const PSView& ps, const FemView& f, const OwnersView& own,
                                                 const MaterialView& mt, const Grad& grad, const int4 blk,
                                                 float (*part)[9], AfterOwner&& after_owner, AfterTopo&& after_topo) {
    static_assert(kFusedRuns == 1, "one incidence per thread");
    const int first_own = blk.z, n_inc = blk.w;
    const int t0 = threadIdx.x;
    if (__builtin_amdgcn_readfirstlane(t0 & ~63) >= n_inc) return;
    const bool has = t0 < n_inc;
    const int pk = gload(own.own, first_own + (has ? t0 : n_inc - 1));
    after_owner();
    __builtin_amdgcn_sched_barrier(0);
    int nid[S];
#pragma unroll
    for (int k = 0; k < S; ++k) nid[k] = gload(f.topo, static_cast<size_t>(pk >> 5) * S + k);
    after_topo();
    float4 h[S];
#pragma unroll
    for (int k = 0; k < S; ++k) h[k] = gload(static_cast<const float4*>(ps.p), nid[k]);
    float xs[S][3];
#pragma unroll
    for (int k = 0; k < S; ++k) rel_node(h[k], h[0], ps.lo_unit, xs[k]);
    const int eid = pk >> 5, r = pk & 31;
    Acc9 acc;
    elem_partial<S, Q, false>(f, mt, grad, eid, r, xs, nullptr, acc);
    if (has) {
#pragma unroll
        for (int i = 0; i < 9; ++i) part[t0][i] = acc.v[i];
    }
}

template <int S, int Q, bool LAT, class Grad>
__device__ __forceinline__ void fused_block(const PSView& ps, const FemView& f, const OwnersView& own,
                                            const MaterialView& mt, const Grad& grad, const float4* y, float damping,
                                            float dt, const int4 blk, float (*part)[9], float* dnw) {
    const int first_slot = blk.x, n_slots = blk.y, first_own = blk.z;
    const int t0 = threadIdx.x;
    if constexpr (LAT) {
        // the slot state rides along the element chain (a thread past the
        // block's slots loads the last one and does not use it; every vertex
        // block has a slot)
        SlotState st{};
        const int ts = first_slot + min(t0, n_slots - 1);
        fused_phase1_lat<S, Q>(ps, f, own, mt, grad, blk, part, [&] { slot_state_level1(own, ts, st); },
                               [&] { slot_state_level2(ps, y, st); });
        __syncthreads();
        if (t0 < n_slots) {
            slot_state_settle(st);
            if (st.free != 0) solve_slot_from_lds(ps, st, part, first_own, damping, dt);
        }
    } else {
        // the slots' vertex ids and owners ranges, loaded now rather than after
        // the barrier (one dependent load less in phase 2; a thread past the
        // slots loads the last one)
        const int ts = first_slot + min(t0, n_slots - 1);
        const int slot_vid = gload(own.slot_vid, ts);
        const int2 slot_nbo = gload(own.slot_nbo, ts);
        fused_phase1<S, Q, dn_in_lds<S, Q, LAT>()>(ps, f, own, mt, grad, blk, part, dnw);
        __syncthreads();
        // phase 2: one thread per slot, partials in the owners order, local solve
        if (t0 >= n_slots) return;
        solve_slot_at(ps, slot_vid, slot_nbo, y, damping, dt, part, first_own);
    }
}

template <int S, int Q, bool LAT>
__device__ __forceinline__ void fused_dispatch(const PSView& ps, const FemView& f, const OwnersView& own,
                                               const MaterialView& mt, const ShapeGradArg<S, Q>& g, const float4* y,
                                               float damping, float dt, const int4 blk, float (*part)[9],
                                               float* dnw) {
    if constexpr (grad_by_value<S, Q>())
        fused_block<S, Q, LAT>(ps, f, own, mt, GradByValue<S, Q>{g}, y, damping, dt, blk, part, dnw);
    else
        fused_block<S, Q, LAT>(ps, f, own, mt, GradInMemory{f.dN}, y, damping, dt, blk, part, dnw);
}

template <int S, int Q, bool LAT>
__global__ __launch_bounds__(kFusedThreads) void k_fused_tasks(FusedTaskArgs a, ShapeGradArg<S, Q> g) {
    __shared__ float part[LAT ? kFusedIncPerBlockLat : kFusedIncPerBlock][9];
    __shared__ float dn_lds[kFusedThreads / 64][dn_lds_size<S, Q, LAT>()];
    const int4 blk = gload(a.blocks, a.first_block + xcd_block(blockIdx.x, gridDim.x));
    fused_dispatch<S, Q, LAT>(a.ps, a.fem, a.own, a.mt, g, a.y, a.damping, a.dt, blk, part,
                              dn_lds[threadIdx.x >> 6]);
}

template <int S, int Q, bool LAT>
__global__ __launch_bounds__(kFusedThreads) void k_batch_fused_tasks(BatchFusedArgs b, ShapeGradArg<S, Q> g) {
    __shared__ float part[LAT ? kFusedIncPerBlockLat : kFusedIncPerBlock][9];
    __shared__ float dn_lds[kFusedThreads / 64][dn_lds_size<S, Q, LAT>()];
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const BatchSeg seg = find_seg(b.segs, b.seg_first, b.seg_count, bid);
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;
    // the scene's views, read before the first store
    const PSView ps = sd.ps;
    const FemView fem = sd.fem;
    const OwnersView own = sd.own;
    const MaterialView mt = sd.mt;
    const float4* y = sd.y;
    const float damping = sd.damping;
    const int4 blk = gload(LAT ? sd.fblocks_lat : sd.fblocks, seg.first + (bid - seg.block_begin));
    fused_dispatch<S, Q, LAT>(ps, fem, own, mt, g, y, damping, b.dt, blk, part, dn_lds[threadIdx.x >> 6]);
}

#endif

namespace {

template <int S, int Q, int L, int QL, bool CACHED>
void launch_one(const GatherArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    const long long threads = static_cast<long long>(a.count) * L;
    const int block = 256;
    const int grid = static_cast<int>((threads + block - 1) / block);
    hipLaunchKernelGGL((k_vertex_gather<S, Q, L, QL, CACHED>), dim3(grid), dim3(block), 0, s, a);
}

template <int S, int Q, int L, int QL, bool CACHED>
void launch_batch_one(const BatchGatherArgs& b, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL((k_batch_gather<S, Q, L, QL, CACHED>), dim3(nblocks), dim3(256), 0, s, b);
}

template <int S, int Q, int QL, bool CACHED>
bool launch_batch_L(int L, const BatchGatherArgs& b, int nblocks, hipStream_t s) {
    switch (L) {
        case 4: launch_batch_one<S, Q, 4, QL, CACHED>(b, nblocks, s); return true;
        case 8: launch_batch_one<S, Q, 8, QL, CACHED>(b, nblocks, s); return true;
        case 16: launch_batch_one<S, Q, 16, QL, CACHED>(b, nblocks, s); return true;
        case 32: launch_batch_one<S, Q, 32, QL, CACHED>(b, nblocks, s); return true;
        case 64: launch_batch_one<S, Q, 64, QL, CACHED>(b, nblocks, s); return true;
        default: return false;
    }
}

template <int S, int Q, int QL, bool CACHED>
bool launch_L(int L, const GatherArgs& a, hipStream_t s) {
    switch (L) {
        case 4: launch_one<S, Q, 4, QL, CACHED>(a, s); return true;
        case 8: launch_one<S, Q, 8, QL, CACHED>(a, s); return true;
        case 16: launch_one<S, Q, 16, QL, CACHED>(a, s); return true;
        case 32: launch_one<S, Q, 32, QL, CACHED>(a, s); return true;
        case 64: launch_one<S, Q, 64, QL, CACHED>(a, s); return true;
        default: return false;
    }
}

// Element-centric internal forces of the explicit integrator.  The reference
// computes them per vertex (kernel_fem_eval_force,
// src/GPU/Explicit/GPU_FEM_Explicit.cu:5-64), so a vertex rebuilds F at every
// point of each incident element: S times over per (element, point).  Here LPE
// lanes (32, or a whole wave for 27-point elements) take one element:
//   nodes:  lane k < S puts node k's displacement relative to node 0 in LDS;
//   points: lane q < Q forms Ju = sum_k x_k (x) dN_k(q), F, cof F and
//           s = lambda (J - alpha) once, and from them (M = JX_inv^T, so that
//           g = M dN_r):
//             Pm = V (mu F + s C) M    force  -V (mu F g + s C g) = -Pm dN_r
//             G  = V mu M^T M          V mu |g|^2 = dN_r^T G dN_r
//             K  = C M                 C g = K dN_r
//           and V lambda, in LDS point-major per component (lanes write
//           consecutive words);
//   nodes:  lane r < S of each 32-lane half adds the force and Hessian block
//           V mu |g|^2 I + V lambda (C g)(C g)^T of node r over its half of
//           the points (every lane reads the same point words, a broadcast);
//           the halves are added and lane r stores the 9 floats at the
//           owners position of (e, r), so that each vertex's partials are
//           contiguous.
// k_explicit_vertex then adds each vertex's partials in owners order.
constexpr int kExplicitThreads = 256;
constexpr int kExplicitPt = 25;      // Pm 9, K 9, G 6, V lambda

template <int Q>
constexpr int explicit_lanes() {
    return Q >= 16 ? 64 : 32;
}

template <int S, int Q>
__global__ __launch_bounds__(kExplicitThreads) void k_explicit_elem(ExplicitElemArgs a) {
    static_assert(S <= 32 && Q <= 32, "one element per 32 or 64 lanes");
    constexpr int LPE = explicit_lanes<Q>();
    constexpr int EPB = kExplicitThreads / LPE;
    constexpr int QH = (Q + LPE / 32 - 1) / (LPE / 32);   // points per 32-lane half
    __shared__ float dn_s[Q * S * 3];
    __shared__ float xs_s[EPB][S * 3];
    __shared__ float pt_s[EPB][kExplicitPt][Q];
    const int tid = threadIdx.x;
    const int h = tid / LPE, lane = tid % LPE;
    const int r = (lane & 31) < S ? (lane & 31) : S - 1;
    const int half = lane >> 5;
    const int E = a.fem.E;
    const int e = blockIdx.x * EPB + h;
    const int ec = e < E ? e : E - 1;
    const int at = gload(a.inc_at, static_cast<size_t>(ec) * S + r);
    for (int i = tid; i < Q * S * 3; i += kExplicitThreads) dn_s[i] = gload(a.fem.dN, i);
    {
        const int k = lane < S ? lane : S - 1;
        const int node = gload(a.fem.topo, static_cast<size_t>(ec) * S + k);
        const float4 u = gload(static_cast<const float4*>(a.ps.p), node);
        float4 ref;
        ref.x = __shfl(u.x, 0, LPE);
        ref.y = __shfl(u.y, 0, LPE);
        ref.z = __shfl(u.z, 0, LPE);
        ref.w = __shfl(u.w, 0, LPE);
        float x[3];
        rel_node(u, ref, a.ps.lo_unit, x);
        if (lane < S) {
            xs_s[h][3 * lane] = x[0];
            xs_s[h][3 * lane + 1] = x[1];
            xs_s[h][3 * lane + 2] = x[2];
        }
    }
    const int q = lane < Q ? lane : Q - 1;
    float JXi[9], V;
    load_element_record(a.fem, ec, q, JXi, V);
    __syncthreads();
    {
        float Ju[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) Ju[i] = 0.f;
        const float* __restrict__ dn = dn_s + q * S * 3;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const float d0 = dn[3 * k], d1 = dn[3 * k + 1], d2 = dn[3 * k + 2];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float xk = xs_s[h][3 * k + c];
                Ju[0 + c] = fmaf(xk, d0, Ju[0 + c]);
                Ju[3 + c] = fmaf(xk, d1, Ju[3 + c]);
                Ju[6 + c] = fmaf(xk, d2, Ju[6 + c]);
            }
        }
        float F[9], C[9];
        deformation_gradient(Ju, JXi, F);
        cofactor(F, C);
        const float J = det_from_cof(F, C);
        const float sl = a.mt.lambda * (J - a.mt.alpha);
        if (lane < Q) {
            // F, C, JX_inv column-major: X(row, col) = X[col * 3 + row]; M(c, l) = JX_inv[c * 3 + l]
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    float pm = 0.f, km = 0.f;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        pm += (a.mt.mu * F[c * 3 + i] + sl * C[c * 3 + i]) * JXi[c * 3 + l];
                        km += C[c * 3 + i] * JXi[c * 3 + l];
                    }
                    pt_s[h][i * 3 + l][lane] = V * pm;
                    pt_s[h][9 + i * 3 + l][lane] = km;
                }
            const float vm = V * a.mt.mu;
            constexpr int gl[6] = {0, 0, 0, 1, 1, 2}, gm[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int i = 0; i < 6; ++i)
                pt_s[h][18 + i][lane] = vm * (JXi[gl[i]] * JXi[gm[i]] + JXi[3 + gl[i]] * JXi[3 + gm[i]] +
                                              JXi[6 + gl[i]] * JXi[6 + gm[i]]);
            pt_s[h][24][lane] = V * a.mt.lambda;
        }
    }
    __syncthreads();
    float acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.f;
    const int p_end = (half + 1) * QH < Q ? (half + 1) * QH : Q;
    for (int p = half * QH; p < p_end; ++p) {
        const float* __restrict__ dn = dn_s + (p * S + r) * 3;
        const float d0 = dn[0], d1 = dn[1], d2 = dn[2];
        float w[kExplicitPt];
#pragma unroll
        for (int i = 0; i < kExplicitPt; ++i) w[i] = pt_s[h][i][p];
        float c[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            acc[i] -= w[i * 3] * d0 + w[i * 3 + 1] * d1 + w[i * 3 + 2] * d2;
            c[i] = w[9 + i * 3] * d0 + w[9 + i * 3 + 1] * d1 + w[9 + i * 3 + 2] * d2;
        }
        const float gd = w[18] * d0 * d0 + w[21] * d1 * d1 + w[23] * d2 * d2 +
                         2.f * (w[19] * d0 * d1 + w[20] * d0 * d2 + w[22] * d1 * d2);
        const float vl = w[24];
        acc[3] += gd + vl * c[0] * c[0];
        acc[4] += vl * c[1] * c[0];
        acc[5] += vl * c[2] * c[0];
        acc[6] += gd + vl * c[1] * c[1];
        acc[7] += vl * c[2] * c[1];
        acc[8] += gd + vl * c[2] * c[2];
    }
    if constexpr (LPE == 64) {
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += __shfl_xor(acc[i], 32);
    }
    if (e >= E || lane >= S) return;
    float* out = a.partial + static_cast<size_t>(at) * 9;
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = acc[i];
}

}  // namespace

#if defined(OE_GATHER_S) && defined(OE_GATHER_Q)
// OE_FN(launch_gather_) -> launch_gather_<S>_<Q>
#define OE_CAT3(p, s, q) p##s##_##q
#define OE_NAME(p, s, q) OE_CAT3(p, s, q)
#define OE_FN(p) OE_NAME(p, OE_GATHER_S, OE_GATHER_Q)

// Quadrature points per lane that divide Q: 1, Q, and for Q = 4 / 27 the
// intermediate divisors.  The cached form takes QL = 1 only.
bool OE_FN(launch_gather_)(int L, int QL, const GatherArgs& a, hipStream_t s) {
    constexpr int S = OE_GATHER_S, Q = OE_GATHER_Q;
    if (a.ju) return QL == 1 && a.mode == GATHER_VBD_INPLACE && launch_L<S, Q, 1, true>(L, a, s);
    if (QL == 1) return launch_L<S, Q, 1, false>(L, a, s);
    if (QL == Q) return launch_L<S, Q, Q, false>(L, a, s);
#if OE_GATHER_Q == 4
    if (QL == 2) return launch_L<S, Q, 2, false>(L, a, s);
#endif
#if OE_GATHER_Q == 27
    if (QL == 3) return launch_L<S, Q, 3, false>(L, a, s);
    if (QL == 9) return launch_L<S, Q, 9, false>(L, a, s);
#endif
    return false;
}

bool OE_FN(launch_batch_gather_)(int L, int QL, bool cached, const BatchGatherArgs& b, int nblocks, hipStream_t s) {
    constexpr int S = OE_GATHER_S, Q = OE_GATHER_Q;
    if (cached) return QL == 1 && launch_batch_L<S, Q, 1, true>(L, b, nblocks, s);
    if (QL == 1) return launch_batch_L<S, Q, 1, false>(L, b, nblocks, s);
    if (QL == Q) return launch_batch_L<S, Q, Q, false>(L, b, nblocks, s);
#if OE_GATHER_Q == 4
    if (QL == 2) return launch_batch_L<S, Q, 2, false>(L, b, nblocks, s);
#endif
#if OE_GATHER_Q == 27
    if (QL == 3) return launch_batch_L<S, Q, 3, false>(L, b, nblocks, s);
    if (QL == 9) return launch_batch_L<S, Q, 9, false>(L, b, nblocks, s);
#endif
    return false;
}

void OE_FN(launch_ju_refresh_)(const PSView& ps, const FemView& f, float* ju, hipStream_t s) {
    constexpr int S = OE_GATHER_S, Q = OE_GATHER_Q;
    constexpr int EPB = refresh_epb<Q>();
    if (f.E <= 0) return;
    hipLaunchKernelGGL((k_ju_refresh<S, Q>), dim3((f.E + EPB - 1) / EPB), dim3(256), 0, s, ps, f, ju);
}

int OE_FN(refresh_elements_per_block_)() { return refresh_epb<OE_GATHER_Q>(); }

#if OE_GATHER_S != 27
namespace {
// The element type's shape-derivative table as float, the values the scenes
// upload as their dN (FemConstants), built once.
const ShapeGradArg<OE_GATHER_S, OE_GATHER_Q>& shape_grad_arg() {
    constexpr int S = OE_GATHER_S, Q = OE_GATHER_Q;
    static const ShapeGradArg<S, Q> g = [] {
        ShapeGradArg<S, Q> t{};
        if constexpr (grad_by_value<S, Q>()) {
            const Element el = S == 4 ? Tetra : S == 8 ? Hexa : S == 10 ? Tetra10 : S == 20 ? Tetra20 : Hexa27;
            const ShapeTable& st = shape_table(el);
            if (static_cast<int>(st.dN.size()) != Q * S * 3) throw std::logic_error("shape table size mismatch");
            for (int i = 0; i < Q * S * 3; ++i) t.dN[i] = static_cast<float>(st.dN[i]);
        }
        return t;
    }();
    return g;
}
}  // namespace

void OE_FN(launch_fused_tasks_)(const FusedTaskArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    if (a.latency)
        hipLaunchKernelGGL((k_fused_tasks<OE_GATHER_S, OE_GATHER_Q, true>), dim3(a.count), dim3(kFusedThreads), 0, s, a,
                           shape_grad_arg());
    else
        hipLaunchKernelGGL((k_fused_tasks<OE_GATHER_S, OE_GATHER_Q, false>), dim3(a.count), dim3(kFusedThreads), 0, s,
                           a, shape_grad_arg());
}

void OE_FN(launch_batch_fused_tasks_)(const BatchFusedArgs& b, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    if (b.latency)
        hipLaunchKernelGGL((k_batch_fused_tasks<OE_GATHER_S, OE_GATHER_Q, true>), dim3(nblocks), dim3(kFusedThreads), 0,
                           s, b, shape_grad_arg());
    else
        hipLaunchKernelGGL((k_batch_fused_tasks<OE_GATHER_S, OE_GATHER_Q, false>), dim3(nblocks), dim3(kFusedThreads),
                           0, s, b, shape_grad_arg());
}

#endif

void OE_FN(launch_explicit_elem_)(const ExplicitElemArgs& a, hipStream_t s) {
    if (a.fem.E <= 0) return;
    constexpr int EPB = kExplicitThreads / explicit_lanes<OE_GATHER_Q>();
    hipLaunchKernelGGL((k_explicit_elem<OE_GATHER_S, OE_GATHER_Q>), dim3((a.fem.E + EPB - 1) / EPB),
                       dim3(kExplicitThreads), 0, s, a);
}

void OE_FN(launch_batch_ju_refresh_)(const BatchRefreshArgs& b, int nblocks, hipStream_t s) {
    constexpr int S = OE_GATHER_S, Q = OE_GATHER_Q;
    if (nblocks <= 0) return;
    hipLaunchKernelGGL((k_batch_ju_refresh<S, Q>), dim3(nblocks), dim3(256), 0, s, b);
}
#endif

}  // namespace oe
