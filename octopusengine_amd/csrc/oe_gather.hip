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
// double precision.  L and QL are tile parameters chosen per element type by
// the autotuner.
//
// This file is compiled once per element type (OE_GATHER_S / OE_GATHER_Q set
// on the command line) so that the template instances build in parallel.
#include "oe_gather.hpp"
#include "oe_kernels.hpp"

namespace oe {

// Gather + local solve of one vertex slot by a group of L lanes.
template <int S, int Q, int L, int QL>
__device__ __forceinline__ void gather_body(const GatherArgs& a, int grp, int lane) {
    constexpr int QC = Q / QL;
    if (grp >= a.count) return;                 // whole group leaves together
    const int slot = a.first_slot + grp;
    const int vid = a.own.slot_vid[slot];
    if (a.mode != GATHER_EXPLICIT_FORCE && a.ps.mask[vid] == 0) return;
    const int2 nbo = a.own.slot_nbo[slot];
    const int nwork = nbo.x * QC;
    Acc9 acc;
    acc_zero(acc);
    const float4* __restrict__ P = a.ps.p;
    for (int t = lane; t < nwork; t += L) {
        const int j = t / QC;
        const int qc = t - j * QC;
        const int pk = a.own.own[nbo.y + j];
        const int eid = pk >> 5;
        const int r = pk & 31;
        const int* __restrict__ tp = a.fem.topo + static_cast<size_t>(eid) * S;
        float xs[S][3];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const float4 x = P[tp[k]];
            xs[k][0] = x.x; xs[k][1] = x.y; xs[k][2] = x.z;
        }
#pragma unroll
        for (int qq = 0; qq < QL; ++qq) {
            const int q = qc * QL + qq;
            const float* __restrict__ dn = a.fem.dN + q * S * 3;
            float Jx[9];
            jx_from_nodes<S>(xs, dn, Jx);
            nh_contrib(Jx, a.fem.jxv + (static_cast<size_t>(eid) * Q + q) * 10, dn + 3 * r, a.mt, acc);
        }
    }
    group_sum<L>(acc);
    if (lane != 0) return;
    if (a.mode == GATHER_EXPLICIT_FORCE) {
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
    const float4 pv = P[vid];
    double dx[3];
    vbd_local_solve(acc, pv, a.ps.last_p[vid], a.y[vid], a.ps.m[vid], a.damping, a.dt, dx);
    const float4 out = make_float4(pv.x + static_cast<float>(dx[0]), pv.y + static_cast<float>(dx[1]),
                                   pv.z + static_cast<float>(dx[2]), pv.w);
    if (a.mode == GATHER_VBD_JACOBI)
        a.p_new[vid] = out;
    else
        a.ps.p[vid] = out;
}

template <int S, int Q, int L, int QL>
__global__ __launch_bounds__(256) void k_vertex_gather(GatherArgs a) {
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    gather_body<S, Q, L, QL>(a, gt / L, gt & (L - 1));
}

// Batched form: one launch covers one colour of many scenes.  Every block
// belongs to one (scene, colour) segment, so the scene's description is read
// once per block with scalar loads.
template <int S, int Q, int L, int QL>
__global__ __launch_bounds__(256) void k_batch_gather(BatchGatherArgs b) {
    const int blk = b.block_base + blockIdx.x;
    const BatchSeg seg = b.segs[b.block_seg[blk]];
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;
    GatherArgs a;
    a.ps = sd.ps;
    a.fem = sd.fem;
    a.own = sd.own;
    a.mt = sd.mt;
    a.y = sd.y;
    a.p_new = nullptr;
    a.damping = sd.damping;
    a.dt = b.dt;
    a.first_slot = seg.first;
    a.count = seg.count;
    a.mode = GATHER_VBD_INPLACE;
    const int local = (blk - seg.block_begin) * 256 + threadIdx.x;
    gather_body<S, Q, L, QL>(a, local / L, local & (L - 1));
}

namespace {

template <int S, int Q, int L, int QL>
void launch_one(const GatherArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    const long long threads = static_cast<long long>(a.count) * L;
    const int block = 256;
    const int grid = static_cast<int>((threads + block - 1) / block);
    hipLaunchKernelGGL((k_vertex_gather<S, Q, L, QL>), dim3(grid), dim3(block), 0, s, a);
}

template <int S, int Q, int L, int QL>
void launch_batch_one(const BatchGatherArgs& b, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL((k_batch_gather<S, Q, L, QL>), dim3(nblocks), dim3(256), 0, s, b);
}

template <int S, int Q, int QL>
bool launch_batch_L(int L, const BatchGatherArgs& b, int nblocks, hipStream_t s) {
    switch (L) {
        case 4: launch_batch_one<S, Q, 4, QL>(b, nblocks, s); return true;
        case 8: launch_batch_one<S, Q, 8, QL>(b, nblocks, s); return true;
        case 16: launch_batch_one<S, Q, 16, QL>(b, nblocks, s); return true;
        case 32: launch_batch_one<S, Q, 32, QL>(b, nblocks, s); return true;
        case 64: launch_batch_one<S, Q, 64, QL>(b, nblocks, s); return true;
        default: return false;
    }
}

template <int S, int Q, int QL>
bool launch_L(int L, const GatherArgs& a, hipStream_t s) {
    switch (L) {
        case 4: launch_one<S, Q, 4, QL>(a, s); return true;
        case 8: launch_one<S, Q, 8, QL>(a, s); return true;
        case 16: launch_one<S, Q, 16, QL>(a, s); return true;
        case 32: launch_one<S, Q, 32, QL>(a, s); return true;
        case 64: launch_one<S, Q, 64, QL>(a, s); return true;
        default: return false;
    }
}

}  // namespace

#if defined(OE_GATHER_S) && defined(OE_GATHER_Q)
// Quadrature points per lane that divide Q: 1, Q, and for Q = 4 / 27 the
// intermediate divisors.
bool OE_GATHER_FN(int L, int QL, const GatherArgs& a, hipStream_t s) {
    constexpr int S = OE_GATHER_S, Q = OE_GATHER_Q;
    if (QL == 1) return launch_L<S, Q, 1>(L, a, s);
    if (QL == Q) return launch_L<S, Q, Q>(L, a, s);
#if OE_GATHER_Q == 4
    if (QL == 2) return launch_L<S, Q, 2>(L, a, s);
#endif
#if OE_GATHER_Q == 27
    if (QL == 3) return launch_L<S, Q, 3>(L, a, s);
    if (QL == 9) return launch_L<S, Q, 9>(L, a, s);
#endif
    return false;
}

bool OE_GATHER_BATCH_FN(int L, int QL, const BatchGatherArgs& b, int nblocks, hipStream_t s) {
    constexpr int S = OE_GATHER_S, Q = OE_GATHER_Q;
    if (QL == 1) return launch_batch_L<S, Q, 1>(L, b, nblocks, s);
    if (QL == Q) return launch_batch_L<S, Q, Q>(L, b, nblocks, s);
#if OE_GATHER_Q == 4
    if (QL == 2) return launch_batch_L<S, Q, 2>(L, b, nblocks, s);
#endif
#if OE_GATHER_Q == 27
    if (QL == 3) return launch_batch_L<S, Q, 3>(L, b, nblocks, s);
    if (QL == 9) return launch_batch_L<S, Q, 9>(L, b, nblocks, s);
#endif
    return false;
}
#endif

}  // namespace oe
