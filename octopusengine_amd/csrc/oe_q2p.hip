// Throughput form of the Q2 VBD colour pass with "pencil" tasks: a task
// (incidence: element e, local node r) takes 9 lanes, lane m = (j, k) owning
// the three quadrature points (i, j, k), i = 0..2, and a wave runs 7 tasks at
// once (lanes 0-62; lane 63 idle).  A block is one wave: every exchange is
// inside the wave, so the block needs no workgroup barrier.
//
// Per task, with u_abc the node values at tensor place (a, b, c) and l, l'
// the 1D Lagrange factors and their derivatives at the Gauss coordinates:
//   stage 0  lane (b, c) loads nodes (a, b, c), a = 0..2, into LDS, as
//            (hi - hi of node 0) + lo (rel_node)
//   stage A  lane (b, k):  T_ab(k) = sum_c l_c(z_k) u_abc,  Tz_ab(k) = sum_c l_c'(z_k) u_abc
//   stage B  lane (j, k):  X_a = sum_b l_b(y_j) T_ab,  Xy_a = sum_b l_b'(y_j) T_ab,
//                          Xz_a = sum_b l_b(y_j) Tz_ab                    (in registers)
//   stage C  lane (j, k), i = 0..2:  dU/dx = sum_a l_a'(x_i) X_a,  dU/dy = sum_a l_a(x_i) Xy_a,
//                          dU/dz = sum_a l_a(x_i) Xz_a, then the point's force / Hessian
//                          contribution to node r (nh_contrib), summed over the lane's 3 points
//   sum      the group's 9 lanes through LDS, into the block's partials
// - 72 FMAs per point for the gradient (108 in the 27-lane form, whose stage 1
// recomputes each c-contraction on three lanes), and two LDS exchanges per
// task instead of two per point.  After the tasks, one thread per slot adds
// its vertex's partials in the owners order and runs the FP64 local solve.
#include "oe_kernels.hpp"
#include "oe_q2.hpp"

namespace oe {

namespace {

constexpr int kGroups = 7;   // tasks per wave
static_assert(kQ2PencilMaxInc % kGroups == 0, "whole rounds of tasks");
constexpr int kRounds = kQ2PencilMaxInc / kGroups;
static_assert(kQ2PencilMaxSlots <= 64, "phase 2 runs on the block's one wave");

// node index of tensor place t = a * 9 + b * 3 + c (quadrature points use the
// same numbering), and the tensor place of node n
struct PencilTables {
    int node[27];
    int place[27];
};
constexpr PencilTables make_pencil_tables() {
    PencilTables t{};
    for (int p = 0; p < 27; ++p) {
        t.node[p] = q2::tensor_to_node(p);
        t.place[p] = q2::kNodeABC[p];
    }
    return t;
}
__constant__ PencilTables kPencil = make_pencil_tables();

// A task group's LDS: the node values, then (once stage A has read them) the
// lanes' sums; T and Tz.  Components are outermost, so the 9 lanes of a group
// touch 9 consecutive words in every stage (or 3, broadcast), and groups sit
// 265 words apart (9 mod 32), so a half-wave's 32 lanes (3.5 groups) fall in
// distinct banks.  The sums use a row stride of 10 for the same reason.  On
// the 256-scene batch the layout with components innermost and 243-word
// groups took 764 bank-conflict cycles per wave.
constexpr int kSumStride = 10;
struct GroupLDS {
    float r1[89];   // stage 0: d[comp * 27 + a * 9 + b * 3 + c]; sum: s[m * kSumStride + v]
    float t[81];    // T_ab(k) at [comp * 27 + a * 9 + b * 3 + k]
    float tz[81];   // Tz_ab(k)
    float rec[9];   // an affine element's JX_inv, word m stored by lane m
    float pad[5];
};
static_assert(sizeof(GroupLDS) == 265 * sizeof(float) && 265 % 32 == 9, "group stride 9 mod 32 banks");
static_assert(8 * kSumStride + 9 <= 89, "the sums fit the first region");

// Waiting for the wave's own LDS operations orders them for every lane of the
// wave (a wave's LDS operations complete in order); the compiler must not move
// memory accesses across.
__device__ __forceinline__ void wave_sync() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ float l_at(const q2::Basis1D& B, int a, int i) { return B.L[a][i]; }
__device__ __forceinline__ float d_at(const q2::Basis1D& B, int a, int i) { return B.D[a][i]; }

// The 1D factors by Gauss point (row p: l_0, l_1, l_2 at p; the same for l')
// and the tensor place of each node, copied to LDS once per block.  A lane
// indexes them by its own (j, k) and node: from LDS these reads neither go
// through the texture path nor wait behind the vector-memory loads in flight.
struct PencilConst {
    float4 l[3];
    float4 d[3];
    int place[27];
};
__device__ __forceinline__ float sel3(int k, float v0, float v1, float v2) { return k == 0 ? v0 : (k == 1 ? v1 : v2); }

// The node ids of lane m's nodes (a, b, c), a = 0..2, for the group's task.
// own_l: lane l holds the block's owners entry l (a block has at most 35
// incidences; lanes past them hold 0, element 0).  task < 43 for every lane.
__device__ __forceinline__ void pencil_nodes(const FemView& f, int own_l, int task, int m, int (&nid)[3]) {
    const int pk = __shfl(own_l, task);
    const size_t base = static_cast<size_t>(pk >> 5) * 27;
#pragma unroll
    for (int a = 0; a < 3; ++a) nid[a] = gload(f.topo, base + kPencil.node[a * 9 + m]);
}

// One round of the block's tasks: task = round * 7 + g for the lane's group g.
// wq_i: the weights of the lane's points (i, j, k), i = 0..2.  nid: this
// round's node ids on entry, the next round's on return (loaded while this
// round computes, so a round waits on one dependent load, not three).
__device__ __forceinline__ void pencil_round(const PSView& ps, const FemView& f, int own_l, int (&nid)[3],
                                             const MaterialView& mt, const q2::Basis1D& B, const PencilConst& pc,
                                             const float (&wq_i)[3],
                                             int n_inc, int round, int g, int m, GroupLDS& sh, float (*part)[9]) {
    const int task = round * kGroups + g;
    // Every lane computes; only the stores are guarded.  A lane without a task
    // works on valid data: the owners entry of lane task is 0 (element 0,
    // node 0) from the block's last incidence on, and nid holds earlier node
    // ids or 0.  Lane 63 reads group 6's rows and must not write them.  (A
    // lane that skipped the work would need its values zeroed on every round,
    // and lane 63 never has a task.)
    const bool act = g < kGroups && task < n_inc;
    const bool writer = g < kGroups;
    const int p1 = m / 3, p2 = m - 3 * p1;   // stage 0 / A: (b, c) and (b, k); stage B / C: (j, k)
    const int pk = __shfl(own_l, task);
    const int eid = pk >> 5, r = pk & 31;
    float4 u[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = gload(static_cast<const float4*>(ps.p), nid[a]);
    // the element's record: JX_inv and |det J| (per point when not affine:
    // point (0, j, k) here, the others below).  An affine record is the same
    // for the group's 9 lanes: lane m loads word m and passes it on through
    // LDS, so the texture path returns 2 words per lane instead of 10.
    float JXi[9], V0 = 0.f, detB = 0.f, rec_m = 0.f;
    if (f.affine) {
        rec_m = gload(f.jxv, static_cast<size_t>(eid) * 10 + m);
        detB = gload(f.jxv, static_cast<size_t>(eid) * 10 + 9);
    } else {
        load_element_record(f, eid, kPencil.node[m], JXi, V0);
    }
    // stage 0: node 0 (tensor place 0) is the a = 0 node of lane m = 0 of the group
    {
        const int src = (g < kGroups ? g : 0) * 9;
        const float4 ref = make_float4(__shfl(u[0].x, src), __shfl(u[0].y, src), __shfl(u[0].z, src), 0.f);
        if (writer) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float d[3];
                rel_node(u[a], ref, ps.lo_unit, d);
#pragma unroll
                for (int c = 0; c < 3; ++c) sh.r1[c * 27 + a * 9 + m] = d[c];
            }
            if (f.affine) sh.rec[m] = rec_m;
        }
    }
    pencil_nodes(f, own_l, task + kGroups, m, nid);
    wave_sync();
    // stage A: lane (b, k) = (p1, p2)
    if (writer) {
        const float4 L = pc.l[p2], D = pc.d[p2];
        const float lc[3] = {L.x, L.y, L.z};
        const float dc[3] = {D.x, D.y, D.z};
        // all 27 reads first: the compiler cannot tell that the writes to T
        // miss them, and would otherwise wait on LDS after every write
        float x[27];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) x[(a * 3 + c) * 3 + cc] = sh.r1[c * 27 + a * 9 + p1 * 3 + cc];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float* xv = &x[(a * 3 + c) * 3];
                sh.t[c * 27 + a * 9 + m] = lc[0] * xv[0] + lc[1] * xv[1] + lc[2] * xv[2];
                sh.tz[c * 27 + a * 9 + m] = dc[0] * xv[0] + dc[1] * xv[1] + dc[2] * xv[2];
            }
        }
    }
    wave_sync();
    // stage B: lane (j, k) = (p1, p2); X, Xy, Xz for a = 0..2 in registers
    float X[3][3], Xy[3][3], Xz[3][3];
    {
        const float4 L = pc.l[p1], D = pc.d[p1];
        const float lb[3] = {L.x, L.y, L.z};
        const float db[3] = {D.x, D.y, D.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float* t = &sh.t[c * 27 + a * 9 + p2];
                const float* z = &sh.tz[c * 27 + a * 9 + p2];
                X[a][c] = lb[0] * t[0] + lb[1] * t[3] + lb[2] * t[6];
                Xy[a][c] = db[0] * t[0] + db[1] * t[3] + db[2] * t[6];
                Xz[a][c] = lb[0] * z[0] + lb[1] * z[3] + lb[2] * z[6];
            }
        }
    }
    // stage C and the material terms at the lane's points (i, j, k).  dN_r is
    // a product of the 1D factors (the shape table holds the same products:
    // q2_basis_1d checks it): r at tensor place (ra, rb, rc) has
    // dN_r(i, j, k) = (l'_ra(x_i) l_rb(y_j) l_rc(z_k), l_ra l'_rb l_rc, l_ra l_rb l'_rc).
    Acc9 acc;
    acc_zero(acc);
    {
        if (f.affine) {
#pragma unroll
            for (int i = 0; i < 9; ++i) JXi[i] = sh.rec[i];
        }
        const int tr = pc.place[r];
        const int ra = tr / 9, rb = tr / 3 % 3, rc = tr % 3;
        const float4 L1 = pc.l[p1], D1 = pc.d[p1], L2 = pc.l[p2], D2 = pc.d[p2];
        const float lrb = sel3(rb, L1.x, L1.y, L1.z);
        const float drb = sel3(rb, D1.x, D1.y, D1.z);
        const float lrc = sel3(rc, L2.x, L2.y, L2.z);
        const float drc = sel3(rc, D2.x, D2.y, D2.z);
        const float ll = lrb * lrc, dl = drb * lrc, ld = lrb * drc;
        // one point at a time (unrolled, the three points' material terms
        // interleave and hold registers that cost a wave per SIMD)
#pragma unroll 1
        for (int i = 0; i < 3; ++i) {
            const int q = kPencil.node[i * 9 + m];
            float Ju[9];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                Ju[c] = d_at(B, 0, i) * X[0][c] + d_at(B, 1, i) * X[1][c] + d_at(B, 2, i) * X[2][c];
                Ju[3 + c] = l_at(B, 0, i) * Xy[0][c] + l_at(B, 1, i) * Xy[1][c] + l_at(B, 2, i) * Xy[2][c];
                Ju[6 + c] = l_at(B, 0, i) * Xz[0][c] + l_at(B, 1, i) * Xz[1][c] + l_at(B, 2, i) * Xz[2][c];
            }
            float V;
            if (f.affine) {
                V = detB * sel3(i, wq_i[0], wq_i[1], wq_i[2]);
            } else if (i == 0) {
                V = V0;
            } else {
                load_element_record(f, eid, q, JXi, V);
            }
            const float lra = sel3(ra, l_at(B, 0, i), l_at(B, 1, i), l_at(B, 2, i));
            const float dra = sel3(ra, d_at(B, 0, i), d_at(B, 1, i), d_at(B, 2, i));
            const float dnr[3] = {dra * ll, lra * dl, lra * ld};
            nh_contrib(Ju, JXi, V, dnr, mt, acc);
        }
    }
    // sum over the group's 9 lanes: stage A has read r1, so it holds the sums now
    if (writer) {
#pragma unroll
        for (int v = 0; v < 9; ++v) sh.r1[m * kSumStride + v] = acc.v[v];
    }
    wave_sync();
    if (act) {
        float s = 0.f;
#pragma unroll
        for (int l = 0; l < 9; ++l) s += sh.r1[l * kSumStride + m];
        part[task][m] = s;
    }
    wave_sync();   // the next round's stage 0 rewrites r1
}

__device__ __forceinline__ void pencil_block(const PSView& ps, const FemView& fem, const OwnersView& own,
                                             const MaterialView& mt, const q2::Basis1D& basis, const float4* y,
                                             float damping, float dt, const int4 blk) {
    __shared__ GroupLDS sh[kGroups];
    __shared__ float part[kQ2PencilMaxInc][9];
    __shared__ PencilConst pc;
    const int lane = static_cast<int>(threadIdx.x);
    const int g = lane / 9, m = lane - 9 * g;
    // lane 63 (g = 7, m = 0) stores nothing; its stage-B reads use group 6's rows
    GroupLDS& mine = sh[g < kGroups ? g : kGroups - 1];
    // the weights of the lane's points, the same for every task
    const float wq_i[3] = {gload(fem.wq, kPencil.node[m]), gload(fem.wq, kPencil.node[9 + m]),
                           gload(fem.wq, kPencil.node[18 + m])};
    if (lane < 3) {
        pc.l[lane] = make_float4(l_at(basis, 0, lane), l_at(basis, 1, lane), l_at(basis, 2, lane), 0.f);
        pc.d[lane] = make_float4(d_at(basis, 0, lane), d_at(basis, 1, lane), d_at(basis, 2, lane), 0.f);
    }
    if (lane < 27) pc.place[lane] = kPencil.place[lane];
    const int own_l = lane < blk.w ? gload(own.own, blk.z + lane) : 0;
    // the slots' vertex ids and owners ranges, loaded now rather than after
    // the rounds (one dependent load less at the block's end; a lane past the
    // slots loads the last one, and every vertex block has a slot)
    const int my_slot = blk.x + min(lane, blk.y - 1);
    const int slot_vid = gload(own.slot_vid, my_slot);
    const int2 slot_nbo = gload(own.slot_nbo, my_slot);
    int nid[3];
    pencil_nodes(fem, own_l, g, m, nid);
    for (int round = 0; round < kRounds && round * kGroups < blk.w; ++round)
        pencil_round(ps, fem, own_l, nid, mt, basis, pc, wq_i, blk.w, round, g, m, mine, part);
    if (lane < blk.y) solve_slot_at(ps, slot_vid, slot_nbo, y, damping, dt, part, blk.z);
}

__global__ __launch_bounds__(64) void k_q2_pencil(Q2FusedArgs a) {
    const int4 blk = gload(a.blocks, a.first_block + xcd_block(blockIdx.x, gridDim.x));
    pencil_block(a.ps, a.fem, a.own, a.mt, a.basis, a.y, a.damping, a.dt, blk);
}

__global__ __launch_bounds__(64) void k_batch_q2_pencil(BatchQ2FusedArgs b) {
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const BatchSeg seg = find_seg(b.segs, b.seg_first, b.seg_count, bid);
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;   // uniform over the block
    const PSView ps = sd.ps;
    const FemView fem = sd.fem;
    const OwnersView own = sd.own;
    const MaterialView mt = sd.mt;
    const float4* y = sd.y;
    const float damping = sd.damping;
    const int4 blk = gload(sd.fblocks, seg.first + (bid - seg.block_begin));
    pencil_block(ps, fem, own, mt, b.basis, y, damping, b.dt, blk);
}

}  // namespace

void launch_q2_pencil(const Q2FusedArgs& a, hipStream_t s) {
    if (a.count > 0) hipLaunchKernelGGL(k_q2_pencil, dim3(a.count), dim3(64), 0, s, a);
}

void launch_batch_q2_pencil(const BatchQ2FusedArgs& b, int nblocks, hipStream_t s) {
    if (nblocks > 0) hipLaunchKernelGGL(k_batch_q2_pencil, dim3(nblocks), dim3(64), 0, s, b);
}

}  // namespace oe
