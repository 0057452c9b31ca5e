// VBD colour pass for 27-node hexahedra in one kernel: element tasks whose
// displacement gradients at the 27 quadrature points are built by sum
// factorisation, then, in the same block, the vertex solves.
//
// A block (256 threads, 8 half-waves) takes a run of consecutive slots of one
// colour (a vertex block {first slot, slot count, first owners index,
// incidence count}, at most 32 incidences).  A task is one incidence (element
// e, local node r); half h takes the block's tasks h, h + 8, h + 16, ...
// A task uses 27 lanes of its half (one lane per quadrature point; lanes 27-31
// idle) and works through LDS:
//   stage 0  lane n loads node n's displacement, stored at its tensor place
//   stage 1  lane (a, j, k):  with T_b = sum_c l_c(z_k) u_abc, Tz_b = sum_c l_c'(z_k) u_abc,
//                             X = sum_b l_b(y_j) T_b,  Xy = sum_b l_b'(y_j) T_b,
//                             Xz = sum_b l_b(y_j) Tz_b
//   stage 2  lane (i, j, k):  dU/dx = sum_a l_a'(x_i) X_a, dU/dy = sum_a l_a(x_i) Xy_a,
//                             dU/dz = sum_a l_a(x_i) Xz_a
// - 108 FMAs per point and coordinate triple instead of the 243 of the plain
// sum over 27 nodes, one load per node instead of one per (node, point) - then
// the point's force / Hessian contribution to node r.  The pass is bound by
// each half's chain of dependent LDS round trips, not by its arithmetic: the
// two contractions of stage 1 share one stage (36 more FMAs than a stage each,
// one wait less), and u and x have their own LDS rows (one more wait less).
// The half's lanes add their contributions with DPP and the task's 9-float
// partial goes to LDS.  After a block barrier one thread per slot adds its
// vertex's partials in the owners order and runs the FP64 local solve, the
// arithmetic of the gather kernel's final lane.
#include "oe_kernels.hpp"
#include "oe_q2.hpp"

namespace oe {

namespace {

constexpr int kHalves = kQ2FusedThreads / 32;            // half-waves per block
constexpr int kRuns = kQ2FusedIncPerBlock / kHalves;     // tasks a half works through in turn
static_assert(kRuns * kHalves == kQ2FusedIncPerBlock, "tasks per block");
static_assert(kQ2FusedMaxSlots <= 64, "phase 2 runs on the first wave");

// The stages exchange data only between the lanes of one half-wave.  A wave's
// LDS operations are performed in order, so waiting for its own LDS operations
// (and keeping the compiler from moving memory accesses across) is enough; no
// workgroup barrier is needed.
__device__ __forceinline__ void half_sync() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// tensor place (a * 9 + b * 3 + c) of node n, and node of tensor place t;
// entries 27-31 repeat entry 26 for the idle lanes
struct Q2Tables {
    int place[32];
    int node[32];
};
constexpr Q2Tables make_tables() {
    Q2Tables t{};
    for (int n = 0; n < 27; ++n) {
        t.place[n] = q2::kNodeABC[n];
        t.node[q2::kNodeABC[n]] = n;
    }
    for (int n = 27; n < 32; ++n) {
        t.place[n] = t.place[26];
        t.node[n] = t.node[26];
    }
    return t;
}
// The two tables packed 12 entries of 5 bits to a 64-bit word, read with
// shifts: a per-lane index into a table in memory is a vector load, and a wait
// for it would cover every load issued before it.
constexpr unsigned long long pack_word(const int (&v)[32], int w) {
    unsigned long long x = 0;
    for (int i = 0; i < 12 && 12 * w + i < 27; ++i) x |= static_cast<unsigned long long>(v[12 * w + i]) << (5 * i);
    return x;
}
constexpr Q2Tables kQ2Host = make_tables();
constexpr unsigned long long kPlaceW0 = pack_word(kQ2Host.place, 0), kPlaceW1 = pack_word(kQ2Host.place, 1),
                             kPlaceW2 = pack_word(kQ2Host.place, 2);
constexpr unsigned long long kNodeW0 = pack_word(kQ2Host.node, 0), kNodeW1 = pack_word(kQ2Host.node, 1),
                             kNodeW2 = pack_word(kQ2Host.node, 2);
__device__ __forceinline__ int unpack27(unsigned long long w0, unsigned long long w1, unsigned long long w2, int i) {
    const unsigned long long w = i < 12 ? w0 : (i < 24 ? w1 : w2);
    const int sh = 5 * (i < 12 ? i : (i < 24 ? i - 12 : i - 24));
    return static_cast<int>((w >> sh) & 31ull);
}
constexpr bool packed_tables_match() {
    for (int i = 0; i < 27; ++i) {
        const int w = i / 12, sh = 5 * (i % 12);
        const unsigned long long pw = w == 0 ? kPlaceW0 : (w == 1 ? kPlaceW1 : kPlaceW2);
        const unsigned long long nw = w == 0 ? kNodeW0 : (w == 1 ? kNodeW1 : kNodeW2);
        if (static_cast<int>((pw >> sh) & 31ull) != kQ2Host.place[i]) return false;
        if (static_cast<int>((nw >> sh) & 31ull) != kQ2Host.node[i]) return false;
    }
    return true;
}
static_assert(packed_tables_match(), "packed Q2 node tables");

// The three 1D factors and their derivatives at one Gauss coordinate.
struct Basis3 {
    float l0, l1, l2, d0, d1, d2;
};
// The factors are read as scalars and selected per lane: a select of the
// argument words themselves would become a load with a per-lane address.
__device__ __forceinline__ float uniform(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ Basis3 basis_at(const q2::Basis1D& B, int i) {
    const bool i0 = i == 0, i1 = i == 1;
    auto pick = [&](const float (&v)[3]) {
        const float a = uniform(v[0]), b = uniform(v[1]), c = uniform(v[2]);
        return i0 ? a : (i1 ? b : c);
    };
    return Basis3{pick(B.L[0]), pick(B.L[1]), pick(B.L[2]), pick(B.D[0]), pick(B.D[1]), pick(B.D[2])};
}

// One half's LDS.  Rows 27-31 are written by the idle lanes and never read.
// u and x do not overlap, so a task's stage 0 may follow the previous task's
// stage 2 without a wait: a wave's LDS operations are performed in order.
struct TaskLDS {
    float4 u[32];      // tensor place -> displacement
    float4 x[32][3];   // (a, j, k) -> X, Xy, Xz
};

// x + (x at lane ctrl(lane)) for the rows in ROWS; other rows add 0
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_add(float x) {
    return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xF, false));
}

// Sum over a 16-lane row (quad_perm 1,0,3,2; quad_perm 2,3,0,1; mirror in 8
// lanes; mirror in 16 lanes): every lane of the row ends with the row total.
__device__ __forceinline__ float row_sum(float x) {
    x = dpp_add<0xB1>(x);
    x = dpp_add<0x4E>(x);
    x = dpp_add<0x141>(x);
    return dpp_add<0x140>(x);
}

// Totals of a and b over each 32-lane half (two 16-lane rows).
// v_permlane16_swap exchanges the second row of its first operand with the
// first row of its second, so the sum of the two results holds a's partials
// of both rows (added lane by lane) in the half's first row and b's in its
// second; a row sum then leaves a's half total in every lane of the first row
// and b's in every lane of the second.  Half the cross-lane work of two
// separate half sums.
__device__ __forceinline__ float half_sum_pair(float a, float b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return row_sum(__uint_as_float(r[0]) + __uint_as_float(r[1]));
}

// Lane 0 of each half, in every lane of that half: v_mov_b32_dpp
// row_newbcast:0 gives each 16-lane row its first lane's value, then
// v_permlane16_swap of that with itself puts rows (0, 0, 2, 2) in its first
// result (lanes 0 and 32).  Two VALU instructions, no branch, no LDS.
__device__ __forceinline__ float half_bcast0(float x) {
    const int r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x150, 0xF, 0xF, false);
    return __uint_as_float(__builtin_amdgcn_permlane16_swap(r, r, false, false)[0]);
}

// Lane-constant part of a half's work: its point's tensor indices, the 1D
// factors of the three stages and its LDS rows.
struct LanePlan {
    int l, node, place, q;
    int a9, x0;   // stage-1 u rows a * 9 + ..., stage-2 x rows x0 + 9 a
    Basis3 s1, s2, s3;
};
__device__ __forceinline__ LanePlan lane_plan(const q2::Basis1D& B) {
    LanePlan p;
    p.l = threadIdx.x & 31;
    const int lc = p.l < 27 ? p.l : 26;
    const int p0 = lc / 9, p1 = (lc / 3) % 3, p2 = lc % 3;
    p.node = lc;
    p.place = unpack27(kPlaceW0, kPlaceW1, kPlaceW2, lc);
    p.q = unpack27(kNodeW0, kNodeW1, kNodeW2, lc);
    p.a9 = p0 * 9;
    p.x0 = p1 * 3 + p2;
    p.s1 = basis_at(B, p2);
    p.s2 = basis_at(B, p1);
    p.s3 = basis_at(B, p0);
    return p;
}

// One task: incidence (eid, r), the lane's node displacement u already loaded.
// Returns the lane's contribution to node r (zero on the idle lanes).
__device__ __forceinline__ void q2_task_body(const FemView& f, const MaterialView& mt, const LanePlan& P, int eid,
                                             int r, float4 u, float unit, TaskLDS& sh, Acc9& acc) {
    // the point's element record and dN_r, loaded while the stages run
    float JXi[9], V;
    load_element_record(f, eid, P.q, JXi, V);
    if (P.l >= 27) V = 0.f;
    const float dnr[3] = {gload(f.dN, (P.q * 27 + r) * 3), gload(f.dN, (P.q * 27 + r) * 3 + 1),
                          gload(f.dN, (P.q * 27 + r) * 3 + 2)};
    // stage 0: the node's value relative to node 0 of the element (lane 0 of
    // the half; half_bcast0)
    {
        const float4 ref = make_float4(half_bcast0(u.x), half_bcast0(u.y), half_bcast0(u.z), 0.f);
        float d[3];
        rel_node(u, ref, unit, d);
        sh.u[P.place] = make_float4(d[0], d[1], d[2], 0.f);
    }
    half_sync();
    // stage 1: lane (a, j, k); u_abc sits at row a * 9 + b * 3 + c
    {
        const Basis3& bz = P.s1;
        const Basis3& by = P.s2;
        float X[3] = {0.f, 0.f, 0.f}, Xy[3] = {0.f, 0.f, 0.f}, Xz[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float4 u0 = sh.u[P.a9 + b * 3], u1 = sh.u[P.a9 + b * 3 + 1], u2 = sh.u[P.a9 + b * 3 + 2];
            const float T[3] = {bz.l0 * u0.x + bz.l1 * u1.x + bz.l2 * u2.x, bz.l0 * u0.y + bz.l1 * u1.y + bz.l2 * u2.y,
                                bz.l0 * u0.z + bz.l1 * u1.z + bz.l2 * u2.z};
            const float Tz[3] = {bz.d0 * u0.x + bz.d1 * u1.x + bz.d2 * u2.x,
                                 bz.d0 * u0.y + bz.d1 * u1.y + bz.d2 * u2.y,
                                 bz.d0 * u0.z + bz.d1 * u1.z + bz.d2 * u2.z};
            const float lb = b == 0 ? by.l0 : (b == 1 ? by.l1 : by.l2);
            const float db = b == 0 ? by.d0 : (b == 1 ? by.d1 : by.d2);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                X[c] += lb * T[c];
                Xy[c] += db * T[c];
                Xz[c] += lb * Tz[c];
            }
        }
        sh.x[P.l][0] = make_float4(X[0], X[1], X[2], 0.f);
        sh.x[P.l][1] = make_float4(Xy[0], Xy[1], Xy[2], 0.f);
        sh.x[P.l][2] = make_float4(Xz[0], Xz[1], Xz[2], 0.f);
    }
    half_sync();
    // stage 2: point (i, j, k); X of (a, j, k) sits at row a * 9 + j * 3 + k
    float Ju[9];   // column-major: Ju[c * 3 + comp] = d u_comp / d xi_c
    {
        const float4 xa = sh.x[P.x0][0], xb = sh.x[P.x0 + 9][0], xc = sh.x[P.x0 + 18][0];
        const float4 ya = sh.x[P.x0][1], yb = sh.x[P.x0 + 9][1], yc = sh.x[P.x0 + 18][1];
        const float4 za = sh.x[P.x0][2], zb = sh.x[P.x0 + 9][2], zc = sh.x[P.x0 + 18][2];
        const Basis3& b = P.s3;
        Ju[0] = b.d0 * xa.x + b.d1 * xb.x + b.d2 * xc.x;
        Ju[1] = b.d0 * xa.y + b.d1 * xb.y + b.d2 * xc.y;
        Ju[2] = b.d0 * xa.z + b.d1 * xb.z + b.d2 * xc.z;
        Ju[3] = b.l0 * ya.x + b.l1 * yb.x + b.l2 * yc.x;
        Ju[4] = b.l0 * ya.y + b.l1 * yb.y + b.l2 * yc.y;
        Ju[5] = b.l0 * ya.z + b.l1 * yb.z + b.l2 * yc.z;
        Ju[6] = b.l0 * za.x + b.l1 * zb.x + b.l2 * zc.x;
        Ju[7] = b.l0 * za.y + b.l1 * zb.y + b.l2 * zc.y;
        Ju[8] = b.l0 * za.z + b.l1 * zb.z + b.l2 * zc.z;
    }
    nh_contrib(Ju, JXi, V, dnr, mt, acc);
}

// A half's tasks task0, task0 + kHalves, ... (fewer than kRuns when the block
// has fewer incidences).  The loads run ahead of the work: task t's
// displacement, the next task's node index and the one after's owners entry
// are issued before task t's stages; indices past the block's last incidence
// are clamped to it.  Partials go to part[task] (nine lanes, one float each).
// after_owner / after_topo run right after the first owners-entry and node-id
// loads: the latency form issues its slot state there.  A wave whose two
// halves have no task returns at once (it holds no slot either).  Past that the
// code is straight-line: a half without a task works on the block's last one
// and stores nothing, so that no wait for a load has to cover the loads of
// another path, which would make the compiler wait for every load in flight.
template <class AfterOwner, class AfterTopo>
__device__ __forceinline__ void q2_task_run(const PSView& ps, const FemView& f, const MaterialView& mt,
                                           const LanePlan& P, const int* __restrict__ own, int first, int count,
                                           int task0, TaskLDS& sh, float (*part)[9], AfterOwner&& after_owner,
                                           AfterTopo&& after_topo) {
    if (__builtin_amdgcn_readfirstlane(task0 & ~1) >= count) return;
    const bool idle = task0 >= count;   // uniform over the half
    const int start = idle ? count - 1 : task0;
    const int last = first + count - 1;
    int pk = gload(own, first + start);
    int pk_next = gload(own, min(first + start + kHalves, last));
    // the scheduler would hoist the hooks' loads (which do not depend on the
    // chain) above the chain's own, and every wait of the chain would then
    // cover them
    __builtin_amdgcn_sched_barrier(0);
    after_owner();
    __builtin_amdgcn_sched_barrier(0);
    int vid = gload(f.topo, static_cast<size_t>(pk >> 5) * 27 + P.node);
    __builtin_amdgcn_sched_barrier(0);
    after_topo();
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll 1
    for (int task = start; task < count; task += kHalves) {
        const float4 u = gload(static_cast<const float4*>(ps.p), vid);
        const int eid = pk >> 5, r = pk & 31;
        pk = pk_next;
        vid = gload(f.topo, static_cast<size_t>(pk >> 5) * 27 + P.node);
        pk_next = gload(own, min(first + task + 2 * kHalves, last));
        Acc9 acc;
        acc_zero(acc);
        q2_task_body(f, mt, P, eid, r, u, ps.lo_unit, sh, acc);
        // the half's totals: value 2k in row 0 and value 2k + 1 in row 1 of s[k]
        float s[5];
#pragma unroll
        for (int k = 0; k < 4; ++k) s[k] = half_sum_pair(acc.v[2 * k], acc.v[2 * k + 1]);
        s[4] = half_sum_pair(acc.v[8], 0.f);
        // lane k of row w writes value 2k + w
        const int k = P.l & 15, i = 2 * k + (P.l >> 4);
        if (!idle && k < 5 && i < 9) {
            float v = s[0];
#pragma unroll
            for (int j = 1; j < 5; ++j) v = k == j ? s[j] : v;
            part[task][i] = v;
        }
    }
}

// Phase 2: thread t < slot count adds the partials of slot first_slot + t in
// the owners order and solves.  t is rebuilt from the wave's first thread (a
// scalar) and the lane count, which the compiler recomputes where threadIdx.x
// would otherwise be kept in scratch across the task loop.
__device__ __forceinline__ void q2_solve_slots(const PSView& ps, const OwnersView& own, const float4* y,
                                              float damping, float dt, const int4 blk, const float (*part)[9],
                                              int wave_first) {
    const int t = wave_first + static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
    if (t >= blk.y) return;
    solve_slot_direct(ps, own, y, damping, dt, blk.x + t, part, blk.z);
}

template <bool LAT>
__device__ __forceinline__ void q2_block(const PSView& ps, const FemView& fem, const OwnersView& own,
                                         const MaterialView& mt, const q2::Basis1D& basis, const float4* y,
                                         float damping, float dt, const int4 blk) {
    __shared__ TaskLDS sh[kHalves];
    __shared__ float part[LAT ? kQ2FusedIncPerBlockLat : kQ2FusedIncPerBlock][9];
    const int half = threadIdx.x >> 5;
    if constexpr (LAT) {
        // the slot state rides along the element chain (a thread past the
        // block's slots loads the last one and does not use it; every vertex
        // block has a slot)
        const int t = static_cast<int>(threadIdx.x);
        SlotState st{};
        const int ts = blk.x + min(t, blk.y - 1);
        const LanePlan P = lane_plan(basis);
        q2_task_run(ps, fem, mt, P, own.own, blk.z, blk.w, half, sh[half], part,
                    [&] { slot_state_level1(own, ts, st); }, [&] { slot_state_level2(ps, y, st); });
        __syncthreads();
        if (t < blk.y) {
            slot_state_settle(st);
            if (st.free != 0) solve_slot_from_lds(ps, st, part, blk.z, damping, dt);
        }
    } else {
        const int wave_first = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) & ~63);
        const LanePlan P = lane_plan(basis);
        q2_task_run(ps, fem, mt, P, own.own, blk.z, blk.w, half, sh[half], part, [] {}, [] {});
        __syncthreads();
        q2_solve_slots(ps, own, y, damping, dt, blk, part, wave_first);
    }
}

template <bool LAT>
__global__ __launch_bounds__(kQ2FusedThreads, LAT ? 1 : 6) void k_q2_fused(Q2FusedArgs a) {
    const int4 blk = gload(a.blocks, a.first_block + xcd_block(blockIdx.x, gridDim.x));
    q2_block<LAT>(a.ps, a.fem, a.own, a.mt, a.basis, a.y, a.damping, a.dt, blk);
}

// The throughput form is held to 80 VGPRs (6 waves per SIMD), which it fits
// without scratch: the element-record index is formed in 32 bits and the
// phase-2 thread index is rebuilt from a scalar (q2_solve_slots).  At 5 waves
// the 256-scene batch stepped 3% slower.  The latency form keeps its
// registers.
template <bool LAT>
__global__ __launch_bounds__(kQ2FusedThreads, LAT ? 1 : 6) void k_batch_q2_fused(BatchQ2FusedArgs b) {
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const BatchSeg seg = find_seg(b.segs, b.seg_first, b.seg_count, bid);
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;     // uniform over the block
    // the scene's views, read before the first store (which the compiler
    // could not otherwise tell apart from the scene table)
    const PSView ps = sd.ps;
    const FemView fem = sd.fem;
    const OwnersView own = sd.own;
    const MaterialView mt = sd.mt;
    const float4* y = sd.y;
    const float damping = sd.damping;
    const int4 blk = gload(LAT ? sd.fblocks_lat : sd.fblocks, seg.first + (bid - seg.block_begin));
    q2_block<LAT>(ps, fem, own, mt, b.basis, y, damping, b.dt, blk);
}

}  // namespace

void launch_q2_fused(const Q2FusedArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    if (a.latency)
        hipLaunchKernelGGL(k_q2_fused<true>, dim3(a.count), dim3(kQ2FusedThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_q2_fused<false>, dim3(a.count), dim3(kQ2FusedThreads), 0, s, a);
}

void launch_batch_q2_fused(const BatchQ2FusedArgs& b, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    if (b.latency)
        hipLaunchKernelGGL(k_batch_q2_fused<true>, dim3(nblocks), dim3(kQ2FusedThreads), 0, s, b);
    else
        hipLaunchKernelGGL(k_batch_q2_fused<false>, dim3(nblocks), dim3(kQ2FusedThreads), 0, s, b);
}

}  // namespace oe
