// VBD colour pass for 27-node hexahedra as two kernels, element tasks then
// vertex solves, with the displacement gradients of an element's 27
// quadrature points built by sum factorisation.
//
// A task is one incidence (element e, local node r) of a vertex of the colour;
// the colour's tasks are the contiguous run of the owners array that its
// vertices own.  A task takes 27 lanes of one 32-lane half of a wave (one
// lane per quadrature point; lanes 27-31 idle) and works through LDS:
//   stage 0  lane n loads node n's displacement, stored at its tensor place
//   stage 1  lane (a, b, k):  T  = sum_c l_c(z_k)  u_abc,  Tz = sum_c l_c'(z_k) u_abc
//   stage 2  lane (a, j, k):  X  = sum_b l_b(y_j) T_ab,  Xy = sum_b l_b'(y_j) T_ab,
//                             Xz = sum_b l_b(y_j) Tz_ab
//   stage 3  lane (i, j, k):  dU/dx = sum_a l_a'(x_i) X_a, dU/dy = sum_a l_a(x_i) Xy_a,
//                             dU/dz = sum_a l_a(x_i) Xz_a
// - 72 FMAs per point and coordinate triple instead of the 243 of the plain sum
// over 27 nodes, one load per node instead of one per (node, point) - then the
// point's force / Hessian contribution to node r.  The half's lanes add their
// contributions with shuffles and lane 0 writes the task's 9-float partial at
// the task's owners index.  The vertex kernel sums a vertex's partials in the
// owners order and runs the FP64 local solve, the arithmetic of the gather
// kernel's final lane.
#include "oe_kernels.hpp"
#include "oe_q2.hpp"

namespace oe {

namespace {

constexpr int kTasksPerBlock = 2;   // 64-thread blocks, one task per 32-lane half

// tensor place (a * 9 + b * 3 + c) of node n, and node of tensor place t
struct Q2Tables {
    int place[27];
    int node[27];
};
constexpr Q2Tables make_tables() {
    Q2Tables t{};
    for (int n = 0; n < 27; ++n) {
        t.place[n] = q2::kNodeABC[n];
        t.node[q2::kNodeABC[n]] = n;
    }
    return t;
}
__constant__ Q2Tables kQ2 = make_tables();

__device__ __forceinline__ float sel3(float v0, float v1, float v2, int k) {
    return k == 0 ? v0 : (k == 1 ? v1 : v2);
}

struct TaskLDS {
    float u[27][3];    // tensor place -> displacement
    float t[27][6];    // (a, b, k) -> T, Tz
    float x[27][9];    // (a, j, k) -> X, Xy, Xz
};

__device__ __forceinline__ void q2_task(const PSView& ps, const FemView& f, const MaterialView& mt,
                                        const q2::Basis1D& B, const int* __restrict__ own, float* partial, int first,
                                        int count, int task, TaskLDS& sh) {
    const int l = threadIdx.x & 31;
    bool valid = task < count;
    int eid = 0, r = 0;
    if (valid) {
        const int pk = own[first + task];
        eid = pk >> 5;
        r = pk & 31;
        valid = ps.mask[f.topo[static_cast<size_t>(eid) * 27 + r]] != 0;
    }
    const bool act = valid && l < 27;
    // lane l takes tensor indices (p0, p1, p2) = (l / 9, l / 3 % 3, l % 3) in every stage
    const int p0 = l / 9, p1 = (l / 3) % 3, p2 = l % 3;
    // stage 0
    if (act) {
        const float4 u = ps.p[f.topo[static_cast<size_t>(eid) * 27 + l]];
        const int t = kQ2.place[l];
        sh.u[t][0] = u.x;
        sh.u[t][1] = u.y;
        sh.u[t][2] = u.z;
    }
    __syncthreads();
    // stage 1: (a, b, k) = (p0, p1, p2)
    if (act) {
        const float l0 = sel3(B.L[0][0], B.L[0][1], B.L[0][2], p2), l1 = sel3(B.L[1][0], B.L[1][1], B.L[1][2], p2),
                    l2 = sel3(B.L[2][0], B.L[2][1], B.L[2][2], p2);
        const float d0 = sel3(B.D[0][0], B.D[0][1], B.D[0][2], p2), d1 = sel3(B.D[1][0], B.D[1][1], B.D[1][2], p2),
                    d2 = sel3(B.D[2][0], B.D[2][1], B.D[2][2], p2);
        const int ab = p0 * 3 + p1;
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            const float u0 = sh.u[ab * 3][comp], u1 = sh.u[ab * 3 + 1][comp], u2 = sh.u[ab * 3 + 2][comp];
            sh.t[l][comp] = l0 * u0 + l1 * u1 + l2 * u2;
            sh.t[l][3 + comp] = d0 * u0 + d1 * u1 + d2 * u2;
        }
    }
    __syncthreads();
    // stage 2: (a, j, k) = (p0, p1, p2); T of (a, b, k) sits at lane a * 9 + b * 3 + k
    if (act) {
        const float l0 = sel3(B.L[0][0], B.L[0][1], B.L[0][2], p1), l1 = sel3(B.L[1][0], B.L[1][1], B.L[1][2], p1),
                    l2 = sel3(B.L[2][0], B.L[2][1], B.L[2][2], p1);
        const float d0 = sel3(B.D[0][0], B.D[0][1], B.D[0][2], p1), d1 = sel3(B.D[1][0], B.D[1][1], B.D[1][2], p1),
                    d2 = sel3(B.D[2][0], B.D[2][1], B.D[2][2], p1);
        const int b0 = p0 * 9 + p2;
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            const float t0 = sh.t[b0][comp], t1 = sh.t[b0 + 3][comp], t2 = sh.t[b0 + 6][comp];
            const float z0 = sh.t[b0][3 + comp], z1 = sh.t[b0 + 3][3 + comp], z2 = sh.t[b0 + 6][3 + comp];
            sh.x[l][comp] = l0 * t0 + l1 * t1 + l2 * t2;
            sh.x[l][3 + comp] = d0 * t0 + d1 * t1 + d2 * t2;
            sh.x[l][6 + comp] = l0 * z0 + l1 * z1 + l2 * z2;
        }
    }
    __syncthreads();
    // stage 3: point (i, j, k) = (p0, p1, p2); X of (a, j, k) sits at lane a * 9 + j * 3 + k
    Acc9 acc;
    acc_zero(acc);
    if (act) {
        const float l0 = sel3(B.L[0][0], B.L[0][1], B.L[0][2], p0), l1 = sel3(B.L[1][0], B.L[1][1], B.L[1][2], p0),
                    l2 = sel3(B.L[2][0], B.L[2][1], B.L[2][2], p0);
        const float d0 = sel3(B.D[0][0], B.D[0][1], B.D[0][2], p0), d1 = sel3(B.D[1][0], B.D[1][1], B.D[1][2], p0),
                    d2 = sel3(B.D[2][0], B.D[2][1], B.D[2][2], p0);
        const int x0 = p1 * 3 + p2;
        float Ju[9];   // column-major: Ju[c * 3 + comp] = d u_comp / d xi_c
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            Ju[comp] = d0 * sh.x[x0][comp] + d1 * sh.x[x0 + 9][comp] + d2 * sh.x[x0 + 18][comp];
            Ju[3 + comp] = l0 * sh.x[x0][3 + comp] + l1 * sh.x[x0 + 9][3 + comp] + l2 * sh.x[x0 + 18][3 + comp];
            Ju[6 + comp] = l0 * sh.x[x0][6 + comp] + l1 * sh.x[x0 + 9][6 + comp] + l2 * sh.x[x0 + 18][6 + comp];
        }
        const int q = kQ2.node[l];
        float JXi[9], V;
        load_element_record(f, eid, q, JXi, V);
        nh_contrib(Ju, JXi, V, f.dN + (q * 27 + r) * 3, mt, acc);
    }
    group_sum<32>(acc);
    if (valid && l == 0) {
        float* out = partial + static_cast<size_t>(first + task) * 9;
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = acc.v[i];
    }
}

__global__ __launch_bounds__(64) void k_q2_tasks(Q2TaskArgs a) {
    __shared__ TaskLDS sh[kTasksPerBlock];
    const int half = threadIdx.x >> 5;
    const int task = xcd_block(blockIdx.x, gridDim.x) * kTasksPerBlock + half;
    q2_task(a.ps, a.fem, a.mt, a.basis, a.own, a.partial, a.first, a.count, task, sh[half]);
}

__global__ __launch_bounds__(64) void k_batch_q2_tasks(BatchQ2TaskArgs b) {
    __shared__ TaskLDS sh[kTasksPerBlock];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const BatchSeg seg = find_seg(b.segs, b.seg_first, b.seg_count, blk);
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;     // uniform over the block
    const int half = threadIdx.x >> 5;
    const int task = (blk - seg.block_begin) * kTasksPerBlock + half;
    q2_task(sd.ps, sd.fem, sd.mt, b.basis, sd.own.own, sd.partial, seg.first, seg.count, task, sh[half]);
}

// Sum of a vertex's partials (owners order) and the local solve.
__device__ __forceinline__ void solve_slot(const PSView& ps, const OwnersView& own, const float4* y,
                                           const float* __restrict__ partial, float damping, float dt, int slot) {
    const int vid = own.slot_vid[slot];
    if (ps.mask[vid] == 0) return;
    const int2 nbo = own.slot_nbo[slot];
    Acc9 acc;
    acc_zero(acc);
    for (int j = 0; j < nbo.x; ++j) {
        const float* p = partial + static_cast<size_t>(nbo.y + j) * 9;
#pragma unroll
        for (int i = 0; i < 9; ++i) acc.v[i] += p[i];
    }
    const float4 pv = ps.p[vid];
    double dx[3];
    vbd_local_solve(acc, pv, ps.last_p[vid], y[vid], ps.m[vid], damping, dt, dx);
    ps.p[vid] = make_float4(pv.x + static_cast<float>(dx[0]), pv.y + static_cast<float>(dx[1]),
                            pv.z + static_cast<float>(dx[2]), pv.w);
}

__global__ __launch_bounds__(256) void k_vertex_solve(VertexSolveArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    solve_slot(a.ps, a.own, a.y, a.partial, a.damping, a.dt, a.first_slot + i);
}

__global__ __launch_bounds__(256) void k_batch_vertex_solve(BatchGatherArgs b) {
    const int blk = static_cast<int>(blockIdx.x);
    const BatchSeg seg = find_seg(b.segs, b.seg_first, b.seg_count, blk);
    const SceneDev& sd = b.scenes[seg.scene];
    if (b.iteration >= sd.iterations) return;
    const int i = (blk - seg.block_begin) * 256 + threadIdx.x;
    if (i >= seg.count) return;
    solve_slot(sd.ps, sd.own, sd.y, sd.partial, sd.damping, b.dt, seg.first + i);
}

}  // namespace

int q2_tasks_per_block() { return kTasksPerBlock; }

void launch_q2_tasks(const Q2TaskArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    const int grid = (a.count + kTasksPerBlock - 1) / kTasksPerBlock;
    hipLaunchKernelGGL(k_q2_tasks, dim3(grid), dim3(64), 0, s, a);
}

void launch_batch_q2_tasks(const BatchQ2TaskArgs& b, int nblocks, hipStream_t s) {
    if (nblocks > 0) hipLaunchKernelGGL(k_batch_q2_tasks, dim3(nblocks), dim3(64), 0, s, b);
}

void launch_vertex_solve(const VertexSolveArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    hipLaunchKernelGGL(k_vertex_solve, dim3((a.count + 255) / 256), dim3(256), 0, s, a);
}

void launch_batch_vertex_solve(const BatchGatherArgs& b, int nblocks, hipStream_t s) {
    if (nblocks > 0) hipLaunchKernelGGL(k_batch_vertex_solve, dim3(nblocks), dim3(256), 0, s, b);
}

}  // namespace oe
