// Per-vertex kernels of the integrators and constraints.  Each is one pass
// over SoA particle state with coalesced 16-byte accesses; they are launch- or
// HBM-bound and carry no reuse.  Positions are stored as displacements from
// the rest position (init_p); the integrators are linear in the positions and
// read them unchanged, the constraints convert to absolute positions.
//   VBD predictor             src/GPU/VBD/GPU_VBD.cu:6-18
//   VBD velocity              src/GPU/VBD/GPU_VBD.cu:20-25
//   Chebyshev acceleration    src/GPU/VBD/GPU_VBD.cu:27-35
//   semi-implicit Euler       src/GPU/GPU_Integrator.cu:27-34
//   XPBD velocity + damping   src/GPU/PBD/GPU_PBD.cu:8-18
//   fixed vertices            src/GPU/GPU_Constraint.cu:8-19
//   box limit                 src/GPU/GPU_Constraint.cu:70-81
//   Jacobi colour copy        src/GPU/VBD/GPU_VBD_FEM.cu:402-413
#include "oe_elementwise.hpp"
#include "oe_gather.hpp"
#include "oe_kernels.hpp"

namespace oe {

namespace {

constexpr int kBlock = 256;
inline int grid_for(long long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ float4 f4(float x, float y, float z, float w) { return make_float4(x, y, z, w); }

// y = p + ((p - last_p) / dt + (g + f w) dt) dt ; last_p = p ; p = y ; f = 0,
// on the compensated state in FP64.
__device__ __forceinline__ void vbd_predict_one(const PSView& ps, float4* y, int i, float dt, float gx, float gy,
                                                float gz, float4* prev_it_p) {
    const float4 p = ps.p[i];
    const float4 f = ps.f[i];
    const float w = ps.w[i];
    const dvec3 P = dunpack(p, ps.lo_unit), LP = dunpack(ps.last_p[i], ps.lo_unit);
    if (prev_it_p) prev_it_p[i] = p;
    const double ddt = dt;
    const double ax = gx + static_cast<double>(f.x) * w, ay = gy + static_cast<double>(f.y) * w,
                 az = gz + static_cast<double>(f.z) * w;
    const float4 Y = dpack(dvec3{P.x + ((P.x - LP.x) / ddt + ax * ddt) * ddt, P.y + ((P.y - LP.y) / ddt + ay * ddt) * ddt,
                                 P.z + ((P.z - LP.z) / ddt + az * ddt) * ddt},
                           ps.lo_inv);
    y[i] = Y;
    ps.last_p[i] = p;
    ps.p[i] = Y;
    ps.f[i] = f4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(kBlock) void k_vbd_predict(PSView ps, float4* y, float4* prev_it_p, float dt, float gx,
                                                        float gy, float gz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.n) return;
    vbd_predict_one(ps, y, i, dt, gx, gy, gz, prev_it_p);
}

// v = (p - last_p) / dt
__device__ __forceinline__ void vbd_velocity_one(const PSView& ps, int i, float dt) {
    const dvec3 P = dunpack(ps.p[i], ps.lo_unit), LP = dunpack(ps.last_p[i], ps.lo_unit);
    const double ddt = dt;
    ps.v[i] = f4(static_cast<float>((P.x - LP.x) / ddt), static_cast<float>((P.y - LP.y) / ddt),
                 static_cast<float>((P.z - LP.z) / ddt), 0.f);
}

__global__ __launch_bounds__(kBlock) void k_vbd_velocity(PSView ps, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.n) return;
    vbd_velocity_one(ps, i, dt);
}

// p = prev2 + omega (p - prev2) from the third iteration on; prev2 = prev1, prev1 = p
__global__ __launch_bounds__(kBlock) void k_chebyshev(PSView ps, float4* prev_it_p, float4* prev_it2_p, int it,
                                                      float omega) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.n) return;
    float4 p = ps.p[i];
    if (it >= 2) {
        const dvec3 P = dunpack(p, ps.lo_unit), Q = dunpack(prev_it2_p[i], ps.lo_unit);
        const double om = omega;
        p = dpack(dvec3{Q.x + om * (P.x - Q.x), Q.y + om * (P.y - Q.y), Q.z + om * (P.z - Q.z)}, ps.lo_inv);
        ps.p[i] = p;
    }
    prev_it2_p[i] = prev_it_p[i];
    prev_it_p[i] = p;
}

// last_p = p ; v += (g + f w) dt ; p += v dt ; f = 0
__global__ __launch_bounds__(kBlock) void k_semi_implicit(PSView ps, float dt, float gx, float gy, float gz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.n) return;
    float4 p = ps.p[i];
    float4 v = ps.v[i];
    const float4 f = ps.f[i];
    const float w = ps.w[i];
    ps.last_p[i] = p;
    v.x += (gx + f.x * w) * dt;
    v.y += (gy + f.y * w) * dt;
    v.z += (gz + f.z * w) * dt;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    ps.v[i] = v;
    ps.p[i] = p;
    ps.f[i] = f4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(kBlock) void k_pbd_velocity(PSView ps, float dt, float global_damping) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.n) return;
    const float4 p = ps.p[i], lp = ps.last_p[i];
    float vx = (p.x - lp.x) / dt, vy = (p.y - lp.y) / dt, vz = (p.z - lp.z) / dt;
    const float nv = sqrtf(vx * vx + vy * vy + vz * vz);
    if (nv > 1e-7f) {
        const float coef = global_damping * dt * ps.w[i];
        const float damp = -nv * (coef > 1.f ? 1.f : coef);
        const float inv = 1.f / nv;
        vx += vx * inv * damp;
        vy += vy * inv * damp;
        vz += vz * inv * damp;
    }
    ps.v[i] = f4(vx, vy, vz, 0.f);
}

// target = offset + com + R (X - com), X = init_p (rest position);  p = target,
// v = f = 0, mask = 0.  p holds the displacement x - X, so the stored value is
// offset + (R d - d) with d = X - com (exactly offset when R = I), with no
// low-order part (w = 0).
__device__ __forceinline__ float4 fix_displacement(float4 X, const float com[3], const float off[3], const float R[9]) {
    const float dx = X.x - com[0], dy = X.y - com[1], dz = X.z - com[2];
    // R is column-major; R d = col0 dx + col1 dy + col2 dz
    const float rx = R[0] * dx + R[3] * dy + R[6] * dz;
    const float ry = R[1] * dx + R[4] * dy + R[7] * dz;
    const float rz = R[2] * dx + R[5] * dy + R[8] * dz;
    return f4(off[0] + (rx - dx), off[1] + (ry - dy), off[2] + (rz - dz), 0.f);
}

__global__ __launch_bounds__(kBlock) void k_fix(PSView ps, const int* ids, int n, FixTransform t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int id = ids[i];
    ps.p[id] = fix_displacement(ps.init_p[id], t.com, t.offset, t.R);
    ps.v[id] = f4(0.f, 0.f, 0.f, 0.f);
    ps.f[id] = f4(0.f, 0.f, 0.f, 0.f);
    ps.mask[id] = 0;
}

// clamp the absolute position X + u into [pmin, pmax]; u is rewritten only
// when a component is clamped, and the return value says whether one was
__device__ __forceinline__ bool box_clamp(float4& u, float4 X, const float pmin[3], const float pmax[3]) {
    const float x = X.x + u.x, y = X.y + u.y, z = X.z + u.z;
    const bool cx = x < pmin[0] || x > pmax[0], cy = y < pmin[1] || y > pmax[1], cz = z < pmin[2] || z > pmax[2];
    if (cx) u.x = fminf(fmaxf(x, pmin[0]), pmax[0]) - X.x;
    if (cy) u.y = fminf(fmaxf(y, pmin[1]), pmax[1]) - X.y;
    if (cz) u.z = fminf(fmaxf(z, pmin[2]), pmax[2]) - X.z;
    return cx || cy || cz;
}

// The compensated form: the absolute position X + (hi + lo) in FP64 against the
// box; a clamped component becomes bound - X (packed again); written only when
// a component is clamped.
__device__ __forceinline__ void box_clamp_comp(const PSView& ps, int i, float4 X, const float pmin[3],
                                               const float pmax[3]) {
    const float4 h = ps.p[i];
    // the low-order part (|lo| <= half an ulp of hi) can only matter within a
    // relative 1e-4 of a bound: elsewhere the float sum decides alone
    const float xf[3] = {X.x + h.x, X.y + h.y, X.z + h.z};
    bool near = false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        near = near || xf[k] < pmin[k] + 1e-4f * (1.f + fabsf(pmin[k])) ||
               xf[k] > pmax[k] - 1e-4f * (1.f + fabsf(pmax[k]));
    if (!near) return;
    dvec3 U = dunpack(h, ps.lo_unit);
    const double x = X.x + U.x, y = X.y + U.y, z = X.z + U.z;
    const bool cx = x < pmin[0] || x > pmax[0], cy = y < pmin[1] || y > pmax[1], cz = z < pmin[2] || z > pmax[2];
    if (!(cx || cy || cz)) return;
    if (cx) U.x = fmin(fmax(x, static_cast<double>(pmin[0])), static_cast<double>(pmax[0])) - X.x;
    if (cy) U.y = fmin(fmax(y, static_cast<double>(pmin[1])), static_cast<double>(pmax[1])) - X.y;
    if (cz) U.z = fmin(fmax(z, static_cast<double>(pmin[2])), static_cast<double>(pmax[2])) - X.z;
    ps.p[i] = dpack(U, ps.lo_inv);
}

__global__ __launch_bounds__(kBlock) void k_box(PSView ps, float3 pmin, float3 pmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.n) return;
    if (ps.mask[i] == 0) return;
    const float lo[3] = {pmin.x, pmin.y, pmin.z}, hi[3] = {pmax.x, pmax.y, pmax.z};
    if (ps.comp) {
        box_clamp_comp(ps, i, ps.init_p[i], lo, hi);
        return;
    }
    float4 u = ps.p[i];
    if (box_clamp(u, ps.init_p[i], lo, hi)) ps.p[i] = u;   // written only when clamped
}

// Fixation then box limit over one scene's vertices, fused as in the batched
// form: a fixed vertex (flag set) is snapped and masked, every other vertex
// with mask 1 is clamped; the same result as k_fix then k_box, since the box
// skips masked vertices.
__global__ __launch_bounds__(kBlock) void k_fix_box(PSView ps, const unsigned char* fixed, FixTransform t, float3 pmin,
                                                    float3 pmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.n) return;
    const bool is_fixed = gload(fixed, i) != 0;
    const int mask = gload(static_cast<const int*>(ps.mask), i);
    float4 u = gload(static_cast<const float4*>(ps.p), i);
    const float4 X = gload(static_cast<const float4*>(ps.init_p), i);
    if (is_fixed) {
        gstore(ps.p, i, fix_displacement(X, t.com, t.offset, t.R));
        gstore(ps.v, i, f4(0.f, 0.f, 0.f, 0.f));
        gstore(ps.f, i, f4(0.f, 0.f, 0.f, 0.f));
        gstore(ps.mask, i, 0);
        return;
    }
    if (mask == 0) return;
    const float lo[3] = {pmin.x, pmin.y, pmin.z}, hi[3] = {pmax.x, pmax.y, pmax.z};
    if (ps.comp) {
        box_clamp_comp(ps, i, X, lo, hi);
        return;
    }
    if (box_clamp(u, X, lo, hi)) gstore(ps.p, i, u);   // written only when clamped
}

__global__ __launch_bounds__(kBlock) void k_copy_colour(PSView ps, const int* slot_vid, int first, int count,
                                                        const float4* p_new) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int vid = slot_vid[first + i];
    if (ps.mask[vid] == 0) return;
    ps.p[vid] = p_new[vid];
}

// Internal force of each vertex from the element partials of k_explicit_elem
// (9 floats per (element, local node), in owners order, so a vertex's partials
// are contiguous): they are added in owners order, then the damping term of
// the reference's force pass, f = sum(-V P g) - k H v
// (src/GPU/Explicit/GPU_FEM_Explicit.cu:45-62), overwrites ps.f.  Four
// incidences are loaded at a time so that their loads overlap.  With
// integrate set (a semi-implicit integrator whose only dynamic is this force
// pass) the thread then runs k_semi_implicit's update of its vertex, the same
// operations on the same values, and the separate pass is skipped.
constexpr int kExplicitVertexBlock = 64;
__global__ __launch_bounds__(kExplicitVertexBlock) void k_explicit_vertex(PSView ps, OwnersView own,
                                                                          const float* partial, int count,
                                                                          float damping, int integrate, float dt,
                                                                          float gx, float gy, float gz) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= count) return;
    const int vid = gload(own.slot_vid, slot);
    const int2 nbo = gload(own.slot_nbo, slot);
    const float* __restrict__ src = partial + static_cast<size_t>(nbo.y) * 9;
    float a[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int j = 0;
    for (; j + 4 <= nbo.x; j += 4) {
        float t[36];
#pragma unroll
        for (int k = 0; k < 36; ++k) t[k] = gload(src, j * 9 + k);
#pragma unroll
        for (int k = 0; k < 36; ++k) a[k % 9] += t[k];
    }
    for (; j < nbo.x; ++j)
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] += gload(src, j * 9 + i);
    float4 f = make_float4(a[0], a[1], a[2], 0.f);
    if (damping != 0.f) {
        const float4 v = ps.v[vid];
        f.x -= damping * (a[3] * v.x + a[4] * v.y + a[5] * v.z);
        f.y -= damping * (a[4] * v.x + a[6] * v.y + a[7] * v.z);
        f.z -= damping * (a[5] * v.x + a[7] * v.y + a[8] * v.z);
    }
    if (!integrate) {
        ps.f[vid] = f;
        return;
    }
    float4 p = ps.p[vid];
    float4 v = ps.v[vid];
    const float w = ps.w[vid];
    ps.last_p[vid] = p;
    v.x += (gx + f.x * w) * dt;
    v.y += (gy + f.y * w) * dt;
    v.z += (gz + f.z * w) * dt;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    ps.v[vid] = v;
    ps.p[vid] = p;
    ps.f[vid] = f4(0.f, 0.f, 0.f, 0.f);
}

// ------------------------------------------------------------ batched forms
// One launch covers every vertex of every scene of a batch; each block belongs
// to one scene (VSeg), whose description is read with scalar loads.
__device__ __forceinline__ const SceneDev& vseg_scene(const BatchVArgs& b, int& i, int& ok) {
    const int blk = static_cast<int>(blockIdx.x);
    const VSeg seg = find_seg(b.segs, 0, b.nsegs, blk);
    i = (blk - seg.block_begin) * kBlock + threadIdx.x;
    ok = i < seg.count;
    return b.scenes[seg.scene];
}

__global__ __launch_bounds__(kBlock) void k_batch_predict(BatchVArgs b) {
    int i, ok;
    const SceneDev& sd = vseg_scene(b, i, ok);
    if (!ok) return;
    const PSView ps = sd.ps;
    vbd_predict_one(ps, sd.y, i, b.dt, sd.g[0], sd.g[1], sd.g[2], nullptr);
}

// fixation then box limit, fused: a fixed vertex is snapped and masked, every
// other vertex (mask 1) is clamped - the same result as the two kernels in
// sequence, since the box skips masked vertices.
__global__ __launch_bounds__(kBlock) void k_batch_fix_box(BatchVArgs b) {
    int i, ok;
    const SceneDev& sd = vseg_scene(b, i, ok);
    if (!ok || b.iteration >= sd.iterations) return;
    const PSView ps = sd.ps;
    const unsigned char* fixed = sd.fixed;
    // every load at once (global address space): the flag, the mask, the
    // displacement and the rest position
    const bool is_fixed = fixed != nullptr && gload(fixed, i) != 0;
    const int mask = gload(static_cast<const int*>(ps.mask), i);
    float4 u = gload(static_cast<const float4*>(ps.p), i);
    const float4 X = gload(static_cast<const float4*>(ps.init_p), i);
    if (is_fixed) {
        gstore(ps.p, i, fix_displacement(X, sd.fix_com, sd.fix_offset, sd.fix_R));
        gstore(ps.v, i, f4(0.f, 0.f, 0.f, 0.f));
        gstore(ps.f, i, f4(0.f, 0.f, 0.f, 0.f));
        gstore(ps.mask, i, 0);
        return;
    }
    if (!sd.has_box || mask == 0) return;
    if (ps.comp) {
        box_clamp_comp(ps, i, X, sd.box_min, sd.box_max);
        return;
    }
    if (box_clamp(u, X, sd.box_min, sd.box_max)) gstore(ps.p, i, u);   // written only when clamped
}

__global__ __launch_bounds__(kBlock) void k_batch_velocity(BatchVArgs b) {
    int i, ok;
    const SceneDev& sd = vseg_scene(b, i, ok);
    if (!ok) return;
    const PSView ps = sd.ps;
    vbd_velocity_one(ps, i, b.dt);
}

}  // namespace

void launch_batch_predict(const BatchVArgs& b, int nblocks, hipStream_t s) {
    if (nblocks > 0) hipLaunchKernelGGL(k_batch_predict, dim3(nblocks), dim3(kBlock), 0, s, b);
}
void launch_batch_fix_box(const BatchVArgs& b, int nblocks, hipStream_t s) {
    if (nblocks > 0) hipLaunchKernelGGL(k_batch_fix_box, dim3(nblocks), dim3(kBlock), 0, s, b);
}
void launch_batch_velocity(const BatchVArgs& b, int nblocks, hipStream_t s) {
    if (nblocks > 0) hipLaunchKernelGGL(k_batch_velocity, dim3(nblocks), dim3(kBlock), 0, s, b);
}

namespace {
}  // namespace

void launch_vbd_predict(const PSView& ps, float4* y, float4* prev_it_p, float dt, const float g[3], hipStream_t s) {
    if (ps.n <= 0) return;
    hipLaunchKernelGGL(k_vbd_predict, dim3(grid_for(ps.n)), dim3(kBlock), 0, s, ps, y, prev_it_p, dt, g[0], g[1], g[2]);
}

void launch_vbd_velocity(const PSView& ps, float dt, hipStream_t s) {
    if (ps.n <= 0) return;
    hipLaunchKernelGGL(k_vbd_velocity, dim3(grid_for(ps.n)), dim3(kBlock), 0, s, ps, dt);
}

void launch_chebyshev(const PSView& ps, float4* prev_it_p, float4* prev_it2_p, int it, float omega, hipStream_t s) {
    if (ps.n <= 0) return;
    hipLaunchKernelGGL(k_chebyshev, dim3(grid_for(ps.n)), dim3(kBlock), 0, s, ps, prev_it_p, prev_it2_p, it, omega);
}

void launch_semi_implicit(const PSView& ps, float dt, const float g[3], hipStream_t s) {
    if (ps.n <= 0) return;
    hipLaunchKernelGGL(k_semi_implicit, dim3(grid_for(ps.n)), dim3(kBlock), 0, s, ps, dt, g[0], g[1], g[2]);
}

void launch_pbd_velocity(const PSView& ps, float dt, float global_damping, hipStream_t s) {
    if (ps.n <= 0) return;
    hipLaunchKernelGGL(k_pbd_velocity, dim3(grid_for(ps.n)), dim3(kBlock), 0, s, ps, dt, global_damping);
}

void launch_fix(const PSView& ps, const int* ids, int n, const FixTransform& t, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fix, dim3(grid_for(n)), dim3(kBlock), 0, s, ps, ids, n, t);
}

void launch_fix_box(const PSView& ps, const unsigned char* fixed, const FixTransform& t, const float pmin[3],
                    const float pmax[3], hipStream_t s) {
    if (ps.n <= 0) return;
    hipLaunchKernelGGL(k_fix_box, dim3(grid_for(ps.n)), dim3(kBlock), 0, s, ps, fixed, t,
                       make_float3(pmin[0], pmin[1], pmin[2]), make_float3(pmax[0], pmax[1], pmax[2]));
}

void launch_box(const PSView& ps, const float pmin[3], const float pmax[3], hipStream_t s) {
    if (ps.n <= 0) return;
    hipLaunchKernelGGL(k_box, dim3(grid_for(ps.n)), dim3(kBlock), 0, s, ps, make_float3(pmin[0], pmin[1], pmin[2]),
                       make_float3(pmax[0], pmax[1], pmax[2]));
}

void launch_explicit_vertex(const PSView& ps, const OwnersView& own, const float* partial, int count, float damping,
                            const float* gravity, float dt, hipStream_t s) {
    if (count <= 0) return;
    const float g0 = gravity ? gravity[0] : 0.f, g1 = gravity ? gravity[1] : 0.f, g2 = gravity ? gravity[2] : 0.f;
    hipLaunchKernelGGL(k_explicit_vertex, dim3((count + kExplicitVertexBlock - 1) / kExplicitVertexBlock),
                       dim3(kExplicitVertexBlock), 0, s, ps, own, partial, count, damping, gravity ? 1 : 0, dt, g0,
                       g1, g2);
}

void launch_copy_colour(const PSView& ps, const int* slot_vid, int first, int count, const float4* p_new,
                        hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_copy_colour, dim3(grid_for(count)), dim3(kBlock), 0, s, ps, slot_vid, first, count, p_new);
}

}  // namespace oe
