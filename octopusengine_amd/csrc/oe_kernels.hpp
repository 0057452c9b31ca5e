// Device code shared by the VBD, XPBD and explicit paths: the per-quadrature
// kinematics (Jx, F, dF/dx), the hyperelastic stress and the per-vertex
// Hessian block, cross-lane sums, and the FP64 3x3 solve.
//
// Conventions: 3x3 matrices are float[9] in column-major order (m[c*3 + r]),
// as the reference's glm matrices are, so JX_inv records built on the host can
// be read directly.
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"

namespace oe {

// Block order for the 8 XCDs of the chip: the dispatcher hands work-group i to
// XCD i % 8, so consecutive block ids land in different L2 caches.  This maps
// the hardware block id to a logical one such that each XCD walks one
// contiguous range of logical blocks (a bijection on [0, n)); neighbouring
// vertices of a scene, which share elements and nodes, then share an L2.
__device__ __forceinline__ int xcd_block(int bid, int n) {
    constexpr int X = 8;
    const int q = n / X, r = n - (n / X) * X;
    const int xcd = bid % X, idx = bid / X;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Segment of a batched launch that holds launch-local block b: the launch's
// segments segs[first .. first + n) have ascending block_begin (the first is
// 0); binary search, uniform over the block.
template <typename Seg>
__device__ __forceinline__ const Seg& find_seg(const Seg* __restrict__ segs, int first, int n, int b) {
    int lo = first, hi = first + n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].block_begin <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return segs[lo];
}

// Per-vertex accumulator: force (3) + lower triangle of the symmetric 3x3
// Hessian block (H00, H10, H20, H11, H21, H22).
struct Acc9 {
    float v[9];
};

__device__ __forceinline__ void acc_zero(Acc9& a) {
#pragma unroll
    for (int i = 0; i < 9; ++i) a.v[i] = 0.f;
}

// Sum over groups of L consecutive lanes (L a power of two, L <= 64).
template <int L>
__device__ __forceinline__ float group_sum(float x) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, L);
    return x;
}

template <int L>
__device__ __forceinline__ void group_sum(Acc9& a) {
#pragma unroll
    for (int i = 0; i < 9; ++i) a.v[i] = group_sum<L>(a.v[i]);
}

// Jx = sum_k x_k (x) dN_k  (column-major: Jx[c*3+r] = sum_k x_k[r] dN_k[c]);
// dn points at the element's shape-derivative table in device memory
template <int S>
__device__ __forceinline__ void jx_from_nodes(const float (&xs)[S][3], const float* __restrict__ dn, float Jx[9]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) Jx[i] = 0.f;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const float d0 = gload(dn, 3 * k), d1 = gload(dn, 3 * k + 1), d2 = gload(dn, 3 * k + 2);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            Jx[0 + r] = fmaf(xs[k][r], d0, Jx[0 + r]);
            Jx[3 + r] = fmaf(xs[k][r], d1, Jx[3 + r]);
            Jx[6 + r] = fmaf(xs[k][r], d2, Jx[6 + r]);
        }
    }
}

// F = Jx * JX_inv (both column-major)
__device__ __forceinline__ void mat_mul(const float A[9], const float B[9], float C[9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            C[c * 3 + r] = A[0 * 3 + r] * B[c * 3 + 0] + A[1 * 3 + r] * B[c * 3 + 1] + A[2 * 3 + r] * B[c * 3 + 2];
}

// Deformation gradient from nodal displacements u = x - X:
//   F = (sum_k x_k (x) dN_k) JX_inv = I + (sum_k u_k (x) dN_k) JX_inv,
// since sum_k X_k (x) dN_k = J_X.  Working on displacements keeps F accurate
// to the displacement's precision instead of the absolute position's.
__device__ __forceinline__ void deformation_gradient(const float Ju[9], const float JXi[9], float F[9]) {
    mat_mul(Ju, JXi, F);
    F[0] += 1.f;
    F[4] += 1.f;
    F[8] += 1.f;
}

// cofactor matrix: column c = cross of the two other columns (cyclic)
__device__ __forceinline__ void cofactor(const float F[9], float C[9]) {
    const float* a = F;       // column 0
    const float* b = F + 3;   // column 1
    const float* c = F + 6;   // column 2
    C[0] = b[1] * c[2] - b[2] * c[1];
    C[1] = b[2] * c[0] - b[0] * c[2];
    C[2] = b[0] * c[1] - b[1] * c[0];
    C[3] = c[1] * a[2] - c[2] * a[1];
    C[4] = c[2] * a[0] - c[0] * a[2];
    C[5] = c[0] * a[1] - c[1] * a[0];
    C[6] = a[1] * b[2] - a[2] * b[1];
    C[7] = a[2] * b[0] - a[0] * b[2];
    C[8] = a[0] * b[1] - a[1] * b[0];
}

// det F = column0 . cof column0.  Elements close to inversion (J < 0.1) get
// the determinant re-evaluated in double precision.
__device__ __forceinline__ float det_from_cof(const float F[9], const float C[9]) {
    float J = F[0] * C[0] + F[1] * C[1] + F[2] * C[2];
    if (J < 0.1f) {
        const double a0 = F[0], a1 = F[1], a2 = F[2], b0 = F[3], b1 = F[4], b2 = F[5], c0 = F[6], c1 = F[7], c2 = F[8];
        const double d = a0 * (b1 * c2 - b2 * c1) + a1 * (b2 * c0 - b0 * c2) + a2 * (b0 * c1 - b1 * c0);
        J = static_cast<float>(d);
    }
    return J;
}

// One (element, quadrature point) contribution (Jx built from displacements) to the force on node r and to
// its 3x3 Hessian block, for the Neo-Hookean family used by the reference's
// GPU VBD / explicit paths (src/GPU/GPU_FEM_Material.cu:25-36,167-207,250-261):
//   P = mu F + lambda (J - alpha) cof F
//   f_r -= V P g,   g = JX_inv^T dN_r
//   H_r += V (mu |g|^2 I + lambda c c^T),   c = cof F g
// The reference assembles H_r from six 3x3 blocks of d2W/dF2; the
// lambda (J - alpha) d2J/dF2 blocks are antisymmetric and cancel in the
// symmetric assembly, which leaves the closed form above.
__device__ __forceinline__ void nh_node(const float F[9], const float C[9], float s, float V, const float g[3],
                                        const MaterialView& mt, Acc9& a);
__device__ __forceinline__ void nh_contrib(const float Jx[9], const float JXi[9], float V, const float* __restrict__ dnr,
                                           const MaterialView& mt, Acc9& a) {
    float F[9];
    deformation_gradient(Jx, JXi, F);
    // g_c = sum_l JX_inv(l, c) dN_r[l]
    float g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c] = JXi[c * 3 + 0] * dnr[0] + JXi[c * 3 + 1] * dnr[1] + JXi[c * 3 + 2] * dnr[2];
    float C[9];
    cofactor(F, C);
    const float J = det_from_cof(F, C);
    const float s = mt.lambda * (J - mt.alpha);
    nh_node(F, C, s, V, g, mt, a);
}

// The part of nh_contrib that depends on node r (through g = JX_inv^T dN_r),
// given the point's F, cof F and s = lambda (J - alpha): element-centric
// callers form those once per point and add every node's terms.
__device__ __forceinline__ void nh_node(const float F[9], const float C[9], float s, float V, const float g[3],
                                        const MaterialView& mt, Acc9& a) {
    // P g = mu F g + s C g ; c = C g
    float cg[3], Fg[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        cg[r] = C[r] * g[0] + C[3 + r] * g[1] + C[6 + r] * g[2];
        Fg[r] = F[r] * g[0] + F[3 + r] * g[1] + F[6 + r] * g[2];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) a.v[r] -= V * (mt.mu * Fg[r] + s * cg[r]);
    const float gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    const float d = V * mt.mu * gg;
    const float l = V * mt.lambda;
    a.v[3] += d + l * cg[0] * cg[0];
    a.v[4] += l * cg[1] * cg[0];
    a.v[5] += l * cg[2] * cg[0];
    a.v[6] += d + l * cg[1] * cg[1];
    a.v[7] += l * cg[2] * cg[1];
    a.v[8] += d + l * cg[2] * cg[2];
}

// Residual force of the local VBD problem at the current position
// (reference kernel_vbd_compute_residual, src/GPU/VBD/GPU_VBD_FEM.cu:482-542):
//   r = f - (k/dt) H (p - last_p) - (m/dt^2)(p - y)
__device__ __forceinline__ float4 vbd_residual(const Acc9& a, const dvec3& p, const dvec3& lp, const dvec3& y, float m,
                                               float damping, float dt) {
    double r[3] = {a.v[0], a.v[1], a.v[2]};
    if (damping != 0.f) {
        const double k = static_cast<double>(damping) / dt;
        const double u0 = p.x - lp.x, u1 = p.y - lp.y, u2 = p.z - lp.z;
        r[0] -= k * (a.v[3] * u0 + a.v[4] * u1 + a.v[5] * u2);
        r[1] -= k * (a.v[4] * u0 + a.v[6] * u1 + a.v[7] * u2);
        r[2] -= k * (a.v[5] * u0 + a.v[7] * u1 + a.v[8] * u2);
    }
    const double mh2 = static_cast<double>(m) / (static_cast<double>(dt) * dt);
    r[0] -= mh2 * (p.x - y.x);
    r[1] -= mh2 * (p.y - y.y);
    r[2] -= mh2 * (p.z - y.z);
    return make_float4(static_cast<float>(r[0]), static_cast<float>(r[1]), static_cast<float>(r[2]), 0.f);
}

// Newton step of the local VBD problem in double precision
// (reference compute_correction, src/GPU/VBD/GPU_VBD_FEM.cu:107-122):
//   f -= (k/dt) H (p - last_p);  H += (k/dt) H
//   f -= (m/dt^2)(p - y);        H += (m/dt^2) I
//   dx = |det H| > 1e-6 ? H^-1 f : 0
__device__ __forceinline__ void vbd_local_solve(const Acc9& a, const dvec3& p, const dvec3& lp, const dvec3& y, float m,
                                                float damping, float dt, double dx[3]) {
    double f0 = a.v[0], f1 = a.v[1], f2 = a.v[2];
    double h00 = a.v[3], h10 = a.v[4], h20 = a.v[5], h11 = a.v[6], h21 = a.v[7], h22 = a.v[8];
    const double ddt = dt;
    if (damping > 0.f) {
        const double k = static_cast<double>(damping) / ddt;
        const double u0 = p.x - lp.x, u1 = p.y - lp.y, u2 = p.z - lp.z;
        f0 -= k * (h00 * u0 + h10 * u1 + h20 * u2);
        f1 -= k * (h10 * u0 + h11 * u1 + h21 * u2);
        f2 -= k * (h20 * u0 + h21 * u1 + h22 * u2);
        const double s = 1.0 + k;
        h00 *= s; h10 *= s; h20 *= s; h11 *= s; h21 *= s; h22 *= s;
    }
    const double mh2 = static_cast<double>(m) / (ddt * ddt);
    f0 -= mh2 * (p.x - y.x);
    f1 -= mh2 * (p.y - y.y);
    f2 -= mh2 * (p.z - y.z);
    h00 += mh2; h11 += mh2; h22 += mh2;
    const double c00 = h11 * h22 - h21 * h21;
    const double c01 = h21 * h20 - h10 * h22;
    const double c02 = h10 * h21 - h11 * h20;
    const double det = h00 * c00 + h10 * c01 + h20 * c02;
    if (fabs(det) > 1e-6) {
        const double c11 = h00 * h22 - h20 * h20;
        const double c12 = h20 * h10 - h00 * h21;
        const double c22 = h00 * h11 - h10 * h10;
        const double inv = 1.0 / det;
        dx[0] = (c00 * f0 + c01 * f1 + c02 * f2) * inv;
        dx[1] = (c01 * f0 + c11 * f1 + c12 * f2) * inv;
        dx[2] = (c02 * f0 + c12 * f1 + c22 * f2) * inv;
    } else {
        dx[0] = dx[1] = dx[2] = 0.0;
    }
}

// Sum of a slot's partials (part[j], j = its owners index - first_own, in
// the owners order) and the local solve; writes the new position.
__device__ __forceinline__ void solve_slot_from_lds(const PSView& ps, const SlotState& st, const float (*part)[9],
                                                    int first_own, float damping, float dt) {
    Acc9 acc;
    acc_zero(acc);
    for (int j = st.first_inc - first_own, e = j + st.n_inc; j < e; ++j) {
#pragma unroll
        for (int i = 0; i < 9; ++i) acc.v[i] += part[j][i];
    }
    double dx[3];
    const dvec3 p = dunpack(st.p, ps.lo_unit);
    vbd_local_solve(acc, p, dunpack(st.last_p, ps.lo_unit), dunpack(st.y, ps.lo_unit), st.m, damping, dt, dx);
    gstore(ps.p, st.vid, dpack(dvec3{p.x + dx[0], p.y + dx[1], p.z + dx[2]}, ps.lo_inv));
}

// Phase 2 of the fused task kernels for one slot: add the slot's partials
// (part[j], j = its owners index - first_own, in the owners order) and solve.
// The same with the slot's vertex id and (incidence count, first owners
// index) already loaded.
__device__ __forceinline__ void solve_slot_at(const PSView& ps, int vid, int2 nbo, const float4* y, float damping,
                                              float dt, const float (*part)[9], int first_own) {
    if (gload(ps.mask, vid) == 0) return;
    Acc9 acc;
    acc_zero(acc);
    for (int j = nbo.y - first_own, e = j + nbo.x; j < e; ++j) {
#pragma unroll
        for (int i = 0; i < 9; ++i) acc.v[i] += part[j][i];
    }
    const float un = ps.lo_unit;
    const dvec3 p = dunpack(gload(static_cast<const float4*>(ps.p), vid), un);
    double dx[3];
    vbd_local_solve(acc, p, dunpack(gload(static_cast<const float4*>(ps.last_p), vid), un), dunpack(gload(y, vid), un),
                    gload(ps.m, vid), damping, dt, dx);
    gstore(ps.p, vid, dpack(dvec3{p.x + dx[0], p.y + dx[1], p.z + dx[2]}, ps.lo_inv));
}

__device__ __forceinline__ void solve_slot_direct(const PSView& ps, const OwnersView& own, const float4* y,
                                                  float damping, float dt, int slot, const float (*part)[9],
                                                  int first_own) {
    solve_slot_at(ps, gload(own.slot_vid, slot), gload(own.slot_nbo, slot), y, damping, dt, part, first_own);
}

}  // namespace oe
