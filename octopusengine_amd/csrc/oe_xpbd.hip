// XPBD per-colour element projection.  Elements of one colour share no vertex,
// so every element of a colour is projected concurrently and writes its own
// nodes only.  Each element carries two constraints (lambda term, then mu
// term), solved one after the other with the positions updated in between.
//
// Two formulations, as in the reference:
//  * "V0" (src/GPU/PBD/GPU_PBD_FEM.cu:13-93): C = sqrt|E|, grad C = grad E/(2C),
//    dl = -C / (sum w |grad C|^2 + 1/(k dt^2)), fixed vertices are not moved.
//    Used for P1 and when the XPBD version is requested.  One thread per element.
//  * "V2" (src/GPU/PBD/GPU_PBD_FEM.cu:206-263, and V1 :141-204 which has the
//    same arithmetic): beta = -2E / (sum w |grad E|^2 + 4E/(k dt^2)), zero when
//    E < 1e-12, every node moved.  One group of LPE lanes per element: lanes
//    q < Q evaluate the quadrature points, lanes k < S own the nodes; the
//    node positions and per-quadrature stresses are staged in LDS.
// Energies per material: src/GPU/PBD/GPU_PBD_FEM_Materials.cu:31-66,97-118.
#include "oe_kernels.hpp"
#include "oe_xpbd.hpp"

namespace oe {

namespace {

// energy e and dE/dF for constraint m of the XPBD energy split.  The
// NeoHooke split uses alpha = 1 + mu/lambda (PBD_ContinuousMaterial.h:61-65,
// GPU_PBD_FEM_Materials.cu:105-107), unlike the GPU VBD NeoHooke stress.
__device__ __forceinline__ void xpbd_material(int material, int m, float alpha, const float F[9], float P[9],
                                              float& e) {
    float C[9];
    cofactor(F, C);
    const float J = det_from_cof(F, C);
    if (material == Stable_NeoHooke_id) {
        if (m == 0) {
            const float d = J - 1.f;
            e = d * d;
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] = 2.f * d * C[i];
        } else {
            float n2 = 0.f;
#pragma unroll
            for (int i = 0; i < 9; ++i) n2 += F[i] * F[i];
            e = n2 - 3.f - 2.f * (J - 1.f);
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] = 2.f * F[i] - 2.f * C[i];
        }
    } else {  // NeoHooke: (J - alpha)^2 and |F|^2 - 3
        if (m == 0) {
            const float d = J - alpha;
            e = d * d;
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] = 2.f * d * C[i];
        } else {
            float n2 = 0.f;
#pragma unroll
            for (int i = 0; i < 9; ++i) n2 += F[i] * F[i];
            e = n2 - 3.f;
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] = 2.f * F[i];
        }
    }
}

// Ps = P * JX_inv^T * V (column-major): Ps(r,c) = V sum_l P(r,l) JX_inv(c,l)
__device__ __forceinline__ void scale_stress(const float P[9], const float JXi[9], float V, float Ps[9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            Ps[c * 3 + r] = (P[0 * 3 + r] * JXi[0 * 3 + c] + P[1 * 3 + r] * JXi[1 * 3 + c] + P[2 * 3 + r] * JXi[2 * 3 + c]) * V;
}

template <int S, int Q>
__global__ __launch_bounds__(256) void k_xpbd_v0(XpbdArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.count) return;
    const int eid = a.eids[a.first + tid];
    const int* __restrict__ tp = a.fem.topo + static_cast<size_t>(eid) * S;
    for (int m = 0; m < 2; ++m) {
        float xs[S][3];
        float grad[S][3];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const float4 x = a.ps.p[tp[k]];
            xs[k][0] = x.x; xs[k][1] = x.y; xs[k][2] = x.z;
            grad[k][0] = grad[k][1] = grad[k][2] = 0.f;
        }
        float C = 0.f;
        for (int q = 0; q < Q; ++q) {
            const float* dn = a.fem.dN + q * S * 3;
            float Jx[9], JXi[9], F[9], P[9], Ps[9], e, V;
            load_element_record(a.fem, eid, q, JXi, V);
            jx_from_nodes<S>(xs, dn, Jx);
            deformation_gradient(Jx, JXi, F);
            xpbd_material(a.mt.material, m, 1.f + a.mt.mu / a.mt.lambda, F, P, e);
            scale_stress(P, JXi, V, Ps);
            C += e * V;
#pragma unroll
            for (int k = 0; k < S; ++k)
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    grad[k][r] += Ps[r] * dn[3 * k] + Ps[3 + r] * dn[3 * k + 1] + Ps[6 + r] * dn[3 * k + 2];
        }
        C = fabsf(C);
        C = sqrtf(C);
        const float Cinv = C < 1e-12f ? 0.f : 1.f / (2.f * C);
        float sum = 0.f;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            grad[k][0] *= Cinv; grad[k][1] *= Cinv; grad[k][2] *= Cinv;
            sum += (grad[k][0] * grad[k][0] + grad[k][1] * grad[k][1] + grad[k][2] * grad[k][2]) * a.ps.w[tp[k]];
        }
        const float stiff = (m == 0) ? a.mt.lambda : a.mt.mu;
        const float alpha = 1.f / (stiff * a.dt * a.dt);
        const float dl = -C / (sum + alpha);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int vid = tp[k];
            if (a.ps.mask[vid] == 0) continue;
            const float wv = a.ps.w[vid];
            float4 x = a.ps.p[vid];
            x.x += dl * wv * grad[k][0];
            x.y += dl * wv * grad[k][1];
            x.z += dl * wv * grad[k][2];
            a.ps.p[vid] = x;
        }
    }
}

// With LPE = 64 (27-node hexahedra) the two 32-lane halves of the wave split
// the node sum of each point's Jx and the point sum of each node's gradient,
// and add their parts with one cross-half shuffle, halving the chain of
// dependent LDS reads and FMAs a wave runs per projection.
template <int S, int Q, int LPE>
__global__ __launch_bounds__(256) void k_xpbd_v2(XpbdArgs a) {
    constexpr int SPLIT = LPE > 32 ? LPE / 32 : 1;
    constexpr int W = LPE / SPLIT;   // lanes per part
    static_assert(S <= W && Q <= W, "element does not fit the lane group");
    constexpr int EPB = 256 / LPE;
    constexpr int KH = (S + SPLIT - 1) / SPLIT;   // nodes per part (point stage)
    constexpr int QH = (Q + SPLIT - 1) / SPLIT;   // points per part (node stage)
    __shared__ float sx[EPB][S][3];
    __shared__ float sP[EPB][Q][9];
    // Large shape-derivative tables (27-node hexahedra, 8.7 KB) are staged in
    // LDS; smaller ones are read from the cache (measured: the copy costs the
    // 20-node tetrahedra 0.3 us per colour and saves the 27-node hexahedra 2 us).
    constexpr bool kDnLds = Q * S * 3 > 1024;
    __shared__ float dn_s[kDnLds ? Q * S * 3 : 1];
    if constexpr (kDnLds)
        for (int i = threadIdx.x; i < Q * S * 3; i += 256) dn_s[i] = gload(a.fem.dN, i);
    const float* __restrict__ dnt = kDnLds ? static_cast<const float*>(dn_s) : a.fem.dN;
    const int le = threadIdx.x / LPE;
    const int lane = threadIdx.x % LPE;
    const int sub = lane % W, part = lane / W;
    const int ei = blockIdx.x * EPB + le;
    const bool valid = ei < a.count;
    const int eid = valid ? a.eids[a.first + ei] : 0;
    const int* __restrict__ tp = a.fem.topo + static_cast<size_t>(eid) * S;
    int vid = 0;
    float wv = 0.f;
    if (valid && lane < S) {
        vid = tp[lane];
        const float4 x = a.ps.p[vid];
        sx[le][lane][0] = x.x; sx[le][lane][1] = x.y; sx[le][lane][2] = x.z;
        wv = a.ps.w[vid];
    }
    float JXi[9], V;
    load_element_record(a.fem, eid, sub < Q ? sub : Q - 1, JXi, V);
    __syncthreads();
    // constant bounds without a split, so that those loops unroll fully
    const int k_begin = SPLIT == 1 ? 0 : part * KH;
    const int k_end = SPLIT == 1 ? S : ((part + 1) * KH < S ? (part + 1) * KH : S);
    const int q_begin = SPLIT == 1 ? 0 : part * QH;
    const int q_end = SPLIT == 1 ? Q : ((part + 1) * QH < Q ? (part + 1) * QH : Q);
    for (int m = 0; m < 2; ++m) {
        float e_loc = 0.f;
        {
            const int q = sub < Q ? sub : Q - 1;
            const float* dn = dnt + q * S * 3;
            float Jx[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) Jx[i] = 0.f;
#pragma unroll
            for (int k = k_begin; k < k_end; ++k) {
                const float d0 = dn[3 * k], d1 = dn[3 * k + 1], d2 = dn[3 * k + 2];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float xr = sx[le][k][r];
                    Jx[r] += xr * d0; Jx[3 + r] += xr * d1; Jx[6 + r] += xr * d2;
                }
            }
            if constexpr (SPLIT == 2) {
#pragma unroll
                for (int i = 0; i < 9; ++i) Jx[i] += __shfl_xor(Jx[i], 32);
            }
            if (valid && part == 0 && sub < Q) {
                float F[9], P[9], Ps[9], e;
                deformation_gradient(Jx, JXi, F);
                xpbd_material(a.mt.material, m, 1.f + a.mt.mu / a.mt.lambda, F, P, e);
                scale_stress(P, JXi, V, Ps);
                e_loc = e * V;
#pragma unroll
                for (int i = 0; i < 9; ++i) sP[le][q][i] = Ps[i];
            }
        }
        const float C = group_sum<LPE>(e_loc);
        __syncthreads();
        float g0 = 0.f, g1 = 0.f, g2 = 0.f, sn = 0.f;
        {
            const int k = sub < S ? sub : S - 1;
#pragma unroll
            for (int q = q_begin; q < q_end; ++q) {
                const float* dn = dnt + (q * S + k) * 3;
                const float* Ps = sP[le][q];
                g0 += Ps[0] * dn[0] + Ps[3] * dn[1] + Ps[6] * dn[2];
                g1 += Ps[1] * dn[0] + Ps[4] * dn[1] + Ps[7] * dn[2];
                g2 += Ps[2] * dn[0] + Ps[5] * dn[1] + Ps[8] * dn[2];
            }
            if constexpr (SPLIT == 2) {
                g0 += __shfl_xor(g0, 32);
                g1 += __shfl_xor(g1, 32);
                g2 += __shfl_xor(g2, 32);
            }
            if (valid && lane < S) sn = (g0 * g0 + g1 * g1 + g2 * g2) * wv;
        }
        const float ssum = group_sum<LPE>(sn);
        const float stiff = (m == 0) ? a.mt.lambda : a.mt.mu;
        const float alpha = 4.f * C / (stiff * a.dt * a.dt);
        const float beta = C < 1e-12f ? 0.f : -2.f * C / (ssum + alpha);
        if (valid && lane < S) {
            const float nx = sx[le][lane][0] + beta * wv * g0;
            const float ny = sx[le][lane][1] + beta * wv * g1;
            const float nz = sx[le][lane][2] + beta * wv * g2;
            sx[le][lane][0] = nx; sx[le][lane][1] = ny; sx[le][lane][2] = nz;
            const float4 old = a.ps.p[vid];
            a.ps.p[vid] = make_float4(nx, ny, nz, old.w);
        }
        __syncthreads();
    }
}

template <int S, int Q>
void launch_v0(const XpbdArgs& a, hipStream_t s) {
    const int grid = (a.count + 255) / 256;
    hipLaunchKernelGGL((k_xpbd_v0<S, Q>), dim3(grid), dim3(256), 0, s, a);
}

template <int S, int Q, int LPE>
void launch_v2(const XpbdArgs& a, hipStream_t s) {
    constexpr int EPB = 256 / LPE;
    const int grid = (a.count + EPB - 1) / EPB;
    hipLaunchKernelGGL((k_xpbd_v2<S, Q, LPE>), dim3(grid), dim3(256), 0, s, a);
}

}  // namespace

bool launch_xpbd(int S, int Q, bool v0_form, const XpbdArgs& a, hipStream_t s) {
    if (a.count <= 0) return true;
    if (v0_form) {
        if (S == 4 && Q == 1) { launch_v0<4, 1>(a, s); return true; }
        if (S == 8 && Q == 8) { launch_v0<8, 8>(a, s); return true; }
        if (S == 10 && Q == 4) { launch_v0<10, 4>(a, s); return true; }
        if (S == 20 && Q == 11) { launch_v0<20, 11>(a, s); return true; }
        if (S == 27 && Q == 27) { launch_v0<27, 27>(a, s); return true; }
        return false;
    }
    if (S == 4 && Q == 1) { launch_v2<4, 1, 4>(a, s); return true; }
    if (S == 8 && Q == 8) { launch_v2<8, 8, 8>(a, s); return true; }
    if (S == 10 && Q == 4) { launch_v2<10, 4, 16>(a, s); return true; }
    if (S == 20 && Q == 11) { launch_v2<20, 11, 32>(a, s); return true; }
    if (S == 27 && Q == 27) { launch_v2<27, 27, 64>(a, s); return true; }
    return false;
}

}  // namespace oe
