// Per-element diagnostics of the current state (reference GPU_FEM::get_stress,
// get_volume, get_volume_diff, get_inverted, src/GPU/GPU_FEM.cu:20-152), one
// thread per element, all four quantities in one pass:
//   stress      sum_q vonMises(P F^T / det F) V_q     (P of the GPU material)
//   volume      sum_q |det F_q| V_q                   (= sum_q |det Jx_q| w_q)
//   volume_diff sum_q (|det F_q| V_q - V_q)
//   inverted    1 if sum_q det(F_q) w_q <= -1e-6 else 0
// Three defects of the reference are not reproduced: its stress kernel keeps
// accumulating Jx over the quadrature points (:32-40), its inverted kernel
// reads JX_inv[eid + qid] instead of JX_inv[qid + i] (:108), and
// get_volume_diff launches the inverted kernel (:135-143).
#include "oe_diag.hpp"
#include "oe_kernels.hpp"

namespace oe {

namespace {

__device__ __forceinline__ float det3(const float A[9]) {
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) + A[6] * (A[1] * A[5] - A[4] * A[2]);
}

template <int S, int Q>
__global__ __launch_bounds__(256) void k_element_diagnostics(PSView ps, FemView f, MaterialView mt, float* stress,
                                                             float* volume, float* volume_diff, float* inverted) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= f.E) return;
    const int* __restrict__ tp = f.topo + static_cast<size_t>(e) * S;
    float xs[S][3];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const float4 x = ps.p[tp[k]];
        xs[k][0] = x.x; xs[k][1] = x.y; xs[k][2] = x.z;
    }
    float s = 0.f, vol = 0.f, vd = 0.f, sv = 0.f;
    for (int q = 0; q < Q; ++q) {
        float Ju[9], JXi[9], V, F[9], C[9];
        load_element_record(f, e, q, JXi, V);
        jx_from_nodes<S>(xs, f.dN + q * S * 3, Ju);
        deformation_gradient(Ju, JXi, F);
        cofactor(F, C);
        const float J = det_from_cof(F, C);
        // P = mu F + lambda (J - alpha) cof F ; sigma = P F^T / J (column-major)
        const float c = mt.lambda * (J - mt.alpha);
        float P[9], sg[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = mt.mu * F[i] + c * C[i];
#pragma unroll
        for (int col = 0; col < 3; ++col)
#pragma unroll
            for (int row = 0; row < 3; ++row)
                sg[col * 3 + row] =
                    (P[row] * F[col] + P[3 + row] * F[3 + col] + P[6 + row] * F[6 + col]) / J;
        const float d01 = sg[0] - sg[4], d12 = sg[4] - sg[8], d02 = sg[0] - sg[8];
        const float vm2 = 0.5f * (d01 * d01 + d12 * d12 + d02 * d02) +
                          3.f * (sg[1] * sg[1] + sg[5] * sg[5] + sg[2] * sg[2]);
        s += sqrtf(vm2) * V;
        vol += fabsf(J) * V;
        vd += fabsf(J) * V - V;
        // quadrature weight w_q = V_q |det JX_inv_q|
        sv += J * V * fabsf(det3(JXi));
    }
    if (stress) stress[e] = s;
    if (volume) volume[e] = vol;
    if (volume_diff) volume_diff[e] = vd;
    if (inverted) inverted[e] = sv > -1e-6f ? 0.f : 1.f;
}

template <int S, int Q>
void launch_diag(const PSView& ps, const FemView& f, const MaterialView& mt, float* st, float* vo, float* vd, float* in,
                 hipStream_t s) {
    if (f.E <= 0) return;
    hipLaunchKernelGGL((k_element_diagnostics<S, Q>), dim3((f.E + 255) / 256), dim3(256), 0, s, ps, f, mt, st, vo, vd,
                       in);
}

}  // namespace

bool launch_element_diagnostics(const PSView& ps, const FemView& f, const MaterialView& mt, float* stress,
                                float* volume, float* volume_diff, float* inverted, hipStream_t s) {
    if (f.S == 4 && f.Q == 1) launch_diag<4, 1>(ps, f, mt, stress, volume, volume_diff, inverted, s);
    else if (f.S == 8 && f.Q == 8) launch_diag<8, 8>(ps, f, mt, stress, volume, volume_diff, inverted, s);
    else if (f.S == 10 && f.Q == 4) launch_diag<10, 4>(ps, f, mt, stress, volume, volume_diff, inverted, s);
    else if (f.S == 20 && f.Q == 11) launch_diag<20, 11>(ps, f, mt, stress, volume, volume_diff, inverted, s);
    else if (f.S == 27 && f.Q == 27) launch_diag<27, 27>(ps, f, mt, stress, volume, volume_diff, inverted, s);
    else return false;
    return true;
}

}  // namespace oe
