// Tensor structure of the 27-node hexahedron (triquadratic Lagrange element
// on [-1, 1]^3 with the 3 x 3 x 3 Gauss rule).  Node k sits at reference
// coordinates (A_k - 1, B_k - 1, C_k - 1) and its shape function factors as
//   N_k(x, y, z) = l_{A_k}(x) l_{B_k}(y) l_{C_k}(z),
// with l_0, l_1, l_2 the quadratic Lagrange polynomials on {-1, 0, 1}.  The
// quadrature points follow the same pattern (point q at g * (A_q - 1, ...),
// g = sqrt(3/5) rounded to float as in the reference's table), so the point
// with tensor indices (i, j, k) has the index of the node at (i, j, k).
// The node order is the engine's (and the reference's) Hexa27 order:
// corners, edge mid-points, face centres, centre.  The host checks these
// tables against the element's shape-function table before using them.
#pragma once

namespace oe {
namespace q2 {

// tensor indices (a, b, c) of node k, packed as a * 9 + b * 3 + c
constexpr int kNodeABC[27] = {
    0 * 9 + 0 * 3 + 0, 2 * 9 + 0 * 3 + 0, 2 * 9 + 2 * 3 + 0, 0 * 9 + 2 * 3 + 0,
    0 * 9 + 0 * 3 + 2, 2 * 9 + 0 * 3 + 2, 2 * 9 + 2 * 3 + 2, 0 * 9 + 2 * 3 + 2,
    1 * 9 + 0 * 3 + 0, 2 * 9 + 1 * 3 + 0, 1 * 9 + 2 * 3 + 0, 0 * 9 + 1 * 3 + 0,
    0 * 9 + 0 * 3 + 1, 2 * 9 + 0 * 3 + 1, 2 * 9 + 2 * 3 + 1, 0 * 9 + 2 * 3 + 1,
    1 * 9 + 0 * 3 + 2, 2 * 9 + 1 * 3 + 2, 1 * 9 + 2 * 3 + 2, 0 * 9 + 1 * 3 + 2,
    1 * 9 + 1 * 3 + 0, 1 * 9 + 0 * 3 + 1, 2 * 9 + 1 * 3 + 1, 1 * 9 + 2 * 3 + 1,
    0 * 9 + 1 * 3 + 1, 1 * 9 + 1 * 3 + 2, 1 * 9 + 1 * 3 + 1};

// node (= quadrature point) index of tensor position (a, b, c) = t / 9, t / 3 % 3, t % 3
constexpr int tensor_to_node(int t) {
    for (int k = 0; k < 27; ++k)
        if (kNodeABC[k] == t) return k;
    return -1;
}

// 1D factors at the three Gauss coordinates: L[a][i] = l_a(x_i), D[a][i] = l_a'(x_i)
struct Basis1D {
    float L[3][3];
    float D[3][3];
};

}  // namespace q2
}  // namespace oe
