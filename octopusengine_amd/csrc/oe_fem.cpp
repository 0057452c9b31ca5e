// Per-(element, quadrature) FEM constants, lumped masses and the colour-major
// owners CSR.
//
// JX_inv and V follow src/GPU/GPU_FEM.cu:176-214 (J = sum_k X_k (x) dN_k,
// V = |det J| w, JX_inv = J^-1) but are evaluated in double precision and
// rounded once to float.  Masses follow src/Dynamic/FEM/FEM_Shape.cpp:86-121.
// The owners CSR follows src/GPU/VBD/GPU_VBD_FEM.cu:594-690: vertices grouped
// by colour (ascending colour, ascending id inside a colour), each vertex
// listing its incident elements in ascending element id with the local node
// index.
#include "oe_host.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace oe {

FemConstants build_fem_constants(const HostMesh& m) {
    const ShapeTable& st = shape_table(m.element);
    const int S = st.S, Q = st.Q, E = m.nb_elements();
    FemConstants fc;
    fc.S = S; fc.Q = Q; fc.E = E;
    fc.jxv.resize(static_cast<size_t>(E) * Q * 10);
    fc.dN.resize(static_cast<size_t>(Q) * S * 3);
    fc.elem_volume.assign(E, 0.0);
    for (size_t i = 0; i < fc.dN.size(); ++i) fc.dN[i] = static_cast<float>(st.dN[i]);
    for (int e = 0; e < E; ++e) {
        const int32_t* t = &m.topo[static_cast<size_t>(e) * S];
        for (int q = 0; q < Q; ++q) {
            double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int k = 0; k < S; ++k) {
                const double* d = &st.dN[(q * S + k) * 3];
                const float* X = &m.xyz[3 * static_cast<size_t>(t[k])];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) J[i][j] += static_cast<double>(X[i]) * d[j];
            }
            const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
            const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
            if (det == 0.0) throw std::invalid_argument("degenerate element (det J = 0)");
            // inverse = adj(J) / det, adj(J)[i][j] = cofactor(J)[j][i]
            double inv[3][3];
            inv[0][0] = c00 / det;
            inv[1][0] = c01 / det;
            inv[2][0] = c02 / det;
            inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
            float* rec = &fc.jxv[(static_cast<size_t>(e) * Q + q) * 10];
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) rec[j * 3 + i] = static_cast<float>(inv[i][j]);  // column-major
            const double V = std::fabs(det) * st.weights[q];
            rec[9] = static_cast<float>(V);
            fc.elem_volume[e] += V;
        }
    }
    return fc;
}

bool build_affine_records(const HostMesh& m, std::vector<float>& rec, std::vector<float>& wq) {
    const ShapeTable& st = shape_table(m.element);
    const int S = st.S, Q = st.Q, E = m.nb_elements();
    // Least-squares affine fit X ~ a + B xi over the element's nodes: the
    // normal matrix M = sum_k phi_k phi_k^T, phi = (1, xi), is the same for
    // every element; invert it once (Gauss-Jordan, partial pivoting).
    double M[4][8] = {};
    for (int k = 0; k < S; ++k) {
        const double phi[4] = {1.0, st.nodes[3 * k], st.nodes[3 * k + 1], st.nodes[3 * k + 2]};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) M[i][j] += phi[i] * phi[j];
    }
    for (int i = 0; i < 4; ++i) M[i][4 + i] = 1.0;
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
        for (int j = 0; j < 8; ++j) std::swap(M[c][j], M[piv][j]);
        const double d = M[c][c];
        for (int j = 0; j < 8; ++j) M[c][j] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = M[r][c];
            for (int j = 0; j < 8; ++j) M[r][j] -= f * M[c][j];
        }
    }
    std::vector<float> out(static_cast<size_t>(E) * 10);
    for (int e = 0; e < E; ++e) {
        const int32_t* t = &m.topo[static_cast<size_t>(e) * S];
        double rhs[3][4] = {};
        double xmax = 0.0;
        for (int k = 0; k < S; ++k) {
            const double phi[4] = {1.0, st.nodes[3 * k], st.nodes[3 * k + 1], st.nodes[3 * k + 2]};
            const float* X = &m.xyz[3 * static_cast<size_t>(t[k])];
            for (int i = 0; i < 3; ++i) {
                xmax = std::max(xmax, std::fabs(static_cast<double>(X[i])));
                for (int j = 0; j < 4; ++j) rhs[i][j] += phi[j] * X[i];
            }
        }
        double c[3][4];   // c[i] = M^-1 rhs[i]: (a_i, B_i0, B_i1, B_i2)
        for (int i = 0; i < 3; ++i)
            for (int r = 0; r < 4; ++r) {
                c[i][r] = 0.0;
                for (int j = 0; j < 4; ++j) c[i][r] += M[r][4 + j] * rhs[i][j];
            }
        const double tol = 8.0 * std::ldexp(std::max(xmax, 1e-30), -24);
        for (int k = 0; k < S; ++k) {
            const float* X = &m.xyz[3 * static_cast<size_t>(t[k])];
            for (int i = 0; i < 3; ++i) {
                const double fit = c[i][0] + c[i][1] * st.nodes[3 * k] + c[i][2] * st.nodes[3 * k + 1] +
                                   c[i][3] * st.nodes[3 * k + 2];
                if (std::fabs(fit - X[i]) > tol) return false;
            }
        }
        double B[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) B[i][j] = c[i][1 + j];
        const double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
        const double c01 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
        const double c02 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
        const double det = B[0][0] * c00 + B[0][1] * c01 + B[0][2] * c02;
        if (det == 0.0) throw std::invalid_argument("degenerate element (det J = 0)");
        double inv[3][3];
        inv[0][0] = c00 / det;
        inv[1][0] = c01 / det;
        inv[2][0] = c02 / det;
        inv[0][1] = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) / det;
        inv[1][1] = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) / det;
        inv[2][1] = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) / det;
        inv[0][2] = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) / det;
        inv[1][2] = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) / det;
        inv[2][2] = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) / det;
        float* r = &out[static_cast<size_t>(e) * 10];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) r[j * 3 + i] = static_cast<float>(inv[i][j]);   // column-major
        r[9] = static_cast<float>(std::fabs(det));
    }
    rec = std::move(out);
    wq.resize(Q);
    for (int q = 0; q < Q; ++q) wq[q] = static_cast<float>(st.weights[q]);
    return true;
}

bool q2_basis_1d(q2::Basis1D& out) {
    const ShapeTable& st = shape_table(Hexa27);
    if (st.S != 27 || st.Q != 27) return false;
    for (int k = 0; k < 27; ++k) {
        const int t = q2::kNodeABC[k];
        const int abc[3] = {t / 9, t / 3 % 3, t % 3};
        for (int d = 0; d < 3; ++d)
            if (st.nodes[3 * k + d] != static_cast<double>(abc[d] - 1)) return false;
    }
    // Gauss coordinates along x from the points (i, 1, 1); the same along y, z
    double x[3];
    for (int i = 0; i < 3; ++i) x[i] = st.qpoints[3 * q2::tensor_to_node(i * 9 + 4)];
    double L[3][3], D[3][3];
    for (int i = 0; i < 3; ++i) {
        const double v = x[i];
        L[0][i] = 0.5 * v * (v - 1.0);
        L[1][i] = 1.0 - v * v;
        L[2][i] = 0.5 * v * (v + 1.0);
        D[0][i] = v - 0.5;
        D[1][i] = -2.0 * v;
        D[2][i] = v + 0.5;
    }
    for (int q = 0; q < 27; ++q) {
        const int tq = q2::kNodeABC[q];
        const int ijk[3] = {tq / 9, tq / 3 % 3, tq % 3};
        for (int d = 0; d < 3; ++d)
            if (st.qpoints[3 * q + d] != x[ijk[d]]) return false;
        for (int k = 0; k < 27; ++k) {
            const int t = q2::kNodeABC[k];
            const int a = t / 9, b = t / 3 % 3, c = t % 3;
            const double g[3] = {D[a][ijk[0]] * L[b][ijk[1]] * L[c][ijk[2]], L[a][ijk[0]] * D[b][ijk[1]] * L[c][ijk[2]],
                                 L[a][ijk[0]] * L[b][ijk[1]] * D[c][ijk[2]]};
            for (int d = 0; d < 3; ++d)
                if (std::fabs(g[d] - st.dN[(q * 27 + k) * 3 + d]) > 1e-12) return false;
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 3; ++i) {
            out.L[a][i] = static_cast<float>(L[a][i]);
            out.D[a][i] = static_cast<float>(D[a][i]);
        }
    return true;
}

std::vector<double> element_volumes(const HostMesh& m) {
    const ShapeTable& st = shape_table(m.element);
    const int S = st.S, Q = st.Q, E = m.nb_elements();
    std::vector<double> vol(E, 0.0);
    for (int e = 0; e < E; ++e) {
        const int32_t* t = &m.topo[static_cast<size_t>(e) * S];
        for (int q = 0; q < Q; ++q) {
            double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int k = 0; k < S; ++k) {
                const double* d = &st.dN[(q * S + k) * 3];
                const float* X = &m.xyz[3 * static_cast<size_t>(t[k])];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) J[i][j] += static_cast<double>(X[i]) * d[j];
            }
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                               J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            if (det == 0.0) throw std::invalid_argument("degenerate element (det J = 0)");
            vol[e] += std::fabs(det) * st.weights[q];
        }
    }
    return vol;
}

std::vector<float> compute_fem_mass(const HostMesh& m, const FemConstants& fc, float density,
                                    Mass_Distribution distrib) {
    const int N = m.nb_vertices();
    const int S = fc.S;
    std::vector<double> mass(N, 0.0);
    if (distrib == FemShape) {
        const double vd = static_cast<double>(density) / S;
        for (int e = 0; e < fc.E; ++e)
            for (int j = 0; j < S; ++j) mass[m.topo[static_cast<size_t>(e) * S + j]] += vd * fc.elem_volume[e];
    } else {
        double vol = 0.0;
        for (int e = 0; e < fc.E; ++e) vol += fc.elem_volume[e];
        const double pm = vol * density / N;
        for (int i = 0; i < N; ++i) mass[i] = pm;
    }
    std::vector<float> out(N);
    for (int i = 0; i < N; ++i) out[i] = static_cast<float>(mass[i]);
    return out;
}

void lame_parameters(float young, float poisson, float* lambda, float* mu) {
    *lambda = young * poisson / ((1.f + poisson) * (1.f - 2.f * poisson));
    *mu = young / (2.f * (1.f + poisson));
}

OwnersCSR build_owners(const HostMesh& m, const Coloration& col, bool by_valence) {
    const int N = m.nb_vertices();
    const int S = elem_nb_vertices(m.element);
    const int E = m.nb_elements();
    if (static_cast<int>(col.colors.size()) != N) throw std::invalid_argument("colouring size mismatch");
    // owner records pack (eid << 5 | local node): eid must fit 26 bits, S <= 32
    if (E >= (1 << 26) || S > 32) throw std::invalid_argument("mesh too large for the packed owners CSR");
    // per-vertex incident (eid, ref) lists in ascending eid
    std::vector<int64_t> cnt(N + 1, 0);
    for (int32_t v : m.topo) cnt[v + 1]++;
    for (int i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> inc(m.topo.size());
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < E; ++e)
        for (int j = 0; j < S; ++j) {
            const int32_t v = m.topo[static_cast<size_t>(e) * S + j];
            inc[fill[v]++] = (e << 5) | j;
        }
    OwnersCSR o;
    const int nc = col.nb_color;
    o.color_begin.assign(nc + 1, 0);
    o.color_max_nb.assign(nc, 0);
    o.color_incidences.assign(nc, 0);
    for (int v = 0; v < N; ++v) {
        const int c = col.colors[v];
        if (c < 0 || c >= nc) throw std::invalid_argument("colour out of range");
        o.color_begin[c + 1]++;
    }
    for (int c = 0; c < nc; ++c) o.color_begin[c + 1] += o.color_begin[c];
    std::vector<int32_t> cur(o.color_begin.begin(), o.color_begin.end() - 1);
    o.slot_vid.resize(N);
    for (int v = 0; v < N; ++v) o.slot_vid[cur[col.colors[v]]++] = v;
    if (by_valence) {
        // Within a colour the vertices are independent, so their order does not
        // change any result; most incidences first groups vertices of equal
        // work into the same wavefront.
        for (int c = 0; c < nc; ++c)
            std::stable_sort(o.slot_vid.begin() + o.color_begin[c], o.slot_vid.begin() + o.color_begin[c + 1],
                             [&](int a, int b) { return cnt[a + 1] - cnt[a] > cnt[b + 1] - cnt[b]; });
    }
    o.slot_nb.resize(N);
    o.slot_off.resize(N);
    o.own.reserve(m.topo.size());
    for (int s = 0; s < N; ++s) {
        const int v = o.slot_vid[s];
        o.slot_off[s] = static_cast<int32_t>(o.own.size());
        o.slot_nb[s] = static_cast<int32_t>(cnt[v + 1] - cnt[v]);
        for (int64_t i = cnt[v]; i < cnt[v + 1]; ++i) o.own.push_back(inc[i]);
        const int c = col.colors[v];
        o.color_max_nb[c] = std::max(o.color_max_nb[c], o.slot_nb[s]);
        o.color_incidences[c] += o.slot_nb[s];
    }
    return o;
}

void build_vertex_blocks(const OwnersCSR& o, int max_inc, int max_slots, std::vector<int32_t>& blocks,
                         std::vector<int32_t>& color_begin) {
    const int nc = static_cast<int>(o.color_begin.size()) - 1;
    blocks.clear();
    color_begin.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) {
        color_begin[c] = static_cast<int32_t>(blocks.size() / 4);
        int first = -1, n_slots = 0, n_inc = 0;
        auto flush = [&] {
            if (n_slots > 0) blocks.insert(blocks.end(), {first, n_slots, o.slot_off[first], n_inc});
            first = -1;
            n_slots = n_inc = 0;
        };
        for (int s = o.color_begin[c]; s < o.color_begin[c + 1]; ++s) {
            const int nb = o.slot_nb[s];
            if (nb > max_inc)
                throw std::invalid_argument("a vertex has " + std::to_string(nb) + " incident elements, more than the " +
                                            std::to_string(max_inc) + " of a fused task block");
            if (n_slots == max_slots || n_inc + nb > max_inc) flush();
            if (n_slots == 0) first = s;
            n_slots++;
            n_inc += nb;
        }
        flush();
    }
    color_begin[nc] = static_cast<int32_t>(blocks.size() / 4);
}

ElementSchedule build_element_schedule(const Coloration& col, int nb_elements) {
    ElementSchedule s;
    const int nc = col.nb_color;
    s.color_begin.assign(nc + 1, 0);
    for (int e = 0; e < nb_elements; ++e) s.color_begin[col.colors[e] + 1]++;
    for (int c = 0; c < nc; ++c) s.color_begin[c + 1] += s.color_begin[c];
    std::vector<int32_t> cur(s.color_begin.begin(), s.color_begin.end() - 1);
    s.eids.resize(nb_elements);
    for (int e = 0; e < nb_elements; ++e) s.eids[cur[col.colors[e]]++] = e;
    return s;
}

}  // namespace oe
