// Reference elements.
//
// The shape functions are not typed in per element.  Each element is described
// by its reference node coordinates (in the node order the mesh converters
// produce) and by its polynomial space; the Lagrange basis is obtained by
// inverting the Vandermonde matrix of that space on the nodes, in double
// precision.  The interpolant is unique, so this yields the same functions as
// the reference's closed forms (include/Dynamic/FEM/FEM_Shape.h:36-66,175-686)
// while being written once for every element.
//
// Quadrature rules keep the reference's point order and its float-rounded
// coordinates and weights, because the reference evaluates everything at those
// float values:
//   P1  1 point               FEM_Shape.h:41-47
//   Q1  2x2x2 Gauss           FEM_Shape.h:178-188
//   P2  4-point rule          FEM_Shape.h:230-238
//   P3  11-point rule         FEM_Shape.h:277-333 (points read as triples,
//                             FEM_Shape.cpp:7-16)
//   Q2  3x3x3 Gauss           FEM_Shape.h:486-517
#include "oe_host.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>

namespace oe {

bool element_supported(int e) {
    return e == Tetra || e == Hexa || e == Tetra10 || e == Tetra20 || e == Hexa27;
}

int elem_nb_vertices(Element e) {
    switch (e) {
        case Tetra: return 4;
        case Hexa: return 8;
        case Tetra10: return 10;
        case Tetra20: return 20;
        case Hexa27: return 27;
        default: throw std::invalid_argument("unsupported element");
    }
}

int elem_nb_quadrature(Element e) {
    switch (e) {
        case Tetra: return 1;
        case Hexa: return 8;
        case Tetra10: return 4;
        case Tetra20: return 11;
        case Hexa27: return 27;
        default: throw std::invalid_argument("unsupported element");
    }
}

std::string element_name(Element e) {
    switch (e) {
        case Tetra: return "Tetra";
        case Hexa: return "Hexa";
        case Tetra10: return "Tetra10";
        case Tetra20: return "Tetra20";
        case Hexa27: return "Hexa27";
        default: return "Unknown";
    }
}

namespace {

// Polynomial space: total degree <= deg (tets) or degree <= deg per axis (hexes).
std::vector<int> monomials(bool tensor, int deg) {
    std::vector<int> m;
    for (int a = 0; a <= deg; ++a)
        for (int b = 0; b <= deg; ++b)
            for (int c = 0; c <= deg; ++c) {
                if (!tensor && a + b + c > deg) continue;
                m.push_back(a); m.push_back(b); m.push_back(c);
            }
    return m;
}

double ipow(double x, int k) {
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

// Solve A X = I (n x n, row-major) by Gauss-Jordan with partial pivoting.
std::vector<double> invert(std::vector<double> A, int n) {
    std::vector<double> X(n * n, 0.0);
    for (int i = 0; i < n; ++i) X[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
        if (std::fabs(A[piv * n + c]) < 1e-14) throw std::runtime_error("singular Vandermonde");
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[c * n + k], A[piv * n + k]);
                std::swap(X[c * n + k], X[piv * n + k]);
            }
        const double inv = 1.0 / A[c * n + c];
        for (int k = 0; k < n; ++k) { A[c * n + k] *= inv; X[c * n + k] *= inv; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = A[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { A[r * n + k] -= f * A[c * n + k]; X[r * n + k] -= f * X[c * n + k]; }
        }
    }
    return X;
}

void add_node(std::vector<double>& v, double x, double y, double z) {
    v.push_back(x); v.push_back(y); v.push_back(z);
}

ShapeTable make_table(Element e) {
    ShapeTable t;
    t.element = e;
    bool tensor = false;
    int deg = 1;
    std::vector<double>& N = t.nodes;
    std::vector<double>& qp = t.qpoints;
    std::vector<double>& w = t.weights;
    switch (e) {
        case Tetra: {
            deg = 1;
            add_node(N, 0, 0, 0); add_node(N, 1, 0, 0); add_node(N, 0, 1, 0); add_node(N, 0, 0, 1);
            add_node(qp, 0.25f, 0.25f, 0.25f);
            w.push_back(1.f / 6.f);
            break;
        }
        case Hexa: {
            tensor = true; deg = 1;
            const int c8[8][3] = {{-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1}};
            for (auto& c : c8) add_node(N, c[0], c[1], c[2]);
            const float g = 1.f / std::sqrt(3.f);
            for (auto& c : c8) { add_node(qp, c[0] * g, c[1] * g, c[2] * g); w.push_back(1.f); }
            break;
        }
        case Tetra10: {
            deg = 2;
            add_node(N, 0, 0, 0); add_node(N, 1, 0, 0); add_node(N, 0, 1, 0); add_node(N, 0, 0, 1);
            // edge mid-points in converter edge order (0,1) (1,2) (0,2) (0,3) (1,3) (2,3)
            add_node(N, .5, 0, 0); add_node(N, .5, .5, 0); add_node(N, 0, .5, 0);
            add_node(N, 0, 0, .5); add_node(N, .5, 0, .5); add_node(N, 0, .5, .5);
            const float a = (5.f - std::sqrt(5.f)) / 20.f;
            const float b = (5.f + 3 * std::sqrt(5.f)) / 20.f;
            add_node(qp, b, a, a); add_node(qp, a, b, a); add_node(qp, a, a, b); add_node(qp, a, a, a);
            for (int i = 0; i < 4; ++i) w.push_back(1.f / 4.f * 1.f / 6.f);
            break;
        }
        case Tetra20: {
            deg = 3;
            const double a = 1.0 / 3.0, b = 2.0 / 3.0;
            add_node(N, 0, 0, 0); add_node(N, 1, 0, 0); add_node(N, 0, 1, 0); add_node(N, 0, 0, 1);
            // two nodes per edge at 1/3 and 2/3 from the first end, edge order
            // (0,1) (1,2) (2,0) (0,3) (1,3) (2,3)
            add_node(N, a, 0, 0); add_node(N, b, 0, 0);
            add_node(N, b, a, 0); add_node(N, a, b, 0);
            add_node(N, 0, b, 0); add_node(N, 0, a, 0);
            add_node(N, 0, 0, a); add_node(N, 0, 0, b);
            add_node(N, b, 0, a); add_node(N, a, 0, b);
            add_node(N, 0, b, a); add_node(N, 0, a, b);
            // face centroids, faces (0,1,3) (1,2,3) (0,3,2) (0,2,1)
            add_node(N, a, 0, a); add_node(N, a, a, a); add_node(N, 0, a, a); add_node(N, a, a, 0);
            // 11-point rule (data table, kept as in the reference)
            static const double P3Q[33] = {
                0.040490506727590427904, 0.013560701879802881248, 0.77125473269537614396,
                0.75250850700965499218, 0.068099370938206657544, 0.097987203649279111528,
                0.067223294893383397919, 0.035183929773598715540, 0.15636389323939528517,
                0.41926631387951301955, 0.047781435559086662956, 0.47961101102565506515,
                0.45076587609127682921, 0.059456616299433828754, 0.056824017127933668103,
                0.12941137378891040544, 0.33019041483746447430, 0.0023910074574393647140,
                0.12154199133392780408, 0.30649398842969027834, 0.56297276014304609149,
                0.097204644587583266535, 0.68439041545304002412, 0.11180076739738309399,
                0.029569495206479612381, 0.31790356021339460924, 0.32329398483747895376,
                0.43271023904776856339, 0.35382323920929709127, 0.10962240533194123060,
                0.24027666492807261966, 0.12680172591539201621, 0.32847320672203844660};
            static const double P3W[11] = {
                0.039251090924839956986, 0.055273369155936898089, 0.055393798871576367671,
                0.059933185146559528333, 0.069469965937635366759, 0.076162715245558357258,
                0.079426680068025307113, 0.10646803415549009608, 0.11023423242849765546,
                0.15497611601624608491, 0.19341081204963445073};
            for (int i = 0; i < 33; ++i) qp.push_back(static_cast<float>(P3Q[i]));
            for (int i = 0; i < 11; ++i) w.push_back(static_cast<float>(P3W[i]) / 6.f);
            break;
        }
        case Hexa27: {
            tensor = true; deg = 2;
            // corners, edge mid-points in converter edge order, face centres in
            // converter quad order, cell centre
            const int n27[27][3] = {
                {-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1},
                {0,-1,-1},{1,0,-1},{0,1,-1},{-1,0,-1},
                {-1,-1,0},{1,-1,0},{1,1,0},{-1,1,0},
                {0,-1,1},{1,0,1},{0,1,1},{-1,0,1},
                {0,0,-1},{0,-1,0},{1,0,0},{0,1,0},{-1,0,0},{0,0,1},
                {0,0,0}};
            for (auto& c : n27) add_node(N, c[0], c[1], c[2]);
            // Gauss points: same pattern of signs as the nodes (corner, edge,
            // face, volume groups), coordinate sqrt(3/5).
            const int q27[27][3] = {
                {-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1},
                {0,-1,-1},{1,0,-1},{0,1,-1},{-1,0,-1},
                {-1,-1,0},{1,-1,0},{1,1,0},{-1,1,0},
                {0,-1,1},{1,0,1},{0,1,1},{-1,0,1},
                {0,0,-1},{0,-1,0},{1,0,0},{0,1,0},{-1,0,0},{0,0,1},
                {0,0,0}};
            const float v = std::sqrt(3.f / 5.f);
            const float z8 = 8.f / 9.f, v5 = 5.f / 9.f;
            const float w1 = v5 * v5 * v5, w2 = v5 * v5 * z8, w3 = v5 * z8 * z8, w4 = z8 * z8 * z8;
            for (int i = 0; i < 27; ++i) {
                add_node(qp, q27[i][0] * v, q27[i][1] * v, q27[i][2] * v);
                w.push_back(i < 8 ? w1 : (i < 20 ? w2 : (i < 26 ? w3 : w4)));
            }
            break;
        }
        default: throw std::invalid_argument("unsupported element");
    }
    t.S = static_cast<int>(N.size() / 3);
    t.Q = static_cast<int>(w.size());
    t.mono = monomials(tensor, deg);
    const int S = t.S;
    if (static_cast<int>(t.mono.size() / 3) != S) throw std::logic_error("space size != node count");
    // Vandermonde V[i][j] = m_j(node_i); basis coefficients C = V^-1, N_k = sum_j C[j][k] m_j
    std::vector<double> V(S * S);
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
            V[i * S + j] = ipow(N[3 * i], t.mono[3 * j]) * ipow(N[3 * i + 1], t.mono[3 * j + 1]) *
                           ipow(N[3 * i + 2], t.mono[3 * j + 2]);
    t.coeff = invert(V, S);
    t.dN.assign(static_cast<size_t>(t.Q) * S * 3, 0.0);
    std::vector<double> Nv(S), dNv(3 * S);
    for (int q = 0; q < t.Q; ++q) {
        t.eval(&qp[3 * q], Nv.data(), dNv.data());
        for (int k = 0; k < S; ++k)
            for (int c = 0; c < 3; ++c) t.dN[(q * S + k) * 3 + c] = dNv[3 * k + c];
    }
    return t;
}

}  // namespace

void ShapeTable::eval(const double* x, double* Nout, double* dNdx) const {
    const int n = S;
    std::vector<double> m(n), dm(3 * n);
    for (int j = 0; j < n; ++j) {
        const int a = mono[3 * j], b = mono[3 * j + 1], c = mono[3 * j + 2];
        const double px = ipow(x[0], a), py = ipow(x[1], b), pz = ipow(x[2], c);
        m[j] = px * py * pz;
        dm[3 * j + 0] = a ? a * ipow(x[0], a - 1) * py * pz : 0.0;
        dm[3 * j + 1] = b ? b * px * ipow(x[1], b - 1) * pz : 0.0;
        dm[3 * j + 2] = c ? c * px * py * ipow(x[2], c - 1) : 0.0;
    }
    for (int k = 0; k < n; ++k) {
        double s = 0, sx = 0, sy = 0, sz = 0;
        for (int j = 0; j < n; ++j) {
            const double cjk = coeff[j * n + k];
            s += cjk * m[j];
            sx += cjk * dm[3 * j]; sy += cjk * dm[3 * j + 1]; sz += cjk * dm[3 * j + 2];
        }
        if (Nout) Nout[k] = s;
        if (dNdx) { dNdx[3 * k] = sx; dNdx[3 * k + 1] = sy; dNdx[3 * k + 2] = sz; }
    }
}

const ShapeTable& shape_table(Element e) {
    static std::mutex mtx;
    static std::map<int, ShapeTable> cache;
    std::lock_guard<std::mutex> lock(mtx);
    auto it = cache.find(e);
    if (it == cache.end()) it = cache.emplace(e, make_table(e)).first;
    return it->second;
}

}  // namespace oe
