/*
 * FP64 CPU oracle - TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * How each part is pinned (tests/test_oracle.py = O, tests/test_oracle_pins.py
 * = P, tests/test_golden_vtk.py = G):
 *   shapes        O: Kronecker delta at the nodes, partition of unity, exact
 *                 derivatives vs finite differences, quadrature exactness on
 *                 monomials; cross-checked against the product's independent
 *                 Vandermonde derivation (tests/test_host.py).
 *   beam / conv.  O: closed-form counts (425/1536, 140481, 18785, 60625), grid
 *                 coordinates, node = edge midpoint / third / face centroid,
 *                 signed volumes summing to the beam volume, first-encounter
 *                 numbering on a hand-checked 1-cell case.  G: the reference's
 *                 own result files (mesh sections, masses).
 *   DSAT / SLF    O: hand-traced known answers on small graphs (K4, two tets
 *                 sharing a face, a path), a literal sorted-set restatement in
 *                 Python (tests/ref_literal.py) on small meshes, proper
 *                 colouring invariants.
 *   FEM consts    O: sum V = element volume, JX_inv J = I.
 *   masses        O: sum = density * volume for both distributions; G.
 *   materials     O: P = dW/dF and H = d2W/dx_r^2 by finite differences of the
 *                 energies; P_snh(I) = 0, P_nh(I) = mu I.
 *   VBD step      O: free fall (lambda = mu = 0) closed form; one free vertex
 *                 = the brute-force Newton minimiser; Jacobi == Gauss-Seidel
 *                 under a proper colouring; Chebyshev keeps free fall;
 *                 residual vanishes at convergence.  P: the damping branch
 *                 (the damped fixed point solves the damped local equation,
 *                 derivatives by finite differences).
 *   XPBD step     P: free-fall closed form of the semi-implicit frame, the
 *                 global-damping velocity update in closed form, V0 on one
 *                 P1 tet and V2 on one P2 tet against hand derivations with
 *                 finite-difference energy gradients.  O: V0 == V2 when E > 0.
 *   explicit      O: f = -dE/dx by finite differences of the total energy.
 *                 P: free-fall closed form, f = -dE/dx - k H_vv v with H_vv
 *                 by finite differences.
 *   diagnostics   O: rest, uniaxial stretch and reflection closed forms.
 * The restated reference CPU paths (cpu_xpbd_impl.h, cpu_vbd_impl.h) are
 * pinned in P: free fall, a hand-traced two-element XPBD case (insertion
 * order, accumulated multipliers, RB_Fixation, global damping), converged
 * CPU VBD steps equal to this oracle's, float twins within float rounding.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ====================================================================== */
/* element tables                                                         */
/* ====================================================================== */

int orc_nodes(int e) {
    switch (e) {
        case ORC_TETRA: return 4;
        case ORC_HEXA: return 8;
        case ORC_TETRA10: return 10;
        case ORC_TETRA20: return 20;
        case ORC_HEXA27: return 27;
        default: return -1;
    }
}

int orc_quadrature(int e) {
    switch (e) {
        case ORC_TETRA: return 1;
        case ORC_HEXA: return 8;
        case ORC_TETRA10: return 4;
        case ORC_TETRA20: return 11;
        case ORC_HEXA27: return 27;
        default: return -1;
    }
}

/* hexahedron node coordinates in {-1, 0, 1}: FEM_Shape.h:190-196 (Q1) and
 * :519-557 (Q2) */
static const int HEX27[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
    {0, -1, -1},  {1, 0, -1},  {0, 1, -1}, {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0},
    {0, -1, 1},   {1, 0, 1},   {0, 1, 1},  {-1, 0, 1},  {0, 0, -1},  {0, -1, 0}, {1, 0, 0}, {0, 1, 0},
    {-1, 0, 0},   {0, 0, 1},   {0, 0, 0}};

/* 1D quadratic Lagrange on nodes -1, 0, 1 and its derivative */
static double lag2(int a, double x) {
    if (a < 0) return 0.5 * x * (x - 1.0);
    if (a > 0) return 0.5 * x * (x + 1.0);
    return 1.0 - x * x;
}
static double dlag2(int a, double x) {
    if (a < 0) return x - 0.5;
    if (a > 0) return x + 0.5;
    return -2.0 * x;
}

/* tetra edges (P2 node order, BeamGenerator.cpp:178) and P3 edge/face orders
 * (MeshConverter.h:76-82) */
static const int TET10_EDGES[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
static const int TET20_EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static const int TET20_FACES[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};

/* Closed-form Lagrange bases.  Tets use barycentric coordinates
 * L = (1-s-t-l, s, t, l); hexes use tensor products of 1D Lagrange.
 * Follows the node sets of FEM_Shape.h:36-66 (P1), :175-224 (Q1),
 * :227-272 (P2), :274-481 (P3), :483-686 (Q2). */
int orc_shape_eval(int e, const double x[3], double* N, double* dN) {
    const double L[4] = {1.0 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
    static const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    int k, c;
    switch (e) {
        case ORC_TETRA:
            for (k = 0; k < 4; ++k) {
                N[k] = L[k];
                for (c = 0; c < 3; ++c) dN[3 * k + c] = dL[k][c];
            }
            return 4;
        case ORC_TETRA10:
            for (k = 0; k < 4; ++k) {
                N[k] = L[k] * (2.0 * L[k] - 1.0);
                for (c = 0; c < 3; ++c) dN[3 * k + c] = (4.0 * L[k] - 1.0) * dL[k][c];
            }
            for (k = 0; k < 6; ++k) {
                const int a = TET10_EDGES[k][0], b = TET10_EDGES[k][1];
                N[4 + k] = 4.0 * L[a] * L[b];
                for (c = 0; c < 3; ++c) dN[3 * (4 + k) + c] = 4.0 * (dL[a][c] * L[b] + L[a] * dL[b][c]);
            }
            return 10;
        case ORC_TETRA20:
            for (k = 0; k < 4; ++k) {
                const double l = L[k];
                N[k] = 0.5 * l * (3.0 * l - 1.0) * (3.0 * l - 2.0);
                /* d/dl of 0.5 (9 l^3 - 9 l^2 + 2 l) = 0.5 (27 l^2 - 18 l + 2) */
                for (c = 0; c < 3; ++c) dN[3 * k + c] = 0.5 * (27.0 * l * l - 18.0 * l + 2.0) * dL[k][c];
            }
            for (k = 0; k < 6; ++k) {
                const int a = TET20_EDGES[k][0], b = TET20_EDGES[k][1];
                const int n0 = 4 + 2 * k, n1 = 5 + 2 * k;
                const double la = L[a], lb = L[b];
                /* node at 1/3 from a: 9/2 la lb (3 la - 1); at 2/3: 9/2 la lb (3 lb - 1) */
                N[n0] = 4.5 * la * lb * (3.0 * la - 1.0);
                N[n1] = 4.5 * la * lb * (3.0 * lb - 1.0);
                for (c = 0; c < 3; ++c) {
                    const double da = dL[a][c], db = dL[b][c];
                    dN[3 * n0 + c] = 4.5 * ((6.0 * la - 1.0) * lb * da + la * (3.0 * la - 1.0) * db);
                    dN[3 * n1 + c] = 4.5 * ((6.0 * lb - 1.0) * la * db + lb * (3.0 * lb - 1.0) * da);
                }
            }
            for (k = 0; k < 4; ++k) {
                const int a = TET20_FACES[k][0], b = TET20_FACES[k][1], d = TET20_FACES[k][2];
                N[16 + k] = 27.0 * L[a] * L[b] * L[d];
                for (c = 0; c < 3; ++c)
                    dN[3 * (16 + k) + c] =
                        27.0 * (dL[a][c] * L[b] * L[d] + L[a] * dL[b][c] * L[d] + L[a] * L[b] * dL[d][c]);
            }
            return 20;
        case ORC_HEXA:
            for (k = 0; k < 8; ++k) {
                const double a = HEX27[k][0], b = HEX27[k][1], d = HEX27[k][2];
                const double fa = 0.5 * (1 + a * x[0]), fb = 0.5 * (1 + b * x[1]), fd = 0.5 * (1 + d * x[2]);
                N[k] = fa * fb * fd;
                dN[3 * k] = 0.5 * a * fb * fd;
                dN[3 * k + 1] = 0.5 * b * fa * fd;
                dN[3 * k + 2] = 0.5 * d * fa * fb;
            }
            return 8;
        case ORC_HEXA27:
            for (k = 0; k < 27; ++k) {
                const int a = HEX27[k][0], b = HEX27[k][1], d = HEX27[k][2];
                const double la = lag2(a, x[0]), lb = lag2(b, x[1]), ld = lag2(d, x[2]);
                N[k] = la * lb * ld;
                dN[3 * k] = dlag2(a, x[0]) * lb * ld;
                dN[3 * k + 1] = la * dlag2(b, x[1]) * ld;
                dN[3 * k + 2] = la * lb * dlag2(d, x[2]);
            }
            return 27;
        default: return -1;
    }
}

/* Quadrature rules (data, kept as in the reference including the float
 * rounding of the stored values): FEM_Shape.h:41-47 (P1), :178-188 (Q1),
 * :230-238 (P2), :277-333 (P3, read as point triples by
 * FEM_Shape.cpp:7-16), :486-517 (Q2). */
int orc_quadrature_rule(int e, double* qp, double* w) {
    int i;
    switch (e) {
        case ORC_TETRA:
            qp[0] = qp[1] = qp[2] = 0.25f;
            w[0] = 1.f / 6.f;
            return 1;
        case ORC_HEXA: {
            const float g = 1.f / sqrtf(3.f);
            for (i = 0; i < 8; ++i) {
                qp[3 * i] = HEX27[i][0] * g;
                qp[3 * i + 1] = HEX27[i][1] * g;
                qp[3 * i + 2] = HEX27[i][2] * g;
                w[i] = 1.0;
            }
            return 8;
        }
        case ORC_TETRA10: {
            const float a = (5.f - sqrtf(5.f)) / 20.f, b = (5.f + 3 * sqrtf(5.f)) / 20.f;
            const double pts[4][3] = {{b, a, a}, {a, b, a}, {a, a, b}, {a, a, a}};
            for (i = 0; i < 4; ++i) {
                qp[3 * i] = pts[i][0];
                qp[3 * i + 1] = pts[i][1];
                qp[3 * i + 2] = pts[i][2];
                w[i] = 1.f / 4.f * 1.f / 6.f;
            }
            return 4;
        }
        case ORC_TETRA20: {
            static const double P[33] = {
                0.040490506727590427904, 0.013560701879802881248, 0.77125473269537614396,
                0.75250850700965499218,  0.068099370938206657544, 0.097987203649279111528,
                0.067223294893383397919, 0.035183929773598715540, 0.15636389323939528517,
                0.41926631387951301955,  0.047781435559086662956, 0.47961101102565506515,
                0.45076587609127682921,  0.059456616299433828754, 0.056824017127933668103,
                0.12941137378891040544,  0.33019041483746447430,  0.0023910074574393647140,
                0.12154199133392780408,  0.30649398842969027834,  0.56297276014304609149,
                0.097204644587583266535, 0.68439041545304002412,  0.11180076739738309399,
                0.029569495206479612381, 0.31790356021339460924,  0.32329398483747895376,
                0.43271023904776856339,  0.35382323920929709127,  0.10962240533194123060,
                0.24027666492807261966,  0.12680172591539201621,  0.32847320672203844660};
            static const double W[11] = {
                0.039251090924839956986, 0.055273369155936898089, 0.055393798871576367671, 0.059933185146559528333,
                0.069469965937635366759, 0.076162715245558357258, 0.079426680068025307113, 0.10646803415549009608,
                0.11023423242849765546,  0.15497611601624608491,  0.19341081204963445073};
            for (i = 0; i < 33; ++i) qp[i] = (float)P[i];
            for (i = 0; i < 11; ++i) w[i] = (float)W[i] / 6.f;
            return 11;
        }
        case ORC_HEXA27: {
            const float v = sqrtf(3.f / 5.f);
            const float z = 8.f / 9.f, f5 = 5.f / 9.f;
            const float w1 = f5 * f5 * f5, w2 = f5 * f5 * z, w3 = f5 * z * z, w4 = z * z * z;
            for (i = 0; i < 27; ++i) {
                qp[3 * i] = HEX27[i][0] * v;
                qp[3 * i + 1] = HEX27[i][1] * v;
                qp[3 * i + 2] = HEX27[i][2] * v;
                w[i] = i < 8 ? w1 : (i < 20 ? w2 : (i < 26 ? w3 : w4));
            }
            return 27;
        }
        default: return -1;
    }
}

/* ====================================================================== */
/* beam generation and conversions (BeamGenerator.cpp:5-43,86-92,172-377) */
/* ====================================================================== */

/* open-addressing map from a sorted key (up to 4 ints) to an int payload */
typedef struct {
    int64_t cap;
    int* keys;   /* cap * 4 */
    int* vals;   /* cap * 3 : id or (first, n0, n1) */
    char* used;
} kmap;

static void kmap_init(kmap* m, int64_t expect) {
    int64_t cap = 16;
    while (cap < 2 * expect + 16) cap <<= 1;
    m->cap = cap;
    m->keys = (int*)malloc(sizeof(int) * 4 * cap);
    m->vals = (int*)malloc(sizeof(int) * 3 * cap);
    m->used = (char*)calloc(cap, 1);
}
static void kmap_free(kmap* m) {
    free(m->keys);
    free(m->vals);
    free(m->used);
}
static uint64_t khash(const int k[4]) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    int i;
    for (i = 0; i < 4; ++i) {
        h ^= (uint64_t)(uint32_t)k[i] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    }
    return h;
}
/* returns slot; *found = 1 if present */
static int64_t kmap_find(kmap* m, const int k[4], int* found) {
    int64_t i = (int64_t)(khash(k) & (uint64_t)(m->cap - 1));
    while (m->used[i]) {
        if (memcmp(&m->keys[4 * i], k, sizeof(int) * 4) == 0) {
            *found = 1;
            return i;
        }
        i = (i + 1) & (m->cap - 1);
    }
    *found = 0;
    return i;
}
static void sort_key(int* k, int n) {
    int i, j;
    for (i = 1; i < n; ++i)
        for (j = i; j > 0 && k[j - 1] > k[j]; --j) {
            int t = k[j];
            k[j] = k[j - 1];
            k[j - 1] = t;
        }
    for (i = n; i < 4; ++i) k[i] = -1;
}

typedef struct {
    float* xyz;
    int nv, cap;
} geo;
static int geo_push(geo* g, float x, float y, float z) {
    if (g->nv == g->cap) {
        g->cap = g->cap ? 2 * g->cap : 1024;
        g->xyz = (float*)realloc(g->xyz, sizeof(float) * 3 * g->cap);
    }
    g->xyz[3 * g->nv] = x;
    g->xyz[3 * g->nv + 1] = y;
    g->xyz[3 * g->nv + 2] = z;
    return g->nv++;
}

int orc_beam(int element, int nx, int ny, int nz, float sx, float sy, float sz, int* nv_out, int* ne_out, float* xyz,
             int* topo) {
    const int hexa = (element == ORC_HEXA || element == ORC_HEXA27);
    const int X = nx + 1, Y = ny + 1, Z = nz + 1;
    const float hx = sx / (float)(X - 1), hy = sy / (float)(Y - 1), hz = sz / (float)(Z - 1);
    static const int tet6[24] = {0, 4, 6, 5, 3, 6, 2, 0, 0, 4, 7, 6, 3, 6, 0, 7, 2, 0, 6, 1, 6, 0, 5, 1};
    const int ncell = nx * ny * nz;
    const int n_lin = hexa ? ncell : 6 * ncell;
    const int S_lin = hexa ? 8 : 4;
    int* lin = (int*)malloc(sizeof(int) * S_lin * n_lin);
    geo g = {0, 0, 0};
    int x, y, z, i, e = 0;
    if (orc_nodes(element) < 0 || nx < 1 || ny < 1 || nz < 1) {
        free(lin);
        return -1;
    }
    for (z = 0; z < Z; ++z)
        for (y = 0; y < Y; ++y)
            for (x = 0; x < X; ++x) geo_push(&g, (float)x * hx, (float)y * hy, (float)z * hz);
#define ID(a, b, c) ((a) + (b) * X + (c) * X * Y)
    for (z = 0; z < Z - 1; ++z)
        for (y = 0; y < Y - 1; ++y)
            for (x = 0; x < X - 1; ++x) {
                const int c8[8] = {ID(x, y, z),         ID(x + 1, y, z),     ID(x + 1, y, z + 1),
                                   ID(x, y, z + 1),     ID(x, y + 1, z),     ID(x + 1, y + 1, z),
                                   ID(x + 1, y + 1, z + 1), ID(x, y + 1, z + 1)};
                if (hexa) {
                    for (i = 0; i < 8; ++i) lin[8 * e + i] = c8[i];
                    e++;
                } else {
                    int t;
                    for (t = 0; t < 6; ++t, ++e)
                        for (i = 0; i < 4; ++i) lin[4 * e + i] = c8[tet6[4 * t + i]];
                }
            }
#undef ID
    {
        const int S = orc_nodes(element);
        int* out = (int*)malloc(sizeof(int) * (size_t)S * n_lin);
        kmap em, fm;
        int k, d;
        kmap_init(&em, (int64_t)n_lin * 6);
        kmap_init(&fm, (int64_t)n_lin * 6);
        for (e = 0; e < n_lin; ++e) {
            int* o = out + (size_t)S * e;
            const int* t = lin + (size_t)S_lin * e;
            for (i = 0; i < S_lin; ++i) o[i] = t[i];
            if (element == ORC_TETRA10) {
                for (k = 0; k < 6; ++k) {
                    int a = o[TET10_EDGES[k][0]], b = o[TET10_EDGES[k][1]], key[4] = {a, b}, found;
                    int64_t sl;
                    sort_key(key, 2);
                    sl = kmap_find(&em, key, &found);
                    if (!found) {
                        if (a > b) { int tt = a; a = b; b = tt; }
                        memcpy(&em.keys[4 * sl], key, sizeof key);
                        em.used[sl] = 1;
                        em.vals[3 * sl] = geo_push(&g, 0.5f * (g.xyz[3 * a] + g.xyz[3 * b]),
                                                   0.5f * (g.xyz[3 * a + 1] + g.xyz[3 * b + 1]),
                                                   0.5f * (g.xyz[3 * a + 2] + g.xyz[3 * b + 2]));
                    }
                    o[4 + k] = em.vals[3 * sl];
                }
            } else if (element == ORC_TETRA20) {
                for (k = 0; k < 6; ++k) {
                    const int a = o[TET20_EDGES[k][0]], b = o[TET20_EDGES[k][1]];
                    int key[4] = {a, b}, found;
                    int64_t sl;
                    sort_key(key, 2);
                    sl = kmap_find(&em, key, &found);
                    if (!found) {
                        int w;
                        memcpy(&em.keys[4 * sl], key, sizeof key);
                        em.used[sl] = 1;
                        em.vals[3 * sl] = a;
                        for (w = 0; w < 2; ++w) {
                            const float wt = (float)(w + 1) / 3.f, om = 1.f - wt;
                            float p[3];
                            for (d = 0; d < 3; ++d) p[d] = g.xyz[3 * a + d] * om + g.xyz[3 * b + d] * wt;
                            em.vals[3 * sl + 1 + w] = geo_push(&g, p[0], p[1], p[2]);
                        }
                    }
                    if (em.vals[3 * sl] == a) {
                        o[4 + 2 * k] = em.vals[3 * sl + 1];
                        o[5 + 2 * k] = em.vals[3 * sl + 2];
                    } else {
                        o[4 + 2 * k] = em.vals[3 * sl + 2];
                        o[5 + 2 * k] = em.vals[3 * sl + 1];
                    }
                }
                for (k = 0; k < 4; ++k) {
                    const int a = o[TET20_FACES[k][0]], b = o[TET20_FACES[k][1]], c = o[TET20_FACES[k][2]];
                    int key[4] = {a, b, c}, found;
                    int64_t sl;
                    sort_key(key, 3);
                    sl = kmap_find(&fm, key, &found);
                    if (!found) {
                        const float w = 1.f / 3.f;
                        float p[3];
                        for (d = 0; d < 3; ++d) p[d] = g.xyz[3 * a + d] * w + g.xyz[3 * b + d] * w + g.xyz[3 * c + d] * w;
                        memcpy(&fm.keys[4 * sl], key, sizeof key);
                        fm.used[sl] = 1;
                        fm.vals[3 * sl] = geo_push(&g, p[0], p[1], p[2]);
                    }
                    o[16 + k] = fm.vals[3 * sl];
                }
            } else if (element == ORC_HEXA27) {
                static const int HE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                              {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
                static const int HQ[6][4] = {{0, 1, 2, 3}, {4, 5, 1, 0}, {5, 6, 2, 1},
                                             {3, 2, 6, 7}, {7, 4, 0, 3}, {7, 6, 5, 4}};
                float cc[3] = {0.f, 0.f, 0.f};
                for (k = 0; k < 12; ++k) {
                    const int a = o[HE[k][0]], b = o[HE[k][1]];
                    int key[4] = {a, b}, found;
                    int64_t sl;
                    sort_key(key, 2);
                    sl = kmap_find(&em, key, &found);
                    if (!found) {
                        float p[3];
                        for (d = 0; d < 3; ++d) p[d] = 0.5f * (g.xyz[3 * a + d] + g.xyz[3 * b + d]);
                        memcpy(&em.keys[4 * sl], key, sizeof key);
                        em.used[sl] = 1;
                        em.vals[3 * sl] = geo_push(&g, p[0], p[1], p[2]);
                    }
                    o[8 + k] = em.vals[3 * sl];
                }
                for (k = 0; k < 6; ++k) {
                    const int a = o[HQ[k][0]], b = o[HQ[k][1]], c = o[HQ[k][2]], q = o[HQ[k][3]];
                    int key[4] = {a, b, c, q}, found;
                    int64_t sl;
                    sort_key(key, 4);
                    sl = kmap_find(&fm, key, &found);
                    if (!found) {
                        const float w = 1.f / 4.f;
                        float p[3];
                        for (d = 0; d < 3; ++d)
                            p[d] = g.xyz[3 * a + d] * w + g.xyz[3 * b + d] * w + g.xyz[3 * c + d] * w +
                                   g.xyz[3 * q + d] * w;
                        memcpy(&fm.keys[4 * sl], key, sizeof key);
                        fm.used[sl] = 1;
                        fm.vals[3 * sl] = geo_push(&g, p[0], p[1], p[2]);
                    }
                    o[20 + k] = fm.vals[3 * sl];
                }
                for (k = 0; k < 8; ++k)
                    for (d = 0; d < 3; ++d) cc[d] += g.xyz[3 * o[k] + d];
                for (d = 0; d < 3; ++d) cc[d] /= 8.f;
                o[26] = geo_push(&g, cc[0], cc[1], cc[2]);
            }
        }
        kmap_free(&em);
        kmap_free(&fm);
        if (nv_out) *nv_out = g.nv;
        if (ne_out) *ne_out = n_lin;
        if (xyz) memcpy(xyz, g.xyz, sizeof(float) * 3 * g.nv);
        if (topo) memcpy(topo, out, sizeof(int) * (size_t)S * n_lin);
        free(out);
    }
    free(lin);
    free(g.xyz);
    return 0;
}

/* ====================================================================== */
/* graphs and colourings (include/Tools/Graph.h)                          */
/* ====================================================================== */

typedef struct {
    int n;
    int64_t* off;
    int* adj;
} mgraph;

static int cmp_int(const void* a, const void* b) {
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

static void mgraph_free(mgraph* g) {
    free(g->off);
    free(g->adj);
}

/* Graph::build_primal (Graph.h:36-54): every ordered co-element pair once per
 * shared element; rows sorted with repeats (multiset iteration order). */
static mgraph build_primal(int element, const int* topo, int ne) {
    const int S = orc_nodes(element);
    mgraph g;
    int64_t i, *cur;
    int e, j, k, n = 0;
    for (i = 0; i < (int64_t)S * ne; ++i)
        if (topo[i] + 1 > n) n = topo[i] + 1;
    g.n = n;
    g.off = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (i = 0; i < (int64_t)S * ne; ++i) g.off[topo[i] + 1] += S - 1;
    for (j = 0; j < n; ++j) g.off[j + 1] += g.off[j];
    g.adj = (int*)malloc(sizeof(int) * (g.off[n] ? g.off[n] : 1));
    cur = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1));
    memcpy(cur, g.off, sizeof(int64_t) * n);
    for (e = 0; e < ne; ++e)
        for (j = 0; j < S; ++j)
            for (k = 0; k < S; ++k)
                if (k != j) g.adj[cur[topo[e * S + j]]++] = topo[e * S + k];
    free(cur);
    for (j = 0; j < n; ++j) qsort(g.adj + g.off[j], (size_t)(g.off[j + 1] - g.off[j]), sizeof(int), cmp_int);
    return g;
}

/* Graph::build_dual_vertex (Graph.h:56-85): elements adjacent once per shared
 * vertex. */
static mgraph build_dual(int element, const int* topo, int ne) {
    const int S = orc_nodes(element);
    int nvtx = 0, v, e, j, k;
    int64_t i, *oo, *cur;
    int* owners;
    mgraph g;
    for (i = 0; i < (int64_t)S * ne; ++i)
        if (topo[i] + 1 > nvtx) nvtx = topo[i] + 1;
    oo = (int64_t*)calloc((size_t)nvtx + 1, sizeof(int64_t));
    for (i = 0; i < (int64_t)S * ne; ++i) oo[topo[i] + 1]++;
    for (v = 0; v < nvtx; ++v) oo[v + 1] += oo[v];
    owners = (int*)malloc(sizeof(int) * (size_t)S * ne);
    cur = (int64_t*)malloc(sizeof(int64_t) * (nvtx ? nvtx : 1));
    memcpy(cur, oo, sizeof(int64_t) * nvtx);
    for (e = 0; e < ne; ++e)
        for (j = 0; j < S; ++j) owners[cur[topo[e * S + j]]++] = e;
    g.n = ne;
    g.off = (int64_t*)calloc((size_t)ne + 1, sizeof(int64_t));
    for (v = 0; v < nvtx; ++v) {
        const int64_t d = oo[v + 1] - oo[v];
        for (i = oo[v]; i < oo[v + 1]; ++i) g.off[owners[i] + 1] += d - 1;
    }
    for (e = 0; e < ne; ++e) g.off[e + 1] += g.off[e];
    g.adj = (int*)malloc(sizeof(int) * (g.off[ne] ? g.off[ne] : 1));
    free(cur);
    cur = (int64_t*)malloc(sizeof(int64_t) * (ne ? ne : 1));
    memcpy(cur, g.off, sizeof(int64_t) * ne);
    for (v = 0; v < nvtx; ++v)
        for (i = oo[v]; i < oo[v + 1]; ++i)
            for (k = (int)(i + 1 - oo[v]); k < oo[v + 1] - oo[v]; ++k) {
                const int a = owners[i], b = owners[oo[v] + k];
                g.adj[cur[a]++] = b;
                g.adj[cur[b]++] = a;
            }
    for (e = 0; e < ne; ++e) qsort(g.adj + g.off[e], (size_t)(g.off[e + 1] - g.off[e]), sizeof(int), cmp_int);
    free(cur);
    free(oo);
    free(owners);
    return g;
}

/* Indexed binary heap over vertices with explicit key updates.  key(v) is a
 * triple compared lexicographically; `maxfirst` selects max- or min-heap. */
typedef struct {
    int n, size, maxfirst;
    int* heap;  /* heap position -> vertex */
    int* pos;   /* vertex -> heap position, -1 if absent */
    int (*key)[3];
} iheap;

static int ih_less(const iheap* h, int a, int b) { /* true if a should be above b */
    int i;
    for (i = 0; i < 3; ++i) {
        if (h->key[a][i] != h->key[b][i])
            return h->maxfirst ? h->key[a][i] > h->key[b][i] : h->key[a][i] < h->key[b][i];
    }
    return 0;
}
static void ih_swap(iheap* h, int i, int j) {
    int t = h->heap[i];
    h->heap[i] = h->heap[j];
    h->heap[j] = t;
    h->pos[h->heap[i]] = i;
    h->pos[h->heap[j]] = j;
}
static void ih_up(iheap* h, int i) {
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!ih_less(h, h->heap[i], h->heap[p])) break;
        ih_swap(h, i, p);
        i = p;
    }
}
static void ih_down(iheap* h, int i) {
    while (1) {
        int l = 2 * i + 1, r = l + 1, b = i;
        if (l < h->size && ih_less(h, h->heap[l], h->heap[b])) b = l;
        if (r < h->size && ih_less(h, h->heap[r], h->heap[b])) b = r;
        if (b == i) break;
        ih_swap(h, i, b);
        i = b;
    }
}
static void ih_init(iheap* h, int n, int maxfirst, int (*key)[3]) {
    int v;
    h->n = n;
    h->size = n;
    h->maxfirst = maxfirst;
    h->heap = (int*)malloc(sizeof(int) * (n ? n : 1));
    h->pos = (int*)malloc(sizeof(int) * (n ? n : 1));
    h->key = key;
    for (v = 0; v < n; ++v) {
        h->heap[v] = v;
        h->pos[v] = v;
    }
    for (v = n / 2 - 1; v >= 0; --v) ih_down(h, v);
}
static int ih_pop(iheap* h) {
    const int top = h->heap[0];
    ih_swap(h, 0, h->size - 1);
    h->size--;
    h->pos[top] = -1;
    if (h->size > 0) ih_down(h, 0);
    return top;
}
static void ih_update(iheap* h, int v) {
    const int i = h->pos[v];
    if (i < 0) return;
    ih_up(h, i);
    ih_down(h, h->pos[v]);
}
static void ih_free(iheap* h) {
    free(h->heap);
    free(h->pos);
}

/* DSAT, Graph.h:321-371 with maxSat ordering :897-903: pick the uncoloured
 * vertex with the largest (saturation, remaining degree, id); colour = first
 * colour unused by coloured neighbours; every neighbour visit (repeats
 * included) adds the colour to its saturation set and decrements its
 * remaining degree. */
static int dsat(const mgraph* g, int* colors) {
    const int n = g->n;
    int (*key)[3] = malloc(sizeof(int[3]) * (n ? n : 1));
    uint64_t(*satbits)[4] = calloc((size_t)(n ? n : 1), sizeof(uint64_t[4]));
    char* used = calloc((size_t)n + 2, 1);
    iheap h;
    int v, maxc = -1;
    int64_t i;
    for (v = 0; v < n; ++v) {
        colors[v] = -1;
        key[v][0] = 0;
        key[v][1] = (int)(g->off[v + 1] - g->off[v]);
        key[v][2] = v;
    }
    ih_init(&h, n, 1, key);
    while (h.size > 0) {
        int c = 0;
        v = ih_pop(&h);
        for (i = g->off[v]; i < g->off[v + 1]; ++i)
            if (colors[g->adj[i]] != -1) used[colors[g->adj[i]]] = 1;
        while (used[c]) ++c;
        for (i = g->off[v]; i < g->off[v + 1]; ++i)
            if (colors[g->adj[i]] != -1) used[colors[g->adj[i]]] = 0;
        if (c >= 256) { /* oracle limit: saturation bitset of 256 colours */
            maxc = -2;
            break;
        }
        colors[v] = c;
        if (c > maxc) maxc = c;
        for (i = g->off[v]; i < g->off[v + 1]; ++i) {
            const int u = g->adj[i];
            if (colors[u] != -1) continue;
            if (!(satbits[u][c >> 6] & (1ull << (c & 63)))) {
                satbits[u][c >> 6] |= 1ull << (c & 63);
                key[u][0]++;
            }
            key[u][1]--;
            ih_update(&h, u);
        }
    }
    ih_free(&h);
    free(key);
    free(satbits);
    free(used);
    return maxc < -1 ? -1 : maxc + 1;
}

/* GraphReduction::ColorReduction (Graph.h:941-969): vertices with colour >=
 * target, in ascending id, take the least frequent colour (< target) among
 * their neighbours (first on ties). */
static void color_reduction(const mgraph* g, int* colors, int target) {
    int* occ = (int*)malloc(sizeof(int) * target);
    int v, c;
    int64_t i;
    for (v = 0; v < g->n; ++v) {
        int best = 0;
        if (colors[v] < target) continue;
        memset(occ, 0, sizeof(int) * target);
        for (i = g->off[v]; i < g->off[v + 1]; ++i)
            if (colors[g->adj[i]] < target) occ[colors[g->adj[i]]]++;
        for (c = 1; c < target; ++c)
            if (occ[c] < occ[best]) best = c;
        colors[v] = best;
    }
    free(occ);
}

int orc_color_dsat(int element, const int* topo, int ne, int nv, int target, int* colors) {
    mgraph g = build_primal(element, topo, ne);
    int nc;
    (void)nv;
    nc = dsat(&g, colors);
    if (target > 0 && nc > 0) {
        color_reduction(&g, colors, target);
        nc = target;
    }
    mgraph_free(&g);
    return nc;
}

/* Greedy_SLF (Graph.h:237-277): repeatedly remove the vertex of smallest
 * (remaining degree, id); colour greedily in reverse removal order with the
 * first free colour (find_color_greedy, :909-930). */
int orc_color_slf(int element, const int* topo, int ne, int nv, int* colors) {
    mgraph g = build_dual(element, topo, ne);
    const int n = g.n;
    int (*key)[3] = malloc(sizeof(int[3]) * (n ? n : 1));
    int* order = (int*)malloc(sizeof(int) * (n ? n : 1));
    char* removed = (char*)calloc((size_t)n + 1, 1);
    char* used;
    iheap h;
    int v, i, nc = 0;
    int64_t k;
    (void)nv;
    for (v = 0; v < n; ++v) {
        key[v][0] = (int)(g.off[v + 1] - g.off[v]);
        key[v][1] = v;
        key[v][2] = 0;
    }
    ih_init(&h, n, 0, key);
    for (i = n - 1; i >= 0; --i) {
        v = ih_pop(&h);
        removed[v] = 1;
        order[i] = v;
        for (k = g.off[v]; k < g.off[v + 1]; ++k) {
            const int u = g.adj[k];
            if (!removed[u]) {
                key[u][0]--;
                ih_update(&h, u);
            }
        }
    }
    used = (char*)calloc((size_t)n + 2, 1);
    for (v = 0; v < n; ++v) colors[v] = -1;
    for (i = 0; i < n; ++i) {
        int c = 0;
        v = order[i];
        for (k = g.off[v]; k < g.off[v + 1]; ++k)
            if (colors[g.adj[k]] != -1) used[colors[g.adj[k]]] = 1;
        while (c < nc && used[c]) ++c;
        for (k = g.off[v]; k < g.off[v + 1]; ++k)
            if (colors[g.adj[k]] != -1) used[colors[g.adj[k]]] = 0;
        colors[v] = c;
        if (c == nc) nc++;
    }
    ih_free(&h);
    free(key);
    free(order);
    free(removed);
    free(used);
    mgraph_free(&g);
    return nc;
}

int64_t orc_primal_conflicts(int element, const int* topo, int ne, int nv, const int* colors) {
    mgraph g = build_primal(element, topo, ne);
    int64_t i, c = 0;
    int v;
    (void)nv;
    for (v = 0; v < g.n; ++v)
        for (i = g.off[v]; i < g.off[v + 1]; ++i)
            if (colors[g.adj[i]] == colors[v]) c++;
    mgraph_free(&g);
    return c;
}

/* ====================================================================== */
/* FEM constants and masses (GPU_FEM.cu:176-214, FEM_Shape.cpp:86-121)   */
/* ====================================================================== */

/* 3x3 helpers, column-major m[c*3 + r] */
static double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) + m[6] * (m[1] * m[5] - m[4] * m[2]);
}
static void inv3(const double* m, double* o) {
    /* Gauss-Jordan with partial pivoting on the row-major copy */
    double a[3][6];
    int r, c, k;
    for (r = 0; r < 3; ++r)
        for (c = 0; c < 3; ++c) {
            a[r][c] = m[c * 3 + r];
            a[r][3 + c] = (r == c);
        }
    for (c = 0; c < 3; ++c) {
        int p = c;
        double t;
        for (r = c + 1; r < 3; ++r)
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        for (k = 0; k < 6; ++k) {
            t = a[c][k];
            a[c][k] = a[p][k];
            a[p][k] = t;
        }
        t = a[c][c];
        for (k = 0; k < 6; ++k) a[c][k] /= t;
        for (r = 0; r < 3; ++r)
            if (r != c) {
                t = a[r][c];
                for (k = 0; k < 6; ++k) a[r][k] -= t * a[c][k];
            }
    }
    for (r = 0; r < 3; ++r)
        for (c = 0; c < 3; ++c) o[c * 3 + r] = a[r][3 + c];
}
static void mul3(const double* A, const double* B, double* C) {
    int r, c;
    for (c = 0; c < 3; ++c)
        for (r = 0; r < 3; ++r) C[c * 3 + r] = A[r] * B[c * 3] + A[3 + r] * B[c * 3 + 1] + A[6 + r] * B[c * 3 + 2];
}
static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
/* mat3x3_com (CUMatrix.cu:28-30): columns (F1 x F2, F2 x F0, F0 x F1) */
static void com3(const double* F, double* C) {
    cross3(F + 3, F + 6, C);
    cross3(F + 6, F, C + 3);
    cross3(F, F + 3, C + 6);
}

static void shape_dN_table(int element, double* dN /* Q*S*3 */, double* w) {
    const int S = orc_nodes(element), Q = orc_quadrature(element);
    double qp[27 * 3], N[27];
    int q;
    orc_quadrature_rule(element, qp, w);
    for (q = 0; q < Q; ++q) orc_shape_eval(element, qp + 3 * q, N, dN + (size_t)q * S * 3);
}

int orc_fem_constants(int element, const float* xyz, int nv, const int* topo, int ne, double* jx_inv, double* V) {
    const int S = orc_nodes(element), Q = orc_quadrature(element);
    double dN[27 * 27 * 3], w[27];
    int e, q, k, r, c;
    (void)nv;
    if (S < 0) return -1;
    shape_dN_table(element, dN, w);
    for (e = 0; e < ne; ++e)
        for (q = 0; q < Q; ++q) {
            double J[9] = {0};
            for (k = 0; k < S; ++k) {
                const float* X = xyz + 3 * (size_t)topo[e * S + k];
                const double* d = dN + ((size_t)q * S + k) * 3;
                for (c = 0; c < 3; ++c)
                    for (r = 0; r < 3; ++r) J[c * 3 + r] += (double)X[r] * d[c];
            }
            if (V) V[(size_t)e * Q + q] = fabs(det3(J)) * w[q];
            if (jx_inv) inv3(J, jx_inv + ((size_t)e * Q + q) * 9);
        }
    return 0;
}

int orc_masses(int element, const float* xyz, int nv, const int* topo, int ne, double density, int femshape,
               double* masses) {
    const int S = orc_nodes(element), Q = orc_quadrature(element);
    double* V = (double*)malloc(sizeof(double) * (size_t)ne * Q);
    int e, q, i;
    double tot = 0;
    orc_fem_constants(element, xyz, nv, topo, ne, NULL, V);
    for (i = 0; i < nv; ++i) masses[i] = 0.0;
    for (e = 0; e < ne; ++e) {
        double ve = 0;
        for (q = 0; q < Q; ++q) ve += V[(size_t)e * Q + q];
        tot += ve;
        if (femshape)
            for (i = 0; i < S; ++i) masses[topo[e * S + i]] += density / S * ve;
    }
    if (!femshape)
        for (i = 0; i < nv; ++i) masses[i] = tot * density / nv;
    free(V);
    return 0;
}

/* ====================================================================== */
/* materials                                                              */
/* ====================================================================== */

/* vec_hat (CUMatrix.cu:20-26) as stored by the reference: columns
 * (0,-v.z,v.y), (v.z,0,-v.x), (-v.y,v.x,0) */
static void hat_ref(const double* v, double* M) {
    M[0] = 0;     M[1] = -v[2]; M[2] = v[1];
    M[3] = v[2];  M[4] = 0;     M[5] = -v[0];
    M[6] = -v[1]; M[7] = v[0];  M[8] = 0;
}

/* PK1 of the GPU VBD/explicit materials (GPU_FEM_Material.cu:25-36) and the
 * per-vertex Hessian block assembled from the six symmetric d2W/dF2 blocks
 * (snh_hessian_sym / nh_hessian_sym :167-207, assemble_sub_hessian_sym
 * :250-261), with dF/dx = g and weight V. */
void orc_vbd_material(int material, double lambda, double mu, const double F[9], const double g[3], double V,
                      double P[9], double H[9]) {
    double C[9], B[6][9], hat[9], s, J, alpha;
    int i, j, k, a, b;
    com3(F, C);
    J = det3(F);
    alpha = (material == ORC_STABLE_NEOHOOKE) ? 1.0 + mu / lambda : 1.0;
    s = lambda * (J - alpha);
    for (i = 0; i < 9; ++i) P[i] = s * C[i] + mu * F[i];
    /* blocks k: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2) */
    for (k = 0; k < 6; ++k)
        for (i = 0; i < 9; ++i) B[k][i] = 0.0;
    hat_ref(F + 6, hat);
    for (i = 0; i < 9; ++i) B[1][i] = hat[i] * s;
    hat_ref(F + 3, hat);
    for (i = 0; i < 9; ++i) B[2][i] = -hat[i] * s;
    hat_ref(F, hat);
    for (i = 0; i < 9; ++i) B[4][i] = hat[i] * s;
    for (i = 0; i < 3; ++i) {
        B[0][i * 4] += mu;
        B[3][i * 4] += mu;
        B[5][i * 4] += mu;
    }
    k = 0;
    for (i = 0; i < 3; ++i)
        for (j = i; j < 3; ++j, ++k)
            /* outerProduct(comF[j], comF[i]): element (r, c) = com_j[r] com_i[c] */
            for (a = 0; a < 3; ++a)
                for (b = 0; b < 3; ++b) B[k][b * 3 + a] += C[j * 3 + a] * C[i * 3 + b] * lambda;
    for (i = 0; i < 9; ++i) H[i] = 0.0;
    for (a = 0; a < 3; ++a)
        for (b = 0; b < 3; ++b) {
            const int ab = b * 3 + a, ba = a * 3 + b;
            H[ab] += g[1] * g[0] * (B[1][ab] + B[1][ba]);
            H[ab] += g[2] * g[0] * (B[2][ab] + B[2][ba]);
            H[ab] += g[2] * g[1] * (B[4][ab] + B[4][ba]);
            H[ab] += g[0] * g[0] * B[0][ab];
            H[ab] += g[1] * g[1] * B[3][ab];
            H[ab] += g[2] * g[2] * B[5][ab];
        }
    for (i = 0; i < 9; ++i) H[i] *= V;
}

/* XPBD energy split (GPU_PBD_FEM_Materials.cu:31-66,97-118): constraint m = 0
 * (stiffness lambda) and m = 1 (stiffness mu). */
void orc_xpbd_material(int material, int m, double lambda, double mu, const double F[9], double P[9], double* e) {
    double C[9], J, n2 = 0;
    int i;
    com3(F, C);
    J = det3(F);
    for (i = 0; i < 9; ++i) n2 += F[i] * F[i];
    if (material == ORC_STABLE_NEOHOOKE) {
        if (m == 0) {
            *e = (J - 1) * (J - 1);
            for (i = 0; i < 9; ++i) P[i] = 2 * (J - 1) * C[i];
        } else {
            *e = n2 - 3 - 2 * (J - 1);
            for (i = 0; i < 9; ++i) P[i] = 2 * F[i] - 2 * C[i];
        }
    } else {
        const double alpha = 1.0 + mu / lambda;
        if (m == 0) {
            *e = (J - alpha) * (J - alpha);
            for (i = 0; i < 9; ++i) P[i] = 2 * (J - alpha) * C[i];
        } else {
            *e = n2 - 3;
            for (i = 0; i < 9; ++i) P[i] = 2 * F[i];
        }
    }
}

/* ====================================================================== */
/* scenes                                                                 */
/* ====================================================================== */

struct orc_scene {
    int element, S, Q, N, E, material, solver, iterations, xpbd_version;
    int* topo;
    double *dN, *wq, *jxi, *V;
    double lambda, mu, damping, global_damping, g[3];
    double *p, *last_p, *init_p, *v, *f, *y, *m, *w;
    int* mask;
    int ncolor;
    int* colors;      /* vertex colours (VBD) or element colours (XPBD) */
    int* order;       /* colour order override */
    int64_t* inc_off; /* vertex -> incident (eid, ref) */
    int* inc;
    int nfix;
    int* fix;
    double com[3];
    int has_box;
    double pmin[3], pmax[3];
    /* VBD options: Chebyshev radius, Jacobi colour sweep (v4) */
    double rho;
    int jacobi;
    double *prev1, *prev2, *pnew;
};

orc_scene* orc_scene_create(int element, int nv, const float* xyz, int ne, const int* topo, int material, double young,
                            double poisson, double density, int femshape, int solver, int iterations, double damping,
                            double global_damping, int xpbd_version, const double gravity[3]) {
    orc_scene* s;
    int i, e, k;
    if (orc_nodes(element) < 0) return NULL;
    s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->element = element;
    s->S = orc_nodes(element);
    s->Q = orc_quadrature(element);
    s->N = nv;
    s->E = ne;
    s->material = material;
    s->solver = solver;
    s->iterations = iterations;
    s->xpbd_version = xpbd_version;
    s->damping = damping;
    s->global_damping = global_damping;
    for (k = 0; k < 3; ++k) s->g[k] = gravity[k];
    /* GPU_FEM.cu:162-163 */
    s->lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
    s->mu = young / (2.0 * (1.0 + poisson));
    s->topo = (int*)malloc(sizeof(int) * (size_t)s->S * ne);
    memcpy(s->topo, topo, sizeof(int) * (size_t)s->S * ne);
    s->dN = (double*)malloc(sizeof(double) * s->Q * s->S * 3);
    s->wq = (double*)malloc(sizeof(double) * s->Q);
    shape_dN_table(element, s->dN, s->wq);
    s->jxi = (double*)malloc(sizeof(double) * 9 * (size_t)ne * s->Q);
    s->V = (double*)malloc(sizeof(double) * (size_t)ne * s->Q);
    orc_fem_constants(element, xyz, nv, topo, ne, s->jxi, s->V);
#define ALLOC3(x) x = (double*)calloc((size_t)3 * nv, sizeof(double))
    ALLOC3(s->p);
    ALLOC3(s->last_p);
    ALLOC3(s->init_p);
    ALLOC3(s->v);
    ALLOC3(s->f);
    ALLOC3(s->y);
#undef ALLOC3
    s->m = (double*)malloc(sizeof(double) * nv);
    s->w = (double*)malloc(sizeof(double) * nv);
    s->mask = (int*)malloc(sizeof(int) * nv);
    orc_masses(element, xyz, nv, topo, ne, density, femshape, s->m);
    for (i = 0; i < nv; ++i) {
        for (k = 0; k < 3; ++k) s->p[3 * i + k] = s->last_p[3 * i + k] = s->init_p[3 * i + k] = s->y[3 * i + k] = xyz[3 * i + k];
        s->w[i] = 1.0 / s->m[i];
        s->mask[i] = 1; /* GPU_ParticleSystem.cu:21 */
    }
    /* incident (element, local node) lists in ascending element id
     * (GPU_VBD_FEM.cu:677-690) */
    s->inc_off = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    for (i = 0; i < s->S * ne; ++i) s->inc_off[topo[i] + 1]++;
    for (i = 0; i < nv; ++i) s->inc_off[i + 1] += s->inc_off[i];
    s->inc = (int*)malloc(sizeof(int) * 2 * (size_t)s->S * ne);
    {
        int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * nv);
        memcpy(cur, s->inc_off, sizeof(int64_t) * nv);
        for (e = 0; e < ne; ++e)
            for (k = 0; k < s->S; ++k) {
                const int64_t at = cur[topo[e * s->S + k]]++;
                s->inc[2 * at] = e;
                s->inc[2 * at + 1] = k;
            }
        free(cur);
    }
    if (solver == ORC_VBD) {
        s->colors = (int*)malloc(sizeof(int) * nv);
        s->ncolor = orc_color_dsat(element, topo, ne, nv, -1, s->colors);
    } else if (solver == ORC_XPBD) {
        s->colors = (int*)malloc(sizeof(int) * ne);
        s->ncolor = orc_color_slf(element, topo, ne, nv, s->colors);
    }
    return s;
}

void orc_scene_destroy(orc_scene* s) {
    if (!s) return;
    free(s->topo); free(s->dN); free(s->wq); free(s->jxi); free(s->V);
    free(s->p); free(s->last_p); free(s->init_p); free(s->v); free(s->f); free(s->y);
    free(s->m); free(s->w); free(s->mask); free(s->colors); free(s->order);
    free(s->inc_off); free(s->inc); free(s->fix);
    free(s->prev1); free(s->prev2); free(s->pnew);
    free(s);
}

/* GPU_VBD(ps, it, rho) and GPU_VBD_FEM(..., version, target_color)
 * (include/GPU/VBD/GPU_VBD.h:15, include/GPU/VBD/GPU_VBD_FEM.h:62-65):
 * rho > 0 turns on the Chebyshev acceleration; jacobi = 1 is the v4 sweep
 * (kernel_vbd_solve_v4 + kernel_copy_in_position, GPU_VBD_FEM.cu:345-413);
 * target_color > 0 recolours with ColorReduction (Graph.h:941-969). */
int orc_scene_vbd_options(orc_scene* s, double rho, int jacobi, int target_color) {
    int i;
    if (s->solver != ORC_VBD) return -1;
    s->rho = rho;
    s->jacobi = jacobi;
    if (target_color > 0) {
        s->ncolor = orc_color_dsat(s->element, s->topo, s->E, s->N, target_color, s->colors);
        free(s->order);
        s->order = NULL;
    }
    free(s->prev1); free(s->prev2); free(s->pnew);
    s->prev1 = (double*)malloc(sizeof(double) * 3 * (size_t)s->N);
    s->prev2 = (double*)malloc(sizeof(double) * 3 * (size_t)s->N);
    s->pnew = (double*)malloc(sizeof(double) * 3 * (size_t)s->N);
    for (i = 0; i < 3 * s->N; ++i) s->prev1[i] = s->prev2[i] = s->pnew[i] = s->p[i];
    return s->ncolor;
}

/* GPU_Fix_Constraint ctor (GPU_Constraint.cu:94-107): com = mean of the fixed
 * initial positions. */
int orc_scene_fix(orc_scene* s, const int* ids, int n) {
    int i, k;
    free(s->fix);
    s->fix = (int*)malloc(sizeof(int) * (n ? n : 1));
    memcpy(s->fix, ids, sizeof(int) * n);
    s->nfix = n;
    for (k = 0; k < 3; ++k) s->com[k] = 0;
    for (i = 0; i < n; ++i)
        for (k = 0; k < 3; ++k) s->com[k] += s->init_p[3 * ids[i] + k];
    for (k = 0; k < 3; ++k) s->com[k] = n ? s->com[k] / n : 0;
    return 0;
}

void orc_scene_box(orc_scene* s, const double pmin[3], const double pmax[3]) {
    int k;
    s->has_box = 1;
    for (k = 0; k < 3; ++k) {
        s->pmin[k] = pmin[k];
        s->pmax[k] = pmax[k];
    }
}

void orc_scene_color_order(orc_scene* s, const int* order) {
    free(s->order);
    s->order = NULL;
    if (order) {
        s->order = (int*)malloc(sizeof(int) * s->ncolor);
        memcpy(s->order, order, sizeof(int) * s->ncolor);
    }
}

/* kernel_constraint_fix (GPU_Constraint.cu:8-19) then kernel_constraint_box_limit (:70-81) */
static void apply_constraints(orc_scene* s) {
    int i, k;
    for (i = 0; i < s->nfix; ++i) {
        const int id = s->fix[i];
        for (k = 0; k < 3; ++k) {
            s->p[3 * id + k] = s->com[k] + (s->init_p[3 * id + k] - s->com[k]);
            s->v[3 * id + k] = 0;
            s->f[3 * id + k] = 0;
        }
        s->mask[id] = 0;
    }
    if (s->has_box)
        for (i = 0; i < s->N; ++i) {
            if (s->mask[i] == 0) continue;
            for (k = 0; k < 3; ++k) {
                double x = s->p[3 * i + k];
                if (x < s->pmin[k]) x = s->pmin[k];
                if (x > s->pmax[k]) x = s->pmax[k];
                s->p[3 * i + k] = x;
            }
        }
}

/* F and dF/dx of element e at quadrature q, node r */
static void kinematics(const orc_scene* s, int e, int q, int r, double F[9], double g[3]) {
    const int S = s->S;
    const double* dN = s->dN + (size_t)q * S * 3;
    const double* Ji = s->jxi + ((size_t)e * s->Q + q) * 9;
    double Jx[9] = {0};
    int k, a, c;
    for (k = 0; k < S; ++k) {
        const double* x = s->p + 3 * (size_t)s->topo[e * S + k];
        for (c = 0; c < 3; ++c)
            for (a = 0; a < 3; ++a) Jx[c * 3 + a] += x[a] * dN[3 * k + c];
    }
    mul3(Jx, Ji, F);
    if (g)
        for (c = 0; c < 3; ++c) g[c] = Ji[c * 3] * dN[3 * r] + Ji[c * 3 + 1] * dN[3 * r + 1] + Ji[c * 3 + 2] * dN[3 * r + 2];
}

/* force and Hessian block of vertex v (compute_f_H_sym, GPU_VBD_FEM.cu:42-64) */
static void vertex_fH(const orc_scene* s, int v, double f[3], double H[9]) {
    int64_t i;
    int q, k;
    for (k = 0; k < 3; ++k) f[k] = 0;
    for (k = 0; k < 9; ++k) H[k] = 0;
    for (i = s->inc_off[v]; i < s->inc_off[v + 1]; ++i) {
        const int e = s->inc[2 * i], r = s->inc[2 * i + 1];
        for (q = 0; q < s->Q; ++q) {
            double F[9], g[3], P[9], Hq[9];
            const double V = s->V[(size_t)e * s->Q + q];
            kinematics(s, e, q, r, F, g);
            orc_vbd_material(s->material, s->lambda, s->mu, F, g, V, P, Hq);
            for (k = 0; k < 3; ++k) f[k] -= (P[k] * g[0] + P[3 + k] * g[1] + P[6 + k] * g[2]) * V;
            for (k = 0; k < 9; ++k) H[k] += Hq[k];
        }
    }
}

/* compute_correction (GPU_VBD_FEM.cu:107-122): writes p + dx to out */
static void vbd_solve_vertex(orc_scene* s, int v, double dt, double* out) {
    double f[3], H[9], Hi[9], d;
    const double kd = s->damping / dt, mh2 = s->m[v] / (dt * dt);
    int k;
    vertex_fH(s, v, f, H);
    if (s->damping != 0.0) {
        const double u[3] = {s->p[3 * v] - s->last_p[3 * v], s->p[3 * v + 1] - s->last_p[3 * v + 1],
                             s->p[3 * v + 2] - s->last_p[3 * v + 2]};
        for (k = 0; k < 3; ++k) f[k] -= kd * (H[k] * u[0] + H[3 + k] * u[1] + H[6 + k] * u[2]);
        for (k = 0; k < 9; ++k) H[k] += kd * H[k];
    }
    for (k = 0; k < 3; ++k) f[k] -= mh2 * (s->p[3 * v + k] - s->y[3 * v + k]);
    H[0] += mh2;
    H[4] += mh2;
    H[8] += mh2;
    d = fabs(det3(H));
    for (k = 0; k < 3; ++k) out[k] = s->p[3 * v + k];
    if (!(d > 1e-6)) return;
    inv3(H, Hi);
    for (k = 0; k < 3; ++k) out[k] += Hi[k] * f[0] + Hi[3 + k] * f[1] + Hi[6 + k] * f[2];
}

/* GPU_VBD::step (GPU_VBD.cu:38-61) with GPU_VBD_FEM::step (GPU_VBD_FEM.cu:712-756)
 * in a fixed colour order; vertices of one colour are independent. */
static void vbd_substep(orc_scene* s, double dt) {
    int i, k, j, c;
    const int cheb = s->rho > 1.1754943508222875e-38 * 1e6 && s->prev1; /* rho > eps (Core/Base.h:11) */
    float omega = 1.f;
    /* kernel_integration (GPU_VBD.cu:6-18) */
    for (i = 0; i < s->N; ++i) {
        if (cheb)
            for (k = 0; k < 3; ++k) s->prev1[3 * i + k] = s->p[3 * i + k];
        for (k = 0; k < 3; ++k) {
            const double a = s->g[k] + s->f[3 * i + k] * s->w[i];
            const double y = s->p[3 * i + k] + ((s->p[3 * i + k] - s->last_p[3 * i + k]) / dt + a * dt) * dt;
            s->y[3 * i + k] = y;
            s->last_p[3 * i + k] = s->p[3 * i + k];
            s->p[3 * i + k] = y;
            s->f[3 * i + k] = 0;
        }
    }
    for (j = 0; j < s->iterations; ++j) {
        for (c = 0; c < s->ncolor; ++c) {
            const int col = s->order ? s->order[c] : c;
            if (s->jacobi) {
                /* every vertex of the colour from the same positions, then copied */
                for (i = 0; i < s->N; ++i)
                    if (s->colors[i] == col && s->mask[i] != 0) vbd_solve_vertex(s, i, dt, s->pnew + 3 * (size_t)i);
                for (i = 0; i < s->N; ++i)
                    if (s->colors[i] == col && s->mask[i] != 0)
                        for (k = 0; k < 3; ++k) s->p[3 * i + k] = s->pnew[3 * i + k];
            } else {
                double out[3];
                for (i = 0; i < s->N; ++i) {
                    if (s->colors[i] != col || s->mask[i] == 0) continue;
                    vbd_solve_vertex(s, i, dt, out);
                    for (k = 0; k < 3; ++k) s->p[3 * i + k] = out[k];
                }
            }
        }
        apply_constraints(s);
        /* kernel_chebychev_acceleration (GPU_VBD.cu:27-35), omega in float (:53-55) */
        if (cheb) {
            const float r = (float)s->rho;
            if (j == 1) omega = 2.f / (2.f - r * r);
            else if (j > 1) omega = 4.f / (4.f - r * r * omega);
            for (i = 0; i < 3 * s->N; ++i) {
                if (j >= 2) s->p[i] = s->prev2[i] + (double)omega * (s->p[i] - s->prev2[i]);
                s->prev2[i] = s->prev1[i];
                s->prev1[i] = s->p[i];
            }
        }
    }
    /* kernel_velocity_update (GPU_VBD.cu:20-25) */
    for (i = 0; i < 3 * s->N; ++i) s->v[i] = (s->p[i] - s->last_p[i]) / dt;
}

/* Per-element diagnostics (GPU_FEM.cu:20-115 with Jx rebuilt at every
 * quadrature point and JX_inv indexed per point):
 * stress = sum_q vonMises(P F^T / det F) V_q (von_mises_stress,
 * GPU_FEM_Material.cu:219-230), volume = sum_q |det Jx_q| w_q, volume_diff =
 * volume - sum_q V_q, inverted = sum_q det(F_q) w_q <= -1e-6. */
void orc_scene_diagnostics(orc_scene* s, double* stress, double* volume, double* vdiff, double* inverted) {
    int e, q, a, b, k;
    for (e = 0; e < s->E; ++e) {
        double st = 0, vo = 0, vd = 0, sv = 0;
        for (q = 0; q < s->Q; ++q) {
            double F[9], P[9], H[9], sg[9], g[3] = {0, 0, 0}, J, w;
            const double V = s->V[(size_t)e * s->Q + q];
            kinematics(s, e, q, 0, F, g);
            orc_vbd_material(s->material, s->lambda, s->mu, F, g, V, P, H);
            J = det3(F);
            for (a = 0; a < 3; ++a)       /* column */
                for (b = 0; b < 3; ++b) { /* row */
                    double t = 0;
                    for (k = 0; k < 3; ++k) t += P[k * 3 + b] * F[k * 3 + a];
                    sg[a * 3 + b] = t / J;
                }
            {
                const double d01 = sg[0] - sg[4], d12 = sg[4] - sg[8], d02 = sg[0] - sg[8];
                st += sqrt(0.5 * (d01 * d01 + d12 * d12 + d02 * d02) +
                           3.0 * (sg[1] * sg[1] + sg[5] * sg[5] + sg[2] * sg[2])) * V;
            }
            w = s->wq[q];
            /* |det Jx| w = |det F| |det J_X| w = |det F| V */
            vo += fabs(J) * V;
            vd += fabs(J) * V - V;
            sv += J * w;
        }
        if (stress) stress[e] = st;
        if (volume) volume[e] = vo;
        if (vdiff) vdiff[e] = vd;
        if (inverted) inverted[e] = sv > -1e-6 ? 0.0 : 1.0;
    }
}

/* kernel_vbd_compute_residual (GPU_VBD_FEM.cu:482-542): free vertices get
 * f - (k/dt) H (p - last_p) - (m/dt^2)(p - y); fixed vertices 0. */
void orc_scene_residual(orc_scene* s, double dt, double* r) {
    int v, k;
    for (v = 0; v < s->N; ++v) {
        double f[3], H[9];
        const double kd = s->damping / dt, mh2 = s->m[v] / (dt * dt);
        for (k = 0; k < 3; ++k) r[3 * v + k] = 0;
        if (s->mask[v] == 0) continue;
        vertex_fH(s, v, f, H);
        if (s->damping != 0.0) {
            const double u[3] = {s->p[3 * v] - s->last_p[3 * v], s->p[3 * v + 1] - s->last_p[3 * v + 1],
                                 s->p[3 * v + 2] - s->last_p[3 * v + 2]};
            for (k = 0; k < 3; ++k) f[k] -= kd * (H[k] * u[0] + H[3 + k] * u[1] + H[6 + k] * u[2]);
        }
        for (k = 0; k < 3; ++k) r[3 * v + k] = f[k] - mh2 * (s->p[3 * v + k] - s->y[3 * v + k]);
    }
}

/* kenerl_semi_exicit_integration (GPU_Integrator.cu:27-34) */
static void semi_implicit(orc_scene* s, double dt) {
    int i, k;
    for (i = 0; i < s->N; ++i)
        for (k = 0; k < 3; ++k) {
            s->last_p[3 * i + k] = s->p[3 * i + k];
            s->v[3 * i + k] += (s->g[k] + s->f[3 * i + k] * s->w[i]) * dt;
            s->p[3 * i + k] += s->v[3 * i + k] * dt;
            s->f[3 * i + k] = 0;
        }
}

/* one element, both constraints: V0 (GPU_PBD_FEM.cu:13-93) or V1/V2 (:141-263) */
static void xpbd_element(orc_scene* s, int e, double dt, int v0) {
    const int S = s->S, Q = s->Q;
    const int* t = s->topo + (size_t)e * S;
    int m, q, k, a;
    for (m = 0; m < 2; ++m) {
        double grad[27][3], C = 0, sum = 0;
        const double stiff = m == 0 ? s->lambda : s->mu;
        for (k = 0; k < S; ++k) grad[k][0] = grad[k][1] = grad[k][2] = 0;
        for (q = 0; q < Q; ++q) {
            double F[9], P[9], Ps[9], en;
            const double* Ji = s->jxi + ((size_t)e * Q + q) * 9;
            const double V = s->V[(size_t)e * Q + q];
            const double* dN = s->dN + (size_t)q * S * 3;
            kinematics(s, e, q, 0, F, NULL);
            orc_xpbd_material(s->material, m, s->lambda, s->mu, F, P, &en);
            /* P * JX_inv^T * V */
            for (k = 0; k < 3; ++k)
                for (a = 0; a < 3; ++a)
                    Ps[k * 3 + a] = (P[a] * Ji[k] + P[3 + a] * Ji[3 + k] + P[6 + a] * Ji[6 + k]) * V;
            C += en * V;
            for (k = 0; k < S; ++k)
                for (a = 0; a < 3; ++a)
                    grad[k][a] += Ps[a] * dN[3 * k] + Ps[3 + a] * dN[3 * k + 1] + Ps[6 + a] * dN[3 * k + 2];
        }
        if (v0) {
            double Cinv, dl;
            C = fabs(C);
            C = sqrt(C);
            Cinv = C < 1e-12 ? 0.0 : 1.0 / (2.0 * C);
            for (k = 0; k < S; ++k) {
                for (a = 0; a < 3; ++a) grad[k][a] *= Cinv;
                sum += (grad[k][0] * grad[k][0] + grad[k][1] * grad[k][1] + grad[k][2] * grad[k][2]) * s->w[t[k]];
            }
            dl = -C / (sum + 1.0 / (stiff * dt * dt));
            for (k = 0; k < S; ++k) {
                if (s->mask[t[k]] == 0) continue;
                for (a = 0; a < 3; ++a) s->p[3 * t[k] + a] += dl * s->w[t[k]] * grad[k][a];
            }
        } else {
            double beta;
            for (k = 0; k < S; ++k)
                sum += (grad[k][0] * grad[k][0] + grad[k][1] * grad[k][1] + grad[k][2] * grad[k][2]) * s->w[t[k]];
            beta = C < 1e-12 ? 0.0 : -2.0 * C / (sum + 4.0 * C / (stiff * dt * dt));
            for (k = 0; k < S; ++k)
                for (a = 0; a < 3; ++a) s->p[3 * t[k] + a] += beta * s->w[t[k]] * grad[k][a];
        }
    }
}

/* GPU_PBD::step (GPU_PBD.cu:20-26) with GPU_PBD_FEM::step (GPU_PBD_FEM.cu:640-717) */
static void xpbd_substep(orc_scene* s, double dt) {
    const int v0 = (s->S == 4) || s->xpbd_version == 0;
    int c, e, i, k;
    semi_implicit(s, dt);
    if (v0 || s->xpbd_version != 3)
        for (c = 0; c < s->ncolor; ++c)
            for (e = 0; e < s->E; ++e)
                if (s->colors[e] == c) xpbd_element(s, e, dt, v0);
    apply_constraints(s);
    /* kernel_velocity_update (GPU_PBD.cu:8-18) */
    for (i = 0; i < s->N; ++i) {
        double* v = s->v + 3 * i;
        double nv;
        for (k = 0; k < 3; ++k) v[k] = (s->p[3 * i + k] - s->last_p[3 * i + k]) / dt;
        nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (nv > 1e-7) {
            const double coef = s->global_damping * dt * s->w[i];
            const double damp = -nv * (coef > 1.0 ? 1.0 : coef);
            for (k = 0; k < 3; ++k) v[k] += v[k] / nv * damp;
        }
    }
}

/* kernel_fem_eval_force (GPU_FEM_Explicit.cu:5-64): f_v = sum(-V P g) - k H_v v_v */
static void explicit_forces(orc_scene* s) {
    int v, k;
    for (v = 0; v < s->N; ++v) {
        double f[3], H[9];
        vertex_fH(s, v, f, H);
        if (s->damping != 0.0) {
            const double* u = s->v + 3 * v;
            for (k = 0; k < 3; ++k) f[k] -= s->damping * (H[k] * u[0] + H[3 + k] * u[1] + H[6 + k] * u[2]);
        }
        for (k = 0; k < 3; ++k) s->f[3 * v + k] = f[k];
    }
}

/* GPU_SemiExplicit::step (GPU_Integrator.cu:36-42) */
static void explicit_substep(orc_scene* s, double dt) {
    explicit_forces(s);
    semi_implicit(s, dt);
    apply_constraints(s);
}

/* Cuda_ParticleSystem_Dynamics::update (Cuda_ParticleSystem_Dynamic.cpp:17-31) */
void orc_scene_step(orc_scene* s, double dt, int substeps) {
    const double h = dt / substeps;
    int i;
    for (i = 0; i < substeps; ++i) {
        if (s->solver == ORC_VBD) vbd_substep(s, h);
        else if (s->solver == ORC_XPBD) xpbd_substep(s, h);
        else explicit_substep(s, h);
    }
}

void orc_scene_get(orc_scene* s, int what, double* out) {
    double* src[6] = {s->p, s->last_p, s->v, s->f, s->y, s->init_p};
    int i;
    if (what >= 0 && what < 6) memcpy(out, src[what], sizeof(double) * 3 * s->N);
    if (what == 6) memcpy(out, s->m, sizeof(double) * s->N);
    if (what == 7)
        for (i = 0; i < s->N; ++i) out[i] = s->mask[i];
}

void orc_scene_set(orc_scene* s, int what, const double* in) {
    double* dst[6] = {s->p, s->last_p, s->v, s->f, s->y, s->init_p};
    if (what >= 0 && what < 6) memcpy(dst[what], in, sizeof(double) * 3 * s->N);
}

int orc_scene_ncolors(orc_scene* s) { return s->ncolor; }
void orc_scene_colors(orc_scene* s, int* out) {
    memcpy(out, s->colors, sizeof(int) * (s->solver == ORC_XPBD ? s->E : s->N));
}
