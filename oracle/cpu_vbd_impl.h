/*
 * Reference CPU VBD path, restated in plain C - TEST INFRASTRUCTURE ONLY.
 * Included twice by cpu_vbd.c with REAL = float (the timed CPU baseline; the
 * reference computes in float) and REAL = double (its FP64 twin).
 *
 * Follows (octopus/):
 *   VertexBlockDescent::step            src/Dynamic/VBD/VertexBlockDescent.cpp:28-44
 *   chebyshev_acceleration, omega       src/Dynamic/VBD/VertexBlockDescent.cpp:12-26,
 *                                       include/Dynamic/VBD/VertexBlockDescent.h:14-18
 *   update_velocity                     src/Dynamic/VBD/VertexBlockDescent.cpp:46-51
 *   VBD_FEM::compute_inertia            src/Dynamic/VBD/VBD_FEM.cpp:13-18
 *   VBD_FEM::solve, solve_vertex        src/Dynamic/VBD/VBD_FEM.cpp:46-85
 *   VBD_FEM::solve_element              src/Dynamic/VBD/VBD_FEM.cpp:87-111
 *   VBD_FEM::assemble_hessian           src/Dynamic/VBD/VBD_FEM.cpp:241-254
 *   M_Stable_NeoHooke pk1, sub-Hessian  src/Dynamic/FEM/FEM_ContinuousMaterial.cpp:77-116
 *   Matrix::Hat, Matrix::Com            src/Core/Base.cpp:61-71
 *   ParticleSystem::step_solver         src/Dynamic/Base/ParticleSystem.cpp:21-27
 *   EulerSemiExplicit::integrate        src/Dynamic/Base/Solver.cpp:12-18
 *   RB_Fixation::init, apply            src/Dynamic/Base/Constraint.cpp:20-38
 *   VBD_FEM_Dynamic (masses, solver)    src/Script/Dynamic/VBD_FEM_Dynamic.cpp:11-35
 * Vertices are solved one after another in id order (sequential
 * Gauss-Seidel, no colouring).  Fixed vertices are inactive: neither
 * integrated nor solved; RB_Fixation puts them back on the rest pose after
 * every sweep.  VBD_FEM::solve's residual plot (plot_residual, a debug-UI
 * graph that leaves the state alone) is not restated.  Only Stable
 * Neo-Hookean has a sub-Hessian on this path (M_NeoHooke has none,
 * include/Dynamic/FEM/FEM_ContinuousMaterial.h:32-40), so create returns
 * NULL for any other material.
 * 3x3 matrices are column-major arrays m[c * 3 + r], as glm stores them.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUFFIX)
#define STRUCT CAT(orc_cpu_vbd_, SUFFIX)

struct STRUCT {
    int N, E, S, Q, iterations, substeps, nfix;
    int* topo;
    REAL *x, *last_x, *x0, *v, *f, *m, *w, *y;
    char* active;
    REAL* dN;      /* Q*S*3 */
    REAL* JXi;     /* E*Q*9 column-major */
    REAL* V;       /* E*Q */
    int *own_off, *own;   /* vertex -> incident (element, local node), element order */
    REAL lambda, mu, alpha, damping, rho, g[3];
    REAL *prev, *prev2;   /* Chebyshev history */
    int* fix;
    REAL com[3];
};

static REAL FN(vdet3_)(const REAL* m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) + m[6] * (m[1] * m[5] - m[4] * m[2]);
}

/* adjugate / determinant */
static void FN(vinv3_)(const REAL* m, REAL* o) {
    const REAL id = (REAL)1 / FN(vdet3_)(m);
    o[0] = (m[4] * m[8] - m[7] * m[5]) * id;
    o[3] = -(m[3] * m[8] - m[6] * m[5]) * id;
    o[6] = (m[3] * m[7] - m[6] * m[4]) * id;
    o[1] = -(m[1] * m[8] - m[7] * m[2]) * id;
    o[4] = (m[0] * m[8] - m[6] * m[2]) * id;
    o[7] = -(m[0] * m[7] - m[6] * m[1]) * id;
    o[2] = (m[1] * m[5] - m[4] * m[2]) * id;
    o[5] = -(m[0] * m[5] - m[3] * m[2]) * id;
    o[8] = (m[0] * m[4] - m[3] * m[1]) * id;
}

static void FN(vcross_)(const REAL* a, const REAL* b, REAL* o) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}

/* [v]x as a column-major matrix */
static void FN(vhat_)(const REAL* v, REAL s, REAL* h) {
    h[0] = 0;
    h[1] = v[2] * s;
    h[2] = -v[1] * s;
    h[3] = -v[2] * s;
    h[4] = 0;
    h[5] = v[0] * s;
    h[6] = v[1] * s;
    h[7] = -v[0] * s;
    h[8] = 0;
}

STRUCT* FN(orc_cpu_vbd_create_)(int element, int nv, const float* xyz, int ne, const int* topo, int material,
                                double young, double poisson, double density, int femshape, int iterations,
                                int substeps, double damping, double rho, const int* fixed, int nfixed,
                                const double gravity[3]) {
    STRUCT* s;
    const int S = orc_nodes(element), Q = orc_quadrature(element);
    double qp[27 * 3], wq[27], Nv[27], dNd[27 * 3];
    double* masses;
    int i, e, q, k, r, c;
    if (S < 0 || material != ORC_STABLE_NEOHOOKE) return NULL;
    s = (STRUCT*)calloc(1, sizeof(STRUCT));
    s->N = nv;
    s->E = ne;
    s->S = S;
    s->Q = Q;
    s->iterations = iterations;
    s->substeps = substeps;
    s->damping = (REAL)damping;
    s->rho = (REAL)rho;
    for (k = 0; k < 3; ++k) s->g[k] = (REAL)gravity[k];
    {
        /* ContinuousMaterial: lambda, mu; M_Stable_NeoHooke: alpha = 1 + mu / lambda */
        const REAL Y = (REAL)young, nu = (REAL)poisson;
        s->lambda = Y * nu / (((REAL)1 + nu) * ((REAL)1 - (REAL)2 * nu));
        s->mu = Y / ((REAL)2 * ((REAL)1 + nu));
        s->alpha = (REAL)1 + s->mu / s->lambda;
    }
    s->topo = (int*)malloc(sizeof(int) * (size_t)S * ne);
    memcpy(s->topo, topo, sizeof(int) * (size_t)S * ne);
    s->x = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)nv);
    s->last_x = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)nv);
    s->x0 = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)nv);
    s->y = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)nv);
    s->prev = (REAL*)calloc((size_t)3 * nv, sizeof(REAL));
    s->prev2 = (REAL*)calloc((size_t)3 * nv, sizeof(REAL));
    s->v = (REAL*)calloc((size_t)3 * nv, sizeof(REAL));
    s->f = (REAL*)calloc((size_t)3 * nv, sizeof(REAL));
    s->m = (REAL*)malloc(sizeof(REAL) * nv);
    s->w = (REAL*)malloc(sizeof(REAL) * nv);
    s->active = (char*)malloc((size_t)nv);
    for (i = 0; i < 3 * nv; ++i) s->x[i] = s->last_x[i] = s->x0[i] = s->y[i] = (REAL)xyz[i];
    masses = (double*)malloc(sizeof(double) * nv);
    orc_masses(element, xyz, nv, topo, ne, density, femshape, masses);
    for (i = 0; i < nv; ++i) {
        s->m[i] = (REAL)masses[i];
        s->w[i] = (REAL)1 / s->m[i];
        s->active[i] = 1;
    }
    free(masses);
    orc_quadrature_rule(element, qp, wq);
    s->dN = (REAL*)malloc(sizeof(REAL) * Q * S * 3);
    for (q = 0; q < Q; ++q) {
        orc_shape_eval(element, qp + 3 * q, Nv, dNd);
        for (k = 0; k < 3 * S; ++k) s->dN[q * S * 3 + k] = (REAL)dNd[k];
    }
    /* get_fem_const: J = sum_k X_k (x) dN_k, V = |det J| w, JX_inv = J^-1 */
    s->JXi = (REAL*)malloc(sizeof(REAL) * 9 * (size_t)ne * Q);
    s->V = (REAL*)malloc(sizeof(REAL) * (size_t)ne * Q);
    for (e = 0; e < ne; ++e)
        for (q = 0; q < Q; ++q) {
            REAL J[9] = {0};
            for (k = 0; k < S; ++k) {
                const REAL* X = s->x + 3 * (size_t)topo[e * S + k];
                const REAL* d = s->dN + (q * S + k) * 3;
                for (c = 0; c < 3; ++c)
                    for (r = 0; r < 3; ++r) J[c * 3 + r] += X[r] * d[c];
            }
            {
                const REAL dt = FN(vdet3_)(J);
                s->V[(size_t)e * Q + q] = (dt < 0 ? -dt : dt) * (REAL)wq[q];
            }
            FN(vinv3_)(J, s->JXi + ((size_t)e * Q + q) * 9);
        }
    /* VBD_FEM::build_neighboors: owners in element order */
    s->own_off = (int*)calloc((size_t)nv + 1, sizeof(int));
    for (i = 0; i < S * ne; ++i) s->own_off[topo[i] + 1]++;
    for (i = 0; i < nv; ++i) s->own_off[i + 1] += s->own_off[i];
    s->own = (int*)malloc(sizeof(int) * 2 * (size_t)S * ne);
    {
        int* cur = (int*)malloc(sizeof(int) * (nv ? nv : 1));
        memcpy(cur, s->own_off, sizeof(int) * nv);
        for (e = 0; e < ne; ++e)
            for (k = 0; k < S; ++k) {
                const int at = cur[topo[e * S + k]]++;
                s->own[2 * at] = e;
                s->own[2 * at + 1] = k;
            }
        free(cur);
    }
    /* RB_Fixation::init: the fixed particles become inactive, com = their mean */
    s->nfix = nfixed;
    s->fix = (int*)malloc(sizeof(int) * (nfixed ? nfixed : 1));
    memcpy(s->fix, fixed, sizeof(int) * nfixed);
    s->com[0] = s->com[1] = s->com[2] = 0;
    for (i = 0; i < nfixed; ++i) {
        s->active[fixed[i]] = 0;
        for (k = 0; k < 3; ++k) s->com[k] += s->x[3 * fixed[i] + k];
    }
    for (k = 0; k < 3; ++k) s->com[k] = nfixed ? s->com[k] / (REAL)nfixed : 0;
    return s;
}

/* VBD_FEM::solve_element: force and Hessian block of local node ref of element eid */
static void FN(vsolve_element_)(const STRUCT* s, int eid, int ref, REAL* fi, REAL* Hi) {
    const int S = s->S, Q = s->Q;
    const int* t = s->topo + (size_t)eid * S;
    int q, k, a, c, i, j, l;
    for (q = 0; q < Q; ++q) {
        const REAL* dN = s->dN + (size_t)q * S * 3;
        const REAL* Ji = s->JXi + ((size_t)eid * Q + q) * 9;
        const REAL V = s->V[(size_t)eid * Q + q];
        REAL Jx[9] = {0}, F[9], g[3], P[9], C[9], d2[9][9], H[9], J, sl;
        for (k = 0; k < S; ++k) {
            const REAL* x = s->x + 3 * (size_t)t[k];
            for (c = 0; c < 3; ++c)
                for (a = 0; a < 3; ++a) Jx[c * 3 + a] += x[a] * dN[3 * k + c];
        }
        /* F = Jx JX_inv */
        for (c = 0; c < 3; ++c)
            for (a = 0; a < 3; ++a)
                F[c * 3 + a] = Jx[a] * Ji[c * 3] + Jx[3 + a] * Ji[c * 3 + 1] + Jx[6 + a] * Ji[c * 3 + 2];
        /* dF/dx = JX_inv^T dN_ref */
        for (c = 0; c < 3; ++c)
            g[c] = Ji[c * 3] * dN[3 * ref] + Ji[c * 3 + 1] * dN[3 * ref + 1] + Ji[c * 3 + 2] * dN[3 * ref + 2];
        /* Com(F) = [F1 x F2, F2 x F0, F0 x F1]; P = lambda (J - alpha) Com F + mu F */
        FN(vcross_)(F + 3, F + 6, C);
        FN(vcross_)(F + 6, F, C + 3);
        FN(vcross_)(F, F + 3, C + 6);
        J = FN(vdet3_)(F);
        sl = s->lambda * (J - s->alpha);
        for (k = 0; k < 9; ++k) P[k] = sl * C[k] + s->mu * F[k];
        /* f -= P dF/dx V */
        for (a = 0; a < 3; ++a) fi[a] -= (P[a] * g[0] + P[3 + a] * g[1] + P[6 + a] * g[2]) * V;
        /* M_Stable_NeoHooke::get_sub_hessian: nine blocks */
        for (k = 0; k < 9; ++k)
            for (l = 0; l < 9; ++l) d2[k][l] = 0;
        FN(vhat_)(F + 6, sl, d2[1]);
        FN(vhat_)(F + 3, -sl, d2[2]);
        for (l = 0; l < 9; ++l) d2[3][l] = -d2[1][l];
        FN(vhat_)(F, sl, d2[5]);
        for (l = 0; l < 9; ++l) {
            d2[6][l] = -d2[2][l];
            d2[7][l] = -d2[5][l];
        }
        for (i = 0; i < 3; ++i) {
            d2[0][i * 3 + i] += s->mu;
            d2[4][i * 3 + i] += s->mu;
            d2[8][i * 3 + i] += s->mu;
        }
        /* d2[i*3+j] += outerProduct(Com[j], Com[i]) lambda: column m, row n = Com[j][n] Com[i][m] */
        for (i = 0; i < 3; ++i)
            for (j = 0; j < 3; ++j)
                for (a = 0; a < 3; ++a)
                    for (c = 0; c < 3; ++c) d2[i * 3 + j][a * 3 + c] += C[j * 3 + c] * C[i * 3 + a] * s->lambda;
        /* assemble_hessian: H[i][j] = dot(g, H_kl g), H_kl[k][l] = d2[k + 3 l][i][j] */
        for (j = 0; j < 3; ++j)
            for (i = 0; i < 3; ++i) {
                REAL Hg[3];
                for (l = 0; l < 3; ++l)
                    Hg[l] = d2[0 + l * 3][i * 3 + j] * g[0] + d2[1 + l * 3][i * 3 + j] * g[1] +
                            d2[2 + l * 3][i * 3 + j] * g[2];
                H[i * 3 + j] = g[0] * Hg[0] + g[1] * Hg[1] + g[2] * Hg[2];
            }
        for (k = 0; k < 9; ++k) Hi[k] += H[k] * V;
    }
}

/* VBD_FEM::solve_vertex */
static void FN(vsolve_vertex_)(STRUCT* s, int vid, REAL dt) {
    REAL fi[3] = {0, 0, 0}, Hi[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Hinv[9], kd, mh2, d, u[3];
    REAL* p = s->x + 3 * (size_t)vid;
    int i, k;
    for (i = s->own_off[vid]; i < s->own_off[vid + 1]; ++i) FN(vsolve_element_)(s, s->own[2 * i], s->own[2 * i + 1], fi, Hi);
    /* damping */
    kd = s->damping / dt;
    for (k = 0; k < 3; ++k) u[k] = p[k] - s->last_x[3 * (size_t)vid + k];
    for (k = 0; k < 3; ++k) fi[k] -= kd * (Hi[k] * u[0] + Hi[3 + k] * u[1] + Hi[6 + k] * u[2]);
    for (k = 0; k < 9; ++k) Hi[k] += kd * Hi[k];
    /* inertia */
    mh2 = s->m[vid] / (dt * dt);
    for (k = 0; k < 3; ++k) fi[k] -= mh2 * (p[k] - s->y[3 * (size_t)vid + k]);
    Hi[0] += mh2;
    Hi[4] += mh2;
    Hi[8] += mh2;
    d = FN(vdet3_)(Hi);
    if (d < 0) d = -d;
    if (!(d > (REAL)1.17549435e-38f * (REAL)1e6)) return;   /* eps, include/Core/Base.h:11 */
    FN(vinv3_)(Hi, Hinv);
    for (k = 0; k < 3; ++k) p[k] += Hinv[k] * fi[0] + Hinv[3 + k] * fi[1] + Hinv[6 + k] * fi[2];
}

void FN(orc_cpu_vbd_step_)(STRUCT* s, double dt_d) {
    const REAL h = (REAL)dt_d / (REAL)s->substeps;
    int sub, i, k, j;
    for (sub = 0; sub < s->substeps; ++sub) {
        REAL omega = 0;
        /* compute_inertia: y = p + v h + (f w + g) h^2 */
        for (i = 0; i < s->N; ++i)
            for (k = 0; k < 3; ++k)
                s->y[3 * i + k] = s->x[3 * i + k] + s->v[3 * i + k] * h +
                                  (s->f[3 * i + k] * s->w[i] + s->g[k]) * h * h;
        /* step_solver with EulerSemiExplicit(1) */
        for (i = 0; i < s->N; ++i) {
            for (k = 0; k < 3; ++k) s->last_x[3 * i + k] = s->x[3 * i + k];
            if (!s->active[i]) continue;
            for (k = 0; k < 3; ++k) {
                s->v[3 * i + k] += (s->f[3 * i + k] * s->w[i] + s->g[k]) * h;
                s->v[3 * i + k] *= (REAL)1;
                s->x[3 * i + k] += s->v[3 * i + k] * h;
                s->f[3 * i + k] = 0;
            }
        }
        for (j = 0; j < s->iterations; ++j) {
            for (i = 0; i < s->N; ++i)
                if (s->active[i]) FN(vsolve_vertex_)(s, i, h);
            /* RB_Fixation::apply (identity rotation, no offset, stiffness 1) */
            for (i = 0; i < s->nfix; ++i) {
                const int id = s->fix[i];
                for (k = 0; k < 3; ++k) {
                    const REAL target = s->com[k] + (s->x0[3 * id + k] - s->com[k]);
                    s->x[3 * id + k] += (target - s->x[3 * id + k]) * (REAL)1;
                    s->v[3 * id + k] *= 0;
                    s->f[3 * id + k] *= 0;
                }
            }
            /* chebyshev_acceleration with compute_omega */
            if (j == 0) omega = 1;
            else if (j == 1) omega = (REAL)2 / ((REAL)2 - s->rho * s->rho);
            else omega = (REAL)4 / ((REAL)4 - s->rho * s->rho * omega);
            for (i = 0; i < s->N; ++i) {
                if (!s->active[i]) continue;
                for (k = 0; k < 3; ++k) {
                    if (j >= 2) s->x[3 * i + k] = omega * (s->x[3 * i + k] - s->prev2[3 * i + k]) + s->prev2[3 * i + k];
                    s->prev2[3 * i + k] = s->prev[3 * i + k];
                    s->prev[3 * i + k] = s->x[3 * i + k];
                }
            }
        }
        /* update_velocity */
        for (i = 0; i < s->N; ++i) {
            if (!s->active[i]) continue;
            for (k = 0; k < 3; ++k) s->v[3 * i + k] = (s->x[3 * i + k] - s->last_x[3 * i + k]) / h;
        }
    }
}

void FN(orc_cpu_vbd_positions_)(STRUCT* s, double* out) {
    int i;
    for (i = 0; i < 3 * s->N; ++i) out[i] = (double)s->x[i];
}

void FN(orc_cpu_vbd_destroy_)(STRUCT* s) {
    if (!s) return;
    free(s->topo); free(s->x); free(s->last_x); free(s->x0); free(s->y); free(s->prev); free(s->prev2);
    free(s->v); free(s->f); free(s->m); free(s->w); free(s->active); free(s->dN); free(s->JXi); free(s->V);
    free(s->own_off); free(s->own); free(s->fix);
    free(s);
}

#undef STRUCT
#undef FN
#undef CAT
#undef CAT_
