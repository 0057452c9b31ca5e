/*
 * Reference CPU XPBD path, restated in plain C - TEST INFRASTRUCTURE ONLY.
 * Included twice by cpu_xpbd.c with REAL = float (the timed CPU baseline, the
 * reference computes in float) and REAL = double.
 *
 * Follows (octopus/):
 *   PBD_System::step                 src/Dynamic/PBD/PositionBasedDynamic.cpp:7-32
 *   step_constraint_gauss_seidel     src/Dynamic/PBD/PositionBasedDynamic.cpp:94-109
 *   update_velocity                  src/Dynamic/PBD/PositionBasedDynamic.cpp:76-87
 *   ParticleSystem::step_solver      src/Dynamic/Base/ParticleSystem.cpp:21-27
 *   EulerSemiExplicit::integrate     src/Dynamic/Base/Solver.cpp:12-18
 *   XPBD_Constraint::apply           src/Dynamic/PBD/XPBD_Constraint.cpp:4-28
 *   XPBD_FEM_Generic::project        src/Dynamic/PBD/XPBD_FEM_Generic.cpp:9-45
 *   FEM_Generic::build               src/Dynamic/FEM/FEM_Generic.cpp:95-105
 *   NeoHooke split                   src/Dynamic/PBD/PBD_ContinuousMaterial.cpp:56-74,
 *                                    include/Dynamic/PBD/PBD_ContinuousMaterial.h:50-73
 *   Developed Stable NeoHooke split  src/Dynamic/PBD/PBD_ContinuousMaterial.cpp:176-196
 *   RB_Fixation                      src/Dynamic/Base/Constraint.cpp:20-38
 *   build_dynamic (masses, order)    src/Script/Dynamic/FEM_Dynamic.cpp:183-226
 * Constraints are projected in insertion order: element 0 lambda term,
 * element 0 mu term, element 1 ...; fixed particles are not integrated but are
 * moved by the projections until RB_Fixation snaps them back.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUFFIX)
#define STRUCT CAT(orc_cpu_xpbd_, SUFFIX)

struct STRUCT {
    int N, E, S, Q, material, nb_step, nb_substep, nfix;
    int* topo;
    REAL *x, *last_x, *x0, *v, *f, *m, *w;
    char* active;
    REAL* dN;     /* Q*S*3 */
    REAL* JXi;    /* E*Q*9 col-major */
    REAL* V;      /* E*Q */
    REAL lam[2];  /* accumulated lambda, reset every substep: per constraint */
    REAL* lambda_acc; /* 2E */
    REAL lambda, mu, alpha, global_damping, g[3];
    int* fix;
    REAL com[3];
};

static REAL FN(det3_)(const REAL* m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) + m[6] * (m[1] * m[5] - m[4] * m[2]);
}

/* inverse by the adjugate (the reference uses glm::inverse) */
static void FN(inv3_)(const REAL* m, REAL* o) {
    const REAL d = FN(det3_)(m);
    const REAL id = (REAL)1 / d;
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

STRUCT* FN(orc_cpu_xpbd_create_)(int element, int nv, const float* xyz, int ne, const int* topo, int material,
                                 double young, double poisson, double density, int femshape, int iterations,
                                 int substeps, double global_damping, const int* fixed, int nfixed,
                                 const double gravity[3]) {
    STRUCT* s = (STRUCT*)calloc(1, sizeof(STRUCT));
    const int S = orc_nodes(element), Q = orc_quadrature(element);
    double qp[27 * 3], wq[27], Nv[27], dNd[27 * 3];
    double* masses;
    int i, e, q, k, r, c;
    s->N = nv;
    s->E = ne;
    s->S = S;
    s->Q = Q;
    s->material = material;
    s->nb_step = iterations;
    s->nb_substep = substeps;
    s->global_damping = (REAL)global_damping;
    for (k = 0; k < 3; ++k) s->g[k] = (REAL)gravity[k];
    {
        const REAL Y = (REAL)young, nu = (REAL)poisson;
        s->lambda = Y * nu / (((REAL)1 + nu) * ((REAL)1 - (REAL)2 * nu));
        s->mu = Y / ((REAL)2 * ((REAL)1 + nu));
        s->alpha = (REAL)1 + s->mu / s->lambda;
    }
    s->topo = (int*)malloc(sizeof(int) * (size_t)S * ne);
    memcpy(s->topo, topo, sizeof(int) * (size_t)S * ne);
    s->x = (REAL*)malloc(sizeof(REAL) * 3 * nv);
    s->last_x = (REAL*)malloc(sizeof(REAL) * 3 * nv);
    s->x0 = (REAL*)malloc(sizeof(REAL) * 3 * nv);
    s->v = (REAL*)calloc((size_t)3 * nv, sizeof(REAL));
    s->f = (REAL*)calloc((size_t)3 * nv, sizeof(REAL));
    s->m = (REAL*)malloc(sizeof(REAL) * nv);
    s->w = (REAL*)malloc(sizeof(REAL) * nv);
    s->active = (char*)malloc((size_t)nv);
    for (i = 0; i < 3 * nv; ++i) s->x[i] = s->last_x[i] = s->x0[i] = (REAL)xyz[i];
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
                const REAL dt = FN(det3_)(J);
                s->V[(size_t)e * Q + q] = (dt < 0 ? -dt : dt) * (REAL)wq[q];
            }
            FN(inv3_)(J, s->JXi + ((size_t)e * Q + q) * 9);
        }
    s->lambda_acc = (REAL*)calloc((size_t)2 * ne, sizeof(REAL));
    /* RB_Fixation::init: particles inactive, com = mean position */
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

/* energy and dE/dF of constraint m (stiffness lambda for m = 0, mu for m = 1) */
static void FN(material_)(const STRUCT* s, int m, const REAL* F, REAL* P, REAL* en) {
    REAL C[9], J, n2 = 0;
    int i;
    C[0] = F[4] * F[8] - F[5] * F[7];
    C[1] = F[5] * F[6] - F[3] * F[8];
    C[2] = F[3] * F[7] - F[4] * F[6];
    C[3] = F[7] * F[2] - F[8] * F[1];
    C[4] = F[8] * F[0] - F[6] * F[2];
    C[5] = F[6] * F[1] - F[7] * F[0];
    C[6] = F[1] * F[5] - F[2] * F[4];
    C[7] = F[2] * F[3] - F[0] * F[5];
    C[8] = F[0] * F[4] - F[1] * F[3];
    J = FN(det3_)(F);
    for (i = 0; i < 9; ++i) n2 += F[i] * F[i];
    if (s->material == ORC_STABLE_NEOHOOKE) {
        if (m == 0) {
            const REAL d = J - (REAL)1;
            *en = d * d;
            for (i = 0; i < 9; ++i) P[i] = (REAL)2 * d * C[i];
        } else {
            *en = n2 - (REAL)3 - (REAL)2 * (J - (REAL)1);
            for (i = 0; i < 9; ++i) P[i] = (REAL)2 * F[i] - (REAL)2 * C[i];
        }
    } else {
        if (m == 0) {
            const REAL d = J - s->alpha;
            *en = d * d;
            for (i = 0; i < 9; ++i) P[i] = (REAL)2 * d * C[i];
        } else {
            *en = n2 - (REAL)3;
            for (i = 0; i < 9; ++i) P[i] = (REAL)2 * F[i];
        }
    }
}

/* XPBD_Constraint::apply with XPBD_FEM_Generic::project, then the immediate
 * position update of step_constraint_gauss_seidel */
static void FN(apply_)(STRUCT* s, int e, int m, REAL dt) {
    const int S = s->S, Q = s->Q;
    const int* t = s->topo + (size_t)e * S;
    const REAL stiff = m == 0 ? s->lambda : s->mu;
    REAL grad[27][3], C = 0, sum = 0, alpha, dl, sg, Cinv;
    int q, k, a, c;
    const REAL epsv = (REAL)1.17549435e-38f * (REAL)1e6;
    if (stiff <= 0) return;
    for (k = 0; k < S; ++k) grad[k][0] = grad[k][1] = grad[k][2] = 0;
    for (q = 0; q < Q; ++q) {
        REAL Jx[9] = {0}, F[9], P[9], Ps[9], en;
        const REAL* Ji = s->JXi + ((size_t)e * Q + q) * 9;
        const REAL V = s->V[(size_t)e * Q + q];
        const REAL* dN = s->dN + (size_t)q * S * 3;
        for (k = 0; k < S; ++k) {
            const REAL* x = s->x + 3 * (size_t)t[k];
            for (c = 0; c < 3; ++c)
                for (a = 0; a < 3; ++a) Jx[c * 3 + a] += x[a] * dN[3 * k + c];
        }
        for (c = 0; c < 3; ++c)
            for (a = 0; a < 3; ++a)
                F[c * 3 + a] = Jx[a] * Ji[c * 3] + Jx[3 + a] * Ji[c * 3 + 1] + Jx[6 + a] * Ji[c * 3 + 2];
        FN(material_)(s, m, F, P, &en);
        for (c = 0; c < 3; ++c)
            for (a = 0; a < 3; ++a) Ps[c * 3 + a] = (P[a] * Ji[c] + P[3 + a] * Ji[3 + c] + P[6 + a] * Ji[6 + c]) * V;
        for (k = 0; k < S; ++k)
            for (a = 0; a < 3; ++a) grad[k][a] += Ps[a] * dN[3 * k] + Ps[3 + a] * dN[3 * k + 1] + Ps[6 + a] * dN[3 * k + 2];
        C += en * V;
    }
    if ((C < 0 ? -C : C) <= epsv) return;
    sg = C > 0 ? (REAL)1 : (REAL)-1;
    C = (REAL)sqrt((double)(C < 0 ? -C : C)) * sg;
    Cinv = (REAL)0.5 / C;
    for (k = 0; k < S; ++k)
        for (a = 0; a < 3; ++a) grad[k][a] *= Cinv;
    for (k = 0; k < S; ++k)
        sum += (grad[k][0] * grad[k][0] + grad[k][1] * grad[k][1] + grad[k][2] * grad[k][2]) * s->w[t[k]];
    alpha = (REAL)1 / (stiff * dt * dt);
    dl = -(C + alpha * s->lambda_acc[2 * e + m]) / (sum + alpha);
    s->lambda_acc[2 * e + m] += dl;
    for (k = 0; k < S; ++k)
        for (a = 0; a < 3; ++a) s->x[3 * t[k] + a] += dl * s->w[t[k]] * grad[k][a];
}

void FN(orc_cpu_xpbd_step_)(STRUCT* s, double dt_d) {
    const REAL h = (REAL)dt_d / (REAL)s->nb_substep;
    int sub, i, k, j, e;
    for (sub = 0; sub < s->nb_substep; ++sub) {
        /* step_solver */
        for (i = 0; i < s->N; ++i) {
            for (k = 0; k < 3; ++k) s->last_x[3 * i + k] = s->x[3 * i + k];
            if (!s->active[i]) continue;
            for (k = 0; k < 3; ++k) {
                s->v[3 * i + k] += (s->f[3 * i + k] * s->w[i] + s->g[k]) * h;
                s->x[3 * i + k] += s->v[3 * i + k] * h;
                s->f[3 * i + k] = 0;
            }
        }
        memset(s->lambda_acc, 0, sizeof(REAL) * 2 * (size_t)s->E);
        for (j = 0; j < s->nb_step; ++j)
            for (e = 0; e < s->E; ++e) {
                FN(apply_)(s, e, 0, h);
                FN(apply_)(s, e, 1, h);
            }
        /* RB_Fixation::apply */
        for (i = 0; i < s->nfix; ++i) {
            const int id = s->fix[i];
            for (k = 0; k < 3; ++k) {
                const REAL target = s->com[k] + (s->x0[3 * id + k] - s->com[k]);
                s->x[3 * id + k] += (target - s->x[3 * id + k]) * (REAL)1;
                s->v[3 * id + k] = 0;
                s->f[3 * id + k] = 0;
            }
        }
        /* update_velocity */
        for (i = 0; i < s->N; ++i) {
            REAL nv2 = 0, nv;
            if (!s->active[i]) continue;
            for (k = 0; k < 3; ++k) {
                s->v[3 * i + k] = (s->x[3 * i + k] - s->last_x[3 * i + k]) / h;
                nv2 += s->v[3 * i + k] * s->v[3 * i + k];
            }
            nv = (REAL)sqrt((double)nv2);
            if (nv > (REAL)1e-12) {
                REAL coef = s->global_damping * h * s->w[i];
                const REAL damp = -nv * (coef < (REAL)1 ? coef : (REAL)1);
                for (k = 0; k < 3; ++k) s->v[3 * i + k] += s->v[3 * i + k] / nv * damp;
            }
        }
    }
}

void FN(orc_cpu_xpbd_positions_)(STRUCT* s, double* out) {
    int i;
    for (i = 0; i < 3 * s->N; ++i) out[i] = (double)s->x[i];
}

void FN(orc_cpu_xpbd_destroy_)(STRUCT* s) {
    if (!s) return;
    free(s->topo); free(s->x); free(s->last_x); free(s->x0); free(s->v); free(s->f);
    free(s->m); free(s->w); free(s->active); free(s->dN); free(s->JXi); free(s->V);
    free(s->lambda_acc); free(s->fix);
    free(s);
}

#undef STRUCT
#undef FN
#undef CAT
#undef CAT_
