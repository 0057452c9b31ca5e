/*
 * FP64 CPU oracle of the GPU soft-body path - TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it to compare against the HIP engine.
 * Nothing in the product (octopusengine_amd/, liboe.so) links, loads or calls
 * it.  It restates the reference's algorithms in plain C, in double precision:
 * closed-form shape functions (the product inverts a Vandermonde matrix), the
 * reference's six-block Hessian assembly (the product uses a closed form), and
 * colourings.  The colourings share the product's data-structure design (CSR
 * multigraph, indexed heap); their independent pins are the literal
 * sorted-set restatement in tests/ref_literal.py and hand-traced graphs.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * octopus/ of ShortyBatou/OctopusEngine).  How each one is pinned is listed in
 * oracle.c next to it.
 */
#ifndef OE_ORACLE_H
#define OE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element ids as in the reference enum (include/Mesh/Elements.h:4-15) */
enum { ORC_TETRA = 3, ORC_HEXA = 6, ORC_TETRA10 = 7, ORC_TETRA20 = 8, ORC_HEXA27 = 9 };
enum { ORC_NEOHOOKE = 2, ORC_STABLE_NEOHOOKE = 3 };
enum { ORC_VBD = 0, ORC_XPBD = 1, ORC_EXPLICIT = 2 };

int orc_nodes(int element);
int orc_quadrature(int element);

/* shape functions and quadrature: N (S), dN (S*3) at x; rule points/weights */
int orc_shape_eval(int element, const double x[3], double* N, double* dN);
int orc_quadrature_rule(int element, double* qp /* Q*3 */, double* w /* Q */);

/* beam + conversions; pass NULL arrays to get the sizes */
int orc_beam(int element, int nx, int ny, int nz, float sx, float sy, float sz, int* nv, int* ne, float* xyz,
             int* topo);

/* colourings; return the colour count or -1 */
int orc_color_dsat(int element, const int* topo, int ne, int nv, int target_color, int* colors);
int orc_color_slf(int element, const int* topo, int ne, int nv, int* colors);
int64_t orc_primal_conflicts(int element, const int* topo, int ne, int nv, const int* colors);

/* per (element, quadrature): JX_inv (9, column-major) and V */
int orc_fem_constants(int element, const float* xyz, int nv, const int* topo, int ne, double* jx_inv, double* V);
int orc_masses(int element, const float* xyz, int nv, const int* topo, int ne, double density, int femshape,
               double* masses);

/* material: P (column-major) and the per-vertex 3x3 block assembled as the
 * reference does, for one F and one dF/dx */
void orc_vbd_material(int material, double lambda, double mu, const double F[9], const double g[3], double V,
                      double P[9], double H[9]);
void orc_xpbd_material(int material, int m, double lambda, double mu, const double F[9], double P[9], double* e);

/* scenes */
typedef struct orc_scene orc_scene;
orc_scene* orc_scene_create(int element, int nv, const float* xyz, int ne, const int* topo, int material, double young,
                            double poisson, double density, int femshape, int solver, int iterations, double damping,
                            double global_damping, int xpbd_version, const double gravity[3]);
void orc_scene_destroy(orc_scene* s);
int orc_scene_fix(orc_scene* s, const int* ids, int n);
void orc_scene_box(orc_scene* s, const double pmin[3], const double pmax[3]);
/* VBD options: Chebyshev radius (0 = off), Jacobi colour sweep (v4), colour
 * reduction target (> 0); returns the colour count */
int orc_scene_vbd_options(orc_scene* s, double rho, int jacobi, int target_color);
/* per-element stress (von Mises x V), volume, volume difference, inverted flag */
void orc_scene_diagnostics(orc_scene* s, double* stress, double* volume, double* vdiff, double* inverted);
/* residual force of every free vertex's local VBD problem (3N) */
void orc_scene_residual(orc_scene* s, double dt, double* r);
/* colour-order override (VBD): ncolor entries, NULL = ascending */
void orc_scene_color_order(orc_scene* s, const int* order);
void orc_scene_step(orc_scene* s, double dt, int substeps);
/* what: 0 p, 1 last_p, 2 v, 3 f, 4 y, 5 init_p (3N doubles); 6 masses (N) ; 7 mask (N, as double) */
void orc_scene_get(orc_scene* s, int what, double* out);
void orc_scene_set(orc_scene* s, int what, const double* in);
int orc_scene_ncolors(orc_scene* s);
void orc_scene_colors(orc_scene* s, int* out);

/* Reference CPU XPBD path (PBD_System + XPBD_FEM_Generic + RB_Fixation),
 * in float (the timed CPU baseline) and in double. */
typedef struct orc_cpu_xpbd_f orc_cpu_xpbd_f;
typedef struct orc_cpu_xpbd_d orc_cpu_xpbd_d;
orc_cpu_xpbd_f* orc_cpu_xpbd_create_f(int element, int nv, const float* xyz, int ne, const int* topo, int material,
                                      double young, double poisson, double density, int femshape, int iterations,
                                      int substeps, double global_damping, const int* fixed, int nfixed,
                                      const double gravity[3]);
void orc_cpu_xpbd_step_f(orc_cpu_xpbd_f* s, double dt);
void orc_cpu_xpbd_positions_f(orc_cpu_xpbd_f* s, double* out);
void orc_cpu_xpbd_destroy_f(orc_cpu_xpbd_f* s);
orc_cpu_xpbd_d* orc_cpu_xpbd_create_d(int element, int nv, const float* xyz, int ne, const int* topo, int material,
                                      double young, double poisson, double density, int femshape, int iterations,
                                      int substeps, double global_damping, const int* fixed, int nfixed,
                                      const double gravity[3]);
void orc_cpu_xpbd_step_d(orc_cpu_xpbd_d* s, double dt);
void orc_cpu_xpbd_positions_d(orc_cpu_xpbd_d* s, double* out);
void orc_cpu_xpbd_destroy_d(orc_cpu_xpbd_d* s);

/* Reference CPU VBD path (VertexBlockDescent + VBD_FEM + RB_Fixation,
 * sequential Gauss-Seidel in vertex id order), in float (the timed CPU
 * baseline) and in double.  Stable Neo-Hookean only (NULL otherwise). */
typedef struct orc_cpu_vbd_f orc_cpu_vbd_f;
typedef struct orc_cpu_vbd_d orc_cpu_vbd_d;
orc_cpu_vbd_f* orc_cpu_vbd_create_f(int element, int nv, const float* xyz, int ne, const int* topo, int material,
                                    double young, double poisson, double density, int femshape, int iterations,
                                    int substeps, double damping, double rho, const int* fixed, int nfixed,
                                    const double gravity[3]);
void orc_cpu_vbd_step_f(orc_cpu_vbd_f* s, double dt);
void orc_cpu_vbd_positions_f(orc_cpu_vbd_f* s, double* out);
void orc_cpu_vbd_destroy_f(orc_cpu_vbd_f* s);
orc_cpu_vbd_d* orc_cpu_vbd_create_d(int element, int nv, const float* xyz, int ne, const int* topo, int material,
                                    double young, double poisson, double density, int femshape, int iterations,
                                    int substeps, double damping, double rho, const int* fixed, int nfixed,
                                    const double gravity[3]);
void orc_cpu_vbd_step_d(orc_cpu_vbd_d* s, double dt);
void orc_cpu_vbd_positions_d(orc_cpu_vbd_d* s, double* out);
void orc_cpu_vbd_destroy_d(orc_cpu_vbd_d* s);

#ifdef __cplusplus
}
#endif

#endif
