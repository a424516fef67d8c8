/*
 * DRIFTER DG right-hand-side ORACLE -- test infrastructure only.
 *
 * This is a plain-C, FP64 restatement of the reference CPU algorithm
 * (olelindberg/drifter) for the DG right-hand-side path.  It exists to CHECK
 * the HIP implementation in drifter_amd/; it is never linked into, called by,
 * or used as a fallback for the product.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load liboracle.so.
 *
 * Every function cites the reference file:line it follows.  Layout
 * conventions are the reference's:
 *   - DOF index d = i + n*(j + n*k), xi fastest (include/core/types.hpp:83)
 *   - packed state U = [elem][Hu|Hv|eta|HT|HS][ndof]
 *     (src/solver/simulation_driver.cpp:199-213)
 *   - full per-element state (this oracle's flat form of PrimitiveState,
 *     include/physics/primitive_equations.hpp:82-103):
 *     [elem][Hu|Hv|eta|HT|HS|H|u|v|T|S|omega|rho][ndof]
 *
 * Pinning status of each function is listed in the header comment of the
 * file that implements it.
 */
#ifndef DRIFTER_ORACLE_H
#define DRIFTER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_N 9 /* order <= 8 (MAX_POLYNOMIAL_ORDER, include/core/types.hpp:131) */

enum { ORC_F_HU = 0, ORC_F_HV, ORC_F_ETA, ORC_F_HT, ORC_F_HS, ORC_NCONS = 5 };
enum {
    ORC_S_HU = 0, ORC_S_HV, ORC_S_ETA, ORC_S_HT, ORC_S_HS, ORC_S_H, ORC_S_U,
    ORC_S_V, ORC_S_T, ORC_S_S, ORC_S_OMEGA, ORC_S_RHO, ORC_NFULL = 12
};

/* OceanConstants + DiffusivityParams (primitive_equations.hpp:41-49,106-117). */
typedef struct {
    double g;         /* 9.81 */
    double nu_h;      /* 1.0 */
    double nu_v;      /* 1e-4 */
    double kappa_v_T; /* 1e-5 */
    double kappa_v_S; /* 1e-5 */
    double rho_0;     /* 1025 */
    double alpha;     /* 2e-4 */
    double beta;      /* 7.6e-4 */
} orc_physics;

void orc_default_physics(orc_physics *p);

/* ---- A1: 1D bases (src/dg/basis_hexahedron.cpp) ---- */
int orc_lgl_nodes(int n, double *nodes, double *weights);
int orc_gl_nodes(int n, double *nodes, double *weights);
void orc_barycentric_weights(int n, const double *nodes, double *bw);
void orc_derivative_matrix(int n, const double *nodes, double *D /* n*n row-major */);
/* Dense Kronecker D_xi, D_eta, D_zeta (ndof*ndof row-major each). */
void orc_kron_diff(int order, const double *D1, double *Dxi, double *Deta, double *Dzeta);
void orc_legendre(int n, double x, double *L, double *dL);

/* ---- A2-A8: PrimitiveEquationsSolver::compute_rhs on the packed layout ---- */
/* f: Coriolis parameter at every DOF (E*ndof), h: bathymetry (E*ndof). */
int orc_rhs(int order, int64_t n_elem, const double *U, const double *h, const double *f,
            const orc_physics *phys, double *dUdt);
/* Same arithmetic, element range [e0, e1) only (used for bounded CPU timing). */
int orc_rhs_range(int order, int64_t e0, int64_t e1, const double *U, const double *h,
                  const double *f, const orc_physics *phys, double *dUdt);
/* PrimitiveState::update_derived on the packed layout: writes H,u,v,T,S (E*ndof each). */
void orc_update_derived(int order, int64_t n_elem, const double *U, const double *h, double *H,
                        double *u, double *v, double *T, double *S);

/* ---- A9: time steppers (src/solver/time_stepper.cpp) ---- */
enum { ORC_EULER = 0, ORC_RK2 = 1, ORC_SSPRK3 = 2, ORC_RK4 = 3, ORC_LSRK45 = 4, ORC_DP45 = 5 };
int orc_rk_step(int scheme, int order, int64_t n_elem, double t, double dt, double *U,
                const double *h, const double *f, const orc_physics *phys);

/* ---- A10: compute_max_wave_speed + AdaptiveTimeController::compute_dt ---- */
double orc_max_wave_speed(int order, int64_t n_elem, const double *H, const double *u,
                          const double *v, int64_t elem_stride);
double orc_compute_dt(double max_wave_speed, double dx_min, double cfl, double dt_min,
                      double dt_max);

/* ---- A11-A12: ModeSplittingSolver::step on the full-state layout ---- */
int orc_mode_split_step(int order, int64_t n_elem, double dt_3d, int n_sub, double *full,
                        const double *h, const double *f, const orc_physics *phys);
/* element-level pieces (mode_splitting.cpp:129-221): u3, up have n^3
 * values in the reference's column layout idx = ih * n + k; ub, hub, HU, HV,
 * eta, H, f, div and the tendencies have n^2 */
int orc_ms_depth_average(int order, const double *u3, double *ub);
int orc_ms_depth_integrate(int order, const double *u3, const double *H, double *hub);
int orc_ms_deviation(int order, const double *u3, const double *ub, double *up);
int orc_ms_divergence_2d(int order, const double *HU, const double *HV, double *div);
int orc_ms_barotropic_rhs(int order, const double *eta, const double *HU, const double *HV,
                          const double *H, const double *f, double g, double *de, double *dU,
                          double *dV);

/* ---- A17: octree and face connectivity (src/mesh/octree_adapter.cpp) ---- */
/* FaceConnectionType order, include/dg/face_connection.hpp:15-46 */
enum {
    ORC_CONN_SAME = 0, ORC_CONN_FINE2X1 = 1, ORC_CONN_FINE1X2 = 2, ORC_CONN_FINE3_2P1 = 3,
    ORC_CONN_FINE3_1P2 = 4, ORC_CONN_FINE2X2 = 5, ORC_CONN_BOUNDARY = 6
};
typedef struct orc_mesh orc_mesh;
orc_mesh *orc_mesh_create(double xmin, double xmax, double ymin, double ymax, double zmin,
                          double zmax);
void orc_mesh_destroy(orc_mesh *m);
void orc_mesh_build_uniform(orc_mesh *m, int nx, int ny, int nz);
/* refine leaves elems[i] with masks[i] (bit0 X, bit1 Y, bit2 Z), then balance */
void orc_mesh_refine(orc_mesh *m, int64_t count, const int64_t *elems, const int *masks);
void orc_mesh_balance(orc_mesh *m);
int64_t orc_mesh_num_elements(const orc_mesh *m);
/* bounds E*6, levels E*3, morton E (any may be NULL) */
void orc_mesh_elements(const orc_mesh *m, double *bounds, int *levels, uint64_t *morton);
/* node count and every node (bounds n*6, levels n*3, split mask, child count,
 * element number or -1), in creation order */
int64_t orc_mesh_num_nodes(const orc_mesh *m);
void orc_mesh_nodes(const orc_mesh *m, double *bounds, int *levels, int *mask, int *nchild,
                    int64_t *leaf);
/* per (element, face): type, neighbour count, up to 4 neighbours with their
 * face ids and sub-face indices (unused slots -1) */
void orc_mesh_connections(const orc_mesh *m, uint8_t *type, int8_t *count, int64_t *elems,
                          int8_t *faces, int8_t *subs);
/* the 2:1 mortar connectivity (the descent stopped at the element's size) */
void orc_mesh_connections_2to1(const orc_mesh *m, uint8_t *type, int8_t *count, int64_t *elems,
                               int8_t *faces, int8_t *subs);

/* ---- A14-A16, A18: fluxes, conforming interfaces, mortars ---- */
enum {
    ORC_FLUX_SW_LF = 0,      /* factory shallow_water_lax_friedrichs, p = {g} */
    ORC_FLUX_SW_HLLC = 1,    /* ShallowWaterHLLCFlux, p = {g} */
    ORC_FLUX_SW_ROE = 2,     /* ShallowWaterRoeFlux, p = {g, entropy_fix} */
    ORC_FLUX_SW_CENTRAL = 3, /* CentralFlux with the shallow-water flux, p = {g} */
    ORC_FLUX_PE_LF5 = 4,     /* apply_interface_fluxes lambda, p = {g} */
    ORC_FLUX_UPWIND = 5,     /* UpwindFlux, p = {vx, vy, vz} */
    ORC_FLUX_LF_ADV = 6      /* LaxFriedrichsFlux for F(U) = v U, p = {vx, vy, vz, lambda} */
};
/* returns the number of variables written to out */
int orc_flux(int kind, const double *UL, const double *UR, const double *n, const double *p,
             double *out);
/* face DOF it + n jt -> volume DOF index (LGL face extraction) */
void orc_face_map(int order, int face, int *vol_index);
/* ConformingInterface::compute_flux for a scalar flux; returns #quad points */
int orc_conforming_flux(int order, int kind, const double *p, const double *UL,
                        const double *UR, const double *n, double *rhs_l, double *rhs_r);
/* MortarSpace operators (row-major): M, Minv, Pc, Lc, Pf[f], Lf[f]; returns #fine */
int orc_mortar_operators(int order, int type, double *M, double *Minv, double *Pc, double *Lc,
                         double *Pf, double *Lf);
/* MortarSpace::compute_mortar_flux (scalar); Uf / rhs_f are nfine*(N+1)^2 */
int orc_mortar_flux(int order, int type, int kind, const double *p, const double *Uc,
                    const double *Uf, const double *n, double *rhs_c, double *rhs_f);

/* ---- hydrostatic pressure and pressure gradient (pressure.c) ---- */
int orc_hydrostatic_pressure(int order, int64_t E, double g, const double *rho, const double *H,
                             double *p);
/* parts: 1 barotropic, 2 baroclinic, 3 both; method 0 standard, 1 density Jacobian */
int orc_pressure_gradient(int order, int64_t E, int method, int parts, double g, double rho0,
                          const double *eta, const double *rho, const double *H,
                          const double *sigma, const double *deta_dx, const double *deta_dy,
                          const double *dh_dx, const double *dh_dy, double *pg_x, double *pg_y);

/* ---- face term of the DG right-hand side (faces.c) ---- */
/* interface lists: conf [nc][4], mort [nm][6], groups [ng][5] (sizes up to
 * 6 E); returns -1 for a connection without a 2:1 mortar */
int orc_face_lists(int64_t E, const double *bounds, const int32_t *levels, const uint8_t *type,
                   const int8_t *count, const int64_t *elems, int64_t *conf, int64_t *nconf,
                   int64_t *mort, int64_t *nmort, int64_t *groups, int64_t *ngroups);
/* dU += M^-1 (sum of the lifted face vectors), packed layout */
int orc_face_term(int order, int64_t E, int64_t nconf, const int64_t *conf, int64_t nmort,
                  const int64_t *mort, const int *mort_type, int kind, const double *p,
                  const double *U, double *dU);

/* ---- A19: GradientErrorEstimator + mark_elements ---- */
/* GradientErrorEstimator::estimate for E elements of one nodal field
 * (element e at field + e*elem_stride), element bounds E*6 (xmin xmax ymin
 * ymax zmin zmax).  Outputs value[E], mask[E] (RefineMask bits), refine[E],
 * coarsen[E]. */
void orc_estimate(int order, int64_t E, const double *field, int64_t elem_stride,
                  const double *bounds, double refine_threshold, double coarsen_threshold,
                  double smoothness_factor, double *value, uint8_t *mask, uint8_t *refine,
                  uint8_t *coarsen);
/* AdaptiveMeshRefinement::mark_elements after the estimate: action[E]
 * (0 None, 1 Refine, 2 Coarsen) and mask_out[E]; levels E*3. */
void orc_mark_elements(int64_t E, const uint8_t *refine, const uint8_t *coarsen,
                       const uint8_t *mask, const int *levels, int max_level, int min_level,
                       uint8_t *action, uint8_t *mask_out);

/* ---- A20: nodal -> orthonormal Legendre modal coefficients ---- */
/* V[i][p] = sqrt((2p+1)/2) P_p(x_i) on the LGL nodes (n x n row-major) and
 * its inverse */
int orc_modal_matrices(int order, double *V, double *Vinv);
/* coef[e][p + n(q + n r)] = sum Vinv[p][i] Vinv[q][j] Vinv[r][k] u[e][i + n(j + n k)] */
int orc_modal_coeffs(int order, int64_t E, const double *u, int64_t elem_stride, double *coef);

/* ---- A13: sigma-coordinate metrics (SigmaMetrics::update) ---- */
/* deta_dt may be NULL (update_steady) */
void orc_sigma_metrics(int64_t npts, const double *sigma, const double *eta, const double *h,
                       const double *deta_dx, const double *deta_dy, const double *dh_dx,
                       const double *dh_dy, const double *deta_dt, double *H, double *Hinv,
                       double *dsdx, double *dsdy, double *dsdt);

/* ---- A21: refinement-indicator MLP (128 -> 64 -> 64 -> 1, GELU) ---- */
#define ORC_MLP_IN 128
#define ORC_MLP_HIDDEN 64
#define ORC_MLP_NPARAM 12481
/* BF16 round-to-nearest-even of (float)x, returned as a double */
double orc_bf16_round(double x);
/* mean squared error over B rows of X (B x 128) against y; pred[B] and
 * grad[12481] (d loss / d params) optional; quantize: see mlp.c */
double orc_mlp_loss_grad(const double *params, int64_t B, const double *X, const double *y,
                         int quantize, double *pred, double *grad);

double orc_verify_mortar_conservation(int nm, int nfine, const double *rc, const double *rf,
                                      const double *wc, const double *wf);

/* ---- AMR solution transfer (SolutionProjection, transfer_solution) ---- */
/* children: nchild x ndof, child index cx + nx (cy + ny cz); returns nchild */
int orc_project_to_children(int order, int mask, const double *U, double *children);
int orc_project_from_children(int order, int mask, const double *children, double *U);
/* parents ascending; returns the number of elements written to U_new */
int64_t orc_transfer_solution(int order, int64_t E, const double *U, int64_t n_ref,
                              const int64_t *parents, double *U_new);

/* ---- vertical velocity diagnosis (vertical_velocity.cpp) ---- */
int orc_diagnose_omega(int order, int64_t E, const double *div, const double *H, int64_t es,
                       double *omega);
int orc_diagnose_omega_from_velocity(int order, int64_t E, const double *dHu_dx,
                                     const double *dHv_dy, const double *H, int64_t es,
                                     double *omega);
void orc_omega_to_w(int64_t npts, const double *omega, const double *u, const double *v,
                    const double *H, const double *dz_dx, const double *dz_dy,
                    const double *dz_dt, double *w);
void orc_omega_to_w_simple(int64_t npts, const double *omega, const double *u, const double *v,
                           const double *eta, const double *h, const double *sigma,
                           const double *deta_dx, const double *deta_dy, const double *deta_dt,
                           double *w);
int orc_column_sigma_nodes(int order, double *sigma);
int orc_diagnose_column(int order, int64_t ncol, const double *div, const double *H,
                        const double *omega_bottom, double *omega);
void orc_continuity_residual(int order, int64_t E, const double *div, const double *dHomega,
                             int64_t es, double *out);

/* ---- SigmaCoordinateOperators (operators_3d.cpp:513-662) ---- */
int orc_sigma_ops_metrics(int order, int64_t E, const double *eta, const double *h,
                          const double *deta_dx, const double *deta_dy, const double *dh_dx,
                          const double *dh_dy, const double *deta_dt, int64_t es, double *H,
                          double *Hinv, double *sigma, double *dsdx, double *dsdy, double *dsdt);
int orc_sigma_ops_apply(int order, int op, int64_t E, const double *U, const double *a,
                        const double *b, const double *c, const double *f, const double *g,
                        const double *dsdx, const double *dsdy, const double *Hinv, int64_t es,
                        double *out1, double *out2);

#ifdef __cplusplus
}
#endif
#endif
