/*
 * drf.h -- C ABI of the MI355X (gfx950) DG right-hand side for DRIFTER.
 *
 * Plain pointers and sizes only: no HIP, torch or Eigen types appear in the
 * signatures.  Device pointers are passed as void*; a HIP stream as void*
 * (a hipStream_t; NULL is HIP's null stream).  Device (_dev) entry points are
 * asynchronous on that stream unless they return a host value; host-pointer
 * entry points use the context's own stream and return when done.
 *
 * Layouts (d = i + n(j + n k), xi fastest, include/core/types.hpp:83):
 *   DRF_LAYOUT_PACKED  U[e][field][d]   -- the reference's packed vector
 *                                          (src/solver/simulation_driver.cpp:199-213)
 *   DRF_LAYOUT_SOA     U[field][e][d]   -- structure of arrays
 *   DRF_LAYOUT_LINES   U[field][i][e][L] with d = i + n L -- structure of
 *                      arrays with the xi index outermost: the kernels'
 *                      native layout, fully coalesced loads and stores
 *   DRF_LAYOUT_BLOCKS  U[block][field][i][e % B][L], block = e / B, with
 *                      B = drf_layout_block_elements(order): the LINES
 *                      layout cut into blocks of B elements, so that a
 *                      workgroup's batch reads and writes one contiguous
 *                      region; the last block is allocated whole
 *                      (drf_state_len) and its slots past element E - 1
 *                      are neither read nor written.  The volume kernel
 *                      addresses a whole state in 32-bit byte offsets here
 *                      (under 4 GiB per state, DRF_ERR_UNSUPPORTED above).
 *   fields: 0 Hu, 1 Hv, 2 eta, 3 HT, 4 HS
 *   full state (mode splitting): [e][Hu|Hv|eta|HT|HS|H|u|v|T|S|omega|rho][d]
 *
 * Errors: every call returns a drf_status (0 = DRF_OK); the message of the
 * last failure on the calling thread is returned by drf_last_error().  The C++
 * facade (primitive_equations_gpu.hpp) turns non-zero statuses into
 * std::runtime_error, as the reference reports errors by exception.
 *
 * Threading: one context per host thread (or per stream); calls on one
 * context are stream-ordered and not re-entrant.
 */
#ifndef DRIFTER_HIP_DRF_H
#define DRIFTER_HIP_DRF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct drf_ctx drf_ctx;
typedef int drf_status;

#define DRF_OK 0
#define DRF_ERR_ARG 1         /* invalid argument (std::invalid_argument) */
#define DRF_ERR_HIP 2         /* HIP runtime failure */
#define DRF_ERR_STATE 3       /* call out of order, e.g. RHS before initialize */
#define DRF_ERR_UNSUPPORTED 4 /* order or option not built */
#define DRF_ERR_RANGE 5       /* index out of range (std::out_of_range) */

#define DRF_LAYOUT_PACKED 0
#define DRF_LAYOUT_SOA 1
#define DRF_LAYOUT_LINES 2
#define DRF_LAYOUT_BLOCKS 3

#define DRF_PREC_FP64 0  /* FP64 storage and arithmetic (parity path) */
#define DRF_PREC_MIXED 1 /* FP32 volume arithmetic and storage; FP64 column integrals */

/* Time-stepping schemes of src/solver/time_stepper.cpp */
#define DRF_RK_EULER 0  /* ForwardEuler        :12-25   */
#define DRF_RK_RK2 1    /* RK2Midpoint         :31-54   */
#define DRF_RK_SSPRK3 2 /* SSPRK3              :64-80   */
#define DRF_RK_RK4 3    /* RK4Classic          :86-119  */
#define DRF_RK_LSRK45 4 /* LSRK45              :126-143 */
#define DRF_RK_DP45 5   /* RK45DormandPrince   :149-201 */

/* OceanConstants + DiffusivityParams, include/physics/primitive_equations.hpp:41-49,106-117 */
typedef struct {
    double g;         /* 9.81     */
    double nu_h;      /* 1.0      */
    double nu_v;      /* 1e-4     */
    double kappa_v_T; /* 1e-5     */
    double kappa_v_S; /* 1e-5     */
    double rho_0;     /* 1025.0   */
    double alpha;     /* 2e-4     */
    double beta;      /* 7.6e-4   */
} drf_physics;

/* ------------------------------------------------------------------ */
/* Lifecycle                                                           */
/* ------------------------------------------------------------------ */

/* Replaces PrimitiveEquationsSolver(int order, const OceanConstants&)
 * (include/physics/primitive_equations.hpp:222).  precision: DRF_PREC_*. */
drf_status drf_create(int order, int precision, int device, drf_ctx **out);
void drf_destroy(drf_ctx *ctx);
const char *drf_last_error(void);
void drf_default_physics(drf_physics *p);
drf_status drf_set_physics(drf_ctx *ctx, const drf_physics *p);
/* 1D LGL nodes, weights and derivative matrix (n*n row-major) used by the
 * kernels (HexahedronBasis, src/dg/basis_hexahedron.cpp:257-264). */
drf_status drf_basis(int order, double *nodes, double *weights, double *D);

/* Replaces PrimitiveEquationsSolver::initialize
 * (include/physics/primitive_equations.hpp:229-231): bathymetry h at every
 * DOF (E*ndof, element-major); dh_dx/dh_dy are accepted and ignored as in
 * the reference (stored but unused by the RHS).  Coriolis: the beta-plane
 * f = f0 + beta (y - y0) (CoriolisParameter, :58-62) at the DOF positions
 * y_positions (E*ndof); y_positions may be NULL only when beta == 0. */
drf_status drf_initialize(drf_ctx *ctx, int64_t n_elem, const double *h, const double *dh_dx,
                          const double *dh_dy, double f0, double beta, double y0,
                          const double *y_positions);

/* ------------------------------------------------------------------ */
/* Right-hand side                                                     */
/* ------------------------------------------------------------------ */

/* RHSFunction drop-in (include/solver/time_stepper.hpp:10): host packed U
 * and dUdt of length E*5*ndof, as SimulationDriver::compute_rhs
 * (src/solver/simulation_driver.cpp:170-197) produces it. */
drf_status drf_rhs(drf_ctx *ctx, double t, const double *U, double *dUdt);

/* Elements per block of DRF_LAYOUT_BLOCKS at an order (no counterpart in the
 * reference): the volume kernel's batch at its widest workgroup, so that
 * every batch lies in one block.  0 for an order outside 1-8. */
int drf_layout_block_elements(int order);
/* Values of one state (5 fields) of the context's mesh in `layout`:
 * E*5*(N+1)^3, and for DRF_LAYOUT_BLOCKS whole blocks (E rounded up to a
 * multiple of the block).  0 before drf_initialize or for an unknown layout. */
int64_t drf_state_len(const drf_ctx *ctx, int layout);

/* Device-resident variant: U and dUdt are device pointers in `layout`,
 * element type double (FP64 context) or float (mixed context). */
drf_status drf_rhs_dev(drf_ctx *ctx, double t, const void *U, void *dUdt, int layout,
                       void *stream);

/* Launch shape of the volume kernel (no counterpart in the reference):
 * resident workgroups per CU (0 = the occupancy limit, the default), the
 * XCD-aware batch order (1, the default) and the workgroup size in grid lines
 * (128, the default, 256 or 512; 0 = 128; orders above 4 run at 128 whatever
 * is set: a wider workgroup stages more elements per batch and reads longer
 * contiguous runs of each field plane).  drf_autotune times
 * drf_rhs_dev on U for each shape (reps launches each, on `stream`), keeps
 * the fastest in the context and returns its time per launch.  The timed
 * launches cycle over U / dUdt and three temporary device copies of them, so
 * that the state comes from HBM rather than from the last-level cache; the
 * copies are allocated and freed by the call (on the caller's buffers alone
 * if they cannot be allocated; not for use under stream capture). */
drf_status drf_set_launch(drf_ctx *ctx, int wg_per_cu, int xcd_remap, int wg_lines);
drf_status drf_get_launch(const drf_ctx *ctx, int *wg_per_cu, int *xcd_remap, int *wg_lines);
drf_status drf_autotune(drf_ctx *ctx, const void *U, void *dUdt, int layout, int reps,
                        void *stream, float *best_us);

/* ------------------------------------------------------------------ */
/* Time stepping (TimeStepper::step, include/solver/time_stepper.hpp:18) */
/* ------------------------------------------------------------------ */

/* One step of `scheme` on the host packed state U (in/out). */
drf_status drf_rk_step(drf_ctx *ctx, int scheme, double t, double dt, double *U);
/* Device-resident state; fused != 0 fuses the stage updates into the RHS
 * kernel epilogue, fused == 0 runs the reference's sequence (RHS, then
 * separate vector updates). */
drf_status drf_rk_step_dev(drf_ctx *ctx, int scheme, double t, double dt, void *U, int layout,
                           int fused, void *stream);

/* SimulationDriver::compute_max_wave_speed (src/solver/simulation_driver.cpp:153-168)
 * evaluated on the derived state of U (PrimitiveState::update_derived). */
drf_status drf_max_wave_speed(drf_ctx *ctx, const double *U, double *c_max);
drf_status drf_max_wave_speed_dev(drf_ctx *ctx, const void *U, int layout, double *c_max,
                                  void *stream);
/* Same, on derived H, u, v arrays (device, E*ndof element-major each). */
drf_status drf_max_wave_speed_derived_dev(drf_ctx *ctx, const void *H, const void *u,
                                          const void *v, int64_t elem_stride, double *c_max,
                                          void *stream);
/* No counterpart in the reference: the maximum of drf_max_wave_speed_dev,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  Allocates nothing: the
 * reduction's partial maxima (at most 2,048 doubles) live in a file-scope
 * device array of the library, one per GPU, shared by every call on that GPU
 * (calls that overlap on one GPU race), and the maximum is written to the
 * caller's device double c_max_dev, asynchronously on `stream`.  FP64
 * contexts; bit-identical to drf_max_wave_speed_dev. */
drf_status drf_max_wave_speed_noalloc_dev(drf_ctx *ctx, const void *U, int layout,
                                          void *c_max_dev, void *stream);
/* AdaptiveTimeController::compute_dt (src/solver/time_stepper.cpp:261-268). */
double drf_compute_dt(double max_wave_speed, double dx_min, double cfl, double dt_min,
                      double dt_max);

/* ------------------------------------------------------------------ */
/* Mode splitting (ModeSplittingSolver::step, include/physics/mode_splitting.hpp:241) */
/* ------------------------------------------------------------------ */

/* full: E*12*ndof FP64, [e][Hu|Hv|eta|HT|HS|H|u|v|T|S|omega|rho][d], in/out. */
drf_status drf_mode_split_step(drf_ctx *ctx, double dt_3d, int n_sub, double *full);
drf_status drf_mode_split_step_dev(drf_ctx *ctx, double dt_3d, int n_sub, void *full,
                                   void *stream);

/* ------------------------------------------------------------------ */
/* HIP graphs of repeated steps (no counterpart in the reference: the   */
/* driver calls step() in a loop, simulation_driver.cpp:108-151)       */
/* ------------------------------------------------------------------ */

typedef struct drf_graph drf_graph;
/* Captures n_steps calls of drf_rk_step_dev(ctx, scheme, ., dt, U, layout,
 * fused) on the same state into one HIP graph.  The graph owns the
 * workspace its kernels use; it keeps U's address and the context's
 * coefficients, so both must outlive it.  Nothing is executed here. */
drf_status drf_graph_rk(drf_ctx *ctx, int scheme, double dt, void *U, int layout, int fused,
                        int n_steps, drf_graph **out);

/* SimulationDriver::step (simulation_driver.cpp:108-131) without a host
 * round trip: the maximum wave speed of U (compute_max_wave_speed, :153-168),
 * dt = AdaptiveTimeController::compute_dt(c, dx_min) (time_stepper.cpp:261-268)
 * written to dt_dev, time_dev += dt (time_ += dt_, :130; time_dev may be
 * NULL), then one fused step of `scheme` with that dt, which the stage
 * kernels read from dt_dev.  dt_dev, time_dev: device doubles.  FP64
 * contexts; fused schemes (not DP45).  Results equal drf_max_wave_speed_dev +
 * drf_compute_dt + drf_rk_step_dev(fused) bit for bit.  dx_min is the
 * driver's 2 / (N + 1) (:40-42). */
drf_status drf_rk_step_adaptive_dev(drf_ctx *ctx, int scheme, double cfl, double dx_min,
                                    double dt_min, double dt_max, void *U, int layout,
                                    void *dt_dev, void *time_dev, void *stream);
/* n_steps such steps captured as one graph (drf_graph_launch replays them;
 * the host reads time_dev once per replay to decide whether to go on, as
 * SimulationDriver::run_for's loop, :94-106).  The graph uses the context's
 * reduction buffer: do not run other calls on the context concurrently. */
drf_status drf_graph_rk_adaptive(drf_ctx *ctx, int scheme, double cfl, double dx_min,
                                 double dt_min, double dt_max, void *U, int layout, int n_steps,
                                 void *dt_dev, void *time_dev, drf_graph **out);
/* Same for n_steps mode-split steps (drf_mode_split_step_dev): with the
 * barotropic subcycles inside each step's kernel, one launch replays
 * n_steps x n_sub substeps. */
drf_status drf_graph_mode_split(drf_ctx *ctx, double dt_3d, int n_sub, void *full, int n_steps,
                                drf_graph **out);
/* Replays the captured steps on `stream` (asynchronous).  Every graph keeps
 * the device addresses of its context's buffers (coefficients, reduction
 * buffer): destroy it before drf_initialize or drf_destroy on that context. */
drf_status drf_graph_launch(drf_graph *g, void *stream);
void drf_graph_destroy(drf_graph *g);

/* ------------------------------------------------------------------ */
/* Octree mesh and face connectivity (OctreeAdapter,                   */
/* include/mesh/octree_adapter.hpp:103-273)                            */
/* ------------------------------------------------------------------ */

typedef struct drf_mesh drf_mesh;

/* FaceConnectionType, include/dg/face_connection.hpp:15-46 */
#define DRF_CONN_SAME_LEVEL 0
#define DRF_CONN_FINE_2X1 1
#define DRF_CONN_FINE_1X2 2
#define DRF_CONN_FINE_3_2PLUS1 3
#define DRF_CONN_FINE_3_1PLUS2 4
#define DRF_CONN_FINE_2X2 5
#define DRF_CONN_BOUNDARY 6

/* OctreeAdapter(xmin, xmax, ymin, ymax, zmin, zmax), octree_adapter.cpp:14-21 */
drf_status drf_mesh_create(double xmin, double xmax, double ymin, double ymax, double zmin,
                           double zmax, drf_mesh **out);
void drf_mesh_destroy(drf_mesh *mesh);
/* OctreeAdapter::build_uniform, octree_adapter.cpp:23-84 */
drf_status drf_mesh_build_uniform(drf_mesh *mesh, int nx, int ny, int nz);
/* OctreeAdapter::refine (+ balance), octree_adapter.cpp:184-203; masks are
 * RefineMask bits (1 X, 2 Y, 4 Z), include/mesh/refine_mask.hpp */
drf_status drf_mesh_refine(drf_mesh *mesh, int64_t count, const int64_t *elems,
                           const uint8_t *masks);
/* OctreeAdapter::balance, octree_adapter.cpp:138-182 */
drf_status drf_mesh_balance(drf_mesh *mesh);
/* OctreeAdapter::num_elements */
int64_t drf_mesh_num_elements(const drf_mesh *mesh);
/* element bounds (E*6: xmin xmax ymin ymax zmin zmax), directional levels
 * (E*3) and Morton codes (E) in element order; any output may be NULL */
drf_status drf_mesh_elements(const drf_mesh *mesh, double *bounds, int32_t *levels,
                             uint64_t *morton);
/* OctreeAdapter::element_bounds (octree_adapter.cpp:234-239); DRF_ERR_RANGE
 * for an invalid element (std::out_of_range in the reference) */
drf_status drf_mesh_element_bounds(const drf_mesh *mesh, int64_t elem, double *bounds6);
/* OctreeAdapter::build_face_connections (octree_adapter.cpp:369-389) as flat
 * arrays over (element, face): type[E*6], count[E*6], and up to four
 * neighbours elems/faces/subs[E*6*4] (FaceConnection::fine_elems,
 * fine_face_ids, subface_indices; unused slots -1) */
drf_status drf_mesh_face_connections(const drf_mesh *mesh, uint8_t *type, int8_t *count,
                                     int64_t *elems, int8_t *faces, int8_t *subs);
/* The same arrays with the neighbour search of find_neighbor_same_or_coarser
 * (octree_adapter.cpp:523-580) stopped at the first node of the element's
 * own size.  The reference descends to a leaf every time, so its
 * "refined neighbour" branch (get_neighbor :344-356, find_fine_neighbors
 * :582-639) is never taken and a coarse face next to finer leaves is
 * reported as SameLevel with the one leaf holding the probe point.  Here that
 * branch is taken: a refined same-size neighbour gives its fine leaves
 * sharing the face in breadth-first order (subface index = rank), typed by
 * get_connection_type (face_connection.hpp:110-152).  Not a parity entry
 * point: the 2:1 mortar faces of the face term (drf_face_plan_create) come
 * from it. */
drf_status drf_mesh_face_connections_2to1(const drf_mesh *mesh, uint8_t *type, int8_t *count,
                                          int64_t *elems, int8_t *faces, int8_t *subs);

/* ------------------------------------------------------------------ */
/* Numerical fluxes (NumericalFlux::flux, include/flux/numerical_flux.hpp:41)
 * over a batch of face points                                         */
/* ------------------------------------------------------------------ */

/* kind                      reference                                 params      vars */
#define DRF_FLUX_SW_LF 0      /* shallow_water_lax_friedrichs, numerical_flux.cpp:134-181; {g}       3 */
#define DRF_FLUX_SW_HLLC 1    /* ShallowWaterHLLCFlux, numerical_flux.hpp:110-131; {g}                3 */
#define DRF_FLUX_SW_ROE 2     /* ShallowWaterRoeFlux, numerical_flux.hpp:134-150; {g, entropy_fix}   3 */
#define DRF_FLUX_SW_CENTRAL 3 /* CentralFlux with the shallow-water flux, :78-89; {g}                3 */
#define DRF_FLUX_PE_LF5 4     /* apply_interface_fluxes, primitive_equations.cpp:452-492; {g}        5 */
#define DRF_FLUX_UPWIND 5     /* UpwindFlux, numerical_flux.hpp:92-107; {vx, vy, vz}                 1 */
#define DRF_FLUX_LF_ADV 6     /* LaxFriedrichsFlux of F(U) = v U, :57-75; {vx, vy, vz, lambda}       1 */

/* Replaces flux(U_L, U_R, n) evaluated point by point.  Device pointers,
 * structure of arrays: UL, UR, out are [var][npts], normals [3][npts].
 * params: host array of up to 4 doubles (table above). */
drf_status drf_flux_points(int kind, int64_t npts, const void *UL, const void *UR,
                           const void *normals, const double *params, void *out, void *stream);

/* Face DOF it + n jt of face 0..5 (-xi,+xi,-eta,+eta,-zeta,+zeta) <- volume
 * node vol_index[it + n jt] (LGL face extraction,
 * src/dg/basis_hexahedron.cpp:431-522).  Host call; vol_index holds (N+1)^2. */
drf_status drf_face_map(int order, int face, int32_t *vol_index);

/* Face traces of one nodal field for a batch of (element, face) pairs:
 * out[k][q] = field[elems[k] * elem_stride + vol_index(faces[k])[q]],
 * out is [nfaces][(N+1)^2].  elem_stride = (N+1)^3 for one field [e][d];
 * (number of fields) (N+1)^3 for a field of the packed layout.  An element
 * outside [0, n_elem) or a face outside 0..5 gives NaN for that face.  Device
 * pointers (elems int64, faces int8). */
drf_status drf_gather_face_traces(int order, int64_t n_elem, int64_t elem_stride, int64_t nfaces,
                                  const void *elems, const void *faces, const void *field,
                                  void *out, void *stream);

/* ConformingInterface::compute_flux (src/dg/mortar.cpp:320-343) for a batch
 * of faces with a scalar flux (DRF_FLUX_UPWIND or DRF_FLUX_LF_ADV).
 * UL, UR, rhs_l, rhs_r are [nfaces][(N+1)^2], normals [nfaces][3].  As in the
 * reference, the face rule has N LGL points per direction (FaceQuadrature
 * with `order` points, src/dg/quadrature_3d.cpp:13-31): face values
 * 0..N^2-1 enter with index i + N j; the other entries of rhs_l/rhs_r are 0. */
drf_status drf_conforming_flux_batch(int order, int kind, const double *params, int64_t nfaces,
                                     const void *UL, const void *UR, const void *normals,
                                     void *rhs_l, void *rhs_r, void *stream);

/* MortarSpace (include/dg/mortar.hpp:32-136): projection and lifting
 * operators of one connection type (DRF_CONN_SAME_LEVEL .. DRF_CONN_FINE_2X2)
 * built on the host at creation and kept in device memory (device < 0: host
 * copies only, for drf_mortar_operators; the flux batch then fails with
 * DRF_ERR_STATE). */
typedef struct drf_mortar drf_mortar;
drf_status drf_mortar_create(int order, int type, int device, drf_mortar **out);
void drf_mortar_destroy(drf_mortar *m);
int drf_mortar_num_fine(const drf_mortar *m);
int drf_mortar_num_dofs(const drf_mortar *m);
/* host copies of the operators, each (N+1)^2 x (N+1)^2 row-major (Pf, Lf:
 * num_fine of them); any output may be NULL */
drf_status drf_mortar_operators(const drf_mortar *m, double *M, double *Minv, double *Pc,
                                double *Lc, double *Pf, double *Lf);
/* MortarSpace::compute_mortar_flux (src/dg/mortar.cpp:212-258) for a batch
 * of interfaces with a scalar flux.  Uc, rhs_c [n][(N+1)^2]; Uf, rhs_f
 * [n][num_fine][(N+1)^2]; normals [n][3]. */
drf_status drf_mortar_flux_batch(drf_mortar *m, int kind, const double *params,
                                 int64_t n_interfaces, const void *Uc, const void *Uf,
                                 const void *normals, void *rhs_c, void *rhs_f, void *stream);
/* verify_mortar_conservation (src/dg/mortar.cpp:386-413) for one interface,
 * host: |sum_i w_coarse_i rhs_coarse_i + sum_f sum_i w_fine_fi rhs_fine_fi| /
 * (|coarse integral| + |fine integral|), 0 when that norm is below 1e-15.
 * rhs_fine and w_fine are [nfine][(N+1)^2]. */
drf_status drf_verify_mortar_conservation(const drf_mortar *m, const double *rhs_coarse,
                                          const double *rhs_fine, const double *w_coarse,
                                          const double *w_fine, double *rel_error);
/* The same for a batch (device; rhs_c / rhs_f as drf_mortar_flux_batch writes
 * them, the weights shared by every interface): rel_error[n_interfaces] */
drf_status drf_mortar_conservation_dev(const drf_mortar *m, int64_t n_interfaces,
                                       const void *rhs_c, const void *rhs_f, const void *w_coarse,
                                       const void *w_fine, void *rel_error, void *stream);

/* ------------------------------------------------------------------ */
/* Hydrostatic pressure and pressure gradient in sigma coordinates     */
/* (PressureGradientElement, include/physics/pressure_gradient.hpp:    */
/*  46-367; the reference header does not compile, :183)               */
/* ------------------------------------------------------------------ */

#define DRF_PG_STANDARD 0         /* baroclinic_gradient_standard, :252-294 */
#define DRF_PG_DENSITY_JACOBIAN 1 /* baroclinic_gradient_density_jacobian, :296-338 */

/* hydrostatic_pressure (:233-250) with the vertical direction on the
 * element's zeta index k: p(i,j,k) = sum_{k' >= k} (w_k'/2) g rho H, summed
 * from k up.  Device arrays, element e at e * elem_stride (>= (N+1)^3).
 * Orders 1-7. */
drf_status drf_hydrostatic_pressure_dev(int order, int64_t n_elem, double g, const void *rho,
                                        const void *H, int64_t elem_stride, void *p,
                                        void *stream);
/* full_gradient (:340-367): parts 1 = barotropic -g H grad(eta) (:213-231),
 * 2 = baroclinic of `method`, 3 = their sum.  Standard: dp/dx|_z = D_xi p +
 * (D_zeta p)(sigma dH/dx - deta/dx)/H, dH/dx = deta/dx + dh/dx, pg = -dp/dx|_z
 * / rho0 (deta_dx, deta_dy, dh_dx, dh_dy may be NULL: 0; the reference
 * leaves deta/dx at 0).  Density Jacobian: z = eta + sigma H, J_x = rho_xi
 * z_zeta - rho_zeta z_xi, pg = -(g/rho0) sum_{k' <= k} (w_k'/2) J.
 * Derivatives in reference coordinates (DG3DElementOperator::
 * gradient_reference, src/dg/operators_3d.cpp:58-69). */
drf_status drf_pressure_gradient_dev(int order, int64_t n_elem, int method, int parts, double g,
                                     double rho0, const void *eta, const void *rho, const void *H,
                                     const void *sigma, const void *deta_dx, const void *deta_dy,
                                     const void *dh_dx, const void *dh_dy, int64_t elem_stride,
                                     void *pg_x, void *pg_y, void *stream);

/* ------------------------------------------------------------------ */
/* Face term of the DG right-hand side on an octree mesh               */
/* (DG3DIntegration::compute_rhs face loop and inverse mass,           */
/*  src/dg/operators_3d.cpp:255-281)                                   */
/* ------------------------------------------------------------------ */

typedef struct drf_face_plan drf_face_plan;

/* Interface lists of a mesh, grouped for the face kernels.  From a face
 * connectivity (the arrays of drf_mesh_face_connections or
 * drf_mesh_face_connections_2to1, E*6 and E*6*4), the elements' bounds (E*6)
 * and directional levels (E*3):
 *  - conforming interfaces: SameLevel pairs whose neighbour points back,
 *    once each (from the lower element), normal = outward normal of that
 *    element's face, flux as compute_interface_flux_conforming (:320-343)
 *    computes it through ConformingInterface (src/dg/mortar.cpp:297-343,
 *    N-point face rule included);
 *  - 2:1 mortars: every non-SameLevel, non-Boundary connection from its
 *    coarse side, as update_mortar_interfaces registers them
 *    (src/amr/refinement.cpp:575-596), normal = outward normal of the coarse
 *    face, flux as compute_interface_flux_nonconforming (:345-391) through
 *    MortarSpace::compute_mortar_flux (src/dg/mortar.cpp:212-258), the fine
 *    faces put on the sub-faces of compute_subface_bounds (:27-78) they
 *    cover;
 *  - a SameLevel face against a coarser leaf adds nothing itself (the coarse
 *    side's mortar covers it; with the reference's connectivity nothing
 *    does).
 * The same scalar flux (kind DRF_FLUX_UPWIND or DRF_FLUX_LF_ADV, params as
 * for drf_flux_points) is applied to each of the five fields.  Interfaces are
 * grouped by normal axis, kind (DRF_CONN_SAME_LEVEL for conforming faces,
 * else the mortar's connection type) and level (the largest directional
 * level of the element owning the interface).  drf_face_term_dev runs every
 * conforming interface in one element-ordered launch (each element adds the
 * faces it owns a side of, its block read and rewritten once) and each
 * mortar (axis, kind) as one persistent kernel over its level groups; an
 * attached plan's conforming interfaces are added inside the volume kernel
 * (drf_face_plan_set_conforming_mode).  When the state's element span exceeds
 * 4 GiB the conforming groups run interface by interface, one launch per
 * axis.  Orders 1-4.
 * DRF_ERR_UNSUPPORTED when a mortar's fine faces do not tile its coarse face
 * as its type says (the 1:3 fallback of get_connection_type);
 * DRF_ERR_ARG when a face would belong to two interfaces.  The plan keeps
 * device copies of its lists (the context's device and precision). */
drf_status drf_face_plan_create(const drf_ctx *ctx, int64_t n_elem, const double *bounds,
                                const int32_t *levels, const uint8_t *type, const int8_t *count,
                                const int64_t *elems, const int8_t *faces, int kind,
                                const double *params, drf_face_plan **out);
void drf_face_plan_destroy(drf_face_plan *plan);
/* How the conforming interfaces run (no counterpart in the reference: the
 * launch organisation; every mode gives the face term to rounding):
 * DRF_FACES_IN_VOLUME (the default) adds them in the volume kernel's last
 * phase when the plan is attached to a right-hand side (on a DRF_LAYOUT_BLOCKS
 * state: the element-ordered launch after it), and element by element in
 * drf_face_term_dev; DRF_FACES_BY_ELEMENT uses the element-ordered launch
 * after the volume kernel; DRF_FACES_BY_INTERFACE one launch per normal axis,
 * interface by interface.  Mortars always run as their own launches. */
enum { DRF_FACES_BY_INTERFACE = 0, DRF_FACES_BY_ELEMENT = 1, DRF_FACES_IN_VOLUME = 2 };
drf_status drf_face_plan_set_conforming_mode(drf_face_plan *plan, int mode);
drf_status drf_face_plan_info(const drf_face_plan *plan, int *n_groups, int64_t *n_conforming,
                              int64_t *n_mortar);
/* per group, in launch order: desc[5g .. 5g+4] = axis, kind, level, first
 * index into its kind's list, count */
drf_status drf_face_plan_groups(const drf_face_plan *plan, int64_t *desc);
/* the lists in group order: conforming[4c ..] = left element, left face,
 * right element, right face; mortar[6m ..] = coarse element, coarse face,
 * fine elements of sub-faces 0..3 (-1 past the type's count) */
drf_status drf_face_plan_lists(const drf_face_plan *plan, int64_t *conforming, int64_t *mortar);
/* dU += the face term: each interface's face vectors lifted to the element
 * nodes (I_f^T) and multiplied by the inverse LGL mass matrix 1/(w_i w_j w_k)
 * (DG3DElementOperator::mass_inv, operators_3d.cpp:52-55;
 * basis_hexahedron.cpp:393-420).  U, dU device pointers in `layout`, element
 * type of the plan's context. */
drf_status drf_face_term_dev(drf_face_plan *plan, const void *U, void *dU, int layout,
                             void *stream);
/* The face vectors of one group before lifting: conforming groups write
 * rhs_l to out_a and rhs_r to out_b, [interface][field][(N+1)^2]; mortar
 * groups write rhs_c to out_a [interface][field][(N+1)^2] and rhs_f to out_b
 * [interface][field][fine][(N+1)^2]. */
drf_status drf_face_group_fluxes_dev(drf_face_plan *plan, int group, const void *U, int layout,
                                     void *out_a, void *out_b, void *stream);
/* Attach a plan (NULL detaches): drf_rhs, drf_rhs_dev and the un-fused
 * Runge-Kutta steps then add its face term to every right-hand side; the
 * fused steps refuse (DRF_ERR_UNSUPPORTED).  drf_initialize detaches. */
drf_status drf_set_face_plan(drf_ctx *ctx, drf_face_plan *plan);

/* ------------------------------------------------------------------ */
/* Error estimator and element marking (GradientErrorEstimator,        */
/* include/amr/refinement.hpp; src/amr/refinement.cpp:39-161,488-520)  */
/* ------------------------------------------------------------------ */

typedef struct drf_estimator drf_estimator;

/* GradientErrorEstimator(basis, params) (refinement.cpp:39-41) with the
 * RefinementParams fields it and mark_elements read (refinement.hpp:63-71:
 * defaults 0.1, 0.01, 0.5, 8, 0). */
drf_status drf_estimator_create(int order, double refine_threshold, double coarsen_threshold,
                                double smoothness_factor, int max_level, int min_level,
                                int device, drf_estimator **out);
void drf_estimator_destroy(drf_estimator *est);

/* Characteristic length cbrt(dx dy dz) of each element from its bounds
 * (E*6, as drf_mesh_elements writes them), estimate_element
 * (refinement.cpp:79-82).  Host call. */
drf_status drf_element_char_length(int64_t n_elem, const double *bounds, double *h_char);

/* GradientErrorEstimator::estimate (refinement.cpp:44-68) on one nodal field
 * (element e at field + e*elem_stride, device), then, when levels (E*3
 * int32, device) is not NULL, the marking of mark_elements (:488-517):
 * action 0 None / 1 Refine / 2 Coarsen and the refine mask (XYZ when the
 * suggestion is NONE).  Outputs (device): value[E] double (normalised by the
 * maximum when it exceeds 1e-14), mask[E], refine[E], coarsen[E] uint8,
 * action[E], mask_out[E] uint8, and (when max_value is not NULL) the
 * un-normalised maximum as one double.  Calls on one estimator are
 * stream-ordered (they share its partial-max buffer). */
drf_status drf_estimate_dev(drf_estimator *est, int64_t n_elem, const void *field,
                            int64_t elem_stride, const void *h_char, const void *levels,
                            void *value, void *mask, void *refine, void *coarsen, void *action,
                            void *mask_out, void *max_value, void *stream);

/* ------------------------------------------------------------------ */
/* Solution transfer on refinement (SolutionProjection,                */
/* src/amr/refinement.cpp:217-371; transfer_solution :533-567)         */
/* ------------------------------------------------------------------ */
/* One nodal field per element (element e at U + e*elem_stride, device,
 * FP64).  Values are the reference's interpolate_3d (:344-371) applied by
 * sum factorisation: the same interpolant, summed in another order (within
 * 1e-13 of max |U|). */

/* project_to_children (:217-276) for n_parent parents at once: child c of
 * parent p at children + (p*nchild + c)*child_stride, nchild = 2^popcount(mask),
 * c = cx + nx (cy + ny cz) with cx ... the child's half along each refined
 * axis.  mask: RefineMask bits 1 X, 2 Y, 4 Z (include/mesh/refine_mask.hpp). */
drf_status drf_project_to_children_dev(int order, int mask, int64_t n_parent, const void *U,
                                       int64_t elem_stride, void *children, int64_t child_stride,
                                       void *stream);
/* project_from_children (:278-342): parent node values from the child that
 * holds each node (the upper half from 0 on), at its clamped child
 * coordinate.  Children laid out as drf_project_to_children_dev writes them. */
drf_status drf_project_from_children_dev(int order, int mask, int64_t n_parent,
                                         const void *children, int64_t child_stride, void *U,
                                         int64_t elem_stride, void *stream);
/* transfer_solution (:533-567) into an empty output: the n_elem - n_refine
 * elements that are not parents, in order, then for each parent in order its
 * 8 children of an XYZ projection (the reference projects with
 * RefineMask::XYZ whatever the refine mask, :560).  parents: device int64[n_refine],
 * strictly ascending and < n_elem, as adapt() builds the refinement map
 * (:432-466); U_new holds n_elem + 7 n_refine elements.  (adapt() passes its
 * solution vector as the output without clearing it, :471, so there the
 * result follows the old elements.) */
drf_status drf_transfer_solution_dev(int order, int64_t n_elem, const void *U, int64_t elem_stride,
                                     int64_t n_refine, const void *parents, void *U_new,
                                     int64_t new_stride, void *stream);
/* The same from host memory (the reference's vector of per-element VecX,
 * packed: element e at U + e (N+1)^3), with parents checked on the host
 * (DRF_ERR_ARG unless strictly ascending and < n_elem).  U_new receives
 * n_elem + 7 n_refine elements.  Copies over PCIe; allocates per call. */
drf_status drf_transfer_solution(int order, int64_t n_elem, const double *U, int64_t n_refine,
                                 const int64_t *parents, double *U_new, int device);
/* Transfer onto the refined mesh's own leaf order (not the reference's
 * behaviour: it fixes the always-XYZ projection and the appended-children
 * order of transfer_solution).  Output element k is the source element
 * src[k] (device int64) evaluated on the sub-box boxes[6k..6k+5] (device
 * double: lo, hi per axis, in the source's reference coordinates); a box of
 * exactly [-1, 1]^3 is a copy, a source outside [0, n_elem) gives NaN. */
drf_status drf_interpolate_boxes_dev(int order, int64_t n_out, int64_t n_elem, const void *src,
                                     const void *boxes, const void *U, int64_t elem_stride,
                                     void *U_new, int64_t new_stride, void *stream);
/* The same for n_fields (1 to 8) fields of one mesh in one launch (no
 * counterpart: the reference transfers field by field): field f reads
 * element e at U[f] + e elem_stride[f] and writes output element k at
 * U_new[f] + k new_stride[f] (host arrays of device pointers and strides).
 * Each field's values are those drf_interpolate_boxes_dev gives, bit for
 * bit; the sources, boxes and interval matrices are set up once per batch
 * of outputs for all fields. */
drf_status drf_interpolate_boxes_fields_dev(int order, int64_t n_out, int64_t n_elem,
                                            const void *src, const void *boxes, int n_fields,
                                            const void *const *U, const int64_t *elem_stride,
                                            void *const *U_new, const int64_t *new_stride,
                                            void *stream);
/* Host: src and boxes of drf_interpolate_boxes_dev from the leaf bounds
 * (E*6, as drf_mesh_elements writes them) before and after
 * drf_mesh_refine (refine + balance, octree_adapter.cpp:184-203).  Leaves
 * are numbered depth first, so each new leaf lies in one old leaf and the
 * walk is linear; DRF_ERR_ARG when the new mesh is not a refinement of the
 * old one. */
drf_status drf_refine_transfer_map(int64_t n_old, const double *old_bounds, int64_t n_new,
                                   const double *new_bounds, int64_t *src, double *boxes);

/* ------------------------------------------------------------------ */
/* Modal coefficients (no counterpart in the reference: the modal      */
/* Vandermonde of include/dg/basis.hpp:40-42 is declared only)         */
/* ------------------------------------------------------------------ */

/* V[i][p] = sqrt((2p+1)/2) P_p(x_i) on the LGL nodes and V^-1, n x n
 * row-major (host). */
drf_status drf_modal_matrices(int order, double *V, double *Vinv);
/* coef[e][p + n(q + n r)] = sum Vinv[p][i] Vinv[q][j] Vinv[r][k] u[e][i + n(j + n k)]
 * (device; u element e at u + e*elem_stride, coef dense [E][(N+1)^3]) */
drf_status drf_modal_coeffs_dev(int order, int64_t n_elem, const void *u, int64_t elem_stride,
                                void *coef, void *stream);
/* Same coefficients times scale * elem_scale[e] (elem_scale: device double
 * per element, or NULL for 1), rounded to BF16 (nearest even) into
 * features[e * row_stride + col_offset + m] (uint16 BF16 bits): the input
 * rows of the refinement MLP.  Mode m = 0, the element mean, is written as
 * zero (it carries no variation). */
drf_status drf_modal_features_bf16_dev(int order, int64_t n_elem, const void *u,
                                       int64_t elem_stride, double scale, const void *elem_scale,
                                       void *features, int64_t row_stride, int col_offset,
                                       void *stream);

/* ------------------------------------------------------------------ */
/* Refinement-indicator MLP (no counterpart in the reference): 128 -> 64 */
/* -> 64 -> 1, GELU (tanh form), BF16 GEMM operands with FP32           */
/* accumulation, MSE loss, Adam                                        */
/* ------------------------------------------------------------------ */

#define DRF_MLP_IN 128
#define DRF_MLP_HIDDEN 64
#define DRF_MLP_NPARAM 12481 /* W1[64][128] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3 */

typedef struct drf_mlp drf_mlp;
/* Glorot-uniform weights from a SplitMix64 stream seeded with `seed`, zero
 * biases; FP32 master weights and Adam state on `device`. */
drf_status drf_mlp_create(uint64_t seed, int device, drf_mlp **out);
void drf_mlp_destroy(drf_mlp *m);
/* host FP32 parameter vectors (DRF_MLP_NPARAM); set also clears Adam */
drf_status drf_mlp_set_params(drf_mlp *m, const float *params);
drf_status drf_mlp_get_params(const drf_mlp *m, float *params);
/* pred[n] (FP32) for X[n][128] BF16 (uint16 bits, 16-byte aligned) */
drf_status drf_mlp_forward_dev(drf_mlp *m, int64_t n, const void *X, void *pred, void *stream);
/* action[n] (uint8): 1 (refine) if pred > refine_level, 2 (coarsen) if
 * pred < coarsen_level, else 0 -- RefinementAction numbering */
drf_status drf_mlp_classify_dev(int64_t n, const void *pred, double refine_level,
                                double coarsen_level, void *action, void *stream);
/* gradient (FP32, DRF_MLP_NPARAM, device) and loss (device float) of the
 * mean squared error over the n rows of X against y (FP32) */
drf_status drf_mlp_loss_grad_dev(drf_mlp *m, int64_t n, const void *X, const void *y, void *grad,
                                 void *loss, void *stream);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) on that loss; loss may be NULL.
 * Steps on one model are stream-ordered (they share its scratch buffers). */
drf_status drf_mlp_train_step_dev(drf_mlp *m, int64_t n, const void *X, const void *y, double lr,
                                  void *loss, void *stream);

/* ------------------------------------------------------------------ */
/* Sigma-coordinate metrics (SigmaMetrics::update,                     */
/* include/mesh/sigma_coordinate.hpp:340-354); opt-in, FP64            */
/* ------------------------------------------------------------------ */

/* Per node (device arrays of npts): H = eta + h, Hinv (0 when H <= 1e-10),
 * dsigma/dx = -(deta/dx + sigma (deta/dx + dh/dx)) / H, dsigma/dy likewise,
 * dsigma/dt = -(deta/dt (1 + sigma)) / H (all three 0 when H < 1e-10).
 * deta_dt may be NULL (update_steady, :357-361).  The right-hand side does
 * not use these terms, as in the reference. */
drf_status drf_sigma_metrics_dev(int64_t npts, const void *sigma, const void *eta,
                                 const void *h, const void *deta_dx, const void *deta_dy,
                                 const void *dh_dx, const void *dh_dy, const void *deta_dt,
                                 void *H, void *Hinv, void *dsigma_dx, void *dsigma_dy,
                                 void *dsigma_dt, void *stream);

/* ------------------------------------------------------------------ */
/* SigmaCoordinateOperators (src/dg/operators_3d.cpp:513-662); FP64,   */
/* opt-in, not used by the right-hand side                             */
/* ------------------------------------------------------------------ */
/* All fields one nodal set per element (element e at + e*elem_stride,
 * device), the same stride for every field of a call.  Derivatives are the
 * reference-coordinate D_xi, D_eta, D_zeta (the reference's dense Kronecker
 * products, applied line by line in the same order). */

/* update_sigma_metrics_with_time (:578-615); deta_dt NULL is
 * update_sigma_metrics (:569-576).  Per node: H = eta + h, H_inv (0 when
 * H <= 1e-10), sigma = (zeta - 1) / 2 of the node, dsigma_dx = (sigma dH/dx -
 * deta/dx) H_inv (this class's sign; SigmaCoordinate::dsigma_dx has
 * -(deta/dx + sigma dH/dx) / H), dsigma_dy likewise, dsigma_dt = -(1 + sigma)
 * deta/dt H_inv. */
drf_status drf_sigma_ops_metrics_dev(int order, int64_t n_elem, const void *eta, const void *h,
                                     const void *deta_dx, const void *deta_dy, const void *dh_dx,
                                     const void *dh_dy, const void *deta_dt, int64_t elem_stride,
                                     void *H, void *H_inv, void *sigma, void *dsigma_dx,
                                     void *dsigma_dy, void *dsigma_dt, void *stream);
/* horizontal_gradient (:517-542): dU_dx = D_xi U - dsigma_dx D_zeta U, dU_dy likewise */
drf_status drf_sigma_horizontal_gradient_dev(int order, int64_t n_elem, const void *U,
                                             const void *dsigma_dx, const void *dsigma_dy,
                                             int64_t elem_stride, void *dU_dx, void *dU_dy,
                                             void *stream);
/* vertical_gradient (:544-554): dU_dz = H_inv D_zeta U */
drf_status drf_sigma_vertical_gradient_dev(int order, int64_t n_elem, const void *U,
                                           const void *H_inv, int64_t elem_stride, void *dU_dz,
                                           void *stream);
/* sigma_divergence (:556-567): D_xi Hu + D_eta Hv + D_zeta omega */
drf_status drf_sigma_divergence_dev(int order, int64_t n_elem, const void *Hu, const void *Hv,
                                    const void *omega, int64_t elem_stride, void *div,
                                    void *stream);
/* ale_correction (:617-632): -w_mesh H_inv D_zeta U */
drf_status drf_sigma_ale_correction_dev(int order, int64_t n_elem, const void *U,
                                        const void *w_mesh, const void *H_inv,
                                        int64_t elem_stride, void *correction, void *stream);
/* material_derivative (:634-662): dU_dt + u dU_dx + v dU_dy + (omega - w_mesh H_inv) D_zeta U */
drf_status drf_sigma_material_derivative_dev(int order, int64_t n_elem, const void *U,
                                             const void *dU_dt, const void *u, const void *v,
                                             const void *omega, const void *w_mesh,
                                             const void *dsigma_dx, const void *dsigma_dy,
                                             const void *H_inv, int64_t elem_stride, void *out,
                                             void *stream);

/* ------------------------------------------------------------------ */
/* Vertical velocity diagnosis (VerticalVelocityDiagnosis,             */
/* ColumnVerticalVelocity, src/physics/vertical_velocity.cpp); FP64,   */
/* not used by the right-hand side (omega = 0 there)                   */
/* ------------------------------------------------------------------ */

/* diagnose_omega (:55-97) for n_elem elements (element e at + e*elem_stride,
 * device): per column of N+1 consecutive DOFs (node_index_3d), omega = 0 at
 * the first, then -(trapezoid integral of div over the column, dsigma =
 * dzeta / 2) / mean H of the segment, 0 where that mean H <= 1e-10. */
drf_status drf_diagnose_omega_dev(int order, int64_t n_elem, const void *div, const void *H,
                                  int64_t elem_stride, void *omega, void *stream);
/* diagnose_omega_from_velocity (:99-108): div = dHu_dx + dHv_dy, then as
 * above (the reference's Hu and Hv arguments are unused there) */
drf_status drf_diagnose_omega_from_velocity_dev(int order, int64_t n_elem, const void *dHu_dx,
                                                const void *dHv_dy, const void *H,
                                                int64_t elem_stride, void *omega, void *stream);
/* omega_to_physical_w (:110-121): w = H omega + u dz_dx + v dz_dy + dz_dt per point */
drf_status drf_omega_to_w_dev(int64_t npts, const void *omega, const void *u, const void *v,
                              const void *H, const void *dz_dx, const void *dz_dy,
                              const void *dz_dt, void *w, void *stream);
/* omega_to_w_simple (:123-143): the same with H = eta + h and dz/d(x, y, t) =
 * deta/d(x, y, t) (1 + sigma) */
drf_status drf_omega_to_w_simple_dev(int64_t npts, const void *omega, const void *u,
                                     const void *v, const void *eta, const void *h,
                                     const void *sigma, const void *deta_dx,
                                     const void *deta_dy, const void *deta_dt, void *w,
                                     void *stream);
/* surface_omega (:145-152), host: (deta_dt + u deta_dx + v deta_dy) / H, 0
 * when H < 1e-10.  (bottom_omega, :154-159, is always 0.) */
double drf_surface_omega(double deta_dt, double u_surf, double v_surf, double deta_dx,
                         double deta_dy, double H);
/* ColumnVerticalVelocity(order) sigma nodes (:180-244), host, order+1 values:
 * LGL nodes mapped to [-1, 0] for 3..5 levels, uniform spacing above */
drf_status drf_column_sigma_nodes(int order, double *sigma);
/* diagnose_column (:246-275) for ncol columns of order+1 contiguous levels;
 * omega_bottom: device double per column, or NULL for 0 */
drf_status drf_diagnose_column_dev(int order, int64_t ncol, const void *div, const void *H,
                                   const void *omega_bottom, void *omega, void *stream);
/* continuity_residual (:161-174) per element: sqrt(sum (div + dHomega)^2 / ndof) */
drf_status drf_continuity_residual_dev(int order, int64_t n_elem, const void *div,
                                       const void *dHomega_dsigma, int64_t elem_stride,
                                       void *residual, void *stream);

/* ------------------------------------------------------------------ */
/* VTK output (VTKWriter, include/io/vtk_writer.hpp,                    */
/* src/io/vtk_writer.cpp:47-475): host-side writer of .vtu / .pvd files */
/* byte-identical to the reference's, fields taken from device memory   */
/* ------------------------------------------------------------------ */

typedef struct drf_vtk drf_vtk;
#define DRF_VTK_VTU 0    /* VTKFormat::VTU */
#define DRF_VTK_PVTU 1   /* VTKFormat::PVTU: one piece without MPI (:367-418) */
#define DRF_VTK_ASCII 0  /* VTKEncoding */
#define DRF_VTK_BINARY 1 /* written as base64, as the reference does (:454-458) */
#define DRF_VTK_BASE64 2 /* base64; the header names a zlib compressor it does not use */

/* VTKWriter(basename, format, encoding) (:47-63): creates basename's directory */
drf_status drf_vtk_create(const char *basename, int format, int encoding, drf_vtk **out);
void drf_vtk_destroy(drf_vtk *w);
/* set_polynomial_order (:65); takes effect at the next drf_vtk_set_mesh */
drf_status drf_vtk_set_polynomial_order(drf_vtk *w, int order);
/* set_mesh + build_mesh_geometry (:67-143) from leaf bounds (E*6, as
 * drf_mesh_elements writes them): equispaced nodes per element (at order 1
 * the reference coordinate is the node index, 0 or 1), linear hexahedra in
 * VTK corner order at order 1, Lagrange hexahedra in DG node order above */
drf_status drf_vtk_set_mesh(drf_vtk *w, int64_t n_elem, const double *bounds);
/* add_point_data / add_cell_data (:145-159) (on_cells 0 / 1) */
drf_status drf_vtk_add_field(drf_vtk *w, const char *name, int ncomp, int on_cells);
/* set_point_data / set_cell_data (:161-212) from host memory; DRF_ERR_ARG
 * for an unknown name ("Unknown point field: name", as the reference throws) */
drf_status drf_vtk_set_field(drf_vtk *w, const char *name, int on_cells, int64_t n,
                             const double *data);
/* set_point_data(name, vector<VecX>) (:173-192) from a device field: ndof
 * values per element at field + e*elem_stride (one strided copy) */
drf_status drf_vtk_set_point_field_dev(drf_vtk *w, const char *name, int64_t n_elem,
                                       const void *field, int64_t elem_stride, int64_t ndof,
                                       void *stream);
/* get_filename (:231-249): basename_NNNNNN.vtu */
drf_status drf_vtk_filename(const drf_vtk *w, int64_t time_idx, char *buf, int64_t buf_len);
/* write (:214-227) and write_timestep (:229) */
drf_status drf_vtk_write(drf_vtk *w, int64_t time_idx, double time);
drf_status drf_vtk_write_timestep(drf_vtk *w, double time);
/* finalize: the .pvd collection of the written steps (:420-437) */
drf_status drf_vtk_finalize(drf_vtk *w);

/* ------------------------------------------------------------------ */
/* Introspection for measurement                                       */
/* ------------------------------------------------------------------ */
/* Algorithmic HBM bytes of one drf_rhs_dev launch at the current mesh. */
double drf_rhs_bytes(const drf_ctx *ctx);
int64_t drf_num_elements(const drf_ctx *ctx);
int drf_order(const drf_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
