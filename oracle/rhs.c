/*
 * ORACLE (test infrastructure only) -- A2-A8: the element-local primitive
 * equations right-hand side, computed the way the reference does it: dense
 * Kronecker (ndof x ndof) derivative mat-vecs per element.
 *
 * Follows src/physics/primitive_equations.cpp:12-51 (update_derived/resize),
 * :86-122 (2D operators), :136-393 (element RHS) and :428-442 (element loop),
 * src/dg/operators_3d.cpp:58-79 (gradient/divergence_reference) and
 * src/solver/simulation_driver.cpp:170-235 (packed layout).
 *
 * Pinned by (tests/test_oracle_rhs.py): the reference's own closed-form tests
 * (quiescent state -> zero RHS, zero transport -> zero eta RHS, zero velocity
 * -> zero tracer advection, Coriolis sign, update_derived values); manufactured
 * polynomial states of degree <= N for which every D is exact, whose tendency
 * is written independently in numpy from the PDE terms; and the Kronecker /
 * sum-factorised equivalence.  The mixing of xi/eta/zeta indices in the eta
 * RHS (d % n treated as the vertical index) is pinned by a brute-force numpy
 * evaluation of the written formula on n = 2, 3 and by a manufactured field
 * that varies only along xi.  Eigen's summation order inside a mat-vec is not
 * reproducible without Eigen ("parity unpinned" at the ulp level only).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n, nd, nh;
    double D1[ORC_MAX_N * ORC_MAX_N];
    double sw[ORC_MAX_N];           /* sigma_weights_ = 0.5 * w_lgl, :116-121 */
    double *Dxi, *Deta, *Dzeta;     /* dense ndof^2 */
    double *Dx2d, *Dy2d;            /* dense nh^2, :96-114 */
} orc_ops;

static int ops_init(orc_ops *o, int order)
{
    if (order < 1 || order + 1 > ORC_MAX_N)
        return -1;
    const int n = order + 1;
    o->n = n;
    o->nd = n * n * n;
    o->nh = n * n;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    orc_derivative_matrix(n, x, o->D1);
    for (int k = 0; k < n; ++k)
        o->sw[k] = 0.5 * w[k];
    const size_t nd2 = (size_t)o->nd * o->nd, nh2 = (size_t)o->nh * o->nh;
    o->Dxi = (double *)malloc(sizeof(double) * (3 * nd2 + 2 * nh2));
    if (!o->Dxi)
        return -1;
    o->Deta = o->Dxi + nd2;
    o->Dzeta = o->Deta + nd2;
    o->Dx2d = o->Dzeta + nd2;
    o->Dy2d = o->Dx2d + nh2;
    orc_kron_diff(order, o->D1, o->Dxi, o->Deta, o->Dzeta);
    /* build_2d_operators, src/physics/primitive_equations.cpp:86-122 */
    memset(o->Dx2d, 0, sizeof(double) * 2 * nh2);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int row = i + j * n;
            for (int ii = 0; ii < n; ++ii) {
                o->Dx2d[row * o->nh + (ii + j * n)] = o->D1[i * n + ii];
                o->Dy2d[row * o->nh + (i + ii * n)] = o->D1[j * n + ii];
            }
        }
    return 0;
}

static void ops_free(orc_ops *o) { free(o->Dxi); }

/* dense mat-vec y = A x (Eigen MatX * VecX), rows summed in column order */
static void mv(int m, const double *A, const double *x, double *y)
{
    for (int r = 0; r < m; ++r) {
        double s = 0.0;
        const double *a = A + (size_t)r * m;
        for (int c = 0; c < m; ++c)
            s += a[c] * x[c];
        y[r] = s;
    }
}

#define NDMAX (ORC_MAX_N * ORC_MAX_N * ORC_MAX_N)

/*
 * PrimitiveEquationsElement::compute_rhs (:136-152) for one element with
 * derived state from update_derived (:12-29) and omega = rho = 0 (resize,
 * :31-51, called by unpack_state, simulation_driver.cpp:222).
 */
static void element_rhs(const orc_ops *o, const orc_physics *p, const double *Hu,
                        const double *Hv, const double *eta, const double *HT, const double *HS,
                        const double *h, const double *f, double *dHu, double *dHv,
                        double *deta, double *dHT, double *dHS)
{
    const int nd = o->nd, n = o->n, nh = o->nh;
    double H[NDMAX], Hinv[NDMAX], u[NDMAX], v[NDMAX], T[NDMAX], S[NDMAX];
    double t0[NDMAX], t1[NDMAX], t2[NDMAX], t3[NDMAX], z[NDMAX];
    double a0[NDMAX], a1[NDMAX], a2[NDMAX];

    /* update_derived :21-28 */
    for (int i = 0; i < nd; ++i) {
        H[i] = eta[i] + h[i];
        Hinv[i] = (H[i] > 1e-10) ? 1.0 / H[i] : 0.0;
        u[i] = Hu[i] * Hinv[i];
        v[i] = Hv[i] * Hinv[i];
        T[i] = HT[i] * Hinv[i];
        S[i] = HS[i] * Hinv[i];
    }

    /* ---- compute_momentum_rhs :154-199 ---- */
    for (int i = 0; i < nd; ++i) {
        dHu[i] = 0.0;
        dHv[i] = 0.0;
    }
    /* momentum_advection :275-305; divergence_reference operators_3d.cpp:71-79
     * with omega = 0 so the zeta flux is Hu*0 */
    for (int i = 0; i < nd; ++i) {
        t0[i] = Hu[i] * u[i];
        t1[i] = Hu[i] * v[i];
        z[i] = Hu[i] * 0.0;
    }
    mv(nd, o->Dxi, t0, a0);
    mv(nd, o->Deta, t1, a1);
    mv(nd, o->Dzeta, z, a2);
    for (int i = 0; i < nd; ++i)
        dHu[i] -= (a0[i] + a1[i]) + a2[i];
    for (int i = 0; i < nd; ++i) {
        t0[i] = Hv[i] * u[i];
        t1[i] = Hv[i] * v[i];
        z[i] = Hv[i] * 0.0;
    }
    mv(nd, o->Dxi, t0, a0);
    mv(nd, o->Deta, t1, a1);
    mv(nd, o->Dzeta, z, a2);
    for (int i = 0; i < nd; ++i)
        dHv[i] -= (a0[i] + a1[i]) + a2[i];

    /* coriolis_terms :307-320 */
    for (int i = 0; i < nd; ++i) {
        dHu[i] += f[i] * Hv[i];
        dHv[i] += -f[i] * Hu[i];
    }

    /* barotropic_pressure_gradient :322-337 (gradient_reference computes all
     * three columns; only xi and eta are used) */
    mv(nd, o->Dxi, eta, a0);
    mv(nd, o->Deta, eta, a1);
    for (int i = 0; i < nd; ++i) {
        dHu[i] += -p->g * H[i] * a0[i];
        dHv[i] += -p->g * H[i] * a1[i];
    }

    /* horizontal_diffusion :339-355 on u and v, multiplied by H :185-188 */
    if (p->nu_h > 0.0) {
        mv(nd, o->Dxi, u, t0);
        mv(nd, o->Dxi, t0, t1); /* d2u/dxi2 */
        mv(nd, o->Deta, u, t0);
        mv(nd, o->Deta, t0, t2); /* d2u/deta2 */
        for (int i = 0; i < nd; ++i)
            dHu[i] += H[i] * (p->nu_h * t1[i] + p->nu_h * t2[i]);
        mv(nd, o->Dxi, v, t0);
        mv(nd, o->Dxi, t0, t1);
        mv(nd, o->Deta, v, t0);
        mv(nd, o->Deta, t0, t2);
        for (int i = 0; i < nd; ++i)
            dHv[i] += H[i] * (p->nu_h * t1[i] + p->nu_h * t2[i]);
    }

    /* vertical_diffusion :357-375 */
    if (p->nu_v > 0.0) {
        mv(nd, o->Dzeta, u, t0);
        mv(nd, o->Dzeta, t0, t1);
        mv(nd, o->Dzeta, v, t0);
        mv(nd, o->Dzeta, t0, t2);
        for (int i = 0; i < nd; ++i) {
            dHu[i] += p->nu_v * t1[i] * Hinv[i];
            dHv[i] += p->nu_v * t2[i] * Hinv[i];
        }
    }

    /* ---- compute_tracer_rhs :201-227; tracer_advection :377-393 ---- */
    for (int i = 0; i < nd; ++i) {
        t0[i] = u[i] * HT[i];
        t1[i] = v[i] * HT[i];
        z[i] = 0.0 * HT[i];
    }
    mv(nd, o->Dxi, t0, a0);
    mv(nd, o->Deta, t1, a1);
    mv(nd, o->Dzeta, z, a2);
    for (int i = 0; i < nd; ++i)
        dHT[i] = 0.0 - ((a0[i] + a1[i]) + a2[i]);
    for (int i = 0; i < nd; ++i) {
        t0[i] = u[i] * HS[i];
        t1[i] = v[i] * HS[i];
        z[i] = 0.0 * HS[i];
    }
    mv(nd, o->Dxi, t0, a0);
    mv(nd, o->Deta, t1, a1);
    mv(nd, o->Dzeta, z, a2);
    for (int i = 0; i < nd; ++i)
        dHS[i] = 0.0 - ((a0[i] + a1[i]) + a2[i]);
    if (p->kappa_v_T > 0.0) {
        mv(nd, o->Dzeta, T, t0);
        mv(nd, o->Dzeta, t0, t3);
        for (int i = 0; i < nd; ++i)
            dHT[i] += p->kappa_v_T * t3[i] * Hinv[i];
    }
    if (p->kappa_v_S > 0.0) {
        mv(nd, o->Dzeta, S, t0);
        mv(nd, o->Dzeta, t0, t3);
        for (int i = 0; i < nd; ++i)
            dHS[i] += p->kappa_v_S * t3[i] * Hinv[i];
    }

    /* ---- compute_eta_rhs :229-273: column "vertical" index is d % n ---- */
    double HUb[ORC_MAX_N * ORC_MAX_N], HVb[ORC_MAX_N * ORC_MAX_N];
    double gx[ORC_MAX_N * ORC_MAX_N], gy[ORC_MAX_N * ORC_MAX_N];
    for (int ih = 0; ih < nh; ++ih) {
        double su = 0.0, sv = 0.0;
        for (int k = 0; k < n; ++k) {
            const int idx = ih * n + k;
            su += o->sw[k] * Hu[idx];
            sv += o->sw[k] * Hv[idx];
        }
        HUb[ih] = su;
        HVb[ih] = sv;
    }
    mv(nh, o->Dx2d, HUb, gx);
    mv(nh, o->Dy2d, HVb, gy);
    for (int ih = 0; ih < nh; ++ih) {
        const double d2 = -(gx[ih] + gy[ih]);
        for (int k = 0; k < n; ++k)
            deta[ih * n + k] = d2;
    }
}

void orc_update_derived(int order, int64_t n_elem, const double *U, const double *h, double *H,
                        double *u, double *v, double *T, double *S)
{
    const int n = order + 1, nd = n * n * n;
    for (int64_t e = 0; e < n_elem; ++e) {
        const double *B = U + e * 5 * nd;
        for (int i = 0; i < nd; ++i) {
            const int64_t g = e * nd + i;
            const double Hi = B[2 * nd + i] + h[g];
            const double Hinv = (Hi > 1e-10) ? 1.0 / Hi : 0.0;
            H[g] = Hi;
            u[g] = B[i] * Hinv;
            v[g] = B[nd + i] * Hinv;
            T[g] = B[3 * nd + i] * Hinv;
            S[g] = B[4 * nd + i] * Hinv;
        }
    }
}

int orc_rhs_range(int order, int64_t e0, int64_t e1, const double *U, const double *h,
                  const double *f, const orc_physics *phys, double *dUdt)
{
    orc_ops o;
    if (ops_init(&o, order))
        return -1;
    const int nd = o.nd;
    /* element loop, src/physics/primitive_equations.cpp:436-438; pack,
     * src/solver/simulation_driver.cpp:188-196 */
    for (int64_t e = e0; e < e1; ++e) {
        const double *B = U + e * 5 * nd;
        double *R = dUdt + e * 5 * nd;
        element_rhs(&o, phys, B, B + nd, B + 2 * nd, B + 3 * nd, B + 4 * nd, h + e * nd,
                    f + e * nd, R, R + nd, R + 2 * nd, R + 3 * nd, R + 4 * nd);
    }
    ops_free(&o);
    return 0;
}

int orc_rhs(int order, int64_t n_elem, const double *U, const double *h, const double *f,
            const orc_physics *phys, double *dUdt)
{
    return orc_rhs_range(order, 0, n_elem, U, h, f, phys, dUdt);
}
