/*
 * ORACLE (test infrastructure only) -- A11/A12: one barotropic/baroclinic
 * mode-split step on the full per-element state.
 *
 * Follows src/physics/mode_splitting.cpp: decompose :359-396 (depth_average
 * :136-149, depth_integrate :151-162, compute_deviation :164-175),
 * subcycle_barotropic :475-485 -> ModeSplittingElement::subcycle :319-332 ->
 * barotropic_step_predictor_corrector :293-317 -> barotropic_rhs :213-221
 * (eta :177-184, momentum :186-211), BarotropicState::update_derived :24-36,
 * update_baroclinic :487-600 and recombine :398-432.  The baroclinic forcing
 * is identically zero (:223-249).
 *
 * Index quirks kept on purpose: columns are c = d / n with "vertical" index
 * d % n (:143); the barotropic eta and H come from .head(n_h) (:378, :384);
 * the bathymetry and Coriolis parameter of the 2D state are h(i), f(i) for
 * i < n_h (:31, :204-205, :515).
 *
 * Pinned by (tests/test_oracle_modesplit.py, tests/test_oracle_modesplit_pins.py):
 * - the reference's component tests (tests/unit/test_mode_splitting.cpp:52-186:
 *   depth average of a constant, its linearity, depth integral, deviation and
 *   its zero average, quiescent barotropic RHS, divergence of a constant
 *   transport) on the element-level entry points below, with closed forms of
 *   D_x2d, D_y2d, the pressure gradient and the Coriolis signs on linear fields;
 * - closed forms through the whole step: the gravity term with and without
 *   Coriolis (Heun's map w -> P w + dt F (1 + z/2)), the free-surface
 *   divergence with Coriolis, eta_avg through update_baroclinic's H, the
 *   omega profile and the column its Hinv comes from, the quiescent column,
 *   uniform transport rotated by Coriolis with g = 0, the eta total of a
 *   resting flat surface (:189-213), Hu not refreshed by recombine;
 * - planted mistakes: copies of this file compiled with swapped D_x2d / D_y2d,
 *   flipped signs, the wrong Hinv column, the final eta in place of eta_avg or
 *   a one-stage substep each fail at least one of those closed forms.
 * What none of these fixes: the order in which Eigen sums the dense 2D
 * mat-vecs (a few ulp).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

#define NHMAX (ORC_MAX_N * ORC_MAX_N)

typedef struct {
    int n, nh, nd;
    double D1[ORC_MAX_N * ORC_MAX_N];
    double sw[ORC_MAX_N];
    double Dx[NHMAX * NHMAX], Dy[NHMAX * NHMAX];
} ms_ops;

static int ms_init(ms_ops *o, int order)
{
    if (order < 1 || order + 1 > ORC_MAX_N)
        return -1;
    const int n = order + 1;
    o->n = n;
    o->nh = n * n;
    o->nd = n * n * n;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    orc_derivative_matrix(n, x, o->D1);
    for (int k = 0; k < n; ++k)
        o->sw[k] = 0.5 * w[k]; /* :80-85 */
    memset(o->Dx, 0, sizeof(o->Dx));
    memset(o->Dy, 0, sizeof(o->Dy));
    /* build_2d_operators :90-119 */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int row = i + j * n;
            for (int ii = 0; ii < n; ++ii) {
                o->Dx[row * o->nh + (ii + j * n)] = o->D1[i * n + ii];
                o->Dy[row * o->nh + (i + ii * n)] = o->D1[j * n + ii];
            }
        }
    return 0;
}

static void mvh(const ms_ops *o, const double *A, const double *x, double *y)
{
    for (int r = 0; r < o->nh; ++r) {
        double s = 0.0;
        for (int c = 0; c < o->nh; ++c)
            s += A[r * o->nh + c] * x[c];
        y[r] = s;
    }
}

/* depth_average :136-149 */
static void depth_average(const ms_ops *o, const double *u3, double *ub)
{
    for (int ih = 0; ih < o->nh; ++ih) {
        double s = 0.0, ws = 0.0;
        for (int k = 0; k < o->n; ++k) {
            s += o->sw[k] * u3[ih * o->n + k];
            ws += o->sw[k];
        }
        ub[ih] = s / ws;
    }
}

/* depth_integrate :151-162 */
static void depth_integrate(const ms_ops *o, const double *u3, const double *H, double *hub)
{
    for (int ih = 0; ih < o->nh; ++ih) {
        double s = 0.0;
        for (int k = 0; k < o->n; ++k) {
            const int idx = ih * o->n + k;
            s += o->sw[k] * H[idx] * u3[idx];
        }
        hub[ih] = s;
    }
}

/* compute_deviation :164-175 */
static void deviation(const ms_ops *o, const double *u3, const double *ub, double *up)
{
    for (int ih = 0; ih < o->nh; ++ih)
        for (int k = 0; k < o->n; ++k) {
            const int idx = ih * o->n + k;
            up[idx] = u3[idx] - ub[ih];
        }
}

/* compute_2d_divergence :129-134 */
static void divergence_2d(const ms_ops *o, const double *HU, const double *HV, double *div)
{
    double a[NHMAX], b[NHMAX];
    mvh(o, o->Dx, HU, a);
    mvh(o, o->Dy, HV, b);
    for (int i = 0; i < o->nh; ++i)
        div[i] = a[i] + b[i];
}

typedef struct {
    double eta[NHMAX], HU[NHMAX], HV[NHMAX], H[NHMAX], U[NHMAX], V[NHMAX];
} baro;

/* BarotropicState::update_derived :24-36 with h(i), i < n_h */
static void baro_update(const ms_ops *o, baro *s, const double *h)
{
    for (int i = 0; i < o->nh; ++i) {
        s->H[i] = s->eta[i] + h[i];
        const double Hinv = (s->H[i] > 1e-10) ? 1.0 / s->H[i] : 0.0;
        s->U[i] = s->HU[i] * Hinv;
        s->V[i] = s->HV[i] * Hinv;
    }
}

/* barotropic_rhs :213-221 with zero forcing */
static void baro_rhs(const ms_ops *o, const baro *s, const double *f, double g, double *de,
                     double *dU, double *dV)
{
    double a[NHMAX], b[NHMAX], ex[NHMAX], ey[NHMAX];
    mvh(o, o->Dx, s->HU, a);
    mvh(o, o->Dy, s->HV, b);
    for (int i = 0; i < o->nh; ++i)
        de[i] = -(a[i] + b[i]);
    mvh(o, o->Dx, s->eta, ex);
    mvh(o, o->Dy, s->eta, ey);
    for (int i = 0; i < o->nh; ++i) {
        const double pgx = -g * s->H[i] * ex[i];
        const double pgy = -g * s->H[i] * ey[i];
        const double cx = f[i] * s->HV[i];
        const double cy = -f[i] * s->HU[i];
        dU[i] = pgx + cx + 0.0;
        dV[i] = pgy + cy + 0.0;
    }
}

int orc_mode_split_step(int order, int64_t E, double dt3, int nsub, double *full,
                        const double *h, const double *f, const orc_physics *p)
{
    ms_ops o;
    if (ms_init(&o, order) || nsub < 1)
        return -1;
    const int n = o.n, nh = o.nh, nd = o.nd;
    const double dtb = dt3 / nsub;
    for (int64_t e = 0; e < E; ++e) {
        double *F = full + e * (int64_t)ORC_NFULL * nd;
        double *Hu = F + ORC_S_HU * nd, *Hv = F + ORC_S_HV * nd, *eta = F + ORC_S_ETA * nd;
        double *HT = F + ORC_S_HT * nd, *HS = F + ORC_S_HS * nd, *H = F + ORC_S_H * nd;
        double *u = F + ORC_S_U * nd, *v = F + ORC_S_V * nd, *T = F + ORC_S_T * nd;
        double *S = F + ORC_S_S * nd, *om = F + ORC_S_OMEGA * nd, *rho = F + ORC_S_RHO * nd;
        const double *he = h + e * nd, *fe = f + e * nd;

        /* ---- decompose :376-394 ---- */
        baro b;
        double up[ORC_MAX_N * NHMAX], vp[ORC_MAX_N * NHMAX];
        depth_average(&o, u, b.U);
        depth_average(&o, v, b.V);
        memcpy(b.eta, eta, sizeof(double) * nh);
        depth_integrate(&o, u, H, b.HU);
        depth_integrate(&o, v, H, b.HV);
        memcpy(b.H, H, sizeof(double) * nh);
        deviation(&o, u, b.U, up);
        deviation(&o, v, b.V, vp);

        /* ---- subcycle :319-332 (predictor-corrector :293-317) ---- */
        double eavg[NHMAX];
        for (int i = 0; i < nh; ++i)
            eavg[i] = 0.0;
        for (int sub = 0; sub < nsub; ++sub) {
            double de0[NHMAX], dU0[NHMAX], dV0[NHMAX], de1[NHMAX], dU1[NHMAX], dV1[NHMAX];
            baro_rhs(&o, &b, fe, p->g, de0, dU0, dV0);
            baro st;
            for (int i = 0; i < nh; ++i) {
                st.eta[i] = b.eta[i] + dtb * de0[i];
                st.HU[i] = b.HU[i] + dtb * dU0[i];
                st.HV[i] = b.HV[i] + dtb * dV0[i];
            }
            baro_update(&o, &st, he);
            baro_rhs(&o, &st, fe, p->g, de1, dU1, dV1);
            for (int i = 0; i < nh; ++i) {
                b.eta[i] += 0.5 * dtb * (de0[i] + de1[i]);
                b.HU[i] += 0.5 * dtb * (dU0[i] + dU1[i]);
                b.HV[i] += 0.5 * dtb * (dV0[i] + dV1[i]);
            }
            baro_update(&o, &b, he);
            for (int i = 0; i < nh; ++i)
                eavg[i] += b.eta[i];
        }
        for (int i = 0; i < nh; ++i)
            eavg[i] /= (double)nsub;

        /* ---- update_baroclinic :504-598 ---- */
        for (int ih = 0; ih < nh; ++ih)
            for (int k = 0; k < n; ++k)
                eta[ih * n + k] = eavg[ih];
        for (int ih = 0; ih < nh; ++ih) {
            const double Hv_ = eavg[ih] + he[ih];
            for (int k = 0; k < n; ++k)
                H[ih * n + k] = Hv_;
        }
        double Ub[NHMAX], Vb[NHMAX];
        depth_average(&o, u, Ub);
        depth_average(&o, v, Vb);
        for (int ih = 0; ih < nh; ++ih)
            for (int k = 0; k < n; ++k) {
                const int idx = ih * n + k;
                const double ut = Ub[ih] + up[idx];
                const double vt = Vb[ih] + vp[idx];
                u[idx] = ut;
                v[idx] = vt;
                Hu[idx] = H[idx] * ut;
                Hv[idx] = H[idx] * vt;
            }
        for (int idx = 0; idx < nd; ++idx) {
            const double Hinv = (H[idx] > 1e-10) ? 1.0 / H[idx] : 0.0;
            T[idx] = HT[idx] * Hinv;
            S[idx] = HS[idx] * Hinv;
        }
        for (int idx = 0; idx < nd; ++idx) {
            HT[idx] = H[idx] * T[idx];
            HS[idx] = H[idx] * S[idx];
        }
        double HUb[NHMAX], HVb[NHMAX], dv[NHMAX];
        depth_integrate(&o, u, H, HUb);
        depth_integrate(&o, v, H, HVb);
        divergence_2d(&o, HUb, HVb, dv);
        for (int ih = 0; ih < nh; ++ih) {
            const double div = dv[ih];
            const double Hc = H[ih * n];
            const double Hinv = (Hc > 1e-10) ? 1.0 / Hc : 0.0;
            for (int k = 0; k < n; ++k) {
                const double sf = (double)k / (double)(n - 1);
                om[ih * n + k] = -sf * div * Hinv;
            }
        }
        /* linear EOS :591-598 */
        for (int idx = 0; idx < nd; ++idx)
            rho[idx] = p->rho_0 * (1.0 - p->alpha * (T[idx] - 10.0) + p->beta * (S[idx] - 35.0));

        /* ---- recombine :411-430: u from the final barotropic U_bar, eta from
         * the final barotropic eta; Hu, Hv are not refreshed ---- */
        for (int ih = 0; ih < nh; ++ih)
            for (int k = 0; k < n; ++k) {
                const int idx = ih * n + k;
                u[idx] = b.U[ih] + up[idx];
                v[idx] = b.V[ih] + vp[idx];
            }
        for (int ih = 0; ih < nh; ++ih)
            for (int k = 0; k < n; ++k)
                eta[ih * n + k] = b.eta[ih];
    }
    return 0;
}

/* ---- the element-level pieces on their own, for the reference's
 * component tests (tests/unit/test_mode_splitting.cpp:52-186) ---- */

int orc_ms_depth_average(int order, const double *u3, double *ub)
{
    ms_ops o;
    if (ms_init(&o, order))
        return -1;
    depth_average(&o, u3, ub);
    return 0;
}

int orc_ms_depth_integrate(int order, const double *u3, const double *H, double *hub)
{
    ms_ops o;
    if (ms_init(&o, order))
        return -1;
    depth_integrate(&o, u3, H, hub);
    return 0;
}

int orc_ms_deviation(int order, const double *u3, const double *ub, double *up)
{
    ms_ops o;
    if (ms_init(&o, order))
        return -1;
    deviation(&o, u3, ub, up);
    return 0;
}

int orc_ms_divergence_2d(int order, const double *HU, const double *HV, double *div)
{
    ms_ops o;
    if (ms_init(&o, order))
        return -1;
    divergence_2d(&o, HU, HV, div);
    return 0;
}

int orc_ms_barotropic_rhs(int order, const double *eta, const double *HU, const double *HV,
                          const double *H, const double *f, double g, double *de, double *dU,
                          double *dV)
{
    ms_ops o;
    if (ms_init(&o, order))
        return -1;
    baro s;
    memcpy(s.eta, eta, sizeof(double) * o.nh);
    memcpy(s.HU, HU, sizeof(double) * o.nh);
    memcpy(s.HV, HV, sizeof(double) * o.nh);
    memcpy(s.H, H, sizeof(double) * o.nh);
    baro_rhs(&o, &s, f, g, de, dU, dV);
    return 0;
}
