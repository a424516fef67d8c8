/*
 * ORACLE (test infrastructure only) -- vertical velocity diagnosis:
 * VerticalVelocityDiagnosis::diagnose_omega / diagnose_omega_from_velocity /
 * omega_to_physical_w / omega_to_w_simple (src/physics/vertical_velocity.cpp:55-143)
 * and ColumnVerticalVelocity (ctor :180-244, diagnose_column :246-275).
 *
 * The reference has no tests of these classes; they are pinned by closed
 * forms in tests/test_oracle_omega.py (trapezoidal integration exact for a
 * divergence linear in sigma, the dry guard, the column layout, the
 * hard-coded node tables, the w formulas).
 */
#include "oracle.h"

#include <math.h>

/* diagnose_omega, :55-97: column i of an element holds its nodes i*n + k
 * (node_index_3d, include/physics/vertical_velocity.hpp), k = 0 at the bottom */
static void omega_element(int n, const double *x, const double *div, const double *H,
                          double *omega)
{
    for (int i = 0; i < n * n; ++i) {
        omega[i * n] = 0.0;
        double integral = 0.0;
        for (int k = 1; k < n; ++k) {
            const int p = i * n + k - 1, c = i * n + k;
            const double H_avg = 0.5 * (H[p] + H[c]);
            const double dsigma = (x[k] - x[k - 1]) * 0.5;
            const double div_avg = 0.5 * (div[p] + div[c]);
            integral += div_avg * dsigma;
            omega[c] = (H_avg > 1e-10) ? -integral / H_avg : 0.0;
        }
    }
}

int orc_diagnose_omega(int order, int64_t E, const double *div, const double *H, int64_t es,
                       double *omega)
{
    const int n = order + 1;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    for (int64_t e = 0; e < E; ++e)
        omega_element(n, x, div + e * es, H + e * es, omega + e * es);
    return 0;
}

/* diagnose_omega_from_velocity, :99-108: div = dHu_dx + dHv_dy first */
int orc_diagnose_omega_from_velocity(int order, int64_t E, const double *dHu_dx,
                                     const double *dHv_dy, const double *H, int64_t es,
                                     double *omega)
{
    const int n = order + 1, nd = n * n * n;
    double x[ORC_MAX_N], w[ORC_MAX_N], div[ORC_MAX_N * ORC_MAX_N * ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    for (int64_t e = 0; e < E; ++e) {
        for (int d = 0; d < nd; ++d)
            div[d] = dHu_dx[e * es + d] + dHv_dy[e * es + d];
        omega_element(n, x, div, H + e * es, omega + e * es);
    }
    return 0;
}

/* omega_to_physical_w, :110-121 */
void orc_omega_to_w(int64_t npts, const double *omega, const double *u, const double *v,
                    const double *H, const double *dz_dx, const double *dz_dy,
                    const double *dz_dt, double *w)
{
    for (int64_t i = 0; i < npts; ++i)
        w[i] = H[i] * omega[i] + u[i] * dz_dx[i] + v[i] * dz_dy[i] + dz_dt[i];
}

/* omega_to_w_simple, :123-143 */
void orc_omega_to_w_simple(int64_t npts, const double *omega, const double *u, const double *v,
                           const double *eta, const double *h, const double *sigma,
                           const double *deta_dx, const double *deta_dy, const double *deta_dt,
                           double *w)
{
    for (int64_t i = 0; i < npts; ++i) {
        const double H = eta[i] + h[i];
        const double dz_dx = deta_dx[i] * (1.0 + sigma[i]);
        const double dz_dy = deta_dy[i] * (1.0 + sigma[i]);
        const double dz_dt = deta_dt[i] * (1.0 + sigma[i]);
        w[i] = H * omega[i] + u[i] * dz_dx + v[i] * dz_dy + dz_dt;
    }
}

/* ColumnVerticalVelocity(order), :180-244: sigma nodes of its levels
 * (hard-coded LGL nodes for 3..5 levels, uniform spacing above) */
int orc_column_sigma_nodes(int order, double *sigma)
{
    const int n = order + 1;
    if (n < 1 || n > ORC_MAX_N)
        return -1;
    if (n == 1) {
        sigma[0] = -0.5;
        return 0;
    }
    if (n == 2) {
        sigma[0] = -1.0;
        sigma[1] = 0.0;
        return 0;
    }
    double z[ORC_MAX_N];
    if (n == 3) {
        z[0] = -1.0, z[1] = 0.0, z[2] = 1.0;
    } else if (n == 4) {
        z[0] = -1.0, z[1] = -sqrt(1.0 / 5.0), z[2] = sqrt(1.0 / 5.0), z[3] = 1.0;
    } else if (n == 5) {
        z[0] = -1.0, z[1] = -sqrt(3.0 / 7.0), z[2] = 0.0, z[3] = sqrt(3.0 / 7.0), z[4] = 1.0;
    } else {
        for (int k = 0; k < n; ++k)
            z[k] = -1.0 + 2.0 * k / (n - 1);
    }
    for (int k = 0; k < n; ++k)
        sigma[k] = 0.5 * (z[k] - 1.0);
    return 0;
}

/* diagnose_column, :246-275, for ncol contiguous columns of order+1 levels;
 * omega_bottom per column (NULL: 0) */
int orc_diagnose_column(int order, int64_t ncol, const double *div, const double *H,
                        const double *omega_bottom, double *omega)
{
    const int n = order + 1;
    double s[ORC_MAX_N];
    if (orc_column_sigma_nodes(order, s))
        return -1;
    for (int64_t c = 0; c < ncol; ++c) {
        const double *dv = div + c * n, *Hc = H + c * n;
        double *om = omega + c * n;
        const double ob = omega_bottom ? omega_bottom[c] : 0.0;
        om[0] = ob;
        double integral = 0.0;
        for (int k = 1; k < n; ++k) {
            const double dsigma = s[k] - s[k - 1];
            const double div_avg = 0.5 * (dv[k - 1] + dv[k]);
            integral += div_avg * dsigma;
            const double H_avg = 0.5 * (Hc[k - 1] + Hc[k]);
            om[k] = (H_avg > 1e-10) ? ob - integral / H_avg : ob;
        }
    }
    return 0;
}

/* continuity_residual, :161-174, per element: sqrt(sum (div + dHomega)^2 / ndof)
 * in DOF order (H and omega are not used by the reference) */
void orc_continuity_residual(int order, int64_t E, const double *div, const double *dHomega,
                             int64_t es, double *out)
{
    const int n = order + 1, nd = n * n * n;
    for (int64_t e = 0; e < E; ++e) {
        double r2 = 0.0;
        for (int d = 0; d < nd; ++d) {
            const double r = div[e * es + d] + dHomega[e * es + d];
            r2 += r * r;
        }
        out[e] = sqrt(r2 / nd);
    }
}
