/*
 * ORACLE (test infrastructure only) -- hydrostatic pressure and pressure
 * gradient in sigma coordinates, element by element.
 *
 * The reference's PressureGradientElement (include/physics/pressure_gradient.hpp:
 * 46-367) does not compile (it calls HexahedronBasis::lgl_weights_1d, :183,
 * which does not exist), so no reference values exist.  This is the build's
 * statement of its algorithm from the mathematics it documents, with one
 * convention fixed: the vertical (sigma) direction is the element's zeta
 * index k (node d = i + n (j + n k)), so that xi and eta derivatives are
 * horizontal.  (The reference indexes a column as n consecutive DOFs,
 * i_h * n_vert + k, the xi line, while it differentiates along zeta.)
 *
 *   hydrostatic_pressure (:233-250): p(i,j,k) = sum_{k'=k}^{n-1}
 *       sw_k' g rho H, sw = w/2 (the sigma weights, :180-187), summed from k up;
 *   barotropic_gradient (:213-231): pg = -g H (D_xi eta, D_eta eta);
 *   baroclinic_gradient_standard (:252-294): dp/dx|_z = D_xi p +
 *       (D_zeta p) dsigma/dx, dsigma/dx = (sigma dH/dx - deta/dx) / H with
 *       dH/dx = deta/dx + dh/dx (the reference leaves deta/dx at 0 as a
 *       stub; here it is an input), pg = -(dp/dx|_z) / rho0;
 *   baroclinic_gradient_density_jacobian (:296-338): z = eta + sigma H
 *       (:352-357), J_x = rho_xi z_zeta - rho_zeta z_xi (J_y with eta),
 *       accumulated from the bottom, I_x(k) = sum_{k'<=k} sw_k' J_x, pg =
 *       -(g / rho0) I;
 *   full_gradient (:340-367): barotropic + baroclinic of the chosen method.
 * Derivatives are the dense Kronecker products of the reference
 * (DG3DElementOperator::gradient_reference, src/dg/operators_3d.cpp:58-69).
 *
 * Pinned by (tests/test_oracle_pressure.py), each against closed forms
 * computed with numpy's Legendre routines: the pressure of a constant and of
 * a zeta-linear density; zero gradient for a uniform density over a flat
 * bottom and surface (both methods); a xi-linear density (standard:
 * -g H a c_k / rho0 with c_k the partial weight sums; density Jacobian:
 * -(g / rho0) sum_{k'<=k} sw_k' a H / 2); the barotropic gradient of a
 * linear surface; the sigma correction over a sloping bottom (sign
 * included); planted mistakes (the correction's sign, the integration
 * direction) each failing one of them.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static void grad3(int order, const double *Dxi, const double *Deta, const double *Dzeta,
                  const double *u, double *gx, double *gy, double *gz)
{
    const int nd = (order + 1) * (order + 1) * (order + 1);
    for (int r = 0; r < nd; ++r) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int s = 0; s < nd; ++s) {
            a += Dxi[r * nd + s] * u[s];
            b += Deta[r * nd + s] * u[s];
            c += Dzeta[r * nd + s] * u[s];
        }
        if (gx)
            gx[r] = a;
        if (gy)
            gy[r] = b;
        if (gz)
            gz[r] = c;
    }
}

int orc_hydrostatic_pressure(int order, int64_t E, double g, const double *rho, const double *H,
                             double *p)
{
    const int n = order + 1, n2 = n * n, nd = n2 * n;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    for (int64_t e = 0; e < E; ++e)
        for (int c = 0; c < n2; ++c)
            for (int k = 0; k < n; ++k) {
                double s = 0.0;
                for (int kk = k; kk < n; ++kk) {
                    const int64_t o = e * nd + c + n2 * kk;
                    s += (0.5 * w[kk]) * g * rho[o] * H[o];
                }
                p[e * nd + c + n2 * k] = s;
            }
    return 0;
}

/* parts: 1 barotropic, 2 baroclinic, 3 both; method 0 standard, 1 density
 * Jacobian.  NULL metric inputs are 0. */
int orc_pressure_gradient(int order, int64_t E, int method, int parts, double g, double rho0,
                          const double *eta, const double *rho, const double *H,
                          const double *sigma, const double *deta_dx, const double *deta_dy,
                          const double *dh_dx, const double *dh_dy, double *pg_x, double *pg_y)
{
    const int n = order + 1, n2 = n * n, nd = n2 * n;
    double x[ORC_MAX_N], w[ORC_MAX_N], D1[ORC_MAX_N * ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    orc_derivative_matrix(n, x, D1);
    double *K = (double *)malloc(sizeof(double) * 3 * nd * nd);
    double *buf = (double *)malloc(sizeof(double) * 12 * nd);
    if (!K || !buf) {
        free(K);
        free(buf);
        return -1;
    }
    double *Dxi = K, *Deta = K + nd * nd, *Dzeta = K + 2 * nd * nd;
    orc_kron_diff(order, D1, Dxi, Deta, Dzeta);
    double *p = buf, *gx = buf + nd, *gy = buf + 2 * nd, *gz = buf + 3 * nd;
    double *rx = buf + 4 * nd, *ry = buf + 5 * nd, *rz = buf + 6 * nd;
    double *z = buf + 7 * nd, *zx = buf + 8 * nd, *zy = buf + 9 * nd, *zz = buf + 10 * nd;
    for (int64_t e = 0; e < E; ++e) {
        const int64_t o = e * nd;
        double *ox = pg_x + o, *oy = pg_y + o;
        for (int d = 0; d < nd; ++d)
            ox[d] = oy[d] = 0.0;
        if (parts & 1) {
            grad3(order, Dxi, Deta, Dzeta, eta + o, gx, gy, NULL);
            for (int d = 0; d < nd; ++d) {
                ox[d] = -g * H[o + d] * gx[d];
                oy[d] = -g * H[o + d] * gy[d];
            }
        }
        if (!(parts & 2))
            continue;
        if (method == 0) {
            orc_hydrostatic_pressure(order, 1, g, rho + o, H + o, p);
            grad3(order, Dxi, Deta, Dzeta, p, gx, gy, gz);
            for (int d = 0; d < nd; ++d) {
                const double ex = deta_dx ? deta_dx[o + d] : 0.0;
                const double ey = deta_dy ? deta_dy[o + d] : 0.0;
                const double dHx = ex + (dh_dx ? dh_dx[o + d] : 0.0);
                const double dHy = ey + (dh_dy ? dh_dy[o + d] : 0.0);
                const double sx = (sigma[o + d] * dHx - ex) / H[o + d];
                const double sy = (sigma[o + d] * dHy - ey) / H[o + d];
                const double px = gx[d] + gz[d] * sx;
                const double py = gy[d] + gz[d] * sy;
                ox[d] += -px / rho0;
                oy[d] += -py / rho0;
            }
        } else {
            for (int d = 0; d < nd; ++d)
                z[d] = eta[o + d] + sigma[o + d] * H[o + d];
            grad3(order, Dxi, Deta, Dzeta, rho + o, rx, ry, rz);
            grad3(order, Dxi, Deta, Dzeta, z, zx, zy, zz);
            for (int c = 0; c < n2; ++c) {
                double ix = 0.0, iy = 0.0;
                for (int k = 0; k < n; ++k) {
                    const int d = c + n2 * k;
                    const double jx = rx[d] * zz[d] - rz[d] * zx[d];
                    const double jy = ry[d] * zz[d] - rz[d] * zy[d];
                    ix += (0.5 * w[k]) * jx;
                    iy += (0.5 * w[k]) * jy;
                    ox[d] += -g / rho0 * ix;
                    oy[d] += -g / rho0 * iy;
                }
            }
        }
    }
    free(K);
    free(buf);
    return 0;
}
