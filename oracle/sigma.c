/*
 * ORACLE (test infrastructure only) -- A13: sigma-coordinate metric terms.
 *
 * Follows SigmaCoordinate::total_depth / dsigma_dx / dsigma_dy / dsigma_dt
 * (include/mesh/sigma_coordinate.hpp:67-103) and SigmaMetrics::update
 * (:340-354): H = eta + h, Hinv = H > 1e-10 ? 1/H : 0,
 * dsigma/dx = -(deta/dx + sigma (deta/dx + dh/dx)) / H (0 when H < 1e-10),
 * dsigma/dt = -(deta/dt (1 + sigma)) / H.  The operators_3d variant with the
 * opposite sign on the sigma term (src/dg/operators_3d.cpp:608) is not the
 * tested formula and is not restated.
 *
 * Pinned by (tests/test_oracle_sigma.py): the reference's closed-form case
 * (tests/unit/test_sigma_coordinate.cpp:71-89), the z(sigma) derivative
 * identity d sigma/dx|_z = -(dz/dx|_sigma) / H checked by finite
 * differences of the transform, and the dry-column guards.
 */
#include "oracle.h"

void orc_sigma_metrics(int64_t npts, const double *sigma, const double *eta, const double *h,
                       const double *deta_dx, const double *deta_dy, const double *dh_dx,
                       const double *dh_dy, const double *deta_dt, double *H, double *Hinv,
                       double *dsdx, double *dsdy, double *dsdt)
{
    for (int64_t i = 0; i < npts; ++i) {
        const double Ht = eta[i] + h[i];
        H[i] = Ht;
        Hinv[i] = (Ht > 1e-10) ? 1.0 / Ht : 0.0;
        const double s = sigma[i];
        if (Ht < 1e-10) {
            dsdx[i] = dsdy[i] = dsdt[i] = 0.0;
            continue;
        }
        const double dHx = deta_dx[i] + dh_dx[i];
        const double dHy = deta_dy[i] + dh_dy[i];
        dsdx[i] = -(deta_dx[i] + s * dHx) / Ht;
        dsdy[i] = -(deta_dy[i] + s * dHy) / Ht;
        const double et = deta_dt ? deta_dt[i] : 0.0;
        dsdt[i] = -(et * (1.0 + s)) / Ht;
    }
}
