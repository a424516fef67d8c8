/*
 * ORACLE (test infrastructure only) -- A9/A10: explicit Runge-Kutta steps on
 * the packed state and the CFL time-step controller.
 *
 * Follows src/solver/time_stepper.cpp:12-201 (schemes), :261-268
 * (AdaptiveTimeController::compute_dt), include/solver/time_stepper.hpp:56-61
 * (LSRK45 coefficients) and src/solver/simulation_driver.cpp:153-168
 * (compute_max_wave_speed).
 *
 * Pinned by (tests/test_oracle_stepper.py): the exact one-step polynomials of
 * each scheme on a linear system du/dt = L u (RK4: 1 + z + z^2/2 + z^3/6 +
 * z^4/24, SSPRK3: 1 + z + z^2/2 + z^3/6, ...) evaluated through a right-hand
 * side with a known linear action, the reference's convergence-ratio tests
 * (tests/unit/test_time_stepper.cpp:17-164), dt clamps (:216-233) and a
 * brute-force numpy max over elements for the wave speed.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int rhs_packed(int order, int64_t E, const double *U, const double *h, const double *f,
                      const orc_physics *p, double *k)
{
    return orc_rhs(order, E, U, h, f, p, k);
}

int orc_rk_step(int scheme, int order, int64_t E, double t, double dt, double *U,
                const double *h, const double *f, const orc_physics *p)
{
    (void)t; /* the reference RHS does not depend on t */
    const int n = order + 1;
    const int64_t len = E * 5 * (int64_t)n * n * n;
    double *w = (double *)malloc(sizeof(double) * len * 7);
    if (!w)
        return -1;
    double *k1 = w, *k2 = w + len, *k3 = w + 2 * len, *k4 = w + 3 * len, *Ut = w + 4 * len;
    double *k5 = w + 5 * len, *k6 = w + 6 * len;
    int rc = 0;
    switch (scheme) {
    case ORC_EULER: /* :14-18 */
        rc |= rhs_packed(order, E, U, h, f, p, k1);
        for (int64_t i = 0; i < len; ++i)
            U[i] += dt * k1[i];
        break;
    case ORC_RK2: /* :33-47 */
        rc |= rhs_packed(order, E, U, h, f, p, k1);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + 0.5 * dt * k1[i];
        rc |= rhs_packed(order, E, Ut, h, f, p, k2);
        for (int64_t i = 0; i < len; ++i)
            U[i] += dt * k2[i];
        break;
    case ORC_SSPRK3: { /* :64-80, U1 in Ut, U2 in k2 */
        rc |= rhs_packed(order, E, U, h, f, p, k1);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + dt * k1[i];
        rc |= rhs_packed(order, E, Ut, h, f, p, k1);
        for (int64_t i = 0; i < len; ++i)
            k2[i] = 0.75 * U[i] + 0.25 * (Ut[i] + dt * k1[i]);
        rc |= rhs_packed(order, E, k2, h, f, p, k1);
        for (int64_t i = 0; i < len; ++i)
            U[i] = (1.0 / 3.0) * U[i] + (2.0 / 3.0) * (k2[i] + dt * k1[i]);
        break;
    }
    case ORC_RK4: /* :88-112 */
        rc |= rhs_packed(order, E, U, h, f, p, k1);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + 0.5 * dt * k1[i];
        rc |= rhs_packed(order, E, Ut, h, f, p, k2);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + 0.5 * dt * k2[i];
        rc |= rhs_packed(order, E, Ut, h, f, p, k3);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + dt * k3[i];
        rc |= rhs_packed(order, E, Ut, h, f, p, k4);
        for (int64_t i = 0; i < len; ++i)
            U[i] += (dt / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
        break;
    case ORC_LSRK45: { /* :126-143, include/solver/time_stepper.hpp:56-61 */
        static const double A[5] = {0.0, -0.4178904745, -1.192151694643, -1.697784692471,
                                    -1.514183444257};
        static const double B[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869,
                                    0.6994504559488, 0.1530572479681};
        double *res = k2;
        memset(res, 0, sizeof(double) * len);
        for (int s = 0; s < 5; ++s) {
            rc |= rhs_packed(order, E, U, h, f, p, k1);
            for (int64_t i = 0; i < len; ++i) {
                res[i] = A[s] * res[i] + dt * k1[i];
                U[i] += B[s] * res[i];
            }
        }
        break;
    }
    case ORC_DP45: /* :151-189 */
        rc |= rhs_packed(order, E, U, h, f, p, k1);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + (dt / 5.0) * k1[i];
        rc |= rhs_packed(order, E, Ut, h, f, p, k2);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + dt * (3.0 / 40.0 * k1[i] + 9.0 / 40.0 * k2[i]);
        rc |= rhs_packed(order, E, Ut, h, f, p, k3);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + dt * (44.0 / 45.0 * k1[i] - 56.0 / 15.0 * k2[i] + 32.0 / 9.0 * k3[i]);
        rc |= rhs_packed(order, E, Ut, h, f, p, k4);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + dt * (19372.0 / 6561.0 * k1[i] - 25360.0 / 2187.0 * k2[i] +
                                 64448.0 / 6561.0 * k3[i] - 212.0 / 729.0 * k4[i]);
        rc |= rhs_packed(order, E, Ut, h, f, p, k5);
        for (int64_t i = 0; i < len; ++i)
            Ut[i] = U[i] + dt * (9017.0 / 3168.0 * k1[i] - 355.0 / 33.0 * k2[i] +
                                 46732.0 / 5247.0 * k3[i] + 49.0 / 176.0 * k4[i] -
                                 5103.0 / 18656.0 * k5[i]);
        rc |= rhs_packed(order, E, Ut, h, f, p, k6);
        for (int64_t i = 0; i < len; ++i)
            U[i] += dt * (35.0 / 384.0 * k1[i] + 500.0 / 1113.0 * k3[i] + 125.0 / 192.0 * k4[i] -
                          2187.0 / 6784.0 * k5[i] + 11.0 / 84.0 * k6[i]);
        break;
    default:
        rc = -1;
    }
    free(w);
    return rc;
}

/*
 * compute_max_wave_speed, src/solver/simulation_driver.cpp:153-168: per
 * element sqrt(9.81 * max H) + max(max|u|, max|v|), then the max over
 * elements.  Arrays are E * elem_stride long; the first ndof entries of each
 * element block are used.
 */
double orc_max_wave_speed(int order, int64_t E, const double *H, const double *u,
                          const double *v, int64_t stride)
{
    const int n = order + 1, nd = n * n * n;
    double c = 0.0;
    for (int64_t e = 0; e < E; ++e) {
        const double *He = H + e * stride, *ue = u + e * stride, *ve = v + e * stride;
        double hm = He[0], um = fabs(ue[0]), vm = fabs(ve[0]);
        for (int i = 1; i < nd; ++i) {
            if (He[i] > hm)
                hm = He[i];
            if (fabs(ue[i]) > um)
                um = fabs(ue[i]);
            if (fabs(ve[i]) > vm)
                vm = fabs(ve[i]);
        }
        const double ce = sqrt(9.81 * hm) + (um > vm ? um : vm);
        if (ce > c)
            c = ce;
    }
    return c;
}

/* AdaptiveTimeController::compute_dt, src/solver/time_stepper.cpp:261-268 */
double orc_compute_dt(double c, double dx_min, double cfl, double dt_min, double dt_max)
{
    if (c < 1e-14)
        return dt_max;
    double dt = cfl * dx_min / c;
    if (dt < dt_min)
        dt = dt_min;
    if (dt > dt_max)
        dt = dt_max;
    return dt;
}
