/*
 * ORACLE (test infrastructure only) -- A19: gradient error estimator and
 * element marking; A20: nodal -> modal (Legendre) coefficients.
 *
 * Follows GradientErrorEstimator::estimate / estimate_element /
 * compute_directional_mask (src/amr/refinement.cpp:44-161),
 * AdaptiveMeshRefinement::mark_elements, can_refine, can_coarsen
 * (:488-520, :598-606), RefinementParams defaults
 * (include/amr/refinement.hpp:63-71) and RefineMask bits
 * (include/mesh/refine_mask.hpp:13-22).  The area-of-interest step of
 * mark_elements (apply_aoi_constraints, :746) is a no-op without a region
 * and is not restated.
 *
 * A20 has no counterpart in the reference (the modal Vandermonde of
 * include/dg/basis.hpp:40-42 is declared, never implemented); the
 * normalisation chosen is the orthonormal Legendre basis
 * sqrt((2p+1)/2) P_p, using the reference's Legendre recurrence
 * (legendre_poly_and_derivative, include/dg/basis_hexahedron.hpp:23-46).
 *
 * Pinned by (tests/test_oracle_estimator.py): the reference's constant and
 * i-ramp cases (tests/unit/test_amr.cpp:321-371); closed forms of
 * (max - min) * cbrt(volume) and the directional mask on separable fields
 * computed in numpy; normalisation/threshold edge cases; modal: a single
 * Legendre mode gives a unit coefficient, V * coef round trip, Parseval
 * against Gauss-Legendre quadrature, and numpy's legvander as an
 * independent Vandermonde.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* compute_directional_mask, refinement.cpp:95-161 */
static uint8_t directional_mask(int n, const double *u, double smooth)
{
    double vx = 0.0, vy = 0.0, vz = 0.0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) {
            double lo = u[k * n * n + j * n], hi = lo;
            for (int i = 1; i < n; ++i) {
                const double v = u[k * n * n + j * n + i];
                lo = dmin(lo, v);
                hi = dmax(hi, v);
            }
            vx = dmax(vx, hi - lo);
        }
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) {
            double lo = u[k * n * n + i], hi = lo;
            for (int j = 1; j < n; ++j) {
                const double v = u[k * n * n + j * n + i];
                lo = dmin(lo, v);
                hi = dmax(hi, v);
            }
            vy = dmax(vy, hi - lo);
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double lo = u[j * n + i], hi = lo;
            for (int k = 1; k < n; ++k) {
                const double v = u[k * n * n + j * n + i];
                lo = dmin(lo, v);
                hi = dmax(hi, v);
            }
            vz = dmax(vz, hi - lo);
        }
    const double mv = dmax(dmax(vx, vy), vz); /* std::max({...}) */
    if (mv < 1e-14)
        return 0;
    const double thr = smooth * mv;
    uint8_t m = 0;
    if (vx > thr)
        m |= 1;
    if (vy > thr)
        m |= 2;
    if (vz > thr)
        m |= 4;
    return m;
}

void orc_estimate(int order, int64_t E, const double *field, int64_t elem_stride,
                  const double *bounds, double refine_threshold, double coarsen_threshold,
                  double smoothness_factor, double *value, uint8_t *mask, uint8_t *refine,
                  uint8_t *coarsen)
{
    const int n = order + 1, nd = n * n * n;
    for (int64_t e = 0; e < E; ++e) {
        const double *u = field + e * elem_stride;
        double hi = u[0], lo = u[0]; /* Eigen maxCoeff / minCoeff */
        for (int d = 1; d < nd; ++d) {
            if (u[d] > hi)
                hi = u[d];
            if (u[d] < lo)
                lo = u[d];
        }
        const double *b = bounds + 6 * e;
        const double h = cbrt((b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]));
        value[e] = (hi - lo) * h;
        mask[e] = directional_mask(n, u, smoothness_factor);
        refine[e] = 0;
        coarsen[e] = 0;
    }
    double mx = 0.0;
    for (int64_t e = 0; e < E; ++e)
        mx = dmax(mx, value[e]);
    if (mx > 1e-14)
        for (int64_t e = 0; e < E; ++e) {
            value[e] /= mx;
            refine[e] = value[e] > refine_threshold;
            coarsen[e] = value[e] < coarsen_threshold;
        }
}

void orc_mark_elements(int64_t E, const uint8_t *refine, const uint8_t *coarsen,
                       const uint8_t *mask, const int *levels, int max_level, int min_level,
                       uint8_t *action, uint8_t *mask_out)
{
    for (int64_t e = 0; e < E; ++e) {
        const int *l = levels + 3 * e;
        const int can_ref = l[0] < max_level && l[1] < max_level && l[2] < max_level;
        const int can_coa = l[0] > min_level || l[1] > min_level || l[2] > min_level;
        action[e] = 0;
        mask_out[e] = 0;
        if (refine[e] && can_ref) {
            action[e] = 1;
            mask_out[e] = mask[e] ? mask[e] : 7;
        } else if (coarsen[e] && can_coa) {
            action[e] = 2;
        }
    }
}

/* ---------------- A20 ---------------- */
static int invert_gj(int m, const double *A, double *Ai)
{
    double *a = (double *)malloc(sizeof(double) * m * m);
    memcpy(a, A, sizeof(double) * m * m);
    for (int i = 0; i < m * m; ++i)
        Ai[i] = 0.0;
    for (int i = 0; i < m; ++i)
        Ai[i * m + i] = 1.0;
    for (int c = 0; c < m; ++c) {
        int piv = c;
        for (int r = c + 1; r < m; ++r)
            if (fabs(a[r * m + c]) > fabs(a[piv * m + c]))
                piv = r;
        if (a[piv * m + c] == 0.0) {
            free(a);
            return -1;
        }
        for (int k = 0; k < m && piv != c; ++k) {
            double t = a[c * m + k];
            a[c * m + k] = a[piv * m + k];
            a[piv * m + k] = t;
            t = Ai[c * m + k];
            Ai[c * m + k] = Ai[piv * m + k];
            Ai[piv * m + k] = t;
        }
        const double d = a[c * m + c];
        for (int k = 0; k < m; ++k) {
            a[c * m + k] /= d;
            Ai[c * m + k] /= d;
        }
        for (int r = 0; r < m; ++r) {
            if (r == c)
                continue;
            const double f = a[r * m + c];
            for (int k = 0; k < m; ++k) {
                a[r * m + k] -= f * a[c * m + k];
                Ai[r * m + k] -= f * Ai[c * m + k];
            }
        }
    }
    free(a);
    return 0;
}

int orc_modal_matrices(int order, double *V, double *Vinv)
{
    const int n = order + 1;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    for (int i = 0; i < n; ++i)
        for (int p = 0; p < n; ++p) {
            double L, dL;
            orc_legendre(p, x[i], &L, &dL);
            V[i * n + p] = sqrt((2.0 * p + 1.0) / 2.0) * L;
        }
    return invert_gj(n, V, Vinv);
}

int orc_modal_coeffs(int order, int64_t E, const double *u, int64_t elem_stride, double *coef)
{
    const int n = order + 1, nd = n * n * n;
    double V[ORC_MAX_N * ORC_MAX_N], Vi[ORC_MAX_N * ORC_MAX_N];
    if (orc_modal_matrices(order, V, Vi))
        return -1;
    for (int64_t e = 0; e < E; ++e) {
        const double *ue = u + e * elem_stride;
        double *ce = coef + e * nd;
        for (int r = 0; r < n; ++r)
            for (int q = 0; q < n; ++q)
                for (int p = 0; p < n; ++p) {
                    double s = 0.0;
                    for (int k = 0; k < n; ++k)
                        for (int j = 0; j < n; ++j)
                            for (int i = 0; i < n; ++i)
                                s += Vi[p * n + i] * Vi[q * n + j] * Vi[r * n + k] *
                                     ue[i + n * (j + n * k)];
                    ce[p + n * (q + n * r)] = s;
                }
    }
    return 0;
}
