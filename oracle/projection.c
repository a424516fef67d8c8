/*
 * ORACLE (test infrastructure only) -- AMR solution transfer:
 * SolutionProjection::project_to_children / project_from_children /
 * interpolate_3d (src/amr/refinement.cpp:217-371) and
 * AdaptiveMeshRefinement::transfer_solution (:533-567).
 *
 * transfer_solution is restated as the function it is on an empty output:
 * the non-refined elements in their old order, then, parent by parent in
 * refinement-map order (ascending element index, :433-466), the 8 children of
 * an XYZ projection whatever the refine mask was (:560).  (adapt() passes the
 * solution vector itself as the output without clearing it, :471, so there the
 * result lands after the old elements.)
 *
 * Pinned by (tests/test_oracle_projection.py): polynomials of degree <= N are
 * reproduced exactly at the children's nodes (closed form); children ->
 * parent inverts parent -> children for such polynomials; child counts per
 * mask (include/mesh/refine_mask.hpp:74); the transfer order on a small list.
 */
#include "oracle.h"

#include <string.h>

/* interpolate_3d, refinement.cpp:344-371 */
static double interp3(int n, const double *x, const double *U, double xi, double eta, double zeta)
{
    double Lx[ORC_MAX_N], Ly[ORC_MAX_N], Lz[ORC_MAX_N];
    for (int i = 0; i < n; ++i) {
        Lx[i] = Ly[i] = Lz[i] = 1.0;
        for (int j = 0; j < n; ++j)
            if (j != i) {
                Lx[i] *= (xi - x[j]) / (x[i] - x[j]);
                Ly[i] *= (eta - x[j]) / (x[i] - x[j]);
                Lz[i] *= (zeta - x[j]) / (x[i] - x[j]);
            }
    }
    double r = 0.0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                r += U[k * n * n + j * n + i] * Lx[i] * Ly[j] * Lz[k];
    return r;
}

static int nchild_of(int mask) { return 1 << (((mask & 1) != 0) + ((mask & 2) != 0) + ((mask & 4) != 0)); }

int orc_project_to_children(int order, int mask, const double *U, double *children)
{
    const int n = order + 1, nd = n * n * n;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    const int rx = mask & 1, ry = (mask >> 1) & 1, rz = (mask >> 2) & 1;
    const int nx = rx ? 2 : 1, ny = ry ? 2 : 1, nz = rz ? 2 : 1;
    int c = 0;
    for (int cz = 0; cz < nz; ++cz)
        for (int cy = 0; cy < ny; ++cy)
            for (int cx = 0; cx < nx; ++cx, ++c) {
                const double x0 = rx ? (-1.0 + cx) : -1.0, x1 = rx ? (double)cx : 1.0;
                const double y0 = ry ? (-1.0 + cy) : -1.0, y1 = ry ? (double)cy : 1.0;
                const double z0 = rz ? (-1.0 + cz) : -1.0, z1 = rz ? (double)cz : 1.0;
                for (int k = 0; k < n; ++k) {
                    const double zp = z0 + 0.5 * (x[k] + 1.0) * (z1 - z0);
                    for (int j = 0; j < n; ++j) {
                        const double yp = y0 + 0.5 * (x[j] + 1.0) * (y1 - y0);
                        for (int i = 0; i < n; ++i) {
                            const double xp = x0 + 0.5 * (x[i] + 1.0) * (x1 - x0);
                            children[c * nd + k * n * n + j * n + i] = interp3(n, x, U, xp, yp, zp);
                        }
                    }
                }
            }
    return c;
}

static double clampd(double v) { return v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v); }

int orc_project_from_children(int order, int mask, const double *children, double *U)
{
    const int n = order + 1, nd = n * n * n;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    const int rx = mask & 1, ry = (mask >> 1) & 1, rz = (mask >> 2) & 1;
    const int nx = rx ? 2 : 1, ny = ry ? 2 : 1;
    for (int pk = 0; pk < n; ++pk) {
        const double zp = x[pk];
        const int cz = rz ? (zp >= 0 ? 1 : 0) : 0;
        for (int pj = 0; pj < n; ++pj) {
            const double yp = x[pj];
            const int cy = ry ? (yp >= 0 ? 1 : 0) : 0;
            for (int pi = 0; pi < n; ++pi) {
                const double xp = x[pi];
                const int cx = rx ? (xp >= 0 ? 1 : 0) : 0;
                const int c = cz * ny * nx + cy * nx + cx;
                const double x0 = rx ? (-1.0 + cx) : -1.0, x1 = rx ? (double)cx : 1.0;
                const double y0 = ry ? (-1.0 + cy) : -1.0, y1 = ry ? (double)cy : 1.0;
                const double z0 = rz ? (-1.0 + cz) : -1.0, z1 = rz ? (double)cz : 1.0;
                const double xc = clampd(2.0 * (xp - x0) / (x1 - x0) - 1.0);
                const double yc = clampd(2.0 * (yp - y0) / (y1 - y0) - 1.0);
                const double zc = clampd(2.0 * (zp - z0) / (z1 - z0) - 1.0);
                U[pk * n * n + pj * n + pi] = interp3(n, x, children + c * nd, xc, yc, zc);
            }
        }
    }
    return nchild_of(mask);
}

int64_t orc_transfer_solution(int order, int64_t E, const double *U, int64_t n_ref,
                              const int64_t *parents, double *U_new)
{
    const int nd = (order + 1) * (order + 1) * (order + 1);
    int64_t o = 0, r = 0;
    for (int64_t e = 0; e < E; ++e) {
        while (r < n_ref && parents[r] < e)
            ++r;
        if (r < n_ref && parents[r] == e)
            continue;
        memcpy(U_new + o * nd, U + e * nd, sizeof(double) * nd);
        ++o;
    }
    for (int64_t p = 0; p < n_ref; ++p) {
        if (parents[p] >= E)
            continue;
        orc_project_to_children(order, 7, U + parents[p] * nd, U_new + o * nd);
        o += 8;
    }
    return o;
}
