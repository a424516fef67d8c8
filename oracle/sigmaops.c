/*
 * ORACLE (test infrastructure only) -- SigmaCoordinateOperators
 * (src/dg/operators_3d.cpp:513-662): metric update (its own sign convention,
 * (sigma dH/dx - deta/dx) / H, :578-615), horizontal and vertical gradients,
 * sigma divergence, ALE correction and material derivative, with the
 * reference-coordinate derivatives applied as the reference does, dense
 * Kronecker matrices D_xi, D_eta, D_zeta times the element vector.
 *
 * Pinned by tests/test_oracle_sigmaops.py: derivatives of polynomials of
 * degree <= N exact, sigma at the nodes, the metric formula's sign on flat
 * and sloping columns, every operator against its closed form on
 * polynomial fields.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    int n, nd;
    double x[ORC_MAX_N];
    double *Dx, *De, *Dz; /* nd x nd */
} Ops;

static int ops_make(int order, Ops *o)
{
    const int n = order + 1, nd = n * n * n;
    double w[ORC_MAX_N], D1[ORC_MAX_N * ORC_MAX_N];
    if (orc_lgl_nodes(n, o->x, w))
        return -1;
    orc_derivative_matrix(n, o->x, D1);
    o->n = n;
    o->nd = nd;
    o->Dx = malloc(sizeof(double) * nd * nd);
    o->De = malloc(sizeof(double) * nd * nd);
    o->Dz = malloc(sizeof(double) * nd * nd);
    if (!o->Dx || !o->De || !o->Dz)
        return -1;
    orc_kron_diff(order, D1, o->Dx, o->De, o->Dz);
    return 0;
}

static void ops_free(Ops *o)
{
    free(o->Dx);
    free(o->De);
    free(o->Dz);
}

static void matvec(int nd, const double *M, const double *u, double *out)
{
    for (int r = 0; r < nd; ++r) {
        double s = 0.0;
        for (int c = 0; c < nd; ++c)
            s += M[(int64_t)r * nd + c] * u[c];
        out[r] = s;
    }
}

/* update_sigma_metrics_with_time, :578-615 (deta_dt NULL: update_sigma_metrics, :569-576) */
int orc_sigma_ops_metrics(int order, int64_t E, const double *eta, const double *h,
                          const double *deta_dx, const double *deta_dy, const double *dh_dx,
                          const double *dh_dy, const double *deta_dt, int64_t es, double *H,
                          double *Hinv, double *sigma, double *dsdx, double *dsdy, double *dsdt)
{
    const int n = order + 1, nd = n * n * n;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    for (int64_t e = 0; e < E; ++e)
        for (int d = 0; d < nd; ++d) {
            const int64_t i = e * es + d;
            H[i] = eta[i] + h[i];
            Hinv[i] = (H[i] > 1e-10) ? 1.0 / H[i] : 0.0;
            sigma[i] = 0.5 * (x[d / (n * n)] - 1.0); /* the node's zeta coordinate */
            const double dH_dx = deta_dx[i] + dh_dx[i];
            const double dH_dy = deta_dy[i] + dh_dy[i];
            dsdx[i] = (sigma[i] * dH_dx - deta_dx[i]) * Hinv[i];
            dsdy[i] = (sigma[i] * dH_dy - deta_dy[i]) * Hinv[i];
            const double et = deta_dt ? deta_dt[i] : 0.0;
            dsdt[i] = -(1.0 + sigma[i]) * et * Hinv[i];
        }
    return 0;
}

/* horizontal_gradient :517-542, vertical_gradient :544-554, sigma_divergence
 * :556-567, ale_correction :617-632, material_derivative :634-662.
 * op: 0 horizontal (out1 = dU_dx, out2 = dU_dy), 1 vertical, 2 divergence
 * (U = Hu, a = Hv, b = omega), 3 ALE (a = w_mesh), 4 material derivative
 * (a = dU_dt, b = u, c = v, f = omega, g = w_mesh). */
int orc_sigma_ops_apply(int order, int op, int64_t E, const double *U, const double *a,
                        const double *b, const double *c, const double *f, const double *g,
                        const double *dsdx, const double *dsdy, const double *Hinv, int64_t es,
                        double *out1, double *out2)
{
    Ops o;
    if (ops_make(order, &o)) {
        ops_free(&o);
        return -1;
    }
    const int nd = o.nd;
    double *t1 = malloc(sizeof(double) * nd * 3), *t2 = t1 + nd, *t3 = t2 + nd;
    for (int64_t e = 0; e < E; ++e) {
        const int64_t b0 = e * es;
        if (op == 2) {
            matvec(nd, o.Dx, U + b0, t1);
            matvec(nd, o.De, a + b0, t2);
            matvec(nd, o.Dz, b + b0, t3);
            for (int d = 0; d < nd; ++d)
                out1[b0 + d] = t1[d] + t2[d] + t3[d];
            continue;
        }
        matvec(nd, o.Dx, U + b0, t1);
        matvec(nd, o.De, U + b0, t2);
        matvec(nd, o.Dz, U + b0, t3);
        for (int d = 0; d < nd; ++d) {
            const int64_t i = b0 + d;
            if (op == 0) {
                out1[i] = t1[d] - dsdx[i] * t3[d];
                out2[i] = t2[d] - dsdy[i] * t3[d];
            } else if (op == 1) {
                out1[i] = Hinv[i] * t3[d];
            } else if (op == 3) {
                const double dU_dz = Hinv[i] * t3[d];
                out1[i] = -a[i] * dU_dz;
            } else {
                const double dU_dx = t1[d] - dsdx[i] * t3[d];
                const double dU_dy = t2[d] - dsdy[i] * t3[d];
                const double om = f[i] - g[i] * Hinv[i];
                out1[i] = a[i] + b[i] * dU_dx + c[i] * dU_dy + om * t3[d];
            }
        }
    }
    free(t1);
    ops_free(&o);
    return 0;
}
