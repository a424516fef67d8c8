/*
 * ORACLE (test infrastructure only) -- A1: 1D Lagrange bases and the dense
 * Kronecker differentiation matrices.
 *
 * Follows src/dg/basis_hexahedron.cpp and include/dg/basis_hexahedron.hpp of
 * the reference.
 *
 * Pinned by (tests/test_oracle_basis.py): closed-form LGL nodes for n=4,5
 * (tests/test_utils.hpp:43-72 of the reference), LGL/GL weight sums (=2),
 * agreement with numpy.polynomial.legendre Gauss nodes, exactness of D on
 * polynomials of degree <= N (tests/unit/test_basis_hexahedron.cpp:63-82),
 * and the Kronecker identity D_xi = D (x) I (x) I checked entry by entry.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* legendre_poly_and_derivative, include/dg/basis_hexahedron.hpp:23-46 */
void orc_legendre(int n, double x, double *L, double *dL)
{
    if (n == 0) {
        *L = 1.0;
        *dL = 0.0;
        return;
    }
    if (n == 1) {
        *L = x;
        *dL = 1.0;
        return;
    }
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0, p = 0.0, d = 0.0;
    for (int k = 2; k <= n; ++k) {
        p = (2.0 * k - 1.0) / k * x * p1 - (k - 1.0) / k * p0;
        d = d0 + (2.0 * k - 1.0) * p1;
        p0 = p1;
        p1 = p;
        d0 = d1;
        d1 = d;
    }
    *L = p;
    *dL = d;
}

/* compute_gauss_legendre_nodes, src/dg/basis_hexahedron.cpp:14-64 */
int orc_gl_nodes(int n, double *x, double *w)
{
    if (n < 1)
        return -1;
    if (n == 1) {
        x[0] = 0.0;
        w[0] = 2.0;
        return 0;
    }
    if (n == 2) {
        x[0] = -sqrt(1.0 / 3.0);
        x[1] = -x[0];
        w[0] = 1.0;
        w[1] = 1.0;
        return 0;
    }
    for (int i = 0; i <= (n - 1) / 2; ++i) {
        double xi = -cos((2.0 * i + 1.0) / (2.0 * n) * M_PI);
        double L, dL;
        for (int it = 0; it < 100; ++it) {
            double prev = xi;
            orc_legendre(n, xi, &L, &dL);
            xi = xi - L / dL;
            if (fabs(xi - prev) < 1e-15)
                break;
        }
        x[i] = xi;
        x[n - 1 - i] = -xi;
        orc_legendre(n, xi, &L, &dL);
        double wi = 2.0 / ((1.0 - xi * xi) * dL * dL);
        w[i] = wi;
        w[n - 1 - i] = wi;
    }
    if (n % 2 == 1) {
        double L, dL;
        orc_legendre(n, 0.0, &L, &dL);
        x[n / 2] = 0.0;
        w[n / 2] = 2.0 / (dL * dL);
    }
    return 0;
}

/*
 * compute_gauss_lobatto_nodes, src/dg/basis_hexahedron.cpp:70-163.
 * Interior nodes: Newton on q = P_{N+1} - P_{N-1} from the Chebyshev-like
 * guess at :98, stop when the step is < 1e-15 or after 100 iterations; weights
 * 2 / (N (N+1) P_N(x)^2).  For n = 3 the reference reads an uninitialised
 * P_{N-1} inside the Newton loop (:114-121 never assign it when N = 2), but the
 * resulting node is overwritten by the centre-node branch (:157-162); here
 * P_{N-1} = P_1 = x is used, which gives the same final nodes.
 */
int orc_lgl_nodes(int n, double *x, double *w)
{
    if (n < 2)
        return -1; /* std::invalid_argument, :74-76 */
    if (n == 2) {
        x[0] = -1.0;
        x[1] = 1.0;
        w[0] = 1.0;
        w[1] = 1.0;
        return 0;
    }
    x[0] = -1.0;
    x[n - 1] = 1.0;
    w[0] = 2.0 / (n * (n - 1));
    w[n - 1] = w[0];
    const int N = n - 1;
    for (int i = 1; i <= (n - 1) / 2; ++i) {
        double xi = -cos((0.25 + i) * M_PI / N - 3.0 / (8.0 * N * M_PI * (0.25 + i)));
        for (int it = 0; it < 100; ++it) {
            double prev = xi;
            /* P_{N-1}, P_N and derivatives by the three-term recurrence */
            double pm2 = 1.0, pm1 = xi, dm2 = 0.0, dm1 = 1.0;
            double pNm1 = xi, dNm1 = 1.0, pN = 0.0;
            for (int k = 2; k <= N; ++k) {
                double pk = (2.0 * k - 1.0) / k * xi * pm1 - (k - 1.0) / k * pm2;
                double dk = dm2 + (2.0 * k - 1.0) * pm1;
                if (k == N - 1) {
                    pNm1 = pk;
                    dNm1 = dk;
                }
                if (k == N)
                    pN = pk;
                pm2 = pm1;
                pm1 = pk;
                dm2 = dm1;
                dm1 = dk;
            }
            const int k = N + 1;
            double pNp1 = (2.0 * k - 1.0) / k * xi * pN - (k - 1.0) / k * pNm1;
            double dNp1 = dNm1 + (2.0 * k - 1.0) * pN;
            double q = pNp1 - pNm1;
            double dq = dNp1 - dNm1;
            xi = xi - q / dq;
            if (fabs(xi - prev) < 1e-15)
                break;
        }
        x[i] = xi;
        x[n - 1 - i] = -xi;
        double LN, dLN;
        orc_legendre(N, xi, &LN, &dLN);
        double wi = 2.0 / (N * (N + 1) * LN * LN);
        w[i] = wi;
        w[n - 1 - i] = wi;
    }
    if (n % 2 == 1) {
        x[n / 2] = 0.0;
        double LN, dLN;
        orc_legendre(N, 0.0, &LN, &dLN);
        w[n / 2] = 2.0 / (N * (N + 1) * LN * LN);
    }
    return 0;
}

/* compute_barycentric_weights, src/dg/basis_hexahedron.cpp:169-184 */
void orc_barycentric_weights(int n, const double *x, double *bw)
{
    for (int j = 0; j < n; ++j) {
        double p = 1.0;
        for (int k = 0; k < n; ++k)
            if (k != j)
                p *= (x[j] - x[k]);
        bw[j] = 1.0 / p;
    }
}

/* compute_derivative_matrix_1d, src/dg/basis_hexahedron.cpp:190-207 */
void orc_derivative_matrix(int n, const double *x, double *D)
{
    double bw[ORC_MAX_N];
    orc_barycentric_weights(n, x, bw);
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            if (i != j) {
                D[i * n + j] = (bw[j] / bw[i]) / (x[i] - x[j]);
                s += D[i * n + j];
            }
        }
        D[i * n + i] = -s;
    }
}

/*
 * build_3d_diff_matrices_from_1d, src/dg/basis_hexahedron.cpp:321-357:
 * D_xi = D (x) I (x) I etc. in the d = i + n(j + n k) ordering, dense
 * ndof x ndof, row-major here.
 */
void orc_kron_diff(int order, const double *D1, double *Dxi, double *Deta, double *Dzeta)
{
    const int n = order + 1, nd = n * n * n;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const int r = i + n * (j + n * k);
                for (int kp = 0; kp < n; ++kp)
                    for (int jp = 0; jp < n; ++jp)
                        for (int ip = 0; ip < n; ++ip) {
                            const int c = ip + n * (jp + n * kp);
                            Dxi[(int64_t)r * nd + c] =
                                D1[i * n + ip] * (j == jp ? 1.0 : 0.0) * (k == kp ? 1.0 : 0.0);
                            Deta[(int64_t)r * nd + c] =
                                (i == ip ? 1.0 : 0.0) * D1[j * n + jp] * (k == kp ? 1.0 : 0.0);
                            Dzeta[(int64_t)r * nd + c] =
                                (i == ip ? 1.0 : 0.0) * (j == jp ? 1.0 : 0.0) * D1[k * n + kp];
                        }
            }
}

void orc_default_physics(orc_physics *p)
{
    p->g = 9.81;
    p->nu_h = 1.0;
    p->nu_v = 1e-4;
    p->kappa_v_T = 1e-5;
    p->kappa_v_S = 1e-5;
    p->rho_0 = 1025.0;
    p->alpha = 2e-4;
    p->beta = 7.6e-4;
}
