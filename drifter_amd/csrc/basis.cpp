// Host-side 1D Legendre-Gauss-Lobatto basis used to build the kernel operators.
//
// The node/weight/derivative arithmetic follows the reference's
// compute_gauss_lobatto_nodes (src/dg/basis_hexahedron.cpp:70-163),
// compute_barycentric_weights (:169-184) and compute_derivative_matrix_1d
// (:190-207) operation for operation, so that the kernels apply bit-for-bit
// the same D matrix as the reference CPU path.
#include <cmath>

#include "drf_internal.hpp"

namespace drf {
namespace {

struct Leg {
    double p, dp;
};

// Three-term Legendre recurrence P_k, P'_k for k = 0..kmax; out[k] filled.
void legendre_table(int kmax, double x, Leg *out)
{
    out[0] = {1.0, 0.0};
    if (kmax >= 1)
        out[1] = {x, 1.0};
    for (int k = 2; k <= kmax; ++k) {
        out[k].p = (2.0 * k - 1.0) / k * x * out[k - 1].p - (k - 1.0) / k * out[k - 2].p;
        out[k].dp = out[k - 2].dp + (2.0 * k - 1.0) * out[k - 1].p;
    }
}

} // namespace

int lgl_nodes(int n, double *x, double *w)
{
    if (n < 2 || n > kMaxN)
        return -1;
    if (n == 2) {
        x[0] = -1.0;
        x[1] = 1.0;
        w[0] = w[1] = 1.0;
        return 0;
    }
    const int N = n - 1;
    x[0] = -1.0;
    x[N] = 1.0;
    w[0] = w[N] = 2.0 / (n * (n - 1));
    Leg tab[kMaxN + 2];
    for (int i = 1; i <= N / 2; ++i) {
        double xi = -std::cos((0.25 + i) * M_PI / N - 3.0 / (8.0 * N * M_PI * (0.25 + i)));
        for (int it = 0; it < 100; ++it) {
            const double prev = xi;
            legendre_table(N, xi, tab);
            const Leg Pm1 = tab[N - 1], P = tab[N];
            const int k = N + 1;
            const double pp1 = (2.0 * k - 1.0) / k * xi * P.p - (k - 1.0) / k * Pm1.p;
            const double dpp1 = Pm1.dp + (2.0 * k - 1.0) * P.p;
            // Newton on q = P_{N+1} - P_{N-1}, whose roots are the interior nodes.
            xi = xi - (pp1 - Pm1.p) / (dpp1 - Pm1.dp);
            if (std::fabs(xi - prev) < 1e-15)
                break;
        }
        legendre_table(N, xi, tab);
        const double wi = 2.0 / (N * (N + 1) * tab[N].p * tab[N].p);
        x[i] = xi;
        x[N - i] = -xi;
        w[i] = w[N - i] = wi;
    }
    if (n % 2 == 1) {
        legendre_table(N, 0.0, tab);
        x[N / 2] = 0.0;
        w[N / 2] = 2.0 / (N * (N + 1) * tab[N].p * tab[N].p);
    }
    return 0;
}

// Gauss-Legendre nodes: Newton on P_n from x = -cos((2i+1) pi / 2n), the
// same iteration and stopping rule as the reference's GaussQuadrature1D
// (src/dg/quadrature_3d.cpp:13-31) via compute_gauss_legendre_nodes
// (src/dg/basis_hexahedron.cpp:14-64); weights 2 / ((1 - x^2) P_n'(x)^2).
int gl_nodes(int n, double *x, double *w)
{
    if (n < 1 || n > 2 * kMaxN)
        return -1;
    if (n == 1) {
        x[0] = 0.0;
        w[0] = 2.0;
        return 0;
    }
    if (n == 2) {
        x[0] = -std::sqrt(1.0 / 3.0);
        x[1] = -x[0];
        w[0] = w[1] = 1.0;
        return 0;
    }
    Leg tab[2 * kMaxN + 2];
    for (int i = 0; i <= (n - 1) / 2; ++i) {
        double xi = -std::cos((2.0 * i + 1.0) / (2.0 * n) * M_PI);
        for (int it = 0; it < 100; ++it) {
            const double prev = xi;
            legendre_table(n, xi, tab);
            xi = xi - tab[n].p / tab[n].dp;
            if (std::fabs(xi - prev) < 1e-15)
                break;
        }
        legendre_table(n, xi, tab);
        const double wi = 2.0 / ((1.0 - xi * xi) * tab[n].dp * tab[n].dp);
        x[i] = xi;
        x[n - 1 - i] = -xi;
        w[i] = w[n - 1 - i] = wi;
    }
    if (n % 2 == 1) {
        legendre_table(n, 0.0, tab);
        x[n / 2] = 0.0;
        w[n / 2] = 2.0 / (tab[n].dp * tab[n].dp);
    }
    return 0;
}

void derivative_matrix(int n, const double *x, double *D)
{
    double bw[kMaxN];
    for (int j = 0; j < n; ++j) {
        double prod = 1.0;
        for (int k = 0; k < n; ++k)
            if (k != j)
                prod *= (x[j] - x[k]);
        bw[j] = 1.0 / prod;
    }
    for (int i = 0; i < n; ++i) {
        double diag = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const double dij = (bw[j] / bw[i]) / (x[i] - x[j]);
            D[i * n + j] = dij;
            diag += dij;
        }
        D[i * n + i] = -diag;
    }
}

} // namespace drf
