// Hydrostatic pressure and the pressure-gradient force in sigma
// coordinates, element by element, on gfx950.
//
// The interface of PressureGradientElement / PressureGradientSolver
// (include/physics/pressure_gradient.hpp:46-367): hydrostatic_pressure
// (:233-250), barotropic_gradient (:213-231), baroclinic_gradient_standard
// (:252-294), baroclinic_gradient_density_jacobian (:296-338) and
// full_gradient (:340-367).  The reference header does not compile
// (HexahedronBasis::lgl_weights_1d, :183, does not exist), so these follow
// its documented algorithm with the vertical (sigma) direction on the
// element's zeta index k, node d = i + n (j + n k); the standard method takes
// d eta / dx as an input where the reference leaves it at 0.
//
// A workgroup stages EPB elements of eta, rho, H and sigma in LDS; each lane
// owns a node.  The pressure (standard method) or z = eta + sigma H (density
// Jacobian) is formed in LDS first, then every lane contracts its xi-, eta-
// and zeta-lines with the 1D D in the order of the reference's dense
// Kronecker rows (contraction off: the reference's rounding).  The density
// Jacobian's bottom-up integral sums the column's Jacobians from LDS.
#include <string>

#include "drf_internal.hpp"

#pragma clang fp contract(off)

namespace drf {
namespace {

struct PgOps {
    double D[kMaxN * kMaxN]; // row-major, stride kMaxN
    double sw[kMaxN];        // w / 2
};

struct PgArgs {
    int64_t E, es;
    int method, parts;
    double g, rho0;
    const double *eta, *rho, *H, *sigma, *detadx, *detady, *dhdx, *dhdy;
    double *px, *py, *p; // p: hydrostatic pressure output (pressure-only launch)
};

template <int N>
__global__ void __launch_bounds__(256) pressure_kernel(PgOps ops, PgArgs a)
{
    constexpr int N2 = N * N, ND = N2 * N;
    constexpr int EPB = ND >= 512 ? 1 : 512 / ND;
    __shared__ double sD[N2], sw[N];
    __shared__ double sEta[EPB * ND], sRho[EPB * ND], sH[EPB * ND], sS[EPB * ND];
    __shared__ double sA[EPB * ND], sJx[EPB * ND], sJy[EPB * ND], sBx[EPB * ND], sBy[EPB * ND];
    for (int q = threadIdx.x; q < N2; q += blockDim.x)
        sD[q] = ops.D[(q / N) * kMaxN + q % N];
    for (int q = threadIdx.x; q < N; q += blockDim.x)
        sw[q] = ops.sw[q];
    const bool pressure_only = a.p != nullptr;
    const int64_t nb = (a.E + EPB - 1) / EPB;
    for (int64_t bt = blockIdx.x; bt < nb; bt += gridDim.x) {
        const int64_t e0 = bt * EPB;
        const int cnt = (a.E - e0 < EPB) ? (int)(a.E - e0) : EPB;
        __syncthreads(); // operators staged; the previous batch is done with LDS
        for (int q = threadIdx.x; q < cnt * ND; q += blockDim.x) {
            const int e = q / ND, d = q - e * ND;
            const int64_t o = (e0 + e) * a.es + d;
            sRho[q] = a.rho[o];
            sH[q] = a.H[o];
            if (!pressure_only) {
                sEta[q] = a.eta[o];
                sS[q] = a.sigma[o];
            }
        }
        __syncthreads();
        // pressure (hydrostatic_pressure, summed from k up) or z = eta + sigma H
        for (int q = threadIdx.x; q < cnt * ND; q += blockDim.x) {
            const int e = q / ND, d = q - e * ND;
            const int c = d % N2, k = d / N2;
            if (pressure_only || a.method == 0) {
                double s = 0.0;
                for (int kk = k; kk < N; ++kk) {
                    const int m = e * ND + c + N2 * kk;
                    s += sw[kk] * a.g * sRho[m] * sH[m];
                }
                sA[q] = s;
            } else {
                sA[q] = sEta[q] + sS[q] * sH[q];
            }
        }
        __syncthreads();
        if (pressure_only) {
            for (int q = threadIdx.x; q < cnt * ND; q += blockDim.x) {
                const int e = q / ND, d = q - e * ND;
                a.p[(e0 + e) * a.es + d] = sA[q];
            }
            continue;
        }
        for (int q = threadIdx.x; q < cnt * ND; q += blockDim.x) {
            const int e = q / ND, d = q - e * ND;
            const int i = d % N, j = (d / N) % N, k = d / N2;
            const double *A = &sA[e * ND], *R = &sRho[e * ND], *Et = &sEta[e * ND];
            const int64_t o = (e0 + e) * a.es + d;
            double ox = 0.0, oy = 0.0;
            if (a.parts & 1) { // barotropic: -g H grad(eta)
                double gx = 0.0, gy = 0.0;
#pragma unroll
                for (int m = 0; m < N; ++m)
                    gx += sD[i * N + m] * Et[m + N * (j + N * k)];
#pragma unroll
                for (int m = 0; m < N; ++m)
                    gy += sD[j * N + m] * Et[i + N * (m + N * k)];
                ox = -a.g * sH[q] * gx;
                oy = -a.g * sH[q] * gy;
            }
            // derivatives of the pressure (standard) or of z (Jacobian)
            double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
            for (int m = 0; m < N; ++m)
                ax += sD[i * N + m] * A[m + N * (j + N * k)];
#pragma unroll
            for (int m = 0; m < N; ++m)
                ay += sD[j * N + m] * A[i + N * (m + N * k)];
#pragma unroll
            for (int m = 0; m < N; ++m)
                az += sD[k * N + m] * A[i + N * (j + N * m)];
            if (a.method == 0) {
                const double ex = a.detadx ? a.detadx[o] : 0.0;
                const double ey = a.detady ? a.detady[o] : 0.0;
                const double dHx = ex + (a.dhdx ? a.dhdx[o] : 0.0);
                const double dHy = ey + (a.dhdy ? a.dhdy[o] : 0.0);
                const double sx = (sS[q] * dHx - ex) / sH[q];
                const double sy = (sS[q] * dHy - ey) / sH[q];
                const double px = ax + az * sx;
                const double py = ay + az * sy;
                if (a.parts & 2) {
                    ox += -px / a.rho0;
                    oy += -py / a.rho0;
                }
                a.px[o] = ox;
                a.py[o] = oy;
            } else {
                double rx = 0.0, ry = 0.0, rz = 0.0;
#pragma unroll
                for (int m = 0; m < N; ++m)
                    rx += sD[i * N + m] * R[m + N * (j + N * k)];
#pragma unroll
                for (int m = 0; m < N; ++m)
                    ry += sD[j * N + m] * R[i + N * (m + N * k)];
#pragma unroll
                for (int m = 0; m < N; ++m)
                    rz += sD[k * N + m] * R[i + N * (j + N * m)];
                sJx[q] = rx * az - rz * ax;
                sJy[q] = ry * az - rz * ay;
                // barotropic part parked until the column integral is known
                sBx[q] = ox;
                sBy[q] = oy;
            }
        }
        if (a.method == 0)
            continue;
        __syncthreads();
        for (int q = threadIdx.x; q < cnt * ND; q += blockDim.x) {
            const int e = q / ND, d = q - e * ND;
            const int c = d % N2, k = d / N2;
            double ix = 0.0, iy = 0.0;
            for (int kk = 0; kk <= k; ++kk) {
                const int m = e * ND + c + N2 * kk;
                ix += sw[kk] * sJx[m];
                iy += sw[kk] * sJy[m];
            }
            double ox = sBx[q], oy = sBy[q];
            if (a.parts & 2) {
                ox += -a.g / a.rho0 * ix;
                oy += -a.g / a.rho0 * iy;
            }
            const int64_t o = (e0 + e) * a.es + d;
            a.px[o] = ox;
            a.py[o] = oy;
        }
    }
}

PgOps pg_ops(int n)
{
    PgOps o{};
    double x[kMaxN], w[kMaxN], D[kMaxN * kMaxN];
    lgl_nodes(n, x, w);
    derivative_matrix(n, x, D);
    for (int r = 0; r < n; ++r) {
        o.sw[r] = 0.5 * w[r];
        for (int c = 0; c < n; ++c)
            o.D[r * kMaxN + c] = D[r * n + c];
    }
    return o;
}

hipError_t launch_pressure(int n, const PgOps &ops, const PgArgs &a, hipStream_t s)
{
    const int nd = n * n * n, epb = nd >= 512 ? 1 : 512 / nd;
    const int64_t nb = (a.E + epb - 1) / epb;
    const dim3 grid((unsigned)(nb < 8192 ? nb : 8192)), blk(256);
    switch (n) {
    case 2: hipLaunchKernelGGL(pressure_kernel<2>, grid, blk, 0, s, ops, a); break;
    case 3: hipLaunchKernelGGL(pressure_kernel<3>, grid, blk, 0, s, ops, a); break;
    case 4: hipLaunchKernelGGL(pressure_kernel<4>, grid, blk, 0, s, ops, a); break;
    case 5: hipLaunchKernelGGL(pressure_kernel<5>, grid, blk, 0, s, ops, a); break;
    case 6: hipLaunchKernelGGL(pressure_kernel<6>, grid, blk, 0, s, ops, a); break;
    case 7: hipLaunchKernelGGL(pressure_kernel<7>, grid, blk, 0, s, ops, a); break;
    case 8: hipLaunchKernelGGL(pressure_kernel<8>, grid, blk, 0, s, ops, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace
} // namespace drf

using namespace drf;

extern "C" {

drf_status drf_hydrostatic_pressure_dev(int order, int64_t n_elem, double g, const void *rho,
                                        const void *H, int64_t elem_stride, void *p,
                                        void *stream)
{
    if (order < 1 || order + 1 > 8)
        return fail(DRF_ERR_UNSUPPORTED, "drf_hydrostatic_pressure_dev: order out of range");
    const int n = order + 1;
    if (n_elem < 0 || elem_stride < (int64_t)n * n * n)
        return fail(DRF_ERR_ARG, "drf_hydrostatic_pressure_dev: bad sizes");
    if (n_elem == 0)
        return DRF_OK;
    if (!rho || !H || !p)
        return fail(DRF_ERR_ARG, "drf_hydrostatic_pressure_dev: null pointer");
    PgArgs a{};
    a.E = n_elem;
    a.es = elem_stride;
    a.g = g;
    a.rho = (const double *)rho;
    a.H = (const double *)H;
    a.p = (double *)p;
    DRF_HIP_CHECK(launch_pressure(n, pg_ops(n), a, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_pressure_gradient_dev(int order, int64_t n_elem, int method, int parts, double g,
                                     double rho0, const void *eta, const void *rho, const void *H,
                                     const void *sigma, const void *deta_dx, const void *deta_dy,
                                     const void *dh_dx, const void *dh_dy, int64_t elem_stride,
                                     void *pg_x, void *pg_y, void *stream)
{
    if (order < 1 || order + 1 > 8)
        return fail(DRF_ERR_UNSUPPORTED, "drf_pressure_gradient_dev: order out of range");
    const int n = order + 1;
    if (n_elem < 0 || elem_stride < (int64_t)n * n * n)
        return fail(DRF_ERR_ARG, "drf_pressure_gradient_dev: bad sizes");
    if (method != DRF_PG_STANDARD && method != DRF_PG_DENSITY_JACOBIAN)
        return fail(DRF_ERR_ARG, "drf_pressure_gradient_dev: unknown method");
    if (parts < 1 || parts > 3)
        return fail(DRF_ERR_ARG, "drf_pressure_gradient_dev: parts must be 1, 2 or 3");
    if (n_elem == 0)
        return DRF_OK;
    if (!eta || !rho || !H || !sigma || !pg_x || !pg_y)
        return fail(DRF_ERR_ARG, "drf_pressure_gradient_dev: null pointer");
    PgArgs a{};
    a.E = n_elem;
    a.es = elem_stride;
    a.method = method;
    a.parts = parts;
    a.g = g;
    a.rho0 = rho0;
    a.eta = (const double *)eta;
    a.rho = (const double *)rho;
    a.H = (const double *)H;
    a.sigma = (const double *)sigma;
    a.detadx = (const double *)deta_dx;
    a.detady = (const double *)deta_dy;
    a.dhdx = (const double *)dh_dx;
    a.dhdy = (const double *)dh_dy;
    a.px = (double *)pg_x;
    a.py = (double *)pg_y;
    DRF_HIP_CHECK(launch_pressure(n, pg_ops(n), a, (hipStream_t)stream));
    return DRF_OK;
}

} // extern "C"
