// Vertical velocity diagnosis on gfx950: VerticalVelocityDiagnosis
// (src/physics/vertical_velocity.cpp:55-174) and ColumnVerticalVelocity
// (:180-275).  Column integrals run one column per lane (the column's nodes
// are consecutive DOFs, node_index_3d), pointwise conversions one node per
// lane; FP64 with contraction off, so every value rounds as the reference's
// sequence of operations.  Not used by the right-hand side (omega = 0 there,
// src/solver/simulation_driver.cpp:222).
#include <cmath>

#include "drf_internal.hpp"

#pragma clang fp contract(off)

namespace drf {
namespace {

struct Levels {
    double s[kMaxN]; // LGL nodes (diagnose_omega) or sigma nodes (diagnose_column)
};

// diagnose_omega (:55-97), with div = A (+ B: diagnose_omega_from_velocity,
// :99-108) formed first.  The column length N is a template argument, so the
// loop unrolls and the node coordinates x.s[k] are read with constant
// indices (kernel-argument registers): indexed at run time they compiled to
// memory loads in every iteration and the kernel ran at 2 TB/s.
template <int N, bool HAS_B>
__global__ void __launch_bounds__(256)
    omega_kernel(int64_t ncol_total, Levels x, const double *__restrict__ A,
                 const double *__restrict__ B, const double *__restrict__ H, int64_t es,
                 double *__restrict__ omega)
{
    constexpr int n2 = N * N;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncol_total;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = c / n2;
        const int i = (int)(c - e * n2);
        const int64_t o = e * es + (int64_t)i * N;
        double dv[N], hv[N], ov[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            dv[k] = HAS_B ? A[o + k] + B[o + k] : A[o + k];
            hv[k] = H[o + k];
        }
        ov[0] = 0.0; // omega(-1) = 0
        double integral = 0.0;
#pragma unroll
        for (int k = 1; k < N; ++k) {
            const double H_avg = 0.5 * (hv[k - 1] + hv[k]);
            const double dsigma = (x.s[k] - x.s[k - 1]) * 0.5;
            const double div_avg = 0.5 * (dv[k - 1] + dv[k]);
            integral += div_avg * dsigma;
            ov[k] = (H_avg > 1e-10) ? -integral / H_avg : 0.0;
        }
#pragma unroll
        for (int k = 0; k < N; ++k)
            omega[o + k] = ov[k];
    }
}

// ColumnVerticalVelocity::diagnose_column (:246-275), columns of N levels
template <int N>
__global__ void __launch_bounds__(256)
    column_kernel(int64_t ncol, Levels sg, const double *__restrict__ div,
                  const double *__restrict__ H, const double *__restrict__ ob,
                  double *__restrict__ omega)
{
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncol;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = c * N;
        double dv[N], hv[N], ov[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            dv[k] = div[o + k];
            hv[k] = H[o + k];
        }
        const double b = ob ? ob[c] : 0.0;
        ov[0] = b;
        double integral = 0.0;
#pragma unroll
        for (int k = 1; k < N; ++k) {
            const double dsigma = sg.s[k] - sg.s[k - 1];
            const double div_avg = 0.5 * (dv[k - 1] + dv[k]);
            integral += div_avg * dsigma;
            const double H_avg = 0.5 * (hv[k - 1] + hv[k]);
            ov[k] = (H_avg > 1e-10) ? b - integral / H_avg : b;
        }
#pragma unroll
        for (int k = 0; k < N; ++k)
            omega[o + k] = ov[k];
    }
}

unsigned grid_for(int64_t items)
{
    const int64_t b = (items + 255) / 256;
    return (unsigned)(b < 65536 ? b : 65536);
}

template <int N>
void launch_omega_n(int64_t ncol, const Levels &x, const double *A, const double *B,
                    const double *H, int64_t es, double *omega, hipStream_t s)
{
    if (B)
        hipLaunchKernelGGL((omega_kernel<N, true>), dim3(grid_for(ncol)), dim3(256), 0, s, ncol, x,
                           A, B, H, es, omega);
    else
        hipLaunchKernelGGL((omega_kernel<N, false>), dim3(grid_for(ncol)), dim3(256), 0, s, ncol,
                           x, A, B, H, es, omega);
}

hipError_t launch_omega(int n, int64_t ncol, const Levels &x, const double *A, const double *B,
                        const double *H, int64_t es, double *omega, hipStream_t s)
{
    switch (n) {
    case 2: launch_omega_n<2>(ncol, x, A, B, H, es, omega, s); break;
    case 3: launch_omega_n<3>(ncol, x, A, B, H, es, omega, s); break;
    case 4: launch_omega_n<4>(ncol, x, A, B, H, es, omega, s); break;
    case 5: launch_omega_n<5>(ncol, x, A, B, H, es, omega, s); break;
    case 6: launch_omega_n<6>(ncol, x, A, B, H, es, omega, s); break;
    case 7: launch_omega_n<7>(ncol, x, A, B, H, es, omega, s); break;
    case 8: launch_omega_n<8>(ncol, x, A, B, H, es, omega, s); break;
    case 9: launch_omega_n<9>(ncol, x, A, B, H, es, omega, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_column(int n, int64_t ncol, const Levels &sg, const double *div,
                         const double *H, const double *ob, double *omega, hipStream_t s)
{
    const dim3 g(grid_for(ncol)), b(256);
    switch (n) {
    case 2: hipLaunchKernelGGL(column_kernel<2>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    case 3: hipLaunchKernelGGL(column_kernel<3>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    case 4: hipLaunchKernelGGL(column_kernel<4>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    case 5: hipLaunchKernelGGL(column_kernel<5>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    case 6: hipLaunchKernelGGL(column_kernel<6>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    case 7: hipLaunchKernelGGL(column_kernel<7>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    case 8: hipLaunchKernelGGL(column_kernel<8>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    case 9: hipLaunchKernelGGL(column_kernel<9>, g, b, 0, s, ncol, sg, div, H, ob, omega); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// omega_to_physical_w (:110-121) / omega_to_w_simple (:123-143)
__global__ void __launch_bounds__(256)
    w_kernel(int64_t npts, int simple, const double *__restrict__ om, const double *__restrict__ u,
             const double *__restrict__ v, const double *__restrict__ p3,
             const double *__restrict__ p4, const double *__restrict__ p5,
             const double *__restrict__ p6, const double *__restrict__ p7,
             const double *__restrict__ p8, double *__restrict__ w)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npts;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (simple) { // p3 eta, p4 h, p5 sigma, p6 deta_dx, p7 deta_dy, p8 deta_dt
            const double H = p3[i] + p4[i];
            const double dz_dx = p6[i] * (1.0 + p5[i]);
            const double dz_dy = p7[i] * (1.0 + p5[i]);
            const double dz_dt = p8[i] * (1.0 + p5[i]);
            w[i] = H * om[i] + u[i] * dz_dx + v[i] * dz_dy + dz_dt;
        } else { // p3 H, p4 dz_dx, p5 dz_dy, p6 dz_dt
            w[i] = p3[i] * om[i] + u[i] * p4[i] + v[i] * p5[i] + p6[i];
        }
    }
}

// continuity_residual (:161-174) per element, DOFs in order
__global__ void __launch_bounds__(256)
    residual_kernel(int nd, int64_t E, const double *__restrict__ div,
                    const double *__restrict__ dHw, int64_t es, double *__restrict__ out)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        double r2 = 0.0;
        for (int d = 0; d < nd; ++d) {
            const double r = div[e * es + d] + dHw[e * es + d];
            r2 += r * r;
        }
        out[e] = sqrt(r2 / nd);
    }
}

// ColumnVerticalVelocity(order) sigma nodes (:180-244): hard-coded LGL
// nodes for 3..5 levels, uniform spacing above (the reference's fallback)
bool column_levels(int n, double *s)
{
    if (n < 2 || n > kMaxN)
        return false;
    if (n == 2) {
        s[0] = -1.0;
        s[1] = 0.0;
        return true;
    }
    double z[kMaxN];
    if (n == 3) {
        z[0] = -1.0, z[1] = 0.0, z[2] = 1.0;
    } else if (n == 4) {
        z[0] = -1.0, z[1] = -std::sqrt(1.0 / 5.0), z[2] = std::sqrt(1.0 / 5.0), z[3] = 1.0;
    } else if (n == 5) {
        z[0] = -1.0, z[1] = -std::sqrt(3.0 / 7.0), z[2] = 0.0, z[3] = std::sqrt(3.0 / 7.0),
        z[4] = 1.0;
    } else {
        for (int k = 0; k < n; ++k)
            z[k] = -1.0 + 2.0 * k / (n - 1);
    }
    for (int k = 0; k < n; ++k)
        s[k] = 0.5 * (z[k] - 1.0);
    return true;
}

} // namespace
} // namespace drf

using namespace drf;

#define DRF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond))                                                                         \
            return ::drf::fail((code), (msg));                                               \
    } while (0)

namespace {

drf_status omega_common(const char *who, int order, int64_t E, const void *A, const void *B,
                        const void *H, int64_t es, void *omega, void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                std::string(who) + ": order out of range");
    const int n = order + 1;
    DRF_REQUIRE(E >= 0 && es >= (int64_t)n * n * n, DRF_ERR_ARG,
                std::string(who) + ": bad sizes (elem_stride must be >= (N+1)^3)");
    if (E == 0)
        return DRF_OK;
    DRF_REQUIRE(A && H && omega, DRF_ERR_ARG, std::string(who) + ": null pointer");
    Levels x{};
    double w[kMaxN];
    DRF_REQUIRE(lgl_nodes(n, x.s, w) == 0, DRF_ERR_ARG, std::string(who) + ": LGL nodes failed");
    const int64_t ncol = E * n * n;
    DRF_HIP_CHECK(launch_omega(n, ncol, x, (const double *)A, (const double *)B,
                               (const double *)H, es, (double *)omega, (hipStream_t)stream));
    return DRF_OK;
}

} // namespace

extern "C" {

drf_status drf_diagnose_omega_dev(int order, int64_t n_elem, const void *div, const void *H,
                                  int64_t elem_stride, void *omega, void *stream)
{
    return omega_common("drf_diagnose_omega_dev", order, n_elem, div, nullptr, H, elem_stride,
                        omega, stream);
}

drf_status drf_diagnose_omega_from_velocity_dev(int order, int64_t n_elem, const void *dHu_dx,
                                                const void *dHv_dy, const void *H,
                                                int64_t elem_stride, void *omega, void *stream)
{
    if (n_elem > 0 && !dHv_dy)
        return fail(DRF_ERR_ARG, "drf_diagnose_omega_from_velocity_dev: null pointer");
    return omega_common("drf_diagnose_omega_from_velocity_dev", order, n_elem, dHu_dx, dHv_dy, H,
                        elem_stride, omega, stream);
}

drf_status drf_omega_to_w_dev(int64_t npts, const void *omega, const void *u, const void *v,
                              const void *H, const void *dz_dx, const void *dz_dy,
                              const void *dz_dt, void *w, void *stream)
{
    DRF_REQUIRE(npts >= 0, DRF_ERR_ARG, "drf_omega_to_w_dev: negative npts");
    if (npts == 0)
        return DRF_OK;
    DRF_REQUIRE(omega && u && v && H && dz_dx && dz_dy && dz_dt && w, DRF_ERR_ARG,
                "drf_omega_to_w_dev: null pointer");
    hipLaunchKernelGGL(w_kernel, dim3(grid_for(npts)), dim3(256), 0, (hipStream_t)stream, npts, 0,
                       (const double *)omega, (const double *)u, (const double *)v,
                       (const double *)H, (const double *)dz_dx, (const double *)dz_dy,
                       (const double *)dz_dt, nullptr, nullptr, (double *)w);
    DRF_HIP_CHECK(hipGetLastError());
    return DRF_OK;
}

drf_status drf_omega_to_w_simple_dev(int64_t npts, const void *omega, const void *u,
                                     const void *v, const void *eta, const void *h,
                                     const void *sigma, const void *deta_dx,
                                     const void *deta_dy, const void *deta_dt, void *w,
                                     void *stream)
{
    DRF_REQUIRE(npts >= 0, DRF_ERR_ARG, "drf_omega_to_w_simple_dev: negative npts");
    if (npts == 0)
        return DRF_OK;
    DRF_REQUIRE(omega && u && v && eta && h && sigma && deta_dx && deta_dy && deta_dt && w,
                DRF_ERR_ARG, "drf_omega_to_w_simple_dev: null pointer");
    hipLaunchKernelGGL(w_kernel, dim3(grid_for(npts)), dim3(256), 0, (hipStream_t)stream, npts, 1,
                       (const double *)omega, (const double *)u, (const double *)v,
                       (const double *)eta, (const double *)h, (const double *)sigma,
                       (const double *)deta_dx, (const double *)deta_dy,
                       (const double *)deta_dt, (double *)w);
    DRF_HIP_CHECK(hipGetLastError());
    return DRF_OK;
}

double drf_surface_omega(double deta_dt, double u_surf, double v_surf, double deta_dx,
                         double deta_dy, double H)
{
    // VerticalVelocityDiagnosis::surface_omega, :145-152
    if (H < 1e-10)
        return 0.0;
    return (deta_dt + u_surf * deta_dx + v_surf * deta_dy) / H;
}

drf_status drf_column_sigma_nodes(int order, double *sigma)
{
    DRF_REQUIRE(sigma, DRF_ERR_ARG, "drf_column_sigma_nodes: null output");
    DRF_REQUIRE(column_levels(order + 1, sigma), DRF_ERR_UNSUPPORTED,
                "drf_column_sigma_nodes: order out of range");
    return DRF_OK;
}

drf_status drf_diagnose_column_dev(int order, int64_t ncol, const void *div, const void *H,
                                   const void *omega_bottom, void *omega, void *stream)
{
    Levels sg{};
    DRF_REQUIRE(column_levels(order + 1, sg.s), DRF_ERR_UNSUPPORTED,
                "drf_diagnose_column_dev: order out of range");
    DRF_REQUIRE(ncol >= 0, DRF_ERR_ARG, "drf_diagnose_column_dev: negative count");
    if (ncol == 0)
        return DRF_OK;
    DRF_REQUIRE(div && H && omega, DRF_ERR_ARG, "drf_diagnose_column_dev: null pointer");
    DRF_HIP_CHECK(launch_column(order + 1, ncol, sg, (const double *)div, (const double *)H,
                                (const double *)omega_bottom, (double *)omega,
                                (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_continuity_residual_dev(int order, int64_t n_elem, const void *div,
                                       const void *dHomega_dsigma, int64_t elem_stride,
                                       void *residual, void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_continuity_residual_dev: order out of range");
    const int n = order + 1, nd = n * n * n;
    DRF_REQUIRE(n_elem >= 0 && elem_stride >= nd, DRF_ERR_ARG,
                "drf_continuity_residual_dev: bad sizes");
    if (n_elem == 0)
        return DRF_OK;
    DRF_REQUIRE(div && dHomega_dsigma && residual, DRF_ERR_ARG,
                "drf_continuity_residual_dev: null pointer");
    hipLaunchKernelGGL(residual_kernel, dim3(grid_for(n_elem)), dim3(256), 0, (hipStream_t)stream,
                       nd, n_elem, (const double *)div, (const double *)dHomega_dsigma,
                       elem_stride, (double *)residual);
    DRF_HIP_CHECK(hipGetLastError());
    return DRF_OK;
}

} // extern "C"
