// Sigma-coordinate metric terms on gfx950: SigmaMetrics::update
// (include/mesh/sigma_coordinate.hpp:340-354) with the tested
// SigmaCoordinate formulas (:67-103), one node per lane, FP64.  An opt-in
// component: the right-hand side never calls it (the reference RHS works in
// reference coordinates without metrics).
#include "drf_internal.hpp"

// Each metric is a handful of adds, one product and a division per node;
// contraction is off so the terms round exactly as written.
#pragma clang fp contract(off)

namespace drf {
namespace {

__global__ void __launch_bounds__(256)
    sigma_metrics_kernel(int64_t npts, const double *__restrict__ sigma,
                         const double *__restrict__ eta, const double *__restrict__ h,
                         const double *__restrict__ detadx, const double *__restrict__ detady,
                         const double *__restrict__ dhdx, const double *__restrict__ dhdy,
                         const double *__restrict__ detadt, double *__restrict__ H,
                         double *__restrict__ Hinv, double *__restrict__ dsdx,
                         double *__restrict__ dsdy, double *__restrict__ dsdt)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npts;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double Ht = eta[i] + h[i];
        H[i] = Ht;
        Hinv[i] = (Ht > 1e-10) ? 1.0 / Ht : 0.0;
        double x = 0.0, y = 0.0, t = 0.0;
        if (!(Ht < 1e-10)) {
            const double s = sigma[i];
            const double ex = detadx[i], ey = detady[i];
            x = -(ex + s * (ex + dhdx[i])) / Ht;
            y = -(ey + s * (ey + dhdy[i])) / Ht;
            const double et = detadt ? detadt[i] : 0.0;
            t = -(et * (1.0 + s)) / Ht;
        }
        dsdx[i] = x;
        dsdy[i] = y;
        dsdt[i] = t;
    }
}

} // namespace
} // namespace drf

using namespace drf;

extern "C" drf_status drf_sigma_metrics_dev(int64_t npts, const void *sigma, const void *eta,
                                            const void *h, const void *deta_dx,
                                            const void *deta_dy, const void *dh_dx,
                                            const void *dh_dy, const void *deta_dt, void *H,
                                            void *Hinv, void *dsigma_dx, void *dsigma_dy,
                                            void *dsigma_dt, void *stream)
{
    if (npts < 0)
        return fail(DRF_ERR_ARG, "drf_sigma_metrics_dev: negative npts");
    if (npts == 0)
        return DRF_OK;
    if (!sigma || !eta || !h || !deta_dx || !deta_dy || !dh_dx || !dh_dy || !H || !Hinv ||
        !dsigma_dx || !dsigma_dy || !dsigma_dt)
        return fail(DRF_ERR_ARG, "drf_sigma_metrics_dev: null pointer");
    const int64_t blocks = (npts + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL(sigma_metrics_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, npts,
                       (const double *)sigma, (const double *)eta, (const double *)h,
                       (const double *)deta_dx, (const double *)deta_dy, (const double *)dh_dx,
                       (const double *)dh_dy, (const double *)deta_dt, (double *)H,
                       (double *)Hinv, (double *)dsigma_dx, (double *)dsigma_dy,
                       (double *)dsigma_dt);
    DRF_HIP_CHECK(hipGetLastError());
    return DRF_OK;
}
