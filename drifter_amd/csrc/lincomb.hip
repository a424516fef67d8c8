// Elementwise Runge-Kutta vector update used by the un-fused stepping path
// (the reference's own sequence of vector operations, src/solver/time_stepper.cpp:12-201)
// and by the schemes whose stages are not fused into the RHS epilogue.
//
//   t = sum_i c_i K_i   (left to right; c_0 K_0 first)
//   t = gamma * t ; if X1: t = X1 + t ; t = beta * t
//   Y = (X0 ? alpha * X0 : 0) + t
//
// With unit coefficients the multiplications are exact, so each scheme's
// expression keeps the reference's association, e.g. RK4's
// U + (dt/6) * (((k1 + 2 k2) + 2 k3) + k4).
#include "drf_internal.hpp"
#include "lincomb.hpp"

namespace drf {
namespace {

__global__ void __launch_bounds__(256) lincomb_kernel(int64_t E, int nd, LinComb lc)
{
    const int64_t total = E * nd;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t / nd, d = t - e * nd;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            double s = lc.c[0] * lc.K[0].p[f * lc.K[0].fs + e * lc.K[0].es + d];
            for (int i = 1; i < lc.nk; ++i)
                s = s + lc.c[i] * lc.K[i].p[f * lc.K[i].fs + e * lc.K[i].es + d];
            s = lc.gamma * s;
            if (lc.X1.p)
                s = lc.X1.p[f * lc.X1.fs + e * lc.X1.es + d] + s;
            s = lc.beta * s;
            double y = s;
            if (lc.X0.p)
                y = lc.alpha * lc.X0.p[f * lc.X0.fs + e * lc.X0.es + d] + s;
            lc.Y.p[f * lc.Y.fs + e * lc.Y.es + d] = y;
        }
    }
}

} // namespace

hipError_t launch_lincomb(int64_t E, int nd, const LinComb &lc, hipStream_t s)
{
    const int64_t total = E * nd;
    if (total <= 0)
        return hipSuccess;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)blocks), dim3(256), 0, s, E, nd, lc);
    return hipGetLastError();
}

} // namespace drf
