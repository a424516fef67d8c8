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

// All views share one layout; thread t walks the nodes in that layout's
// memory order so consecutive lanes touch consecutive addresses.
__global__ void __launch_bounds__(256) lincomb_kernel(int64_t E, int n, LinComb lc)
{
    const int n2 = n * n, nd = n2 * n;
    const int64_t total = E * nd;
    const bool lines = lc.Y.ls == 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t e;
        int L, i;
        if (lines) { // [i][e][L]
            const int64_t q = t / n2;
            L = (int)(t - q * n2);
            i = (int)(q / E);
            e = q - (int64_t)i * E;
        } else { // [e][d]
            e = t / nd;
            const int d = (int)(t - e * nd);
            L = d / n;
            i = d - L * n;
        }
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            double s = lc.c[0] * lc.K[0].p[f * lc.K[0].fs + node_offset(lc.K[0], e, L, i)];
            for (int k = 1; k < lc.nk; ++k)
                s = s + lc.c[k] * lc.K[k].p[f * lc.K[k].fs + node_offset(lc.K[k], e, L, i)];
            s = lc.gamma * s;
            if (lc.X1.p)
                s = lc.X1.p[f * lc.X1.fs + node_offset(lc.X1, e, L, i)] + s;
            s = lc.beta * s;
            double y = s;
            if (lc.X0.p)
                y = lc.alpha * lc.X0.p[f * lc.X0.fs + node_offset(lc.X0, e, L, i)] + s;
            lc.Y.p[f * lc.Y.fs + node_offset(lc.Y, e, L, i)] = y;
        }
    }
}

} // namespace

hipError_t launch_lincomb(int64_t E, int n, const LinComb &lc, hipStream_t s)
{
    const int64_t total = E * (int64_t)n * n * n;
    if (total <= 0)
        return hipSuccess;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)blocks), dim3(256), 0, s, E, n, lc);
    return hipGetLastError();
}

} // namespace drf
