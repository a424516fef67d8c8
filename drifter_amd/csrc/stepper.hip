// CFL wave-speed reduction (SimulationDriver::compute_max_wave_speed,
// src/solver/simulation_driver.cpp:153-168).
#include "drf_internal.hpp"

namespace drf {
namespace {

__device__ __forceinline__ double wave_max(double a, double b) { return a > b ? a : b; }

// Per element: sqrt(9.81 * max H) + max(max|u|, max|v|) on the derived state
// of U (update_derived, src/physics/primitive_equations.cpp:21-28).  One
// thread per grid line (e, L) of n nodes: in the line layout consecutive
// lines of a plane are consecutive addresses, so a wavefront's loads are
// contiguous.  The n^2 line maxima of an element meet in LDS, where one
// thread per element forms its wave speed; partial[block] = max over the
// block's elements.  max and sqrt are exact, so any grouping gives the
// reference's value.
__global__ void __launch_bounds__(256)
    wave_speed_kernel(int64_t E, int n, View<const double> U, View<const double> h,
                      double *partial)
{
    __shared__ double lm[3][256];
    __shared__ double red[4];
    const int nl = n * n, epb = 256 / nl;
    const int t = threadIdx.x;
    const int le = t / nl, L = t - le * nl;
    double cmax = 0.0;
    const int64_t nbatch = (E + epb - 1) / epb;
    for (int64_t b = blockIdx.x; b < nbatch; b += gridDim.x) {
        const int64_t e = b * epb + le;
        if (le < epb && e < E) {
            double hm = -INFINITY, um = 0.0, vm = 0.0;
            // the line's offsets, once per line (the element part of the
            // blocked layout takes a division)
            const int64_t uo = node_offset(U, e, L, 0), ho = node_offset(h, e, L, 0);
            for (int i = 0; i < n; ++i) {
                const int64_t o = uo + (int64_t)i * U.is;
                const double eta = U.p[2 * U.fs + o];
                const double H = eta + h.p[ho + (int64_t)i * h.is];
                const double hi = (H > 1e-10) ? 1.0 / H : 0.0;
                const double u = U.p[0 * U.fs + o] * hi;
                const double v = U.p[1 * U.fs + o] * hi;
                hm = wave_max(hm, H);
                um = wave_max(um, fabs(u));
                vm = wave_max(vm, fabs(v));
            }
            lm[0][t] = hm;
            lm[1][t] = um;
            lm[2][t] = vm;
        }
        __syncthreads();
        if (t < epb && b * epb + t < E) {
            double hm = -INFINITY, um = 0.0, vm = 0.0;
            for (int q = t * nl; q < (t + 1) * nl; ++q) {
                hm = wave_max(hm, lm[0][q]);
                um = wave_max(um, lm[1][q]);
                vm = wave_max(vm, lm[2][q]);
            }
            cmax = wave_max(cmax, sqrt(9.81 * hm) + wave_max(um, vm));
        }
        __syncthreads(); // lm is rewritten by the next batch
    }
    for (int o = 32; o > 0; o >>= 1)
        cmax = wave_max(cmax, __shfl_xor(cmax, o));
    if ((t & 63) == 0)
        red[t >> 6] = cmax;
    __syncthreads();
    if (t == 0)
        partial[blockIdx.x] = wave_max(wave_max(red[0], red[1]), wave_max(red[2], red[3]));
}

template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
    wave_speed_derived_kernel(int64_t E, int nd, const double *H, const double *u,
                              const double *v, int64_t es, double *partial)
{
    __shared__ double red[WAVES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double cmax = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * WAVES + w; e < E; e += (int64_t)gridDim.x * WAVES) {
        double hm = -INFINITY, um = 0.0, vm = 0.0;
        for (int d = lane; d < nd; d += 64) {
            hm = wave_max(hm, H[e * es + d]);
            um = wave_max(um, fabs(u[e * es + d]));
            vm = wave_max(vm, fabs(v[e * es + d]));
        }
        for (int o = 32; o > 0; o >>= 1) {
            hm = wave_max(hm, __shfl_xor(hm, o));
            um = wave_max(um, __shfl_xor(um, o));
            vm = wave_max(vm, __shfl_xor(vm, o));
        }
        cmax = wave_max(cmax, sqrt(9.81 * hm) + wave_max(um, vm));
    }
    if (lane == 0)
        red[w] = cmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int i = 1; i < WAVES; ++i)
            m = wave_max(m, red[i]);
        partial[blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(256) max_reduce_kernel(const double *p, int64_t np, double *out)
{
    __shared__ double red[4];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < np; i += 256)
        m = wave_max(m, p[i]);
    for (int o = 32; o > 0; o >>= 1)
        m = wave_max(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        out[0] = wave_max(wave_max(red[0], red[1]), wave_max(red[2], red[3]));
}

// AdaptiveTimeController::compute_dt (src/solver/time_stepper.cpp:261-268)
// on the device maximum, and the driver's time_ += dt_
// (simulation_driver.cpp:130).  One thread.
__global__ void adaptive_dt_kernel(const double *__restrict__ cmax, double dx_min, double cfl,
                                   double dt_min, double dt_max, double *__restrict__ dt,
                                   double *__restrict__ t)
{
    const double c = *cmax;
    double d = dt_max; // stationary: the largest step
    if (!(c < 1e-14)) {
        d = cfl * dx_min / c;
        d = (d < dt_min) ? dt_min : ((dt_max < d) ? dt_max : d);
    }
    *dt = d;
    if (t)
        *t = *t + d;
}

} // namespace

hipError_t launch_adaptive_dt(const double *cmax, double dx_min, double cfl, double dt_min,
                              double dt_max, double *dt, double *t, hipStream_t s)
{
    hipLaunchKernelGGL(adaptive_dt_kernel, dim3(1), dim3(1), 0, s, cmax, dx_min, cfl, dt_min,
                       dt_max, dt, t);
    return hipGetLastError();
}

hipError_t launch_wave_speed(int order, int64_t E, View<const double> U, View<const double> h,
                             double *partial, int64_t npart, hipStream_t s)
{
    const int n = order + 1;
    if (n * n > 256)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(wave_speed_kernel, dim3((unsigned)npart), dim3(256), 0, s, E, n, U, h,
                       partial);
    return hipGetLastError();
}

hipError_t launch_wave_speed_derived(int order, int64_t E, const double *H, const double *u,
                                     const double *v, int64_t es, double *partial,
                                     int64_t npart, hipStream_t s)
{
    const int n = order + 1, nd = n * n * n;
    hipLaunchKernelGGL(wave_speed_derived_kernel<4>, dim3((unsigned)npart), dim3(256), 0, s, E,
                       nd, H, u, v, es, partial);
    return hipGetLastError();
}

hipError_t launch_max_reduce(const double *partial, int64_t npart, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(256), 0, s, partial, npart, out);
    return hipGetLastError();
}

} // namespace drf
