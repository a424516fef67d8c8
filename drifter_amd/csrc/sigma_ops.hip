// SigmaCoordinateOperators (src/dg/operators_3d.cpp:513-662) on gfx950: the
// sigma-coordinate metric update with that class's own sign convention, and
// the horizontal / vertical gradients, sigma divergence, ALE correction and
// material derivative built on the reference-coordinate derivatives.  Opt-in
// (the right-hand side does not use them).
//
// A workgroup stages EPB elements (EPB (N+1)^3 about 1024 nodes) of the
// differentiated field(s) in LDS with coalesced loads; each lane then owns a
// node and contracts its xi-, eta- and zeta-lines with the 1D D.  The
// reference multiplies by dense Kronecker matrices (D_xi = I x I x D ...),
// whose non-zero terms come in the same order as the line contraction's, and
// contraction is off, so results round as the reference's sequence.
#include "drf_internal.hpp"

#pragma clang fp contract(off)

namespace drf {
namespace {

enum SigmaOp { HGRAD = 0, VGRAD = 1, DIV = 2, ALE = 3, MATDER = 4 };

struct DOps {
    double D[kMaxN * kMaxN]; // row-major n x n
};

template <int N, int OP>
__global__ void __launch_bounds__(256)
    sigma_ops_kernel(int64_t E, int64_t es, DOps ops, const double *__restrict__ U,
                     const double *__restrict__ a, const double *__restrict__ b,
                     const double *__restrict__ c, const double *__restrict__ f,
                     const double *__restrict__ g, const double *__restrict__ dsdx,
                     const double *__restrict__ dsdy, const double *__restrict__ Hinv,
                     double *__restrict__ out1, double *__restrict__ out2)
{
    constexpr int N2 = N * N, ND = N2 * N;
    constexpr int EPB = ND >= 1024 ? 1 : 1024 / ND;
    constexpr int NF = (OP == DIV) ? 3 : 1;
    __shared__ double sD[N2], sF[NF][EPB * ND];
    for (int q = threadIdx.x; q < N2; q += blockDim.x)
        sD[q] = ops.D[(q / N) * kMaxN + q % N];
    const int64_t nb = (E + EPB - 1) / EPB;
    for (int64_t bt = blockIdx.x; bt < nb; bt += gridDim.x) {
        const int64_t e0 = bt * EPB;
        const int cnt = (E - e0 < EPB) ? (int)(E - e0) : EPB;
        __syncthreads(); // D staged; the previous batch is done with sF
        for (int q = threadIdx.x; q < cnt * ND; q += blockDim.x) {
            const int e = q / ND, d = q - e * ND;
            const int64_t o = (e0 + e) * es + d;
            sF[0][q] = U[o];
            if (OP == DIV) {
                sF[NF > 1 ? 1 : 0][q] = a[o];
                sF[NF > 2 ? 2 : 0][q] = b[o];
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < cnt * ND; q += blockDim.x) {
            const int e = q / ND, d = q - e * ND;
            const int i = d % N, j = (d / N) % N, k = d / N2;
            const double *Fx = &sF[0][e * ND];
            const double *Fy = &sF[NF > 1 ? 1 : 0][e * ND];
            const double *Fz = &sF[NF > 2 ? 2 : 0][e * ND];
            double tx = 0.0, ty = 0.0, tz = 0.0;
            if (OP == HGRAD || OP == DIV || OP == MATDER) {
#pragma unroll
                for (int m = 0; m < N; ++m)
                    tx += sD[i * N + m] * Fx[m + N * (j + N * k)];
#pragma unroll
                for (int m = 0; m < N; ++m)
                    ty += sD[j * N + m] * Fy[i + N * (m + N * k)];
            }
#pragma unroll
            for (int m = 0; m < N; ++m)
                tz += sD[k * N + m] * Fz[i + N * (j + N * m)];
            const int64_t o = (e0 + e) * es + d;
            if (OP == HGRAD) {
                out1[o] = tx - dsdx[o] * tz;
                out2[o] = ty - dsdy[o] * tz;
            } else if (OP == VGRAD) {
                out1[o] = Hinv[o] * tz;
            } else if (OP == DIV) {
                out1[o] = tx + ty + tz;
            } else if (OP == ALE) {
                const double dU_dz = Hinv[o] * tz;
                out1[o] = -a[o] * dU_dz;
            } else {
                const double dU_dx = tx - dsdx[o] * tz;
                const double dU_dy = ty - dsdy[o] * tz;
                const double om = f[o] - g[o] * Hinv[o];
                out1[o] = a[o] + b[o] * dU_dx + c[o] * dU_dy + om * tz;
            }
        }
    }
}

// update_sigma_metrics_with_time (:578-615), one node per lane
__global__ void __launch_bounds__(256)
    sigma_ops_metrics_kernel(int n, int64_t E, int64_t es, DOps x, const double *__restrict__ eta,
                             const double *__restrict__ h, const double *__restrict__ detadx,
                             const double *__restrict__ detady, const double *__restrict__ dhdx,
                             const double *__restrict__ dhdy, const double *__restrict__ detadt,
                             double *__restrict__ H, double *__restrict__ Hinv,
                             double *__restrict__ sig, double *__restrict__ dsdx,
                             double *__restrict__ dsdy, double *__restrict__ dsdt)
{
    const int n2 = n * n, nd = n2 * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < E * nd;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t / nd;
        const int d = (int)(t - e * nd);
        const int64_t i = e * es + d;
        const double Hi = eta[i] + h[i];
        const double Hv = (Hi > 1e-10) ? 1.0 / Hi : 0.0;
        const double s = 0.5 * (x.D[d / n2] - 1.0); // the node's zeta coordinate
        const double dH_dx = detadx[i] + dhdx[i];
        const double dH_dy = detady[i] + dhdy[i];
        H[i] = Hi;
        Hinv[i] = Hv;
        sig[i] = s;
        dsdx[i] = (s * dH_dx - detadx[i]) * Hv;
        dsdy[i] = (s * dH_dy - detady[i]) * Hv;
        const double et = detadt ? detadt[i] : 0.0;
        dsdt[i] = -(1.0 + s) * et * Hv;
    }
}

template <int OP>
hipError_t launch_op(int n, int64_t E, int64_t es, const DOps &ops, const double *U,
                     const double *a, const double *b, const double *c, const double *f,
                     const double *g, const double *dsdx, const double *dsdy, const double *Hinv,
                     double *o1, double *o2, hipStream_t s)
{
    const int nd = n * n * n, epb = nd >= 1024 ? 1 : 1024 / nd;
    const int64_t nb = (E + epb - 1) / epb;
    const dim3 grid((unsigned)(nb < 4096 ? nb : 4096)), blk(256);
#define DRF_SOP(NN)                                                                           \
    hipLaunchKernelGGL((sigma_ops_kernel<NN, OP>), grid, blk, 0, s, E, es, ops, U, a, b, c, f, \
                       g, dsdx, dsdy, Hinv, o1, o2)
    switch (n) {
    case 2: DRF_SOP(2); break;
    case 3: DRF_SOP(3); break;
    case 4: DRF_SOP(4); break;
    case 5: DRF_SOP(5); break;
    case 6: DRF_SOP(6); break;
    case 7: DRF_SOP(7); break;
    case 8: DRF_SOP(8); break;
    case 9: DRF_SOP(9); break;
    default: return hipErrorInvalidValue;
    }
#undef DRF_SOP
    return hipGetLastError();
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

drf_status sizes_ok(const char *who, int order, int64_t E, int64_t es)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                std::string(who) + ": order out of range");
    const int n = order + 1;
    DRF_REQUIRE(E >= 0 && es >= (int64_t)n * n * n, DRF_ERR_ARG,
                std::string(who) + ": bad sizes (elem_stride must be >= (N+1)^3)");
    return DRF_OK;
}

DOps d_ops(int n)
{
    DOps o{};
    double x[kMaxN], w[kMaxN], D[kMaxN * kMaxN];
    lgl_nodes(n, x, w);
    derivative_matrix(n, x, D);
    for (int r = 0; r < n; ++r)
        for (int c2 = 0; c2 < n; ++c2)
            o.D[r * kMaxN + c2] = D[r * n + c2];
    return o;
}

template <int OP>
drf_status run_op(const char *who, int order, int64_t E, int64_t es, const void *U,
                  const void *a, const void *b, const void *c, const void *f, const void *g,
                  const void *dsdx, const void *dsdy, const void *Hinv, void *o1, void *o2,
                  void *stream)
{
    drf_status st = sizes_ok(who, order, E, es);
    if (st || E == 0)
        return st;
    const bool ok = U && o1 &&
                    (OP == HGRAD ? (dsdx && dsdy && o2)
                     : OP == VGRAD ? Hinv != nullptr
                     : OP == DIV  ? (a && b)
                     : OP == ALE  ? (a && Hinv)
                                  : (a && b && c && f && g && dsdx && dsdy && Hinv));
    DRF_REQUIRE(ok, DRF_ERR_ARG, std::string(who) + ": null pointer");
    DRF_HIP_CHECK(launch_op<OP>(order + 1, E, es, d_ops(order + 1), (const double *)U,
                                (const double *)a, (const double *)b, (const double *)c,
                                (const double *)f, (const double *)g, (const double *)dsdx,
                                (const double *)dsdy, (const double *)Hinv, (double *)o1,
                                (double *)o2, (hipStream_t)stream));
    return DRF_OK;
}

} // namespace

extern "C" {

drf_status drf_sigma_ops_metrics_dev(int order, int64_t n_elem, const void *eta, const void *h,
                                     const void *deta_dx, const void *deta_dy, const void *dh_dx,
                                     const void *dh_dy, const void *deta_dt, int64_t elem_stride,
                                     void *H, void *H_inv, void *sigma, void *dsigma_dx,
                                     void *dsigma_dy, void *dsigma_dt, void *stream)
{
    drf_status st = sizes_ok("drf_sigma_ops_metrics_dev", order, n_elem, elem_stride);
    if (st || n_elem == 0)
        return st;
    DRF_REQUIRE(eta && h && deta_dx && deta_dy && dh_dx && dh_dy && H && H_inv && sigma &&
                    dsigma_dx && dsigma_dy && dsigma_dt,
                DRF_ERR_ARG, "drf_sigma_ops_metrics_dev: null pointer");
    const int n = order + 1;
    DOps x{};
    double w[kMaxN];
    lgl_nodes(n, x.D, w); // the LGL nodes, in the first n slots
    const int64_t total = n_elem * n * n * n, b = (total + 255) / 256;
    hipLaunchKernelGGL(sigma_ops_metrics_kernel, dim3((unsigned)(b < 65536 ? b : 65536)),
                       dim3(256), 0, (hipStream_t)stream, n, n_elem, elem_stride, x,
                       (const double *)eta, (const double *)h, (const double *)deta_dx,
                       (const double *)deta_dy, (const double *)dh_dx, (const double *)dh_dy,
                       (const double *)deta_dt, (double *)H, (double *)H_inv, (double *)sigma,
                       (double *)dsigma_dx, (double *)dsigma_dy, (double *)dsigma_dt);
    DRF_HIP_CHECK(hipGetLastError());
    return DRF_OK;
}

drf_status drf_sigma_horizontal_gradient_dev(int order, int64_t n_elem, const void *U,
                                             const void *dsigma_dx, const void *dsigma_dy,
                                             int64_t elem_stride, void *dU_dx, void *dU_dy,
                                             void *stream)
{
    return run_op<HGRAD>("drf_sigma_horizontal_gradient_dev", order, n_elem, elem_stride, U,
                         nullptr, nullptr, nullptr, nullptr, nullptr, dsigma_dx, dsigma_dy,
                         nullptr, dU_dx, dU_dy, stream);
}

drf_status drf_sigma_vertical_gradient_dev(int order, int64_t n_elem, const void *U,
                                           const void *H_inv, int64_t elem_stride, void *dU_dz,
                                           void *stream)
{
    return run_op<VGRAD>("drf_sigma_vertical_gradient_dev", order, n_elem, elem_stride, U,
                         nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, H_inv,
                         dU_dz, nullptr, stream);
}

drf_status drf_sigma_divergence_dev(int order, int64_t n_elem, const void *Hu, const void *Hv,
                                    const void *omega, int64_t elem_stride, void *div,
                                    void *stream)
{
    return run_op<DIV>("drf_sigma_divergence_dev", order, n_elem, elem_stride, Hu, Hv, omega,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, div, nullptr,
                       stream);
}

drf_status drf_sigma_ale_correction_dev(int order, int64_t n_elem, const void *U,
                                        const void *w_mesh, const void *H_inv,
                                        int64_t elem_stride, void *correction, void *stream)
{
    return run_op<ALE>("drf_sigma_ale_correction_dev", order, n_elem, elem_stride, U, w_mesh,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, H_inv, correction,
                       nullptr, stream);
}

drf_status drf_sigma_material_derivative_dev(int order, int64_t n_elem, const void *U,
                                             const void *dU_dt, const void *u, const void *v,
                                             const void *omega, const void *w_mesh,
                                             const void *dsigma_dx, const void *dsigma_dy,
                                             const void *H_inv, int64_t elem_stride, void *out,
                                             void *stream)
{
    return run_op<MATDER>("drf_sigma_material_derivative_dev", order, n_elem, elem_stride, U,
                          dU_dt, u, v, omega, w_mesh, dsigma_dx, dsigma_dy, H_inv, out, nullptr,
                          stream);
}

} // extern "C"
