// C ABI of the refinement-side kernels: gradient error estimator with
// element marking, element characteristic lengths, modal coefficients.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "adapt.hpp"
#include "drf_internal.hpp"

namespace drf {

// Orthonormal Legendre Vandermonde on the LGL nodes; inverse by LU with
// partial pivoting (n <= 9).
int modal_matrices(int n, double *V, double *Vinv)
{
    double x[kMaxN], w[kMaxN];
    if (lgl_nodes(n, x, w))
        return -1;
    for (int i = 0; i < n; ++i) {
        double pm1 = 0.0, p = 1.0; // P_{k-1}, P_k
        for (int k = 0; k < n; ++k) {
            V[i * n + k] = std::sqrt((2.0 * k + 1.0) / 2.0) * p;
            const double pn = ((2.0 * k + 1.0) * x[i] * p - k * pm1) / (k + 1.0);
            pm1 = p;
            p = pn;
        }
    }
    double LU[kMaxN * kMaxN];
    int piv[kMaxN];
    std::memcpy(LU, V, sizeof(double) * n * n);
    for (int c = 0; c < n; ++c) {
        int r = c;
        for (int i = c + 1; i < n; ++i)
            if (std::fabs(LU[i * n + c]) > std::fabs(LU[r * n + c]))
                r = i;
        if (LU[r * n + c] == 0.0)
            return -1;
        piv[c] = r;
        if (r != c)
            for (int k = 0; k < n; ++k)
                std::swap(LU[c * n + k], LU[r * n + k]);
        for (int i = c + 1; i < n; ++i) {
            const double f = LU[i * n + c] / LU[c * n + c];
            LU[i * n + c] = f;
            for (int k = c + 1; k < n; ++k)
                LU[i * n + k] -= f * LU[c * n + k];
        }
    }
    for (int col = 0; col < n; ++col) {
        double y[kMaxN];
        for (int i = 0; i < n; ++i)
            y[i] = (i == col) ? 1.0 : 0.0;
        for (int c = 0; c < n; ++c) // apply the row swaps in order
            std::swap(y[c], y[piv[c]]);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k)
                y[i] -= LU[i * n + k] * y[k];
        for (int i = n - 1; i >= 0; --i) {
            for (int k = i + 1; k < n; ++k)
                y[i] -= LU[i * n + k] * y[k];
            y[i] /= LU[i * n + i];
        }
        for (int i = 0; i < n; ++i)
            Vinv[i * n + col] = y[i];
    }
    return 0;
}

} // namespace drf

using namespace drf;

struct drf_estimator {
    int order, device;
    EstimatorParams p;
    double *d_partial; // kEstimatorPartials block maxima
};

namespace {

#define DRF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond))                                                                         \
            return ::drf::fail((code), (msg));                                               \
    } while (0)

ModalOps modal_ops(int n, bool &ok)
{
    ModalOps ops;
    std::memset(&ops, 0, sizeof(ops));
    double V[kMaxN * kMaxN], Vi[kMaxN * kMaxN];
    ok = modal_matrices(n, V, Vi) == 0;
    for (int r = 0; r < n && ok; ++r)
        for (int c = 0; c < n; ++c)
            ops.Vinv[r * kMaxN + c] = Vi[r * n + c];
    return ops;
}

} // namespace

extern "C" {

drf_status drf_element_char_length(int64_t n_elem, const double *bounds, double *h_char)
{
    DRF_REQUIRE(n_elem >= 0, DRF_ERR_ARG, "drf_element_char_length: negative count");
    DRF_REQUIRE(n_elem == 0 || (bounds && h_char), DRF_ERR_ARG,
                "drf_element_char_length: null pointer");
    for (int64_t e = 0; e < n_elem; ++e) {
        const double *b = bounds + 6 * e;
        h_char[e] = std::cbrt((b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]));
    }
    return DRF_OK;
}

drf_status drf_estimator_create(int order, double refine_threshold, double coarsen_threshold,
                                double smoothness_factor, int max_level, int min_level,
                                int device, drf_estimator **out)
{
    DRF_REQUIRE(out, DRF_ERR_ARG, "drf_estimator_create: null output");
    *out = nullptr;
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_estimator_create: order out of range");
    DRF_HIP_CHECK(hipSetDevice(device));
    double *d = nullptr;
    DRF_HIP_CHECK(hipMalloc(&d, sizeof(double) * kEstimatorPartials));
    EstimatorParams p{refine_threshold, coarsen_threshold, smoothness_factor, max_level,
                      min_level};
    *out = new drf_estimator{order, device, p, d};
    return DRF_OK;
}

void drf_estimator_destroy(drf_estimator *est)
{
    if (!est)
        return;
    (void)hipSetDevice(est->device);
    (void)hipFree(est->d_partial);
    delete est;
}

drf_status drf_estimate_dev(drf_estimator *est, int64_t n_elem, const void *field,
                            int64_t elem_stride, const void *h_char, const void *levels,
                            void *value, void *mask, void *refine, void *coarsen, void *action,
                            void *mask_out, void *max_value, void *stream)
{
    DRF_REQUIRE(est, DRF_ERR_ARG, "drf_estimate_dev: null estimator");
    const int n = est->order + 1;
    DRF_REQUIRE(n_elem >= 0 && elem_stride >= (int64_t)n * n * n, DRF_ERR_ARG,
                "drf_estimate_dev: bad sizes (elem_stride must be >= (N+1)^3)");
    if (n_elem == 0)
        return DRF_OK;
    DRF_REQUIRE(field && h_char && value && mask && refine && coarsen, DRF_ERR_ARG,
                "drf_estimate_dev: null pointer");
    DRF_REQUIRE(!levels || (action && mask_out), DRF_ERR_ARG,
                "drf_estimate_dev: marking needs action and mask_out");
    DRF_HIP_CHECK(hipSetDevice(est->device));
    DRF_HIP_CHECK(launch_estimate(n, n_elem, (const double *)field, elem_stride,
                                  (const double *)h_char, (const int32_t *)levels, est->p,
                                  est->d_partial, (double *)max_value, (double *)value,
                                  (uint8_t *)mask,
                                  (uint8_t *)refine, (uint8_t *)coarsen, (uint8_t *)action,
                                  (uint8_t *)mask_out, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_modal_matrices(int order, double *V, double *Vinv)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_modal_matrices: order out of range");
    DRF_REQUIRE(V && Vinv, DRF_ERR_ARG, "drf_modal_matrices: null output");
    DRF_REQUIRE(modal_matrices(order + 1, V, Vinv) == 0, DRF_ERR_ARG,
                "drf_modal_matrices: singular Vandermonde");
    return DRF_OK;
}

drf_status drf_modal_coeffs_dev(int order, int64_t n_elem, const void *u, int64_t elem_stride,
                                void *coef, void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_modal_coeffs_dev: order out of range");
    const int n = order + 1, nd = n * n * n;
    DRF_REQUIRE(n_elem >= 0 && elem_stride >= nd, DRF_ERR_ARG, "drf_modal_coeffs_dev: bad sizes");
    if (n_elem == 0)
        return DRF_OK;
    DRF_REQUIRE(u && coef, DRF_ERR_ARG, "drf_modal_coeffs_dev: null pointer");
    bool ok;
    const ModalOps ops = modal_ops(n, ok);
    DRF_REQUIRE(ok, DRF_ERR_ARG, "drf_modal_coeffs_dev: singular Vandermonde");
    DRF_HIP_CHECK(launch_modal(n, n_elem, (const double *)u, elem_stride, ops, (double *)coef,
                               nullptr, nd, 0, 1.0, nullptr, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_modal_features_bf16_dev(int order, int64_t n_elem, const void *u,
                                       int64_t elem_stride, double scale, const void *elem_scale,
                                       void *features, int64_t row_stride, int col_offset,
                                       void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_modal_features_bf16_dev: order out of range");
    const int n = order + 1, nd = n * n * n;
    DRF_REQUIRE(n_elem >= 0 && elem_stride >= nd && col_offset >= 0 &&
                    row_stride >= (int64_t)col_offset + nd,
                DRF_ERR_ARG, "drf_modal_features_bf16_dev: bad sizes");
    if (n_elem == 0)
        return DRF_OK;
    DRF_REQUIRE(u && features, DRF_ERR_ARG, "drf_modal_features_bf16_dev: null pointer");
    bool ok;
    const ModalOps ops = modal_ops(n, ok);
    DRF_REQUIRE(ok, DRF_ERR_ARG, "drf_modal_features_bf16_dev: singular Vandermonde");
    DRF_HIP_CHECK(launch_modal(n, n_elem, (const double *)u, elem_stride, ops, nullptr,
                               (uint16_t *)features, row_stride, col_offset, scale,
                               (const double *)elem_scale, (hipStream_t)stream));
    return DRF_OK;
}

} // extern "C"
