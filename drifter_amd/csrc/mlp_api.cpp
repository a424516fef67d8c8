// C ABI of the refinement-indicator MLP: lifecycle, parameters, inference,
// gradients and Adam training steps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "drf_internal.hpp"
#include "mlp.hpp"

using namespace drf;

struct drf_mlp {
    int device;
    float *d_w, *d_m, *d_v;    // master weights and Adam moments (kMlpParams each)
    uint16_t *d_wq;            // BF16 W1 | W2 | W3
    float *d_partial, *d_lpart; // kMlpMaxParts partial gradients / losses
    int64_t step;
};

namespace {

#define DRF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond))                                                                         \
            return ::drf::fail((code), (msg));                                               \
    } while (0)

uint16_t host_bf16(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}

// Upload master weights, their BF16 copies, and clear the Adam state.
drf_status upload(drf_mlp *m, const float *w)
{
    std::vector<uint16_t> q(kMlpQuant);
    const int nw1 = kMlpHidden * kMlpIn, nw2 = kMlpHidden * kMlpHidden;
    const int ow2 = nw1 + kMlpHidden, ow3 = ow2 + nw2 + kMlpHidden;
    for (int i = 0; i < nw1; ++i)
        q[i] = host_bf16(w[i]);
    for (int i = 0; i < nw2; ++i)
        q[nw1 + i] = host_bf16(w[ow2 + i]);
    for (int i = 0; i < kMlpHidden; ++i)
        q[nw1 + nw2 + i] = host_bf16(w[ow3 + i]);
    DRF_HIP_CHECK(hipSetDevice(m->device));
    DRF_HIP_CHECK(hipMemcpy(m->d_w, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    DRF_HIP_CHECK(hipMemcpy(m->d_wq, q.data(), sizeof(uint16_t) * kMlpQuant,
                            hipMemcpyHostToDevice));
    DRF_HIP_CHECK(hipMemset(m->d_m, 0, sizeof(float) * kMlpParams));
    DRF_HIP_CHECK(hipMemset(m->d_v, 0, sizeof(float) * kMlpParams));
    m->step = 0;
    return DRF_OK;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

} // namespace

extern "C" {

drf_status drf_mlp_create(uint64_t seed, int device, drf_mlp **out)
{
    DRF_REQUIRE(out, DRF_ERR_ARG, "drf_mlp_create: null output");
    *out = nullptr;
    DRF_HIP_CHECK(hipSetDevice(device));
    drf_mlp *m = new drf_mlp{device, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc(&m->d_w, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMalloc(&m->d_m, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMalloc(&m->d_v, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMalloc(&m->d_wq, sizeof(uint16_t) * kMlpQuant);
    if (e == hipSuccess) e = hipMalloc(&m->d_partial, sizeof(float) * kMlpParams * kMlpMaxParts);
    if (e == hipSuccess) e = hipMalloc(&m->d_lpart, sizeof(float) * kMlpMaxParts);
    if (e != hipSuccess) {
        drf_mlp_destroy(m);
        return fail(DRF_ERR_HIP, std::string("drf_mlp_create: ") + hipGetErrorString(e));
    }
    // Glorot-uniform weights from SplitMix64, zero biases
    std::vector<float> w(kMlpParams, 0.0f);
    uint64_t st = seed;
    auto next = [&st]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto fill = [&](int off, int count, int fan_in, int fan_out) {
        const double lim = std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < count; ++i)
            w[off + i] = (float)(lim * (2.0 * ((next() >> 11) * 0x1.0p-53) - 1.0));
    };
    const int nw1 = kMlpHidden * kMlpIn, nw2 = kMlpHidden * kMlpHidden;
    fill(0, nw1, kMlpIn, kMlpHidden);
    fill(nw1 + kMlpHidden, nw2, kMlpHidden, kMlpHidden);
    fill(nw1 + kMlpHidden + nw2 + kMlpHidden, kMlpHidden, kMlpHidden, 1);
    const drf_status st_ = upload(m, w.data());
    if (st_ != DRF_OK) {
        drf_mlp_destroy(m);
        return st_;
    }
    *out = m;
    return DRF_OK;
}

void drf_mlp_destroy(drf_mlp *m)
{
    if (!m)
        return;
    (void)hipSetDevice(m->device);
    for (void *p : {(void *)m->d_w, (void *)m->d_m, (void *)m->d_v, (void *)m->d_wq,
                    (void *)m->d_partial, (void *)m->d_lpart})
        if (p)
            (void)hipFree(p);
    delete m;
}

drf_status drf_mlp_set_params(drf_mlp *m, const float *params)
{
    DRF_REQUIRE(m && params, DRF_ERR_ARG, "drf_mlp_set_params: null argument");
    return upload(m, params);
}

drf_status drf_mlp_get_params(const drf_mlp *m, float *params)
{
    DRF_REQUIRE(m && params, DRF_ERR_ARG, "drf_mlp_get_params: null argument");
    DRF_HIP_CHECK(hipSetDevice(m->device));
    DRF_HIP_CHECK(hipMemcpy(params, m->d_w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return DRF_OK;
}

drf_status drf_mlp_forward_dev(drf_mlp *m, int64_t n, const void *X, void *pred, void *stream)
{
    DRF_REQUIRE(m && n >= 0, DRF_ERR_ARG, "drf_mlp_forward_dev: bad arguments");
    if (n == 0)
        return DRF_OK;
    DRF_REQUIRE(X && pred, DRF_ERR_ARG, "drf_mlp_forward_dev: null pointer");
    DRF_REQUIRE(aligned16(X), DRF_ERR_ARG, "drf_mlp_forward_dev: X must be 16-byte aligned");
    DRF_HIP_CHECK(hipSetDevice(m->device));
    DRF_HIP_CHECK(launch_mlp_forward(n, (const uint16_t *)X, m->d_w, m->d_wq, (float *)pred,
                                     (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_mlp_classify_dev(int64_t n, const void *pred, double refine_level,
                                double coarsen_level, void *action, void *stream)
{
    DRF_REQUIRE(n >= 0, DRF_ERR_ARG, "drf_mlp_classify_dev: negative count");
    if (n == 0)
        return DRF_OK;
    DRF_REQUIRE(pred && action, DRF_ERR_ARG, "drf_mlp_classify_dev: null pointer");
    DRF_HIP_CHECK(launch_mlp_classify(n, (const float *)pred, (float)refine_level,
                                      (float)coarsen_level, (uint8_t *)action,
                                      (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_mlp_loss_grad_dev(drf_mlp *m, int64_t n, const void *X, const void *y, void *grad,
                                 void *loss, void *stream)
{
    DRF_REQUIRE(m && n > 0, DRF_ERR_ARG, "drf_mlp_loss_grad_dev: need a model and n > 0");
    DRF_REQUIRE(X && y && grad, DRF_ERR_ARG, "drf_mlp_loss_grad_dev: null pointer");
    DRF_REQUIRE(aligned16(X), DRF_ERR_ARG, "drf_mlp_loss_grad_dev: X must be 16-byte aligned");
    DRF_HIP_CHECK(hipSetDevice(m->device));
    const AdamStep none{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    DRF_HIP_CHECK(launch_mlp_grad(n, (const uint16_t *)X, (const float *)y, m->d_w, m->d_wq,
                                  m->d_m, m->d_v, m->d_partial, m->d_lpart, (float *)grad,
                                  (float *)loss, none, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_mlp_train_step_dev(drf_mlp *m, int64_t n, const void *X, const void *y, double lr,
                                  void *loss, void *stream)
{
    DRF_REQUIRE(m && n > 0 && lr > 0.0, DRF_ERR_ARG,
                "drf_mlp_train_step_dev: need a model, n > 0 and lr > 0");
    DRF_REQUIRE(X && y, DRF_ERR_ARG, "drf_mlp_train_step_dev: null pointer");
    DRF_REQUIRE(aligned16(X), DRF_ERR_ARG, "drf_mlp_train_step_dev: X must be 16-byte aligned");
    DRF_HIP_CHECK(hipSetDevice(m->device));
    m->step += 1;
    const double b1 = 0.9, b2 = 0.999;
    const AdamStep adam{(float)lr, (float)b1, (float)b2, 1e-8f,
                        (float)(1.0 / (1.0 - std::pow(b1, (double)m->step))),
                        (float)(1.0 / (1.0 - std::pow(b2, (double)m->step)))};
    DRF_HIP_CHECK(launch_mlp_grad(n, (const uint16_t *)X, (const float *)y, m->d_w, m->d_wq,
                                  m->d_m, m->d_v, m->d_partial, m->d_lpart, nullptr,
                                  (float *)loss, adam, (hipStream_t)stream));
    return DRF_OK;
}

} // extern "C"
