// Internal declarations shared by the HIP translation units of libdrifter_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "drifter_hip/drf.h"

namespace drf {

constexpr int kMaxN = 9; // n = order + 1 <= 9 (MAX_POLYNOMIAL_ORDER = 8)

// Error plumbing: thread-local message returned by drf_last_error().
void set_error(const std::string &msg);
drf_status fail(drf_status code, const std::string &msg);

#define DRF_HIP_CHECK(expr)                                                                  \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::drf::fail(DRF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Host-side 1D LGL basis (basis.cpp).
int lgl_nodes(int n, double *x, double *w);
void derivative_matrix(int n, const double *x, double *D);

// Operator constants passed to every kernel by value (they land in SGPRs).
template <class T> struct OpsT {
    T D[kMaxN * kMaxN];      // 1D derivative matrix, row-major
    T sw[kMaxN];             // 0.5 * LGL weights (sigma weights)
    double D64[kMaxN * kMaxN]; // FP64 copies: column integrals of the free-surface
    double sw64[kMaxN];        // equation stay in FP64 at every precision
};
using Ops = OpsT<double>;

struct Phys {
    double g, nu_h, nu_v, kT, kS, rho0, alpha, beta;
};

// Strided view of a 5-field state: element e, field f, node d lives at
// base[f * fs + e * es + d].
template <class T> struct View {
    T *p;
    int64_t fs, es;
};

inline void layout_strides(int layout, int64_t E, int nd, int nfields, int64_t &fs,
                           int64_t &es)
{
    if (layout == DRF_LAYOUT_SOA) {
        fs = E * nd;
        es = nd;
    } else {
        fs = nd;
        es = (int64_t)nfields * nd;
    }
}

// Epilogue of the volume kernel: what is done with the tendency k of each node.
enum EpiMode : int {
    EPI_RHS = 0,     // R = k
    EPI_STAGE = 1,   // Y = U0 + c1 k ; A: acc_mode 0 none, 1 A = k, 2 A = A + c2 k
    EPI_RK4FIN = 2,  // Y = U0 + c1 (A + k)
    EPI_COMB = 3,    // Y = c0 U0 + c1 (X + c2 k), X = the RHS input state
    EPI_LSRK = 4,    // A = c2 A + c1 k (acc_mode 1: A = c1 k) ; Y = X + c3 A
};

template <class T> struct Epi {
    int mode;
    int acc_mode;
    View<T> R;        // EPI_RHS output
    View<const T> U0; // base state
    View<T> Y;        // stage output
    View<T> A;        // accumulator
    View<const T> X;  // RHS input state (re-read in the epilogue)
    double c0, c1, c2, c3;
};

template <class T> struct VolGeom {
    int64_t E;
    const T *h; // E*nd bathymetry (element-major)
    const T *f; // E*nd Coriolis or nullptr
    double f0;  // used when f == nullptr
};

// Launchers (volume.hip).  Return hipError_t of the launch.
// ops is always given in FP64; the FP32 launcher rounds it once on the host.
hipError_t launch_volume_fp64(int order, const VolGeom<double> &geo, const Ops &ops,
                              const Phys &ph, View<const double> U, const Epi<double> &epi,
                              hipStream_t s);
hipError_t launch_volume_fp32(int order, const VolGeom<float> &geo, const Ops &ops,
                              const Phys &ph, View<const float> U, const Epi<float> &epi,
                              hipStream_t s);

// Vector kernels (stepper.hip)
hipError_t launch_axpy_fields(int64_t E, int nd, View<double> Y, View<const double> X, double a,
                              View<const double> K, hipStream_t s); // Y = X + a K
hipError_t launch_wave_speed(int order, int64_t E, View<const double> U, const double *h,
                             double *partial, int64_t npart, hipStream_t s);
hipError_t launch_wave_speed_derived(int order, int64_t E, const double *H, const double *u,
                                     const double *v, int64_t es, double *partial,
                                     int64_t npart, hipStream_t s);
hipError_t launch_max_reduce(const double *partial, int64_t npart, double *out, hipStream_t s);

// Mode split (modesplit.hip)
hipError_t launch_mode_split(int order, int64_t E, double dt3, int nsub, double *full,
                             const double *h, const double *f, double f0, const Ops &ops,
                             const Phys &ph, hipStream_t s);

} // namespace drf

// The context owned by the C ABI.
struct drf_ctx {
    int order = 0, n = 0, nd = 0, nh = 0;
    int precision = DRF_PREC_FP64;
    int device = 0;
    hipStream_t stream = nullptr;
    drf::Ops ops{};
    drf::Phys phys{};
    int64_t E = 0;
    double *d_h = nullptr;      // E*nd
    double *d_f = nullptr;      // E*nd or null when uniform
    float *d_h32 = nullptr;     // mixed precision copy
    float *d_f32 = nullptr;
    double f0 = 0.0;
    // host-packed staging and RK work space (grown on demand)
    void *d_work = nullptr;
    size_t work_bytes = 0;
    double *d_red = nullptr; // reduction scratch
    int64_t red_len = 0;
};
