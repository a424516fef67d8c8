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
    T D[kMaxN * kMaxN]; // 1D derivative matrix, row-major (n x n used)
    T sw[kMaxN];        // 0.5 * LGL weights (sigma weights)
};
using Ops = OpsT<double>;

struct Phys {
    double g, nu_h, nu_v, kT, kS, rho0, alpha, beta;
};

// Strided view of a multi-field nodal state.  Node d = i + n L of element e
// (i the xi index, L = j + n k the grid line) in field f lives at
// p[f * fs + e * es + L * ls + i * is].
// A field array in one of the public layouts: value (field f, element e,
// line L, xi index i) at f fs + elem(e) + L ls + i is, elem(e) = e es, or in
// the blocked layout (bb > 0) (e / bb) bs + (e % bb) es.
template <class T> struct View {
    T *p;
    int64_t fs, es, ls, is;
    int64_t bs = 0; // stride between blocks (blocked layout)
    int64_t bb = 0; // elements per block, 0: not blocked
};

// Volume-kernel batch geometry, shared by the kernels and the host: seven
// LDS slots per element node plus the double-buffered column integrals, and
// the LDS a workgroup may give its elements (gfx950: 160 KB per workgroup,
// less room for the staged operators).
constexpr int kVolSlots = 7;
constexpr int kVolLdsBudget = 160 * 1024 - 1024;
// elements per batch of a workgroup of `lines` grid lines at `order`
constexpr int vol_batch_elements(int order, int lines)
{
    const int n = order + 1, n2 = n * n, n3 = n2 * n;
    const int lds = (kVolSlots * n3 + 4 * n2) * 8;
    const int by_lines = lines / n2 > 0 ? lines / n2 : 1;
    const int by_lds = kVolLdsBudget / lds > 0 ? kVolLdsBudget / lds : 1;
    return by_lines < by_lds ? by_lines : by_lds;
}
// elements per block of DRF_LAYOUT_BLOCKS: the batch of the widest workgroup
// (512 lines up to order 4, 128 above), which every narrower batch divides
inline int layout_block_elements(int order)
{
    return (order >= 1 && order <= 8) ? vol_batch_elements(order, order <= 4 ? 512 : 128) : 0;
}

// Strides of the public layouts (drf.h).  LINES is the kernels' native
// layout: for a fixed xi index, consecutive (element, line) pairs are
// consecutive in memory, so a wavefront whose lanes own consecutive lines
// touches one contiguous span per load or store.
inline void layout_strides(int layout, int64_t E, int n, int nfields, int64_t &fs, int64_t &es,
                           int64_t &ls, int64_t &is, int64_t &bs, int64_t &bb)
{
    const int64_t nd = (int64_t)n * n * n, n2 = (int64_t)n * n;
    bs = 0;
    bb = 0;
    if (layout == DRF_LAYOUT_BLOCKS) {
        bb = layout_block_elements(n - 1);
        es = n2;
        ls = 1;
        is = bb * n2;
        fs = (int64_t)n * is;
        bs = (int64_t)nfields * fs;
    } else if (layout == DRF_LAYOUT_LINES) {
        fs = E * nd;
        es = n2;
        ls = 1;
        is = E * n2;
    } else if (layout == DRF_LAYOUT_SOA) {
        fs = E * nd;
        es = nd;
        ls = n;
        is = 1;
    } else {
        fs = nd;
        es = (int64_t)nfields * nd;
        ls = n;
        is = 1;
    }
}

// the element's part of an offset (every layout): element stride es, and in
// the blocked layout block stride bs and bb elements per block
__host__ __device__ inline int64_t elem_offset(int64_t e, int64_t es, int64_t bs, int64_t bb)
{
    if (bb) {
        // 32-bit division where the element index allows it (every mesh here)
        const int64_t blk = (e >> 31) ? e / bb : (int64_t)((uint32_t)e / (uint32_t)bb);
        return blk * bs + (e - blk * bb) * es;
    }
    return e * es;
}
template <class T> __host__ __device__ inline int64_t elem_offset(const View<T> &v, int64_t e)
{
    return elem_offset(e, v.es, v.bs, v.bb);
}

template <class T> __host__ __device__ inline int64_t node_offset(const View<T> &v, int64_t e, int L,
                                                                  int i)
{
    return elem_offset(v, e) + (int64_t)L * v.ls + (int64_t)i * v.is;
}

// values of one state of E elements in a layout (whole blocks when blocked)
inline int64_t layout_state_len(int layout, int64_t E, int n, int nfields)
{
    const int64_t nd = (int64_t)n * n * n;
    if (layout == DRF_LAYOUT_BLOCKS) {
        const int64_t B = layout_block_elements(n - 1);
        return B > 0 ? (E + B - 1) / B * B * nfields * nd : 0;
    }
    return E * nfields * nd;
}

// Epilogue of the volume kernel: what is done with the tendency k of each node.
enum EpiMode : int {
    EPI_RHS = 0,       // R = k
    EPI_STAGE = 1,     // Y = U0 + c1 k ; acc_mode 1 also writes A = k
    EPI_STAGE_ACC = 2, // A = A + c2 k ; Y = U0 + c1 k
    EPI_RK4FIN = 3,    // Y = U0 + c1 (A + k)
    EPI_COMB = 4,      // Y = c0 U0 + c1 (X + c2 k), X = the RHS input state
    EPI_LSRK = 5,      // A = c2 A + c1 k (acc_mode 1: A = c1 k) ; Y = X + c3 A
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
    // Device time step: when dtp is set, the stage's dt-proportional
    // coefficient (c2 of EPI_COMB, c1 otherwise) holds a multiplier m and
    // becomes (m * dt) / dtq with dt = *dtp, the host's arithmetic.
    const double *dtp;
    double dtq;
};

// The conforming face term fused into the volume kernel's last phase (a
// face plan's element table; eface == nullptr: none).  Per element face:
// (neighbour << 4) | (neighbour's face << 1) | side, -1 where not conforming.
template <class T> struct VolFaces {
    const int64_t *eface;
    const T *wq; // face-rule weight by face DOF, (N+1)^2 values, 0 past the N-point rule
    T p[4];      // scalar flux parameters (vx, vy, vz, lambda)
    int kind;    // DRF_FLUX_UPWIND or DRF_FLUX_LF_ADV
    int nq;      // face DOFs that carry the N-point rule
};

template <class T> struct VolGeom {
    int64_t E;
    View<const T> h; // bathymetry at every node (fs unused)
    View<const T> f; // Coriolis parameter at every node; f.p == nullptr -> f0
    double f0;
    int wg_per_cu;   // launch shape: resident workgroups per CU (0: occupancy limit)
    int xcd_remap;   // renumber workgroups so neighbouring batches share an XCD
    int wg_lines;    // grid lines per workgroup: 128, 256 or 512 (orders above 4: 128)
    VolFaces<T> faces; // conforming face term added in the plain RHS (orders 1-4)
};

// Resident workgroups per CU of each plain-RHS volume kernel, by [FP64 / FP32]
// [128 / 256 / 512 lines][order][with the fused face phase]; set by the
// kernel's first launch (0 before), read by the autotuner (api.cpp).
extern int g_rhs_per_cu[2][3][kMaxN][2];

// Launchers, one translation unit per workgroup size (volume_fp64*.hip,
// volume_fp32*.hip).  Return hipError_t of the launch.  ops is always given
// in FP64; the FP32 launchers round it once on the host.
hipError_t launch_volume_fp64_w128(int order, const VolGeom<double> &geo, const Ops &ops,
                                   const Phys &ph, View<const double> U, const Epi<double> &epi,
                                   hipStream_t s);
hipError_t launch_volume_fp64_w256(int order, const VolGeom<double> &geo, const Ops &ops,
                                   const Phys &ph, View<const double> U, const Epi<double> &epi,
                                   hipStream_t s);
hipError_t launch_volume_fp64_w512(int order, const VolGeom<double> &geo, const Ops &ops,
                                   const Phys &ph, View<const double> U, const Epi<double> &epi,
                                   hipStream_t s);
hipError_t launch_volume_fp32_w128(int order, const VolGeom<float> &geo, const Ops &ops,
                                   const Phys &ph, View<const float> U, const Epi<float> &epi,
                                   hipStream_t s);
hipError_t launch_volume_fp32_w256(int order, const VolGeom<float> &geo, const Ops &ops,
                                   const Phys &ph, View<const float> U, const Epi<float> &epi,
                                   hipStream_t s);
hipError_t launch_volume_fp32_w512(int order, const VolGeom<float> &geo, const Ops &ops,
                                   const Phys &ph, View<const float> U, const Epi<float> &epi,
                                   hipStream_t s);

inline hipError_t launch_volume_fp64(int order, const VolGeom<double> &geo, const Ops &ops,
                                   const Phys &ph, View<const double> U, const Epi<double> &epi,
                                   hipStream_t s)
{
    switch (order <= 4 ? geo.wg_lines : 128) {
    case 256: return launch_volume_fp64_w256(order, geo, ops, ph, U, epi, s);
    case 512: return launch_volume_fp64_w512(order, geo, ops, ph, U, epi, s);
    default: return launch_volume_fp64_w128(order, geo, ops, ph, U, epi, s);
    }
}

inline hipError_t launch_volume_fp32(int order, const VolGeom<float> &geo, const Ops &ops,
                                   const Phys &ph, View<const float> U, const Epi<float> &epi,
                                   hipStream_t s)
{
    switch (order <= 4 ? geo.wg_lines : 128) {
    case 256: return launch_volume_fp32_w256(order, geo, ops, ph, U, epi, s);
    case 512: return launch_volume_fp32_w512(order, geo, ops, ph, U, epi, s);
    default: return launch_volume_fp32_w128(order, geo, ops, ph, U, epi, s);
    }
}

// Vector kernels (stepper.hip)
hipError_t launch_wave_speed(int order, int64_t E, View<const double> U, View<const double> h,
                             double *partial, int64_t npart, hipStream_t s);
hipError_t launch_wave_speed_derived(int order, int64_t E, const double *H, const double *u,
                                     const double *v, int64_t es, double *partial,
                                     int64_t npart, hipStream_t s);
hipError_t launch_max_reduce(const double *partial, int64_t npart, double *out, hipStream_t s);
// AdaptiveTimeController::compute_dt of the device maximum *cmax into *dt,
// and *t += dt when t is not null
hipError_t launch_adaptive_dt(const double *cmax, double dx_min, double cfl, double dt_min,
                              double dt_max, double *dt, double *t, hipStream_t s);

// Face term of an attached face plan (faces_api.cpp): dU += the plan's
// conforming and mortar contributions.
drf_status face_term(const drf_face_plan *p, const void *U, void *dU, int layout, hipStream_t s);
// The conforming part as the volume kernel's fused phase (false when the
// plan keeps it out of the volume kernel), and the mortar groups alone.
template <class T> bool volume_faces(const drf_face_plan *p, int layout, VolFaces<T> *out);
drf_status face_term_mortars(const drf_face_plan *p, const void *U, void *dU, int layout,
                             hipStream_t s);

// Device address of the file-scope scratch array of noalloc.hip, looked up
// once per device at context creation (the device must be current).
drf_status register_scratch(int device);

// Mode split (modesplit.hip)
hipError_t launch_mode_split(int order, int64_t E, double dt3, int nsub, double *full,
                             View<const double> h, View<const double> f, double f0,
                             const Ops &ops, const Phys &ph, hipStream_t s);

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
    // bathymetry / Coriolis at every node, stored in the LINES layout
    double *d_h = nullptr;
    double *d_f = nullptr;      // null when f is uniform (beta == 0)
    float *d_h32 = nullptr;     // mixed precision copies
    float *d_f32 = nullptr;
    double f0 = 0.0;
    // host-packed staging and RK work space (grown on demand)
    void *d_work = nullptr;
    size_t work_bytes = 0;
    double *d_red = nullptr; // reduction scratch
    int64_t red_len = 0;
    // volume-kernel launch shape (drf_set_launch / drf_autotune)
    int wg_per_cu = 0;
    int xcd_remap = 1;
    int wg_lines = 128;
    // face term added to every right-hand side (drf_set_face_plan), or null
    drf_face_plan *faces = nullptr;
};
