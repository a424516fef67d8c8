// Refinement-side kernels on gfx950: the gradient error estimator with
// element marking, and nodal -> orthonormal Legendre modal coefficients.
//
// Replaces GradientErrorEstimator::estimate / estimate_element /
// compute_directional_mask (src/amr/refinement.cpp:44-161) and the marking
// loop of AdaptiveMeshRefinement::mark_elements (:488-520, can_refine /
// can_coarsen :598-606).  A wavefront takes 64 / P elements at a time (P =
// n^2 rounded up to a power of two, at most 64): their nodes are staged in
// the wave's slice of LDS with coalesced loads, and each lane owns grid lines
// of one element, so line variations along xi, eta and zeta read LDS and the
// per-element reductions stay within a group of P lanes.  Every value is a
// max, min, difference or single product, so results are bit-identical to
// the sequential reference loop.
#include "adapt.hpp"
#include "drf_internal.hpp"

namespace drf {
namespace {

constexpr int kWaves = 4; // wavefronts (elements in flight) per workgroup

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }

__device__ __forceinline__ double wave_reduce_max(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = dmax(v, __shfl_xor(v, o));
    return v;
}

// Lanes of a wavefront are split into groups of P = min(64, 2^ceil(log2 n^2))
// lanes, one group per element; lane L of a group owns grid lines L, L + P, ...
// (< n^2) of its element.  Element values are staged in the wave's LDS slice
// with coalesced loads.
struct LineMap {
    int P, epw, lpl; // group size, elements per wave, lines per lane
};

__host__ __device__ constexpr LineMap line_map(int n)
{
    const int n2 = n * n;
    int P = 1;
    while (P < n2 && P < 64)
        P <<= 1;
    return LineMap{P, 64 / P, (n2 + P - 1) / P};
}

__device__ __forceinline__ double group_max(double v, int P)
{
    for (int o = P >> 1; o > 0; o >>= 1)
        v = dmax(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double group_min(double v, int P)
{
    for (int o = P >> 1; o > 0; o >>= 1)
        v = dmin(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A wave's elements e0 .. e0 + epw - 1 (cnt of them), nodal values in
// registers: lane l holds values l, l + 64, ... of the wave's slice.  The
// kernels load the next elements this way while they work on the current
// ones, then store them to the slice.  Per-element scalars (h_char, the
// feature scale) are loaded with them: a load issued after the prefetch
// would make the wave wait for the prefetch too (gfx9 counts vector-memory
// loads in one in-order counter).
template <int n> struct Slice {
    static constexpr int nd = n * n * n;
    static constexpr int VPL = (line_map(n).epw * nd + 63) / 64;
    double v[VPL];
    __device__ __forceinline__ void load(const double *__restrict__ f, int64_t es, int64_t e0,
                                         int cnt, int lane)
    {
#pragma unroll
        for (int k = 0; k < VPL; ++k) {
            const int idx = lane + 64 * k;
            if (idx < cnt * nd) {
                const int el = idx / nd, d = idx - el * nd;
                v[k] = f[(e0 + el) * es + d];
            }
        }
    }
    __device__ __forceinline__ void store(double *u, int cnt, int lane) const
    {
#pragma unroll
        for (int k = 0; k < VPL; ++k) {
            const int idx = lane + 64 * k;
            if (idx < cnt * nd)
                u[idx] = v[k];
        }
    }
};

// Pass 1: raw indicator value[e] = (max - min) * h_char[e] and the
// directional mask; partial[block] = max of the block's values.
template <int n>
__global__ void __launch_bounds__(64 * kWaves)
    estimate_kernel(int64_t E, const double *__restrict__ field, int64_t es,
                    const double *__restrict__ hchar, double smooth, double *__restrict__ value,
                    uint8_t *__restrict__ mask, double *__restrict__ partial)
{
    extern __shared__ double sm[];
    __shared__ double red[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n2 = n * n, nd = n2 * n;
    const LineMap lm = line_map(n);
    const int g = lane / lm.P, gl = lane % lm.P; // element within the wave, lane within it
    double *u = sm + (size_t)w * lm.epw * nd;
    double bmax = 0.0;
    const int64_t waves = (int64_t)gridDim.x * kWaves;
    const int64_t step = waves * lm.epw;
    auto count = [&](int64_t e) { return (E - e < lm.epw) ? (int)(E - e) : lm.epw; };
    Slice<n> nx;
    int64_t e0 = ((int64_t)blockIdx.x * kWaves + w) * lm.epw;
    double hc_next = 0.0; // h_char of the lane's element in the next slice
    auto load_next = [&](int64_t e) {
        nx.load(field, es, e, count(e), lane);
        hc_next = (g < count(e)) ? hchar[e + g] : 0.0;
    };
    if (e0 < E)
        load_next(e0);
    for (; e0 < E; e0 += step) {
        const int cnt = count(e0);
        nx.store(u, cnt, lane);
        const double hc = hc_next;
        if (e0 + step < E)
            load_next(e0 + step);
        wave_sync_lds();
        double hi = -INFINITY, lo = INFINITY, vx = 0.0, vy = 0.0, vz = 0.0;
        if (g < cnt) {
            const double *ue = u + g * nd;
            for (int t = 0; t < lm.lpl; ++t) {
                const int L = gl + t * lm.P;
                if (L >= n2)
                    break;
                const int a = L % n, b = L / n;
                // xi-line (j, k) = (a, b): also the element's max and min
                double l0 = ue[b * n2 + a * n], h0 = l0;
                for (int i = 1; i < n; ++i) {
                    const double v = ue[b * n2 + a * n + i];
                    l0 = dmin(l0, v);
                    h0 = dmax(h0, v);
                }
                vx = dmax(vx, h0 - l0);
                hi = dmax(hi, h0);
                lo = dmin(lo, l0);
                // eta-line (i, k) = (a, b)
                l0 = ue[b * n2 + a];
                h0 = l0;
                for (int j = 1; j < n; ++j) {
                    const double v = ue[b * n2 + j * n + a];
                    l0 = dmin(l0, v);
                    h0 = dmax(h0, v);
                }
                vy = dmax(vy, h0 - l0);
                // zeta-line (i, j) = (a, b)
                l0 = ue[b * n + a];
                h0 = l0;
                for (int k = 1; k < n; ++k) {
                    const double v = ue[k * n2 + b * n + a];
                    l0 = dmin(l0, v);
                    h0 = dmax(h0, v);
                }
                vz = dmax(vz, h0 - l0);
            }
        }
        hi = group_max(hi, lm.P);
        lo = group_min(lo, lm.P);
        vx = group_max(vx, lm.P);
        vy = group_max(vy, lm.P);
        vz = group_max(vz, lm.P);
        if (g < cnt) {
            const int64_t e = e0 + g;
            const double val = (hi - lo) * hc;
            bmax = dmax(bmax, val);
            if (gl == 0) {
                const double mv = dmax(dmax(vx, vy), vz);
                uint8_t m = 0;
                if (!(mv < 1e-14)) {
                    const double thr = smooth * mv;
                    m = (vx > thr ? 1 : 0) | (vy > thr ? 2 : 0) | (vz > thr ? 4 : 0);
                }
                value[e] = val;
                mask[e] = m;
            }
        }
        wave_sync_lds(); // the slice is reloaded for the next elements
    }
    bmax = group_max(bmax, 64);
    if (lane == 0)
        red[w] = bmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int i = 1; i < kWaves; ++i)
            m = dmax(m, red[i]);
        partial[blockIdx.x] = m;
    }
}


// Pass 2: global max from the partials (every block), normalisation,
// thresholds, then mark_elements when levels are given.
__global__ void __launch_bounds__(256)
    normalize_mark_kernel(int64_t E, const double *__restrict__ partial, int npart,
                          EstimatorParams p, const int32_t *__restrict__ levels,
                          double *__restrict__ max_out, double *__restrict__ value,
                          const uint8_t *__restrict__ mask,
                          uint8_t *__restrict__ refine, uint8_t *__restrict__ coarsen,
                          uint8_t *__restrict__ action, uint8_t *__restrict__ mask_out)
{
    __shared__ double red[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < npart; i += 256)
        m = dmax(m, partial[i]);
    m = wave_reduce_max(m);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = m;
    __syncthreads();
    const double mx = dmax(dmax(red[0], red[1]), dmax(red[2], red[3]));
    if (max_out && blockIdx.x == 0 && threadIdx.x == 0)
        *max_out = mx;
    const bool norm = mx > 1e-14;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * 256) {
        bool r = false, c = false;
        if (norm) {
            const double v = value[e] / mx;
            value[e] = v;
            r = v > p.refine_threshold;
            c = v < p.coarsen_threshold;
        }
        refine[e] = r;
        coarsen[e] = c;
        if (levels) {
            const int lx = levels[3 * e], ly = levels[3 * e + 1], lz = levels[3 * e + 2];
            const bool can_ref = lx < p.max_level && ly < p.max_level && lz < p.max_level;
            const bool can_coa = lx > p.min_level || ly > p.min_level || lz > p.min_level;
            uint8_t a = 0, mo = 0;
            if (r && can_ref) {
                a = 1;
                mo = mask[e] ? mask[e] : 7;
            } else if (c && can_coa) {
                a = 2;
            }
            action[e] = a;
            mask_out[e] = mo;
        }
    }
}

// Round-to-nearest-even FP32 -> BF16 (NaN -> quiet NaN).
__device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}

// Nodal -> modal by sum factorisation, line-parallel: three passes of n x n
// contractions (xi, then eta, then zeta) through the wave's LDS slice, lanes
// owning grid lines.  OUT = double (coefficients) or uint16_t (BF16
// features, scaled, mean mode zeroed).
template <int n, class OUT>
__global__ void __launch_bounds__(64 * kWaves)
    modal_kernel(int64_t E, const double *__restrict__ u, int64_t es, ModalOps ops,
                 OUT *__restrict__ out, int64_t out_stride, int out_offset, double scale,
                 const double *__restrict__ elem_scale)
{
    extern __shared__ double sm[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n2 = n * n, nd = n2 * n;
    const LineMap lm = line_map(n);
    const int g = lane / lm.P, gl = lane % lm.P;
    double *a = sm + (size_t)w * 2 * lm.epw * nd, *bb = a + lm.epw * nd;
    const int64_t waves = (int64_t)gridDim.x * kWaves;
    const int64_t step = waves * lm.epw;
    auto count = [&](int64_t e) { return (E - e < lm.epw) ? (int)(E - e) : lm.epw; };
    Slice<n> nx;
    int64_t e0 = ((int64_t)blockIdx.x * kWaves + w) * lm.epw;
    double es_next = 1.0; // elem_scale of the lane's element in the next slice
    auto load_next = [&](int64_t e) {
        nx.load(u, es, e, count(e), lane);
        es_next = (elem_scale && g < count(e)) ? elem_scale[e + g] : 1.0;
    };
    if (e0 < E)
        load_next(e0);
    for (; e0 < E; e0 += step) {
        const int cnt = count(e0);
        nx.store(a, cnt, lane);
        const double sc = elem_scale ? scale * es_next : scale; // used after the next load
        if (e0 + step < E)
            load_next(e0 + step);
        wave_sync_lds();
        const bool act = g < cnt;
        double *ae = a + g * nd, *be = bb + g * nd;
        if (act) // xi: b[p + n L] = sum_i Vi[p][i] a[i + n L], L = (j, k)
            for (int L = gl; L < n2; L += lm.P)
                for (int p = 0; p < n; ++p) {
                    double s = 0.0;
                    for (int i = 0; i < n; ++i)
                        s += ops.Vinv[p * kMaxN + i] * ae[L * n + i];
                    be[L * n + p] = s;
                }
        wave_sync_lds();
        if (act) // eta: a[p + n(q + n k)] = sum_j Vi[q][j] b[p + n(j + n k)], L = (p, k)
            for (int L = gl; L < n2; L += lm.P) {
                const int p = L % n, k = L / n;
                for (int q = 0; q < n; ++q) {
                    double s = 0.0;
                    for (int j = 0; j < n; ++j)
                        s += ops.Vinv[q * kMaxN + j] * be[k * n2 + j * n + p];
                    ae[k * n2 + q * n + p] = s;
                }
            }
        wave_sync_lds();
        if (act) { // zeta: coef[L + n2 r] = sum_k Vi[r][k] a[L + n2 k], L = (p, q)
            const int64_t e = e0 + g;
            OUT *dst = out + e * out_stride + out_offset;
            for (int L = gl; L < n2; L += lm.P)
                for (int r = 0; r < n; ++r) {
                    double s = 0.0;
                    for (int k = 0; k < n; ++k)
                        s += ops.Vinv[r * kMaxN + k] * ae[k * n2 + L];
                    const int d = L + n2 * r;
                    if constexpr (sizeof(OUT) == 8)
                        dst[d] = s;
                    else // features: the element mean (mode 0) carries no variation
                        dst[d] = (d == 0) ? (uint16_t)0 : to_bf16((float)(s * sc));
                }
        }
        wave_sync_lds();
    }
}

unsigned grid_for(int n, int64_t E, int64_t cap)
{
    const int64_t per_block = (int64_t)kWaves * line_map(n).epw;
    const int64_t need = (E + per_block - 1) / per_block;
    return (unsigned)(need < cap ? need : cap);
}

} // namespace

hipError_t launch_estimate(int n, int64_t E, const double *field, int64_t es, const double *hchar,
                           const int32_t *levels, const EstimatorParams &p, double *partial,
                           double *max_out, double *value, uint8_t *mask, uint8_t *refine, uint8_t *coarsen,
                           uint8_t *action, uint8_t *mask_out, hipStream_t s)
{
    if (E <= 0)
        return hipSuccess;
    const unsigned g1 = grid_for(n, E, kEstimatorPartials);
    const size_t lds = sizeof(double) * (size_t)kWaves * line_map(n).epw * n * n * n;
#define DRF_EST(NN)                                                                           \
    hipLaunchKernelGGL(estimate_kernel<NN>, dim3(g1), dim3(64 * kWaves), lds, s, E, field, es,  \
                       hchar, p.smoothness_factor, value, mask, partial)
    switch (n) {
    case 2: DRF_EST(2); break;
    case 3: DRF_EST(3); break;
    case 4: DRF_EST(4); break;
    case 5: DRF_EST(5); break;
    case 6: DRF_EST(6); break;
    case 7: DRF_EST(7); break;
    case 8: DRF_EST(8); break;
    case 9: DRF_EST(9); break;
    default: return hipErrorInvalidValue;
    }
#undef DRF_EST
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return err;
    const int64_t g2n = (E + 255) / 256;
    const unsigned g2 = (unsigned)(g2n < 1024 ? g2n : 1024);
    hipLaunchKernelGGL(normalize_mark_kernel, dim3(g2), dim3(256), 0, s, E, partial, (int)g1, p,
                       levels, max_out, value, mask, refine, coarsen, action, mask_out);
    return hipGetLastError();
}

hipError_t launch_modal(int n, int64_t E, const double *u, int64_t es, const ModalOps &ops,
                        double *coef, uint16_t *bf16, int64_t out_stride, int out_offset,
                        double scale, const double *elem_scale, hipStream_t s)
{
    if (E <= 0)
        return hipSuccess;
    const unsigned g = grid_for(n, E, 8192);
    const size_t lds = sizeof(double) * (size_t)kWaves * 2 * line_map(n).epw * n * n * n;
#define DRF_MODAL(NN)                                                                         \
    do {                                                                                      \
        if (coef)                                                                             \
            hipLaunchKernelGGL((modal_kernel<NN, double>), dim3(g), dim3(64 * kWaves), lds, s, E, \
                               u, es, ops, coef, out_stride, out_offset, scale, elem_scale);  \
        else                                                                                  \
            hipLaunchKernelGGL((modal_kernel<NN, uint16_t>), dim3(g), dim3(64 * kWaves), lds, s,  \
                               E, u, es, ops, bf16, out_stride, out_offset, scale, elem_scale); \
    } while (0)
    switch (n) {
    case 2: DRF_MODAL(2); break;
    case 3: DRF_MODAL(3); break;
    case 4: DRF_MODAL(4); break;
    case 5: DRF_MODAL(5); break;
    case 6: DRF_MODAL(6); break;
    case 7: DRF_MODAL(7); break;
    case 8: DRF_MODAL(8); break;
    case 9: DRF_MODAL(9); break;
    default: return hipErrorInvalidValue;
    }
#undef DRF_MODAL
    return hipGetLastError();
}

} // namespace drf
