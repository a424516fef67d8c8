// Refinement-side kernels on gfx950: the gradient error estimator with
// element marking, and nodal -> orthonormal Legendre modal coefficients.
//
// Replaces GradientErrorEstimator::estimate / estimate_element /
// compute_directional_mask (src/amr/refinement.cpp:44-161) and the marking
// loop of AdaptiveMeshRefinement::mark_elements (:488-520, can_refine /
// can_coarsen :598-606).  One wavefront per element, grid-stride over
// elements; the element's nodes are staged in the wave's slice of LDS so
// that line variations along xi, eta and zeta read LDS, not HBM.  Every
// value is a max, min, difference or single product, so results are
// bit-identical to the sequential reference loop.
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
__device__ __forceinline__ double wave_reduce_min(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = dmin(v, __shfl_xor(v, o));
    return v;
}

// Pass 1: raw indicator value[e] = (max - min) * h_char[e] and the
// directional mask; partial[block] = max of the block's values.
__global__ void __launch_bounds__(64 * kWaves)
    estimate_kernel(int n, int64_t E, const double *__restrict__ field, int64_t es,
                    const double *__restrict__ hchar, double smooth, double *__restrict__ value,
                    uint8_t *__restrict__ mask, double *__restrict__ partial)
{
    extern __shared__ double sm[];
    __shared__ double red[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n2 = n * n, nd = n2 * n;
    double *u = sm + (size_t)w * nd;
    double bmax = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * kWaves + w; e < E; e += (int64_t)gridDim.x * kWaves) {
        const double *src = field + e * es;
        double hi = -INFINITY, lo = INFINITY;
        for (int d = lane; d < nd; d += 64) {
            const double v = src[d];
            u[d] = v;
            hi = dmax(hi, v);
            lo = dmin(lo, v);
        }
        hi = wave_reduce_max(hi);
        lo = wave_reduce_min(lo);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // lines: 0..n2-1 along xi, n2..2n2-1 along eta, 2n2..3n2-1 along zeta
        double vx = 0.0, vy = 0.0, vz = 0.0;
        for (int L = lane; L < 3 * n2; L += 64) {
            const int dir = L / n2, r = L - dir * n2;
            const int a = r % n, b = r / n;
            int base, stride;
            if (dir == 0) { // (j, k) = (a, b)
                base = b * n2 + a * n;
                stride = 1;
            } else if (dir == 1) { // (i, k) = (a, b)
                base = b * n2 + a;
                stride = n;
            } else { // (i, j) = (a, b)
                base = b * n + a;
                stride = n2;
            }
            double l0 = u[base], h0 = l0;
            for (int t = 1; t < n; ++t) {
                const double v = u[base + t * stride];
                l0 = dmin(l0, v);
                h0 = dmax(h0, v);
            }
            const double var = h0 - l0;
            if (dir == 0)
                vx = dmax(vx, var);
            else if (dir == 1)
                vy = dmax(vy, var);
            else
                vz = dmax(vz, var);
        }
        vx = wave_reduce_max(vx);
        vy = wave_reduce_max(vy);
        vz = wave_reduce_max(vz);
        const double val = (hi - lo) * hchar[e];
        bmax = dmax(bmax, val);
        if (lane == 0) {
            const double mv = dmax(dmax(vx, vy), vz);
            uint8_t m = 0;
            if (!(mv < 1e-14)) {
                const double thr = smooth * mv;
                m = (vx > thr ? 1 : 0) | (vy > thr ? 2 : 0) | (vz > thr ? 4 : 0);
            }
            value[e] = val;
            mask[e] = m;
        }
        __builtin_amdgcn_wave_barrier(); // LDS slice reused by the next element
    }
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

// Nodal -> modal by sum factorisation: three passes of n x n contractions
// (xi, then eta, then zeta) through the wave's LDS slice; one wavefront per
// element.  OUT = double (coefficients) or uint16_t (BF16 features, scaled).
template <class OUT>
__global__ void __launch_bounds__(64 * kWaves)
    modal_kernel(int n, int64_t E, const double *__restrict__ u, int64_t es, ModalOps ops,
                 OUT *__restrict__ out, int64_t out_stride, int out_offset, double scale,
                 const double *__restrict__ elem_scale)
{
    extern __shared__ double sm[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n2 = n * n, nd = n2 * n;
    double *a = sm + (size_t)w * 2 * nd, *b = a + nd;
    for (int64_t e = (int64_t)blockIdx.x * kWaves + w; e < E; e += (int64_t)gridDim.x * kWaves) {
        const double *src = u + e * es;
        for (int d = lane; d < nd; d += 64)
            a[d] = src[d];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int d = lane; d < nd; d += 64) { // b[p + n(j + n k)] = sum_i Vi[p][i] a[i + ...]
            const int p = d % n, jk = d / n;
            double s = 0.0;
            for (int i = 0; i < n; ++i)
                s += ops.Vinv[p * kMaxN + i] * a[jk * n + i];
            b[d] = s;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int d = lane; d < nd; d += 64) { // a[p + n(q + n k)] = sum_j Vi[q][j] b[p + n(j + n k)]
            const int p = d % n, q = (d / n) % n, k = d / n2;
            double s = 0.0;
            for (int j = 0; j < n; ++j)
                s += ops.Vinv[q * kMaxN + j] * b[k * n2 + j * n + p];
            a[d] = s;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        OUT *dst = out + e * out_stride + out_offset;
        const double sc = elem_scale ? scale * elem_scale[e] : scale;
        for (int d = lane; d < nd; d += 64) { // coef[p + n(q + n r)] = sum_k Vi[r][k] a[p + n(q + n k)]
            const int pq = d % n2, r = d / n2;
            double s = 0.0;
            for (int k = 0; k < n; ++k)
                s += ops.Vinv[r * kMaxN + k] * a[k * n2 + pq];
            if constexpr (sizeof(OUT) == 8)
                dst[d] = s;
            else // features: the element mean (mode 0) carries no variation
                dst[d] = (d == 0) ? (uint16_t)0 : to_bf16((float)(s * sc));
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

unsigned grid_for(int64_t E, int64_t cap)
{
    const int64_t need = (E + kWaves - 1) / kWaves;
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
    const unsigned g1 = grid_for(E, kEstimatorPartials);
    const size_t lds = sizeof(double) * (size_t)kWaves * n * n * n;
    hipLaunchKernelGGL(estimate_kernel, dim3(g1), dim3(64 * kWaves), lds, s, n, E, field, es, hchar,
                       p.smoothness_factor, value, mask, partial);
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
    const unsigned g = grid_for(E, 8192);
    const size_t lds = sizeof(double) * (size_t)kWaves * 2 * n * n * n;
    if (coef)
        hipLaunchKernelGGL(modal_kernel<double>, dim3(g), dim3(64 * kWaves), lds, s, n, E, u, es,
                           ops, coef, out_stride, out_offset, scale, elem_scale);
    else
        hipLaunchKernelGGL(modal_kernel<uint16_t>, dim3(g), dim3(64 * kWaves), lds, s, n, E, u, es,
                           ops, bf16, out_stride, out_offset, scale, elem_scale);
    return hipGetLastError();
}

} // namespace drf
