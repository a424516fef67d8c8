// AMR solution transfer on gfx950: evaluation of an element's nodal
// polynomial at the LGL nodes of sub-boxes of its reference cube
// (SolutionProjection::project_to_children, src/amr/refinement.cpp:217-276,
// as AdaptiveMeshRefinement::transfer_solution :533-567 applies it) and the
// reverse gather from children (project_from_children :278-342).
//
// The reference evaluates interpolate_3d (:344-371), a sum over all (N+1)^3
// nodes, at every child node.  Here the same tensor-product interpolant is
// applied by sum factorisation through LDS.  Children of one parent share
// their passes: the xi pass makes both x-halves at once, the eta pass the four
// (x, y) quarters, the zeta pass the eight children, (2 + 4 + 8) (N+1)^4
// FMAs per parent instead of (N+1)^6 per child node set.  The 1D matrices of
// the whole interval and its two halves are built once on the host with the
// reference's own expressions (coordinate lo + (x + 1) (hi - lo) / 2,
// Lagrange products of quotients in node order), so they match the
// reference bit for bit; only the summation order of the interpolant differs.
// Loads and stores are coalesced along the node index; a parent's children
// are consecutive elements, so its stores form one contiguous span.
#include "amr.hpp"

#include <cmath>

namespace drf {
namespace {

constexpr int kThreads = 256;
constexpr int kCopyChunk = 16; // elements per chunk of the ranked copy (many workgroups)

template <int N>
struct Shape {
    static constexpr int N2 = N * N, ND = N2 * N;
    static constexpr int EPB = ND >= 1024 ? 1 : 1024 / ND;       // outputs per batch (boxes)
    static constexpr int PPB = 8 * ND >= 2048 ? 1 : 2048 / (8 * ND); // parents per batch
    // outputs per batch of the box interpolation: as many elements as the
    // workgroup has lanes for their grid lines (10 of 25 lines at N = 5,
    // 250 lanes busy, where 1,024 nodes per batch kept 200)
    static constexpr int BPB = N2 >= kThreads ? 1 : kThreads / N2;
};

__device__ __forceinline__ int64_t parent_source(BoxMode mode, int64_t p,
                                                 const int64_t *__restrict__ idx)
{
    return mode == BoxMode::Transfer ? idx[p] : p;
}

// Children of PPB parents per batch.  Transfer: parents idx[p], XYZ, written
// after the n_keep copies; Children: parent p = element p, the given mask.
template <int N, BoxMode MODE>
__device__ __forceinline__ void children_body(int64_t bid, int64_t nblk, int mask,
                                              int64_t n_parent, int64_t n_src,
                                              const double *__restrict__ U, int64_t es,
                                              int64_t base, const int64_t *__restrict__ idx,
                                              const NodeSet &ns, double *__restrict__ out,
                                              int64_t os)
{
    using S = Shape<N>;
    constexpr int N2 = S::N2, ND = S::ND, PPB = S::PPB;
    constexpr int VPT = (PPB * ND + kThreads - 1) / kThreads; // parent values per thread
    __shared__ double sM[3 * N2], sU[PPB * ND], sT1[PPB * 2 * ND], sT2[PPB * 4 * ND];
    const int t = threadIdx.x;
    const int rx = mask & 1, ry = (mask >> 1) & 1, rz = (mask >> 2) & 1;
    const int nx = 1 + rx, ny = 1 + ry, nz = 1 + rz, nxy = nx * ny, nch = nxy * nz;
    for (int q = t; q < 3 * N2; q += kThreads) {
        const int b = q / N2, r = q - b * N2;
        sM[q] = ns.M[b][(r / N) * kMaxN + r % N];
    }
    const int64_t nb = (n_parent + PPB - 1) / PPB;
    // the parents' values of batch b, loaded ahead into registers
    double nxt[VPT];
    auto fetch = [&](int64_t b) {
        const int64_t p0 = b * PPB;
        const int cnt = (n_parent - p0 < PPB) ? (int)(n_parent - p0) : PPB;
#pragma unroll
        for (int v = 0; v < VPT; ++v) {
            const int q = t + v * kThreads;
            nxt[v] = 0.0;
            if (q < cnt * ND) {
                const int e = q / ND, d = q - e * ND;
                const int64_t src = parent_source(MODE, p0 + e, idx);
                nxt[v] = (src >= 0 && src < n_src) ? U[src * es + d] : __builtin_nan("");
            }
        }
    };
    int64_t b = bid;
    if (b < nb)
        fetch(b);
    for (; b < nb; b += nblk) {
        const int64_t p0 = b * PPB;
        const int cnt = (n_parent - p0 < PPB) ? (int)(n_parent - p0) : PPB;
        // sU was last read before two barriers of the previous batch
#pragma unroll
        for (int v = 0; v < VPT; ++v) {
            const int q = t + v * kThreads;
            if (q < cnt * ND)
                sU[q] = nxt[v];
        }
        if (b + nblk < nb)
            fetch(b + nblk);
        __syncthreads();
        // xi: T1[e][hx][k][j][p] = sum_i Mx_hx[p][i] u[e][k][j][i]
        // nx, ny, nz are 1 or 2: unit u = q / ND splits by shifts and masks
        for (int q = t; q < cnt * nx * ND; q += kThreads) {
            const int u = q / ND, d = q - u * ND, e = u >> rx, hx = u & rx;
            const int line = d / N, p = d - line * N;
            const double *M = sM + (rx ? 1 + hx : 0) * N2 + p * N, *x = sU + e * ND + line * N;
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
                s += M[i] * x[i];
            sT1[q] = s;
        }
        __syncthreads();
        // eta: T2[e][hy][hx][k][q][p] = sum_j My_hy[q][j] T1[e][hx][k][j][p]
        for (int q = t; q < cnt * nxy * ND; q += kThreads) {
            const int u = q / ND, d = q - u * ND, e = u >> (rx + ry), hyx = u & (nxy - 1);
            const int hy = hyx >> rx, hx = hyx & rx;
            const int k = d / N2, rr = d - k * N2, qq = rr / N, p = rr - qq * N;
            const double *M = sM + (ry ? 1 + hy : 0) * N2 + qq * N;
            const double *T = sT1 + (e * nx + hx) * ND + k * N2 + p;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j)
                s += M[j] * T[j * N];
            sT2[q] = s;
        }
        __syncthreads();
        // zeta: child c = hx + nx (hy + ny hz): out[r][L] = sum_k Mz_hz[r][k] T2[e][hy][hx][k][L]
        for (int q = t; q < cnt * nch * ND; q += kThreads) {
            const int u = q / ND, d = q - u * ND, e = u >> (rx + ry + rz), c = u & (nch - 1);
            const int hz = c >> (rx + ry), hyx = c & (nxy - 1);
            const int rr = d / N2, L = d - rr * N2;
            const double *M = sM + (rz ? 1 + hz : 0) * N2 + rr * N;
            const double *T = sT2 + (e * nxy + hyx) * ND + L;
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k)
                s += M[k] * T[k * N2];
            out[(base + (p0 + e) * nch + c) * os + d] = s;
        }
        // no barrier: the next batch writes sU (read only before the eta
        // pass's barrier), then waits before touching sT1 and sT2
    }
}

template <int N, BoxMode MODE>
__global__ void __launch_bounds__(kThreads)
    children_kernel(int mask, int64_t n_parent, int64_t n_src, const double *__restrict__ U,
                    int64_t es, int64_t base, const int64_t *__restrict__ idx, NodeSet ns,
                    double *__restrict__ out, int64_t os)
{
    children_body<N, MODE>(blockIdx.x, gridDim.x, mask, n_parent, n_src, U, es, base, idx, ns,
                           out, os);
}

// transfer_solution's copies: output o < n_keep is the o-th element that is
// not a parent, o + #{j : idx[j] - j <= o} (idx strictly ascending, so
// idx[j] - j is non-decreasing).  One binary search per element.
template <int N>
__device__ __forceinline__ void copy_ranked_body(int64_t bid, int64_t nblk, int64_t n_keep,
                                                 int64_t n_src, const double *__restrict__ U,
                                                 int64_t es, int64_t n_parent,
                                                 const int64_t *__restrict__ idx,
                                                 double *__restrict__ out, int64_t os)
{
    constexpr int ND = N * N * N;
    __shared__ int64_t sSrc[kCopyChunk];
    const int t = threadIdx.x;
    const int64_t nb = (n_keep + kCopyChunk - 1) / kCopyChunk;
    for (int64_t b = bid; b < nb; b += nblk) {
        const int64_t o0 = b * kCopyChunk;
        const int cnt = (n_keep - o0 < kCopyChunk) ? (int)(n_keep - o0) : kCopyChunk;
        if (t < cnt) {
            const int64_t o = o0 + t;
            int64_t lo = 0, hi = n_parent;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (idx[mid] - mid <= o)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            const int64_t src = o + lo;
            sSrc[t] = src < n_src ? src : -1;
        }
        __syncthreads();
        for (int q = t; q < cnt * ND; q += kThreads) {
            const int e = q / ND, d = q - e * ND;
            const int64_t src = sSrc[e];
            out[(o0 + e) * os + d] = (src >= 0) ? U[src * es + d] : __builtin_nan("");
        }
        __syncthreads();
    }
}

// transfer_solution in one launch: the first n_copy_blk workgroups make the
// copies, the others the parents' children.  The two parts write disjoint
// outputs and are each too small to fill the GPU alone, so running them side
// by side instead of one after the other halves the transfer's time.
template <int N>
__global__ void __launch_bounds__(kThreads)
    transfer_kernel(int64_t n_copy_blk, int64_t n_keep, int64_t n_parent, int64_t n_src,
                    const double *__restrict__ U, int64_t es, const int64_t *__restrict__ idx,
                    NodeSet ns, double *__restrict__ out, int64_t os)
{
    if ((int64_t)blockIdx.x < n_copy_blk)
        copy_ranked_body<N>(blockIdx.x, n_copy_blk, n_keep, n_src, U, es, n_parent, idx, out, os);
    else
        children_body<N, BoxMode::Transfer>((int64_t)blockIdx.x - n_copy_blk,
                                            (int64_t)gridDim.x - n_copy_blk, 7, n_parent, n_src,
                                            U, es, n_keep, idx, ns, out, os);
}

// M[p][i] = L_i(lo + (x_p + 1) (hi - lo) / 2), the reference's arithmetic.
__device__ __forceinline__ double lagrange_entry(const NodeSet &ns, int n, int p, int i, double lo,
                                                 double hi)
{
#pragma clang fp contract(off)
    const double xi = lo + 0.5 * (ns.x[p] + 1.0) * (hi - lo);
    double L = 1.0;
    for (int j = 0; j < n; ++j)
        if (j != i)
            L *= (xi - ns.x[j]) / (ns.x[i] - ns.x[j]);
    return L;
}

// 0, 1, 2: the whole interval, its lower or upper half; 3: any other
__device__ __forceinline__ int interval_kind(double lo, double hi)
{
    if (lo == -1.0 && hi == 1.0)
        return 0;
    if (lo == -1.0 && hi == 0.0)
        return 1;
    if (lo == 0.0 && hi == 1.0)
        return 2;
    return 3;
}

// Output element o = source idx[o] on sub-box boxes[6o..6o+5], for each of
// the nf fields; a whole cube is a copy, the standard halves use the shared
// matrices, any other interval gets its own matrix built in LDS.  A batch's
// sources, kinds and matrices are set up once for all its fields.  Each
// thread owns one grid line of an element per slot and makes its N outputs
// per pass: the xi pass reads the line straight from the source field (the
// next field's line is loaded while the current field's other passes run),
// the eta and zeta passes read the line they contract from LDS, and the
// matrices are read from LDS (holding an element's three matrices in
// registers for the batch measured slower at N = 5: 204 VGPRs and 2 waves
// per SIMD, against 70 VGPRs and the 5 workgroups per CU that LDS allows).  Every output is the same sum, in the same
// order, as a pass that makes one output per thread.
template <int N>
__global__ void __launch_bounds__(kThreads)
    interp_boxes_kernel(int64_t n_out, int64_t n_src, BoxFields bf,
                        const int64_t *__restrict__ idx, const double *__restrict__ boxes,
                        NodeSet ns)
{
    using S = Shape<N>;
    constexpr int N2 = S::N2, ND = S::ND, EPB = S::BPB;
    constexpr int SL = (EPB * N2 + kThreads - 1) / kThreads; // grid lines per thread
    __shared__ double sT[EPB * ND], sV[EPB * ND], sM[EPB * 3 * N2], sStd[3 * N2], sBox[EPB * 6];
    __shared__ int64_t sSrc[EPB];
    __shared__ int sKind[EPB * 3], sCopy[EPB];
    const int t = threadIdx.x;
    const int nf = bf.nf;
    for (int q = t; q < 3 * N2; q += kThreads) {
        const int b = q / N2, r = q - b * N2;
        sStd[q] = ns.M[b][(r / N) * kMaxN + r % N];
    }
    const int64_t nb = (n_out + EPB - 1) / EPB;
    for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t o0 = b * EPB;
        const int cnt = (n_out - o0 < EPB) ? (int)(n_out - o0) : EPB;
        const int nl = cnt * N2; // grid lines of the batch
        if (t < cnt) {
            const int64_t o = o0 + t, src = idx[o];
            const bool ok = src >= 0 && src < n_src;
            int whole = 1;
            for (int a = 0; a < 3; ++a) {
                const double lo = boxes[6 * o + 2 * a], hi = boxes[6 * o + 2 * a + 1];
                sBox[t * 6 + 2 * a] = lo;
                sBox[t * 6 + 2 * a + 1] = hi;
                const int kd = interval_kind(lo, hi);
                sKind[t * 3 + a] = kd;
                whole &= (kd == 0);
            }
            sSrc[t] = ok ? src : -1;
            sCopy[t] = (whole || !ok) ? 1 : 0; // an invalid source is written as NaN
        }
        __syncthreads();
        for (int q = t; q < cnt * 3 * N2; q += kThreads) {
            const int e = q / (3 * N2), r = q - e * 3 * N2, a = r / N2, pi = r - a * N2;
            if (sKind[e * 3 + a] != 3)
                continue;
            const int p = pi / N, i = pi - p * N;
            sM[q] = lagrange_entry(ns, N, p, i, sBox[e * 6 + 2 * a], sBox[e * 6 + 2 * a + 1]);
        }
        auto mat = [&](int e, int a) {
            const int kd = sKind[e * 3 + a];
            return kd < 3 ? sStd + kd * N2 : sM + (e * 3 + a) * N2;
        };
        double u[SL][N]; // each slot's source line
        auto gather = [&](int f) {
#pragma unroll
            for (int sl = 0; sl < SL; ++sl) {
                const int q = t + sl * kThreads;
#pragma unroll
                for (int i = 0; i < N; ++i)
                    u[sl][i] = 0.0;
                if (q < nl) {
                    const int e = q / N2, L = q - e * N2;
                    const int64_t src = sSrc[e];
                    const double *x = bf.U[f] + src * bf.es[f] + L * N;
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        u[sl][i] = (src >= 0) ? x[i] : __builtin_nan("");
                }
            }
        };
        gather(0);
        __syncthreads(); // the interval matrices complete before the xi pass reads them
        for (int f = 0; f < nf; ++f) {
            double *__restrict__ out = bf.out[f];
            const int64_t os = bf.os[f];
            // xi: T[e][k][j][p] = sum_i Mx[p][i] u[e][k][j][i], line L = j + N k
            // (sT was last read by the previous field's eta pass, a barrier ago)
#pragma unroll
            for (int sl = 0; sl < SL; ++sl) {
                const int q = t + sl * kThreads;
                if (q >= nl)
                    continue;
                const int e = q / N2, L = q - e * N2;
                if (sCopy[e]) {
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        out[(o0 + e) * os + L * N + i] = u[sl][i];
                    continue;
                }
                const double *Mx = mat(e, 0);
#pragma unroll
                for (int p = 0; p < N; ++p) {
                    double s = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        s += Mx[p * N + i] * u[sl][i];
                    sT[e * ND + L * N + p] = s;
                }
            }
            if (f + 1 < nf)
                gather(f + 1);
            __syncthreads();
            // eta: V[e][k][q][p] = sum_j My[q][j] T[e][k][j][p], thread (k, p) = (L / N, L % N)
#pragma unroll
            for (int sl = 0; sl < SL; ++sl) {
                const int q = t + sl * kThreads;
                if (q >= nl)
                    continue;
                const int e = q / N2, L = q - e * N2;
                if (sCopy[e])
                    continue;
                const int k = L / N, p = L - k * N;
                const double *T = sT + e * ND + k * N2 + p, *My = mat(e, 1);
                double x[N];
#pragma unroll
                for (int j = 0; j < N; ++j)
                    x[j] = T[j * N];
#pragma unroll
                for (int qq = 0; qq < N; ++qq) {
                    double s = 0.0;
#pragma unroll
                    for (int j = 0; j < N; ++j)
                        s += My[qq * N + j] * x[j];
                    sV[e * ND + k * N2 + qq * N + p] = s;
                }
            }
            __syncthreads();
            // zeta: out[e][r][L] = sum_k Mz[r][k] V[e][k][L], L = q + N p in the (eta, xi) plane
#pragma unroll
            for (int sl = 0; sl < SL; ++sl) {
                const int q = t + sl * kThreads;
                if (q >= nl)
                    continue;
                const int e = q / N2, L = q - e * N2;
                if (sCopy[e])
                    continue;
                const double *V = sV + e * ND + L, *Mz = mat(e, 2);
                double x[N];
#pragma unroll
                for (int k = 0; k < N; ++k)
                    x[k] = V[k * N2];
#pragma unroll
                for (int r = 0; r < N; ++r) {
                    double s = 0.0;
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        s += Mz[r * N + k] * x[k];
                    out[(o0 + e) * os + r * N2 + L] = s;
                }
            }
            // sV is rewritten by the next field's eta pass, a barrier on
        }
        __syncthreads(); // the next batch rewrites the sources, kinds and matrices
    }
}

// project_from_children: parent node (pk, pj, pi) takes the child holding it
// (half[a] of each axis) evaluated at its local coordinate; the passes carry
// the (cz, cy) and then cz child index alongside the contracted node index.
template <int N>
__global__ void __launch_bounds__(kThreads)
    from_children_kernel(int mask, int64_t n_parent, const double *__restrict__ C, int64_t cs,
                         FromChildrenOps ops, double *__restrict__ U, int64_t es)
{
    using S = Shape<N>;
    constexpr int N2 = S::N2, ND = S::ND, EPB = S::EPB;
    __shared__ double sX[EPB * 4 * ND], sY[EPB * 2 * ND], sQ[3 * N2];
    __shared__ int sH[3 * N];
    const int t = threadIdx.x;
    const int rx = mask & 1, ry = (mask >> 1) & 1, rz = (mask >> 2) & 1;
    const int nx = 1 + rx, ny = 1 + ry, nz = 1 + rz;
    const int nch = nx * ny * nz, nzy = nz * ny;
    for (int q = t; q < 3 * N2; q += kThreads) {
        const int a = q / N2, r = q - a * N2;
        sQ[q] = ops.Q[a][r];
    }
    for (int q = t; q < 3 * N; q += kThreads)
        sH[q] = ops.half[q / N][q % N];
    __syncthreads();
    const int64_t nb = (n_parent + EPB - 1) / EPB;
    for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t p0 = b * EPB;
        const int cnt = (n_parent - p0 < EPB) ? (int)(n_parent - p0) : EPB;
        for (int q = t; q < cnt * nzy * ND; q += kThreads) {
            const int u = q / ND, d = q - u * ND, e = u >> (ry + rz), czy = u & (nzy - 1);
            const int line = d / N, pi = d - line * N;
            const int cz = czy >> ry, cy = czy & ry;
            const int c = (cz * ny + cy) * nx + sH[pi];
            const double *src = C + ((p0 + e) * nch + c) * cs + line * N;
            const double *Q = sQ + pi * N;
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
                s += Q[i] * src[i];
            sX[q] = s;
        }
        __syncthreads();
        for (int q = t; q < cnt * nz * ND; q += kThreads) {
            const int u = q / ND, d = q - u * ND, e = u >> rz, cz = u & rz;
            const int k = d / N2, rr = d - k * N2, pj = rr / N, pi = rr - pj * N;
            const double *X = sX + e * nzy * ND + (cz * ny + sH[N + pj]) * ND + k * N2 + pi;
            const double *Q = sQ + N2 + pj * N;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j)
                s += Q[j] * X[j * N];
            sY[q] = s;
        }
        __syncthreads();
        for (int q = t; q < cnt * ND; q += kThreads) {
            const int e = q / ND, d = q - e * ND, pk = d / N2, L = d - pk * N2;
            const double *Y = sY + e * nz * ND + sH[2 * N + pk] * ND + L;
            const double *Q = sQ + 2 * N2 + pk * N;
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k)
                s += Q[k] * Y[k * N2];
            U[(p0 + e) * es + d] = s;
        }
        __syncthreads();
    }
}

template <int64_t PER>
unsigned grid_for(int64_t units)
{
    const int64_t nb = (units + PER - 1) / PER;
    return (unsigned)(nb < 4096 ? nb : 4096);
}

template <int N>
hipError_t interp_n(BoxMode mode, int mask, int64_t n_out, int64_t n_src, const BoxFields &bf,
                    int64_t n_keep, int64_t n_parent, const int64_t *idx, const double *boxes,
                    const NodeSet &ns, hipStream_t s)
{
    const double *U = bf.U[0];
    const int64_t es = bf.es[0], os = bf.os[0];
    double *out = bf.out[0];
    using S = Shape<N>;
    const dim3 blk(kThreads);
    switch (mode) {
    case BoxMode::Transfer: {
        const int64_t gc = n_keep > 0 ? grid_for<kCopyChunk>(n_keep) : 0;
        const int64_t gp = n_parent > 0 ? grid_for<S::PPB>(n_parent) : 0;
        if (gc + gp > 0)
            hipLaunchKernelGGL((transfer_kernel<N>), dim3((unsigned)(gc + gp)), blk, 0, s, gc,
                               n_keep, n_parent, n_src, U, es, idx, ns, out, os);
        break;
    }
    case BoxMode::Children: // every source element is a parent
        hipLaunchKernelGGL((children_kernel<N, BoxMode::Children>), dim3(grid_for<S::PPB>(n_src)),
                           blk, 0, s, mask, n_src, n_src, U, es, (int64_t)0, idx, ns, out, os);
        break;
    default:
        hipLaunchKernelGGL((interp_boxes_kernel<N>), dim3(grid_for<S::BPB>(n_out)), blk, 0, s,
                           n_out, n_src, bf, idx, boxes, ns);
        break;
    }
    return hipGetLastError();
}

template <int N>
hipError_t from_children_n(int mask, int64_t n_parent, const double *C, int64_t cs,
                           const FromChildrenOps &ops, double *U, int64_t es, hipStream_t s)
{
    hipLaunchKernelGGL((from_children_kernel<N>), dim3(grid_for<Shape<N>::EPB>(n_parent)),
                       dim3(kThreads), 0, s, mask, n_parent, C, cs, ops, U, es);
    return hipGetLastError();
}

} // namespace

hipError_t launch_interp_boxes(int n, BoxMode mode, int mask, int64_t n_out, int64_t n_src,
                               const BoxFields &bf, int64_t n_keep, int64_t n_parent,
                               const int64_t *idx, const double *boxes, const NodeSet &nodes,
                               hipStream_t s)
{
    if (n_out <= 0)
        return hipSuccess;
    if (bf.nf < 1 || bf.nf > kMaxBoxFields || (mode != BoxMode::Boxes && bf.nf != 1))
        return hipErrorInvalidValue;
#define DRF_INTERP(NN)                                                                        \
    return interp_n<NN>(mode, mask, n_out, n_src, bf, n_keep, n_parent, idx, boxes, nodes, s)
    switch (n) {
    case 2: DRF_INTERP(2);
    case 3: DRF_INTERP(3);
    case 4: DRF_INTERP(4);
    case 5: DRF_INTERP(5);
    case 6: DRF_INTERP(6);
    case 7: DRF_INTERP(7);
    case 8: DRF_INTERP(8);
    case 9: DRF_INTERP(9);
    default: return hipErrorInvalidValue;
    }
#undef DRF_INTERP
}

hipError_t launch_from_children(int n, int mask, int64_t n_parent, const double *children,
                                int64_t cs, const FromChildrenOps &ops, double *U, int64_t es,
                                hipStream_t s)
{
    if (n_parent <= 0)
        return hipSuccess;
#define DRF_FROM(NN) return from_children_n<NN>(mask, n_parent, children, cs, ops, U, es, s)
    switch (n) {
    case 2: DRF_FROM(2);
    case 3: DRF_FROM(3);
    case 4: DRF_FROM(4);
    case 5: DRF_FROM(5);
    case 6: DRF_FROM(6);
    case 7: DRF_FROM(7);
    case 8: DRF_FROM(8);
    case 9: DRF_FROM(9);
    default: return hipErrorInvalidValue;
    }
#undef DRF_FROM
}

} // namespace drf
