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

// Every layout is a sequence of blocks, each holding the five fields' spans
// of fs values one after the other: LINES and SOA one block (fs = E nd),
// PACKED one per element (fs = nd), BLOCKS one per B elements (fs = B nd).
// All views share the layout, so position t of field 0 in memory order is
// (t / fs) 5 fs + t % fs and field f adds f fs: thread t walks positions
// consecutively, and consecutive lanes touch consecutive addresses.  Slots
// past element E - 1 in the last block of BLOCKS are skipped.
template <class I>
__global__ void __launch_bounds__(256) lincomb_kernel(I npos, I fs, int n2, int bb, int64_t E,
                                                     LinComb lc)
{
    const I stride = (I)gridDim.x * blockDim.x;
    const bool one_block = npos == fs;
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < npos; t += stride) {
        const I blk = one_block ? (I)0 : t / fs;
        const I r = t - blk * fs;
        if (bb && (int64_t)(blk + 1) * bb > E && // the last block, partly filled
            (int64_t)blk * bb + (int64_t)((r / (I)n2) % (I)bb) >= E)
            continue;
        const int64_t o = (int64_t)blk * 5 * (int64_t)fs + (int64_t)r;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const int64_t of = o + (int64_t)f * fs;
            double s = lc.c[0] * lc.K[0].p[of];
            for (int k = 1; k < lc.nk; ++k)
                s = s + lc.c[k] * lc.K[k].p[of];
            s = lc.gamma * s;
            if (lc.X1.p)
                s = lc.X1.p[of] + s;
            s = lc.beta * s;
            double y = s;
            if (lc.X0.p)
                y = lc.alpha * lc.X0.p[of] + s;
            lc.Y.p[of] = y;
        }
    }
}

} // namespace

hipError_t launch_lincomb(int64_t E, int n, const LinComb &lc, hipStream_t s)
{
    const int64_t nd = (int64_t)n * n * n, fs = lc.Y.fs;
    if (E <= 0)
        return hipSuccess;
    // positions of one field: whole blocks (the views' block size, 1 element
    // for PACKED, all of them for LINES and SOA)
    const int64_t npos = (fs == E * nd) ? fs : ((E * nd + fs - 1) / fs) * fs;
    int64_t blocks = (npos + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    const int bb = (int)lc.Y.bb;
    if (npos < ((int64_t)1 << 31))
        hipLaunchKernelGGL(lincomb_kernel<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, s,
                           (uint32_t)npos, (uint32_t)fs, n * n, bb, E, lc);
    else
        hipLaunchKernelGGL(lincomb_kernel<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, s,
                           (uint64_t)npos, (uint64_t)fs, n * n, bb, E, lc);
    return hipGetLastError();
}

} // namespace drf
