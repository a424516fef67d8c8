// Barotropic/baroclinic mode-split step, fused into one kernel on gfx950.
//
// Replaces ModeSplittingSolver::step (src/physics/mode_splitting.cpp:434-453):
// decompose (:359-396), subcycle_barotropic (:475-485) with the Heun
// predictor-corrector (:293-317) for all n_sub substeps, update_baroclinic
// (:487-600) and recombine (:398-432).  The reference is element-local, so a
// workgroup keeps the 2D barotropic state of its elements in registers and
// LDS for every substep: the 3D state is read once and written once per
// baroclinic step, and nothing touches HBM between substeps.  All arithmetic
// is FP64 (free-surface and transport accumulation across substeps and the
// depth-integrated velocity correction stay in FP64 at every precision).
//
// Thread mapping: one thread per 2D node ih of an element; in the reference's
// quirky column layout the column of ih is the contiguous DOF run
// ih*n .. ih*n+n-1 (:143), so each thread streams exactly one run.  The 2D
// operators D_x2d / D_y2d (:90-119) act on the fast / slow digit of ih.
// The state and the predictor stage of a substep live in separate LDS
// buffers, two barriers per substep.
#include "drf_internal.hpp"

namespace drf {
namespace {

template <int N> struct MsCfg {
    static constexpr int n = N + 1, nh = n * n, nd = nh * n;
    static constexpr int EPB = (256 / nh) > 0 ? 256 / nh : 1;
    static constexpr int NT = EPB * nh;
    static constexpr int BLOCK = ((NT + 63) / 64) * 64;
};

__device__ __forceinline__ double hinv(double H) { return (H > 1e-10) ? 1.0 / H : 0.0; }

// Four waves per SIMD up to N = 4 (the D rows in registers put N = 4 two
// VGPRs over 128; the spill sits outside the substep loop and measured
// 251 us instead of 271 us per step at 40 substeps, 64x64x8 elements).
template <int N>
__global__ void __launch_bounds__(MsCfg<N>::BLOCK)
    __attribute__((amdgpu_waves_per_eu(N <= 4 ? 4 : 1)))
    mode_split_kernel(int64_t E, double dt3, int nsub, double *__restrict__ full,
                      const View<const double> hb, const View<const double> fb, double f0,
                      const Ops ops, const Phys ph)
{
    using C = MsCfg<N>;
    constexpr int n = C::n, nh = C::nh, nd = C::nd, EPB = C::EPB;
    // state (s_a, s_b, s_c) and predictor stage (t_a, t_b, t_c) of the substep
    __shared__ double s_a[EPB][nh], s_b[EPB][nh], s_c[EPB][nh];
    __shared__ double t_a[EPB][nh], t_b[EPB][nh], t_c[EPB][nh];
    // D staged once: read from the kernel arguments with a lane-dependent
    // index it became 12 global loads in every substep's dependent chain
    __shared__ double sD[n * n];

    const int tid = threadIdx.x;
    for (int q = tid; q < n * n; q += blockDim.x)
        sD[q] = ops.D[q];
    const bool in_tile = tid < C::NT;
    const int le = in_tile ? tid / nh : 0;
    const int ih = in_tile ? tid - le * nh : 0;
    const int64_t e = (int64_t)blockIdx.x * EPB + le;
    const bool act = in_tile && e < E;
    const int a = ih % n, b = ih / n; // 2D fast / slow digit
    // The two rows of D this thread's stencils use, in registers for all
    // substeps: the substep loop is bound by LDS bandwidth, and D was a
    // third of its LDS reads.
    __syncthreads();
    double Da[n], Db[n];
#pragma unroll
    for (int m = 0; m < n; ++m) {
        Da[m] = sD[a * n + m];
        Db[m] = sD[b * n + m];
    }

    double *F = full + (act ? e : 0) * (int64_t)12 * nd;
    double *Hu = F + 0 * nd, *Hv = F + 1 * nd, *eta = F + 2 * nd, *HT = F + 3 * nd,
           *HS = F + 4 * nd, *H = F + 5 * nd, *u = F + 6 * nd, *v = F + 7 * nd,
           *T = F + 8 * nd, *S = F + 9 * nd, *om = F + 10 * nd, *rho = F + 11 * nd;
    const int c0 = ih * n;
    const double g = ph.g;

    // Column sums of the weights (depth_average :142-147)
    double wsum = 0.0;
#pragma unroll
    for (int k = 0; k < n; ++k)
        wsum += ops.sw[k];

    // ---------------- decompose :376-394 ----------------
    double up[n], vp[n], Ub = 0.0, Vb = 0.0, HUb = 0.0, HVb = 0.0;
    double eb = 0.0, Hb = 0.0, hh = 0.0, fh = f0;
    if (act) {
        double uc[n], vc[n], Hc[n];
#pragma unroll
        for (int k = 0; k < n; ++k) {
            uc[k] = u[c0 + k];
            vc[k] = v[c0 + k];
            Hc[k] = H[c0 + k];
        }
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int k = 0; k < n; ++k) {
            su += ops.sw[k] * uc[k];
            sv += ops.sw[k] * vc[k];
        }
        Ub = su / wsum;
        Vb = sv / wsum;
#pragma unroll
        for (int k = 0; k < n; ++k) {
            HUb += ops.sw[k] * Hc[k] * uc[k];
            HVb += ops.sw[k] * Hc[k] * vc[k];
        }
#pragma unroll
        for (int k = 0; k < n; ++k) {
            up[k] = uc[k] - Ub;
            vp[k] = vc[k] - Vb;
        }
        eb = eta[ih]; // .head(n_h), :378
        Hb = H[ih];   // .head(n_h), :384
        // bathymetry and Coriolis at DOF index ih (= node i + n L)
        hh = hb.p[node_offset(hb, e, ih / n, ih % n)];
        if (fb.p)
            fh = fb.p[node_offset(fb, e, ih / n, ih % n)];
    }

    // ---------------- subcycle :319-332 ----------------
    const double dtb = dt3 / nsub;
    double eavg = 0.0;
    double Ufin = 0.0, Vfin = 0.0;
    for (int sub = 0; sub < nsub; ++sub) {
        double de0 = 0.0, dU0 = 0.0, dV0 = 0.0;
        // barotropic_rhs on the current state (:177-211).  s_* was last read
        // before the previous substep's second barrier.
        if (act) {
            s_a[le][ih] = HUb;
            s_b[le][ih] = HVb;
            s_c[le][ih] = eb;
        }
        __syncthreads();
        double es = 0.0, HUs = 0.0, HVs = 0.0, Hs = 0.0;
        if (act) {
            double dx = 0.0, dy = 0.0, ex = 0.0, ey = 0.0;
#pragma unroll
            for (int m = 0; m < n; ++m) {
                dx += Da[m] * s_a[le][m + n * b];
                dy += Db[m] * s_b[le][a + n * m];
                ex += Da[m] * s_c[le][m + n * b];
                ey += Db[m] * s_c[le][a + n * m];
            }
            de0 = -(dx + dy);
            dU0 = -g * Hb * ex + fh * HVb + 0.0;
            dV0 = -g * Hb * ey + -fh * HUb + 0.0;
            es = eb + dtb * de0;
            HUs = HUb + dtb * dU0;
            HVs = HVb + dtb * dV0;
            Hs = es + hh; // BarotropicState::update_derived :31
            // t_* was last read before this substep's first barrier
            t_a[le][ih] = HUs;
            t_b[le][ih] = HVs;
            t_c[le][ih] = es;
        }
        __syncthreads();
        if (act) {
            double dx = 0.0, dy = 0.0, ex = 0.0, ey = 0.0;
#pragma unroll
            for (int m = 0; m < n; ++m) {
                dx += Da[m] * t_a[le][m + n * b];
                dy += Db[m] * t_b[le][a + n * m];
                ex += Da[m] * t_c[le][m + n * b];
                ey += Db[m] * t_c[le][a + n * m];
            }
            const double de1 = -(dx + dy);
            const double dU1 = -g * Hs * ex + fh * HVs + 0.0;
            const double dV1 = -g * Hs * ey + -fh * HUs + 0.0;
            eb += 0.5 * dtb * (de0 + de1);
            HUb += 0.5 * dtb * (dU0 + dU1);
            HVb += 0.5 * dtb * (dV0 + dV1);
            Hb = eb + hh;
            eavg += eb;
        }
    }
    // U_bar = HU / H of the last substep (update_derived :32-34)
    if (act) {
        const double hi = hinv(Hb);
        Ufin = HUb * hi;
        Vfin = HVb * hi;
    }
    eavg /= (double)nsub;

    // ---------------- update_baroclinic :504-598 ----------------
    double HU3 = 0.0, HV3 = 0.0, Hn = 0.0;
    double ut[n], vt[n];
    if (act) {
        Hn = eavg + hh;
        // eta := eta_avg here (:509) is overwritten by recombine's final
        // barotropic eta (:423) before anything reads it: only that is stored
#pragma unroll
        for (int k = 0; k < n; ++k) {
            H[c0 + k] = Hn;
            ut[k] = Ub + up[k];
            vt[k] = Vb + vp[k];
            Hu[c0 + k] = Hn * ut[k];
            Hv[c0 + k] = Hn * vt[k];
        }
        const double hi = hinv(Hn);
#pragma unroll
        for (int k = 0; k < n; ++k) {
            const double Tk = HT[c0 + k] * hi;
            const double Sk = HS[c0 + k] * hi;
            T[c0 + k] = Tk;
            S[c0 + k] = Sk;
            HT[c0 + k] = Hn * Tk;
            HS[c0 + k] = Hn * Sk;
            rho[c0 + k] = ph.rho0 * (1.0 - ph.alpha * (Tk - 10.0) + ph.beta * (Sk - 35.0));
        }
#pragma unroll
        for (int k = 0; k < n; ++k) {
            HU3 += ops.sw[k] * Hn * ut[k];
            HV3 += ops.sw[k] * Hn * vt[k];
        }
        s_a[le][ih] = HU3;
        s_b[le][ih] = HV3;
    }
    __syncthreads();
    if (act) {
        double dx = 0.0, dy = 0.0;
#pragma unroll
        for (int m = 0; m < n; ++m) {
            dx += Da[m] * s_a[le][m + n * b];
            dy += Db[m] * s_b[le][a + n * m];
        }
        const double div = dx + dy;
        const double hi = hinv(Hn); // H(ih * n_vert), :575
#pragma unroll
        for (int k = 0; k < n; ++k) {
            const double sf = (double)k / (double)(n - 1);
            om[c0 + k] = -sf * div * hi;
        }
        // ---------------- recombine :411-425 ----------------
#pragma unroll
        for (int k = 0; k < n; ++k) {
            u[c0 + k] = Ufin + up[k];
            v[c0 + k] = Vfin + vp[k];
            eta[c0 + k] = eb;
        }
    }
}

template <int N>
hipError_t ms_launch(int64_t E, double dt3, int nsub, double *full, View<const double> h,
                     View<const double> f, double f0, const Ops &ops, const Phys &ph,
                     hipStream_t s)
{
    using C = MsCfg<N>;
    if (E <= 0)
        return hipSuccess;
    const int64_t blocks = (E + C::EPB - 1) / C::EPB;
    hipLaunchKernelGGL(mode_split_kernel<N>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, E, dt3,
                       nsub, full, h, f, f0, ops, ph);
    return hipGetLastError();
}

} // namespace

hipError_t launch_mode_split(int order, int64_t E, double dt3, int nsub, double *full,
                             View<const double> h, View<const double> f, double f0,
                             const Ops &ops, const Phys &ph, hipStream_t s)
{
    switch (order) {
    case 1: return ms_launch<1>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    case 2: return ms_launch<2>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    case 3: return ms_launch<3>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    case 4: return ms_launch<4>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    case 5: return ms_launch<5>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    case 6: return ms_launch<6>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    case 7: return ms_launch<7>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    case 8: return ms_launch<8>(E, dt3, nsub, full, h, f, f0, ops, ph, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace drf
