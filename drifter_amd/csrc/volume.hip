// Element-local volume right-hand side of the hydrostatic primitive
// equations on gfx950 (MI355X), sum-factorised.
//
// Replaces PrimitiveEquationsSolver::compute_rhs
// (src/physics/primitive_equations.cpp:428-442) and everything below it:
// update_derived (:12-29), compute_momentum_rhs (:154-199), compute_tracer_rhs
// (:201-227), compute_eta_rhs (:229-273), with the reference-coordinate
// derivatives of DG3DElementOperator::{gradient,divergence}_reference
// (src/dg/operators_3d.cpp:58-79) applied by sum factorisation instead of
// dense Kronecker mat-vecs (D_xi = D (x) I (x) I, src/dg/basis_hexahedron.cpp:321-357).
//
// Work decomposition.  A workgroup owns EPB elements; every element is
// handled by n^2 threads, each owning one grid line of n nodes:
//   phase 1  thread (j,k) owns the xi-line   -> loads the element from HBM,
//            forms the derived state, all xi-derivatives in registers, the
//            xi-column integrals of the eta equation, and stages the fields
//            needed by the other two directions in LDS;
//   phase 2  thread (i,k) owns the eta-line  -> eta-derivatives from LDS,
//            writes its contributions in place over the consumed fluxes;
//   phase 3  thread (i,j) owns the zeta-line -> zeta second derivatives,
//            adds its contributions in place;
//   phase 4  thread (j,k) again              -> adds the staged contributions
//            to its register partials and writes the tendency (or a fused
//            Runge-Kutta stage update) back to HBM.
// Each value crosses LDS once per direction instead of once per derivative
// tap, which keeps LDS traffic below the HBM time of the element.
//
// Index quirk reproduced on purpose: the free-surface tendency integrates
// Hu, Hv over the xi index (d % n) and differentiates the column integrals
// along eta and zeta, replicating the result along xi
// (primitive_equations.cpp:244-272).  omega is identically 0 in the RHS
// (src/solver/simulation_driver.cpp:222 -> primitive_equations.cpp:47-48), so
// zeta fluxes vanish.
#include "drf_internal.hpp"

namespace drf {
namespace {

// Staged LDS fields per element node.
enum : int {
    S_FUV = 0, // Hu*v  -> eta/zeta contribution to dHu
    S_FVV,     // Hv*v  -> ... dHv
    S_FTV,     // v*HT  -> ... dHT
    S_FSV,     // v*HS  -> ... dHS
    S_ETA,
    S_U,
    S_V,
    S_H,
    S_T,
    S_S,
    S_NF
};

template <int N> struct Cfg {
    static constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
    static constexpr int lds_per_elem_fp64 = (S_NF * n3 + 2 * n2) * 8;
    static constexpr int epb_thr = (128 / n2) > 0 ? 128 / n2 : 1;
    static constexpr int epb_lds = (65536 / lds_per_elem_fp64) > 0 ? 65536 / lds_per_elem_fp64 : 1;
    static constexpr int EPB = epb_thr < epb_lds ? epb_thr : epb_lds;
    static constexpr int NT = EPB * n2;
    static constexpr int BLOCK = ((NT + 63) / 64) * 64;
};

template <class TS, class TC> __device__ __forceinline__ TC ld(const TS *p) { return (TC)(*p); }

template <class TC> __device__ __forceinline__ TC hinv_of(TC H)
{
    // PrimitiveState::update_derived, primitive_equations.cpp:23
    return (H > (TC)1e-10) ? (TC)1 / H : (TC)0;
}

// y = D x along one line (n taps), D row-major in kernel arguments.
template <int n, class TC>
__device__ __forceinline__ void dline(const TC *__restrict__ D, const TC (&x)[n], TC (&y)[n])
{
#pragma unroll
    for (int i = 0; i < n; ++i) {
        TC s = D[i * n] * x[0];
#pragma unroll
        for (int m = 1; m < n; ++m)
            s += D[i * n + m] * x[m];
        y[i] = s;
    }
}

template <int N, class TS, class TC>
__global__ void __launch_bounds__(Cfg<N>::BLOCK)
    volume_kernel(const VolGeom<TS> geo, const OpsT<TC> ops, const Phys ph, const View<const TS> U,
                  const Epi<TS> epi)
{
    using C = Cfg<N>;
    constexpr int n = C::n, n2 = C::n2, n3 = C::n3, EPB = C::EPB;
    __shared__ TC sm[EPB][S_NF][n3];
    __shared__ double bar[EPB][2][n2];

    const int tid = threadIdx.x;
    const bool in_tile = tid < C::NT;
    const int le = in_tile ? tid / n2 : 0;
    const int L = in_tile ? tid - le * n2 : 0;
    const int64_t e = (int64_t)blockIdx.x * EPB + le;
    const bool act = in_tile && e < geo.E;

    const TC g = (TC)ph.g, nu_h = (TC)ph.nu_h, nu_v = (TC)ph.nu_v, kT = (TC)ph.kT, kS = (TC)ph.kS;
    const TC *__restrict__ D = ops.D;

    // Register partial tendencies of the xi-line owned in phases 1 and 4.
    TC pHu[n], pHv[n], pHT[n], pHS[n];

    // ---------------- phase 1: xi-line (j, k) = (L % n, L / n) ----------------
    if (act) {
        const int64_t base = e * U.es + (int64_t)L * n;
        const int64_t gbase = e * n3 + (int64_t)L * n;
        TC hu[n], hv[n], et[n], ht[n], hs[n], H[n], u[n], v[n];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            hu[i] = ld<TS, TC>(U.p + 0 * U.fs + base + i);
            hv[i] = ld<TS, TC>(U.p + 1 * U.fs + base + i);
            et[i] = ld<TS, TC>(U.p + 2 * U.fs + base + i);
            ht[i] = ld<TS, TC>(U.p + 3 * U.fs + base + i);
            hs[i] = ld<TS, TC>(U.p + 4 * U.fs + base + i);
        }
        TC T[n], S[n];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            H[i] = et[i] + ld<TS, TC>(geo.h + gbase + i);
            const TC hi = hinv_of(H[i]);
            u[i] = hu[i] * hi;
            v[i] = hv[i] * hi;
            T[i] = ht[i] * hi;
            S[i] = hs[i] * hi;
        }
        TC fc[n];
        if (geo.f) {
#pragma unroll
            for (int i = 0; i < n; ++i)
                fc[i] = ld<TS, TC>(geo.f + gbase + i);
        } else {
#pragma unroll
            for (int i = 0; i < n; ++i)
                fc[i] = (TC)geo.f0;
        }

        // xi-derivatives (momentum_advection :275-305, tracer_advection
        // :377-393, barotropic_pressure_gradient :322-337, horizontal_diffusion
        // :339-355)
        TC x[n], d[n], dd[n];
#pragma unroll
        for (int i = 0; i < n; ++i)
            x[i] = hu[i] * u[i];
        dline<n>(D, x, d);
#pragma unroll
        for (int i = 0; i < n; ++i)
            pHu[i] = -d[i] + fc[i] * hv[i];
#pragma unroll
        for (int i = 0; i < n; ++i)
            x[i] = hv[i] * u[i];
        dline<n>(D, x, d);
#pragma unroll
        for (int i = 0; i < n; ++i)
            pHv[i] = -d[i] + (-fc[i]) * hu[i];
        dline<n>(D, et, d);
#pragma unroll
        for (int i = 0; i < n; ++i)
            pHu[i] += (-g * H[i]) * d[i];
        dline<n>(D, u, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int i = 0; i < n; ++i)
            pHu[i] += H[i] * (nu_h * dd[i]);
        dline<n>(D, v, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int i = 0; i < n; ++i)
            pHv[i] += H[i] * (nu_h * dd[i]);
#pragma unroll
        for (int i = 0; i < n; ++i)
            x[i] = u[i] * ht[i];
        dline<n>(D, x, d);
#pragma unroll
        for (int i = 0; i < n; ++i)
            pHT[i] = -d[i];
#pragma unroll
        for (int i = 0; i < n; ++i)
            x[i] = u[i] * hs[i];
        dline<n>(D, x, d);
#pragma unroll
        for (int i = 0; i < n; ++i)
            pHS[i] = -d[i];

        // Stage the fields the eta- and zeta-line owners need.
        TC *s = &sm[le][0][0];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            const int nd_ = L * n + i;
            s[S_FUV * n3 + nd_] = hu[i] * v[i];
            s[S_FVV * n3 + nd_] = hv[i] * v[i];
            s[S_FTV * n3 + nd_] = v[i] * ht[i];
            s[S_FSV * n3 + nd_] = v[i] * hs[i];
            s[S_ETA * n3 + nd_] = et[i];
            s[S_U * n3 + nd_] = u[i];
            s[S_V * n3 + nd_] = v[i];
            s[S_H * n3 + nd_] = H[i];
            s[S_T * n3 + nd_] = T[i];
            s[S_S * n3 + nd_] = S[i];
        }
        // Column integrals of compute_eta_rhs (:244-254): over the xi index,
        // summed in node order from 0.
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int i = 0; i < n; ++i) {
            su += ops.sw64[i] * (double)hu[i];
            sv += ops.sw64[i] * (double)hv[i];
        }
        bar[le][0][L] = su;
        bar[le][1][L] = sv;
    }
    __syncthreads();

    // ---------------- phase 2: eta-line (i, k) = (L % n, L / n) ----------------
    if (act) {
        const int i0 = L % n, k0 = L / n;
        TC *s = &sm[le][0][0];
        TC fuv[n], fvv[n], ftv[n], fsv[n], et[n], u[n], v[n], H[n];
#pragma unroll
        for (int j = 0; j < n; ++j) {
            const int m = i0 + n * j + n2 * k0;
            fuv[j] = s[S_FUV * n3 + m];
            fvv[j] = s[S_FVV * n3 + m];
            ftv[j] = s[S_FTV * n3 + m];
            fsv[j] = s[S_FSV * n3 + m];
            et[j] = s[S_ETA * n3 + m];
            u[j] = s[S_U * n3 + m];
            v[j] = s[S_V * n3 + m];
            H[j] = s[S_H * n3 + m];
        }
        TC d[n], dd[n], aHu[n], aHv[n];
        dline<n>(D, fuv, d);
#pragma unroll
        for (int j = 0; j < n; ++j)
            aHu[j] = -d[j];
        dline<n>(D, u, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int j = 0; j < n; ++j)
            aHu[j] += H[j] * (nu_h * dd[j]);
        dline<n>(D, fvv, d);
#pragma unroll
        for (int j = 0; j < n; ++j)
            aHv[j] = -d[j];
        dline<n>(D, et, d);
#pragma unroll
        for (int j = 0; j < n; ++j)
            aHv[j] += (-g * H[j]) * d[j];
        dline<n>(D, v, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int j = 0; j < n; ++j)
            aHv[j] += H[j] * (nu_h * dd[j]);
        TC dT[n], dS[n];
        dline<n>(D, ftv, dT);
        dline<n>(D, fsv, dS);
#pragma unroll
        for (int j = 0; j < n; ++j) {
            const int m = i0 + n * j + n2 * k0;
            s[S_FUV * n3 + m] = aHu[j];
            s[S_FVV * n3 + m] = aHv[j];
            s[S_FTV * n3 + m] = -dT[j];
            s[S_FSV * n3 + m] = -dS[j];
        }
    }
    __syncthreads();

    // ---------------- phase 3: zeta-line (i, j) = L ----------------
    if (act) {
        TC *s = &sm[le][0][0];
        TC u[n], v[n], T[n], S[n], hi[n];
#pragma unroll
        for (int k = 0; k < n; ++k) {
            const int m = L + n2 * k;
            u[k] = s[S_U * n3 + m];
            v[k] = s[S_V * n3 + m];
            T[k] = s[S_T * n3 + m];
            S[k] = s[S_S * n3 + m];
            hi[k] = hinv_of(s[S_H * n3 + m]);
        }
        // vertical_diffusion :357-375: kappa * d2/dzeta2 * Hinv
        TC d[n], dd[n];
        dline<n>(D, u, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int k = 0; k < n; ++k)
            s[S_FUV * n3 + L + n2 * k] += nu_v * dd[k] * hi[k];
        dline<n>(D, v, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int k = 0; k < n; ++k)
            s[S_FVV * n3 + L + n2 * k] += nu_v * dd[k] * hi[k];
        dline<n>(D, T, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int k = 0; k < n; ++k)
            s[S_FTV * n3 + L + n2 * k] += kT * dd[k] * hi[k];
        dline<n>(D, S, d);
        dline<n>(D, d, dd);
#pragma unroll
        for (int k = 0; k < n; ++k)
            s[S_FSV * n3 + L + n2 * k] += kS * dd[k] * hi[k];
    }
    __syncthreads();

    // ---------------- phase 4: xi-line (j, k) again; write ----------------
    if (act) {
        const TC *s = &sm[le][0][0];
        const int j0 = L % n, k0 = L / n;
        // eta tendency (:256-272): -(D_x2d HU_bar + D_y2d HV_bar), where the
        // 2D "x" index is j and the 2D "y" index is k.
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int m = 0; m < n; ++m) {
            s1 += ops.D64[j0 * n + m] * bar[le][0][m + n * k0];
            s2 += ops.D64[k0 * n + m] * bar[le][1][j0 + n * m];
        }
        const TC deta = (TC)(-(s1 + s2));
        TC kv[5][n];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            const int nd_ = L * n + i;
            kv[0][i] = pHu[i] + s[S_FUV * n3 + nd_];
            kv[1][i] = pHv[i] + s[S_FVV * n3 + nd_];
            kv[2][i] = deta;
            kv[3][i] = pHT[i] + s[S_FTV * n3 + nd_];
            kv[4][i] = pHS[i] + s[S_FSV * n3 + nd_];
        }
        const int64_t off = (int64_t)L * n;
        switch (epi.mode) {
        case EPI_RHS: {
            TS *R = epi.R.p + e * epi.R.es + off;
#pragma unroll
            for (int f = 0; f < 5; ++f)
#pragma unroll
                for (int i = 0; i < n; ++i)
                    R[f * epi.R.fs + i] = (TS)kv[f][i];
            break;
        }
        case EPI_STAGE: {
            const TC c1 = (TC)epi.c1, c2 = (TC)epi.c2;
            const TS *U0 = epi.U0.p + e * epi.U0.es + off;
            TS *Y = epi.Y.p + e * epi.Y.es + off;
            TS *A = epi.A.p ? epi.A.p + e * epi.A.es + off : nullptr;
#pragma unroll
            for (int f = 0; f < 5; ++f)
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    const TC k = kv[f][i];
                    const TC u0 = ld<TS, TC>(U0 + f * epi.U0.fs + i);
                    if (epi.acc_mode == 1)
                        A[f * epi.A.fs + i] = (TS)k;
                    else if (epi.acc_mode == 2)
                        A[f * epi.A.fs + i] = (TS)(ld<TS, TC>(A + f * epi.A.fs + i) + c2 * k);
                    Y[f * epi.Y.fs + i] = (TS)(u0 + c1 * k);
                }
            break;
        }
        case EPI_RK4FIN: {
            const TC c1 = (TC)epi.c1;
            const TS *U0 = epi.U0.p + e * epi.U0.es + off;
            const TS *A = epi.A.p + e * epi.A.es + off;
            TS *Y = epi.Y.p + e * epi.Y.es + off;
#pragma unroll
            for (int f = 0; f < 5; ++f)
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    const TC a = ld<TS, TC>(A + f * epi.A.fs + i);
                    const TC u0 = ld<TS, TC>(U0 + f * epi.U0.fs + i);
                    Y[f * epi.Y.fs + i] = (TS)(u0 + c1 * (a + kv[f][i]));
                }
            break;
        }
        case EPI_COMB: {
            const TC c0 = (TC)epi.c0, c1 = (TC)epi.c1, c2 = (TC)epi.c2;
            const TS *U0 = epi.U0.p + e * epi.U0.es + off;
            const TS *X = epi.X.p + e * epi.X.es + off;
            TS *Y = epi.Y.p + e * epi.Y.es + off;
#pragma unroll
            for (int f = 0; f < 5; ++f)
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    const TC xv = ld<TS, TC>(X + f * epi.X.fs + i);
                    const TC u0 = ld<TS, TC>(U0 + f * epi.U0.fs + i);
                    Y[f * epi.Y.fs + i] = (TS)(c0 * u0 + c1 * (xv + c2 * kv[f][i]));
                }
            break;
        }
        case EPI_LSRK: {
            const TC c1 = (TC)epi.c1, c2 = (TC)epi.c2, c3 = (TC)epi.c3;
            const TS *X = epi.X.p + e * epi.X.es + off;
            TS *A = epi.A.p + e * epi.A.es + off;
            TS *Y = epi.Y.p + e * epi.Y.es + off;
#pragma unroll
            for (int f = 0; f < 5; ++f)
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    const TC k = kv[f][i];
                    const TC a = (epi.acc_mode == 1) ? c1 * k
                                                     : c2 * ld<TS, TC>(A + f * epi.A.fs + i) + c1 * k;
                    A[f * epi.A.fs + i] = (TS)a;
                    const TC xv = ld<TS, TC>(X + f * epi.X.fs + i);
                    Y[f * epi.Y.fs + i] = (TS)(xv + c3 * a);
                }
            break;
        }
        default:
            break;
        }
    }
}

template <class TC> OpsT<TC> cast_ops(const Ops &o)
{
    OpsT<TC> r{};
    for (int i = 0; i < kMaxN * kMaxN; ++i)
        r.D[i] = (TC)o.D[i];
    for (int i = 0; i < kMaxN; ++i)
        r.sw[i] = (TC)o.sw[i];
    for (int i = 0; i < kMaxN * kMaxN; ++i)
        r.D64[i] = o.D[i];
    for (int i = 0; i < kMaxN; ++i)
        r.sw64[i] = o.sw[i];
    return r;
}

template <int N, class TS, class TC>
hipError_t launch_n(const VolGeom<TS> &geo, const Ops &ops, const Phys &ph, View<const TS> U,
                    const Epi<TS> &epi, hipStream_t s)
{
    using C = Cfg<N>;
    if (geo.E <= 0)
        return hipSuccess;
    const int64_t blocks = (geo.E + C::EPB - 1) / C::EPB;
    hipLaunchKernelGGL((volume_kernel<N, TS, TC>), dim3((unsigned)blocks), dim3(C::BLOCK), 0, s,
                       geo, cast_ops<TC>(ops), ph, U, epi);
    return hipGetLastError();
}

template <class TS, class TC>
hipError_t dispatch(int order, const VolGeom<TS> &geo, const Ops &ops, const Phys &ph,
                    View<const TS> U, const Epi<TS> &epi, hipStream_t s)
{
    switch (order) {
    case 1: return launch_n<1, TS, TC>(geo, ops, ph, U, epi, s);
    case 2: return launch_n<2, TS, TC>(geo, ops, ph, U, epi, s);
    case 3: return launch_n<3, TS, TC>(geo, ops, ph, U, epi, s);
    case 4: return launch_n<4, TS, TC>(geo, ops, ph, U, epi, s);
    case 5: return launch_n<5, TS, TC>(geo, ops, ph, U, epi, s);
    case 6: return launch_n<6, TS, TC>(geo, ops, ph, U, epi, s);
    case 7: return launch_n<7, TS, TC>(geo, ops, ph, U, epi, s);
    case 8: return launch_n<8, TS, TC>(geo, ops, ph, U, epi, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace

hipError_t launch_volume_fp64(int order, const VolGeom<double> &geo, const Ops &ops,
                              const Phys &ph, View<const double> U, const Epi<double> &epi,
                              hipStream_t s)
{
    return dispatch<double, double>(order, geo, ops, ph, U, epi, s);
}

hipError_t launch_volume_fp32(int order, const VolGeom<float> &geo, const Ops &ops,
                              const Phys &ph, View<const float> U, const Epi<float> &epi,
                              hipStream_t s)
{
    return dispatch<float, float>(order, geo, ops, ph, U, epi, s);
}

} // namespace drf
