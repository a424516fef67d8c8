// Face numerical fluxes, conforming interfaces, 2:1 mortars and face-trace
// gathers on gfx950.
//
// Replaces, at their component boundaries:
//   NumericalFlux family   include/flux/numerical_flux.hpp:57-397,
//                          src/flux/numerical_flux.cpp:13-200, and the
//                          5-variable Lax-Friedrichs of
//                          src/physics/primitive_equations.cpp:452-492
//   ConformingInterface::compute_flux   src/dg/mortar.cpp:320-343
//   MortarSpace::compute_mortar_flux    src/dg/mortar.cpp:212-258
//   LGL face extraction                 src/dg/basis_hexahedron.cpp:431-522
// Pointwise fluxes and conforming faces run one face point per lane.  Mortars
// of orders 1-4 run on the FP64 matrix cores: 16 interfaces per wavefront,
// each operator product a (16 MT) x 16 x (N+1)^2 GEMM with the operators in
// LDS.  Higher orders run persistent workgroups over the interfaces of one
// connection type: the type's projection and lift operators (built once on
// the host) are staged in LDS when they fit, and every lane owns one mortar
// DOF of one of the workgroup's interfaces.
#include "drf_internal.hpp"
#include "interface.hpp"

// Products are not fused into FMAs in this file, so that every flux and
// projection rounds exactly as the reference's sequence of operations.
#pragma clang fp contract(off)

namespace drf {
namespace {

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }

// ---- shallow-water fluxes (h, hu, hv) ----

__device__ void sw_normal_flux(const double *U, const double *n, double g, double *Fn)
{
    // physical flux of the factory's Lax-Friedrichs (numerical_flux.cpp:136-161)
    // contracted with n in the order of LaxFriedrichsFlux::flux (:229-234)
    double h = U[0];
    const double hu = U[1], hv = U[2];
    if (h < 1e-10)
        h = 1e-10;
    const double u = hu / h, v = hv / h;
    const double F[3][3] = {{hu, hu * u + 0.5 * g * h * h, hu * v},
                            {hv, hv * u, hv * v + 0.5 * g * h * h},
                            {0.0, 0.0, 0.0}};
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d)
            s += F[d][i] * n[d];
        Fn[i] = s;
    }
}

__device__ void sw_lf(const double *L, const double *R, const double *n, double g, double *out)
{
    double FL[3], FR[3];
    sw_normal_flux(L, n, g, FL);
    sw_normal_flux(R, n, g, FR);
    const double hL = dmax(L[0], 1e-10), hR = dmax(R[0], 1e-10);
    const double unL = (L[1] / hL) * n[0] + (L[2] / hL) * n[1];
    const double unR = (R[1] / hR) * n[0] + (R[2] / hR) * n[1];
    const double lam = dmax(fabs(unL) + sqrt(g * hL), fabs(unR) + sqrt(g * hR));
    for (int i = 0; i < 3; ++i)
        out[i] = 0.5 * (FL[i] + FR[i]) - 0.5 * lam * (R[i] - L[i]);
}

__device__ void sw_central(const double *L, const double *R, const double *n, double g,
                           double *out)
{
    double FL[3], FR[3];
    sw_normal_flux(L, n, g, FL);
    sw_normal_flux(R, n, g, FR);
    for (int i = 0; i < 3; ++i)
        out[i] = 0.5 * (FL[i] + FR[i]);
}

__device__ void hllc_phys(double h, double u, double v, double g, double nx, double ny, double *F)
{
    const double un = u * nx + v * ny;
    F[0] = h * un;
    F[1] = h * u * un + 0.5 * g * h * h * nx;
    F[2] = h * v * un + 0.5 * g * h * h * ny;
}

// ShallowWaterHLLCFlux (numerical_flux.hpp:272-397): Einfeldt speeds and the
// contact speed of the standard shallow-water HLLC.
__device__ void sw_hllc(const double *L, const double *R, const double *n, double g, double *out)
{
    const double hL = L[0] < 1e-10 ? 1e-10 : L[0], hR = R[0] < 1e-10 ? 1e-10 : R[0];
    const double uL = L[1] / hL, vL = L[2] / hL, uR = R[1] / hR, vR = R[2] / hR;
    const double nx = n[0], ny = n[1];
    const double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
    const double cL = sqrt(g * hL), cR = sqrt(g * hR);
    const double SL = dmin(unL - cL, unR - cR), SR = dmax(unL + cL, unR + cR);
    const double den = hL * (SL - unL) - hR * (SR - unR);
    const double Ss = (fabs(den) > 1e-10)
                          ? (hL * unL * (SL - unL) - hR * unR * (SR - unR) +
                             0.5 * g * (hR * hR - hL * hL)) /
                                den
                          : 0.5 * (unL + unR);
    double FL[3], FR[3];
    hllc_phys(hL, uL, vL, g, nx, ny, FL);
    hllc_phys(hR, uR, vR, g, nx, ny, FR);
    if (SL >= 0) {
        for (int i = 0; i < 3; ++i)
            out[i] = FL[i];
    } else if (SR <= 0) {
        for (int i = 0; i < 3; ++i)
            out[i] = FR[i];
    } else if (Ss >= 0) {
        const double hs = hL * (SL - unL) / (SL - Ss);
        const double Us[3] = {hs, hs * (uL + (Ss - unL) * nx), hs * (vL + (Ss - unL) * ny)};
        const double Uv[3] = {hL, L[1], L[2]};
        for (int i = 0; i < 3; ++i)
            out[i] = FL[i] + SL * (Us[i] - Uv[i]);
    } else {
        const double hs = hR * (SR - unR) / (SR - Ss);
        const double Us[3] = {hs, hs * (uR + (Ss - unR) * nx), hs * (vR + (Ss - unR) * ny)};
        const double Uv[3] = {hR, R[1], R[2]};
        for (int i = 0; i < 3; ++i)
            out[i] = FR[i] + SR * (Us[i] - Uv[i]);
    }
}

// ShallowWaterRoeFlux with Harten's entropy fix (numerical_flux.cpp:13-128)
__device__ void sw_roe(const double *L, const double *R, const double *n, double g, double efix,
                       double *out)
{
    const double hL = L[0] < 1e-10 ? 1e-10 : L[0], hR = R[0] < 1e-10 ? 1e-10 : R[0];
    const double uL = L[1] / hL, vL = L[2] / hL, uR = R[1] / hR, vR = R[2] / hR;
    const double nx = n[0], ny = n[1];
    const double unL = uL * nx + vL * ny, utL = -uL * ny + vL * nx;
    const double unR = uR * nx + vR * ny, utR = -uR * ny + vR * nx;
    const double sL = sqrt(hL), sR = sqrt(hR), den = sL + sR;
    const double hr = 0.5 * (hL + hR);
    const double ur = (sL * uL + sR * uR) / den, vr = (sL * vL + sR * vR) / den;
    const double unr = ur * nx + vr * ny, cr = sqrt(g * hr);
    double l1 = unr - cr, l3 = unr + cr;
    const double l2 = unr, delta = efix * cr;
    if (fabs(l1) < delta) {
        l1 = (l1 * l1 + delta * delta) / (2.0 * delta);
        l1 = fabs(l1);
    }
    if (fabs(l3) < delta) {
        l3 = (l3 * l3 + delta * delta) / (2.0 * delta);
        l3 = fabs(l3);
    }
    const double dh = hR - hL, dun = unR - unL, dut = utR - utL;
    const double a1 = 0.5 * (dh - hr / cr * dun), a2 = hr * dut, a3 = 0.5 * (dh + hr / cr * dun);
    const double A1 = fabs(l1), A2 = fabs(l2), A3 = fabs(l3);
    const double F1L = hL * unL, F2L = hL * uL * unL + 0.5 * g * hL * hL * nx,
                 F3L = hL * vL * unL + 0.5 * g * hL * hL * ny;
    const double F1R = hR * unR, F2R = hR * uR * unR + 0.5 * g * hR * hR * nx,
                 F3R = hR * vR * unR + 0.5 * g * hR * hR * ny;
    out[0] = 0.5 * (F1L + F1R) - 0.5 * (A1 * a1 + A2 * 0.0 + A3 * a3);
    out[1] = 0.5 * (F2L + F2R) -
             0.5 * (A1 * (a1 * (ur - cr * nx)) + A2 * (a2 * (-ny)) + A3 * (a3 * (ur + cr * nx)));
    out[2] = 0.5 * (F3L + F3R) -
             0.5 * (A1 * (a1 * (vr - cr * ny)) + A2 * (a2 * nx) + A3 * (a3 * (vr + cr * ny)));
}

// 5-variable Lax-Friedrichs of apply_interface_fluxes (primitive_equations.cpp:452-492)
__device__ void pe_lf5(const double *L, const double *R, const double *n, double g, double *out)
{
    const double cmax = sqrt(g * dmax(L[2] + 10.0, R[2] + 10.0));
    double F[2][5];
    for (int s = 0; s < 2; ++s) {
        const double *U = s ? R : L;
        const double H = U[2] + 10.0;
        const double u = H > 1e-10 ? U[0] / H : 0.0, v = H > 1e-10 ? U[1] / H : 0.0;
        const double un = u * n[0] + v * n[1];
        F[s][2] = U[0] * n[0] + U[1] * n[1];
        F[s][0] = U[0] * un + 0.5 * g * H * H * n[0];
        F[s][1] = U[1] * un + 0.5 * g * H * H * n[1];
        F[s][3] = U[3] * un;
        F[s][4] = U[4] * un;
    }
    for (int i = 0; i < 5; ++i)
        out[i] = 0.5 * (F[0][i] + F[1][i]) - 0.5 * cmax * (R[i] - L[i]);
}

// scalar fluxes carried through ConformingInterface / MortarSpace
__device__ __forceinline__ double scalar_flux(int kind, double a, double b, const double *n,
                                              const FluxParams &p)
{
    return scalar_flux_t<double>(kind, a, b, n, p.v);
}

__global__ void __launch_bounds__(256)
    flux_points_kernel(int kind, int64_t npts, const double *__restrict__ UL,
                       const double *__restrict__ UR, const double *__restrict__ nrm,
                       FluxParams p, double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts)
        return;
    const int nv = flux_num_vars(kind);
    double L[5], R[5], n[3], F[5];
    for (int v = 0; v < nv; ++v) {
        L[v] = UL[v * npts + t];
        R[v] = UR[v * npts + t];
    }
    for (int d = 0; d < 3; ++d)
        n[d] = nrm[d * npts + t];
    switch (kind) {
    case DRF_FLUX_SW_LF: sw_lf(L, R, n, p.v[0], F); break;
    case DRF_FLUX_SW_HLLC: sw_hllc(L, R, n, p.v[0], F); break;
    case DRF_FLUX_SW_ROE: sw_roe(L, R, n, p.v[0], p.v[1], F); break;
    case DRF_FLUX_SW_CENTRAL: sw_central(L, R, n, p.v[0], F); break;
    case DRF_FLUX_PE_LF5: pe_lf5(L, R, n, p.v[0], F); break;
    default: F[0] = scalar_flux(kind, L[0], R[0], n, p); break;
    }
    for (int v = 0; v < nv; ++v)
        out[v * npts + t] = F[v];
}

// ConformingInterface::compute_flux: N-point LGL face rule (the quadrature is
// built with `order` points, quadrature_3d.cpp:13-31).  One face point per
// lane over the flattened (face, point) range.
__global__ void __launch_bounds__(256)
    conforming_kernel(int kind, FluxParams p, int nf, FaceRule rule, int64_t nfaces,
                      const double *__restrict__ UL, const double *__restrict__ UR,
                      const double *__restrict__ nrm, double *__restrict__ rl,
                      double *__restrict__ rr)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfaces * nf)
        return;
    const int64_t face = t / nf;
    const int q = (int)(t - face * nf);
    const int nq1 = rule.nq1;
    double l = 0.0, r = 0.0;
    if (q < nq1 * nq1) {
        const double n[3] = {nrm[3 * face], nrm[3 * face + 1], nrm[3 * face + 2]};
        const int i = q % nq1, j = q / nq1;
        const double w = rule.w[i] * rule.w[j];
        const double F = scalar_flux(kind, UL[t], UR[t], n, p);
        l = 0.0 + w * F;
        r = 0.0 - w * F;
    }
    rl[t] = l;
    rr[t] = r;
}

// MortarSpace::compute_mortar_flux for a batch of interfaces of one type.
// ops = [Minv | Pc | Lc | Pf[0..nfine) | Lf[0..nfine)], each nm x nm
// row-major.  STAGE: the operators are copied to LDS once per workgroup
// (they fit for every type up to order 4); otherwise they are read through
// the caches.  A workgroup takes ipb = 256 / nm interfaces at a time, one
// lane per (interface, mortar DOF), and strides over the batch.
template <bool STAGE>
__global__ void __launch_bounds__(256)
    mortar_kernel(int kind, FluxParams p, int nm, int nfine, const double *__restrict__ ops,
                  int64_t nif, const double *__restrict__ Uc, const double *__restrict__ Uf,
                  const double *__restrict__ nrm, double *__restrict__ rc,
                  double *__restrict__ rf)
{
    extern __shared__ double sm[];
    const int nmat = 3 + 2 * nfine, mm = nm * nm;
    const int ipb = 256 / nm;                  // interfaces per iteration
    const int per = (1 + nfine) * nm + 2 * nm; // inputs, flux, M^-1 flux
    double *sbuf = sm;                         // ipb * per
    const double *op = ops;
    if (STAGE) {
        double *sops = sm + (size_t)ipb * per;
        for (int i = threadIdx.x; i < nmat * mm; i += blockDim.x)
            sops[i] = ops[i];
        op = sops;
    }
    const double *Minv = op, *Pc = op + mm, *Lc = op + 2 * mm, *Pf = op + 3 * mm,
                 *Lf = op + (size_t)(3 + nfine) * mm;
    const int g = threadIdx.x / nm, r = threadIdx.x - g * nm;
    const bool lane_on = g < ipb;
    double *su = sbuf + (size_t)(lane_on ? g : 0) * per; // Uc | Uf[nfine] | F | M^-1 F
    double *sF = su + (1 + nfine) * nm;
    for (int64_t i0 = (int64_t)blockIdx.x * ipb; i0 < nif; i0 += (int64_t)gridDim.x * ipb) {
        const int64_t it = i0 + g;
        const bool act = lane_on && it < nif;
        __syncthreads(); // operators staged; the previous interfaces are done
        if (act) {
            su[r] = Uc[it * nm + r];
            for (int f = 0; f < nfine; ++f)
                su[nm + f * nm + r] = Uf[(it * nfine + f) * nm + r];
        }
        __syncthreads();
        if (act) {
            const double n[3] = {nrm[3 * it], nrm[3 * it + 1], nrm[3 * it + 2]};
            double uc = 0.0;
            for (int c = 0; c < nm; ++c)
                uc += Pc[r * nm + c] * su[c];
            double uf = 0.0;
            for (int f = 0; f < nfine; ++f) {
                double t = 0.0;
                for (int c = 0; c < nm; ++c)
                    t += Pf[(size_t)f * mm + r * nm + c] * su[nm + f * nm + c];
                uf += t;
            }
            sF[r] = scalar_flux(kind, uc, uf, n, p);
        }
        __syncthreads();
        if (act) {
            double mf = 0.0;
            for (int c = 0; c < nm; ++c)
                mf += Minv[r * nm + c] * sF[c];
            sF[nm + r] = mf;
        }
        __syncthreads();
        if (act) {
            double t = 0.0;
            for (int c = 0; c < nm; ++c)
                t += Lc[r * nm + c] * sF[nm + c];
            rc[it * nm + r] = t;
            for (int f = 0; f < nfine; ++f) {
                double v = 0.0;
                for (int c = 0; c < nm; ++c)
                    v += -Lf[(size_t)f * mm + r * nm + c] * sF[nm + c];
                rf[(it * nfine + f) * nm + r] = v;
            }
        }
    }
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MortarSpace::compute_mortar_flux on the FP64 matrix cores, for nm =
// (N+1)^2 <= 25 (orders 1-4).  A wavefront takes 16 interfaces; every
// operator product is a (16 MT) x 16 x nm GEMM, operator (rows = mortar DOFs,
// zero-padded to MT tiles of 16 rows by KP columns in LDS) times the 16
// interfaces' vectors, in v_mfma_f64_16x16x4f64 steps of K = 4, one chain per
// 16-row tile.  Lane l = col + 16 kq holds
//   A: tile row m = col, K index 4c + kq of step c;   B: the same K index of
//   interface col;   D: tile rows kq + 4 i (i < 4) of interface col
// (the f64 variant interleaves its output rows; measured on gfx950).  Two
// permutations make every lane's memory accesses contiguous: step c, lane
// group kq contracts DOF KC kq + c (so a lane's KC input DOFs are
// consecutive), and A row m of tile t is operator row 16 t + 4 (m % 4) + m / 4
// (so a lane's four outputs of tile t are rows 16 t + 4 kq .. 16 t + 4 kq + 3).
// The fine projections accumulate into one chain per tile (K = nfine nm); the
// flux is taken on the D layout, and M^-1 F and the lifts read their B operand
// back from a per-wave (16 MT) x 16 LDS tile.  Products round as the matrix
// cores round them (within 1e-13 of the sequential sums).
//
// Orders 1-3 (one tile) run four wavefronts per workgroup; order 4 (two
// tiles, up to 82 KB of operators for 2 x 2 sub-faces, 144 KB of LDS with
// the flux tiles) runs sixteen, so that one workgroup's operator copy serves
// four wavefronts per SIMD (126 VGPRs fit the 128 that allows; measured
// 62.8 us for 100,000 2 x 2 interfaces, against 76.3 us with eight).
template <int NM> constexpr int mortar_waves() { return NM > 16 ? 16 : 4; }

template <int NM, int NF>
__global__ void __launch_bounds__(64 * mortar_waves<NM>())
    mortar_mfma_kernel(int kind, FluxParams p, const double *__restrict__ ops, int64_t nif,
                       const double *__restrict__ Uc, const double *__restrict__ Uf,
                       const double *__restrict__ nrm, double *__restrict__ rc,
                       double *__restrict__ rf)
{
    constexpr int W = mortar_waves<NM>(), MT = (NM + 15) / 16, NR = 16 * MT;
    constexpr int KC = (NM + 3) / 4, KP = 4 * KC, NMAT = 3 + 2 * NF, MM = NM * NM;
    static_assert(KP <= NR, "the K index of M^-1 F must stay inside the flux tile");
    // one padded, permuted operator: NR rows of KP, row stride KP + 1 (odd,
    // so the 16 rows an A fragment reads fall in different LDS banks)
    constexpr int RS = KP + 1, OPS = NR * RS;
    __shared__ double sOp[NMAT * OPS];
    __shared__ double sT[W][NR * 16];
    // sOp[mat][16 t + m][4c + kq] = Op[16 t + 4 (m % 4) + m / 4][KC kq + c]
    for (int q = threadIdx.x; q < NMAT * OPS; q += 64 * W) {
        const int m = q / OPS, r = (q - m * OPS) / RS, j = q - m * OPS - r * RS;
        const int rr = r & 15;
        const int row = (r & ~15) + 4 * (rr & 3) + (rr >> 2), k = KC * (j & 3) + (j >> 2);
        sOp[q] = (j < KP && row < NM && k < NM) ? ops[m * MM + row * NM + k] : 0.0;
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 15, kq = l >> 4;
    double *T = sT[w];
    const double *Minv = sOp, *Pc = sOp + OPS, *Lc = sOp + 2 * OPS, *Pf = sOp + 3 * OPS,
                 *Lf = sOp + (3 + NF) * OPS;
    const int k0 = KC * kq; // this lane's first K index
    // A fragment of tile t, step c, of the operator at o
#define MORTAR_A(o, t, c) (o)[((t) * 16 + col) * RS + 4 * (c) + kq]
    for (int64_t i0 = ((int64_t)blockIdx.x * W + w) * 16; i0 < nif;
         i0 += (int64_t)gridDim.x * W * 16) {
        const int64_t it = i0 + col;
        const bool valid = it < nif;
        // the lane's KC consecutive DOFs of every input face
        double bc[KC], bf[NF][KC];
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            bc[c] = 0.0;
#pragma unroll
            for (int f = 0; f < NF; ++f)
                bf[f][c] = 0.0;
        }
        if (valid) {
            const double *uc = Uc + it * NM + k0, *uf = Uf + it * NF * NM + k0;
#pragma unroll
            for (int c = 0; c < KC; ++c)
                if (NM % 4 == 0 || k0 + c < NM) {
                    bc[c] = uc[c];
#pragma unroll
                    for (int f = 0; f < NF; ++f)
                        bf[f][c] = uf[f * NM + c];
                }
        }
        dbl4 ac[MT], af[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            ac[t] = dbl4{0.0, 0.0, 0.0, 0.0};
            af[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < KC; ++c)
                ac[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(MORTAR_A(Pc, t, c), bc[c], ac[t], 0,
                                                             0, 0);
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int c = 0; c < KC; ++c)
                    af[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(MORTAR_A(Pf + f * OPS, t, c),
                                                                 bf[f][c], af[t], 0, 0, 0);
        }
        double n3[3] = {0.0, 0.0, 0.0};
        if (valid) {
            n3[0] = nrm[3 * it];
            n3[1] = nrm[3 * it + 1];
            n3[2] = nrm[3 * it + 2];
        }
        // D register i of tile t is mortar DOF 16 t + 4 kq + i
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 16 * t + 4 * kq + i;
                T[row * 16 + col] =
                    (valid && row < NM) ? scalar_flux(kind, ac[t][i], af[t][i], n3, p) : 0.0;
            }
        wave_sync_lds();
        dbl4 am[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            am[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < KC; ++c)
                am[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(MORTAR_A(Minv, t, c),
                                                             T[(k0 + c) * 16 + col], am[t], 0, 0,
                                                             0);
        }
        wave_sync_lds(); // every lane has read F before M^-1 F overwrites it
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                T[(16 * t + 4 * kq + i) * 16 + col] = am[t][i];
        wave_sync_lds();
        double bm[KC];
#pragma unroll
        for (int c = 0; c < KC; ++c)
            bm[c] = T[(k0 + c) * 16 + col];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            dbl4 ar = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < KC; ++c)
                ar = __builtin_amdgcn_mfma_f64_16x16x4f64(MORTAR_A(Lc, t, c), bm[c], ar, 0, 0, 0);
            if (valid)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (16 * t + 4 * kq + i < NM)
                        rc[it * NM + 16 * t + 4 * kq + i] = ar[i];
        }
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                dbl4 a2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int c = 0; c < KC; ++c)
                    a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(MORTAR_A(Lf + f * OPS, t, c), bm[c],
                                                              a2, 0, 0, 0);
                if (valid)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (16 * t + 4 * kq + i < NM)
                            rf[(it * NF + f) * NM + 16 * t + 4 * kq + i] = -a2[i];
            }
        wave_sync_lds(); // the tile is rewritten by the next interfaces
    }
#undef MORTAR_A
}

template <int NM, int NF>
hipError_t launch_mortar_mfma_nf(int kind, const FluxParams &p, const double *ops, int64_t nif,
                                 const double *Uc, const double *Uf, const double *nrm, double *rc,
                                 double *rf, hipStream_t s)
{
    constexpr int W = mortar_waves<NM>();
    // persistent workgroups: as many as are resident at once, so that each
    // stages the operators into LDS once
    static const int cap = [] {
        int dev = 0, per = 1, cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, mortar_mfma_kernel<NM, NF>, 64 * W,
                                                         0) != hipSuccess ||
            per < 1)
            per = 1;
        return cus * per;
    }();
    const int64_t need = (nif + 16 * W - 1) / (16 * W);
    hipLaunchKernelGGL((mortar_mfma_kernel<NM, NF>), dim3((unsigned)(need < cap ? need : cap)),
                       dim3(64 * W), 0, s, kind, p, ops, nif, Uc, Uf, nrm, rc, rf);
    return hipGetLastError();
}

template <int NM>
hipError_t launch_mortar_mfma(int kind, const FluxParams &p, int nfine, const double *ops,
                              int64_t nif, const double *Uc, const double *Uf, const double *nrm,
                              double *rc, double *rf, hipStream_t s)
{
    switch (nfine) {
    case 1: return launch_mortar_mfma_nf<NM, 1>(kind, p, ops, nif, Uc, Uf, nrm, rc, rf, s);
    case 2: return launch_mortar_mfma_nf<NM, 2>(kind, p, ops, nif, Uc, Uf, nrm, rc, rf, s);
    case 3: return launch_mortar_mfma_nf<NM, 3>(kind, p, ops, nif, Uc, Uf, nrm, rc, rf, s);
    case 4: return launch_mortar_mfma_nf<NM, 4>(kind, p, ops, nif, Uc, Uf, nrm, rc, rf, s);
    default: return hipErrorInvalidValue;
    }
}

// verify_mortar_conservation (src/dg/mortar.cpp:386-413), one interface per
// lane: the weighted coarse and fine integrals in the reference's order.
__global__ void __launch_bounds__(256)
    mortar_conservation_kernel(int nm, int nfine, int64_t nif, const double *__restrict__ rc,
                               const double *__restrict__ rf, const double *__restrict__ wc,
                               const double *__restrict__ wf, double *__restrict__ out)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nif;
         k += (int64_t)gridDim.x * blockDim.x) {
        double ic = 0.0, ifn = 0.0;
        for (int i = 0; i < nm; ++i)
            ic += wc[i] * rc[k * nm + i];
        for (int f = 0; f < nfine; ++f)
            for (int i = 0; i < nm; ++i)
                ifn += wf[f * nm + i] * rf[(k * nfine + f) * nm + i];
        const double total = ic + ifn, norm = fabs(ic) + fabs(ifn);
        out[k] = (norm < 1e-15) ? 0.0 : fabs(total) / norm;
    }
}

// Face traces of a nodal field: out[k][q] = field[elems[k] * es + node(face[k], q)].
// The node count per edge is a template argument, so the index divisions are
// multiplications.
template <int n>
__global__ void __launch_bounds__(256)
    gather_faces_kernel(int64_t nelem, int64_t es, int64_t nfaces,
                        const int64_t *__restrict__ elems, const int8_t *__restrict__ faces,
                        const double *__restrict__ field, double *__restrict__ out)
{
    constexpr int nf = n * n;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfaces * nf)
        return;
    const int64_t k = t / nf;
    const int q = (int)(t - k * nf);
    const int64_t e = elems[k];
    const int f = faces[k];
    if (e < 0 || e >= nelem || f < 0 || f > 5) {
        out[t] = __builtin_nan("");
        return;
    }
    out[t] = field[e * es + face_node(n, f, q % n, q / n)];
}

} // namespace

hipError_t launch_flux_points(int kind, int64_t npts, const double *UL, const double *UR,
                              const double *nrm, const FluxParams &p, double *out, hipStream_t s)
{
    if (npts <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(flux_points_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s,
                       kind, npts, UL, UR, nrm, p, out);
    return hipGetLastError();
}

hipError_t launch_conforming(int kind, const FluxParams &p, int nf, const FaceRule &rule,
                             int64_t nfaces, const double *UL, const double *UR,
                             const double *nrm, double *rl, double *rr, hipStream_t s)
{
    const int64_t total = nfaces * nf;
    if (total <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(conforming_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       kind, p, nf, rule, nfaces, UL, UR, nrm, rl, rr);
    return hipGetLastError();
}

hipError_t launch_mortar(int kind, const FluxParams &p, int nm, int nfine, const double *ops,
                         int64_t nif, const double *Uc, const double *Uf, const double *nrm,
                         double *rc, double *rf, hipStream_t s)
{
    if (nif <= 0)
        return hipSuccess;
    switch (nm) { // orders 1-4 on the matrix cores
    case 4: return launch_mortar_mfma<4>(kind, p, nfine, ops, nif, Uc, Uf, nrm, rc, rf, s);
    case 9: return launch_mortar_mfma<9>(kind, p, nfine, ops, nif, Uc, Uf, nrm, rc, rf, s);
    case 16: return launch_mortar_mfma<16>(kind, p, nfine, ops, nif, Uc, Uf, nrm, rc, rf, s);
    case 25: return launch_mortar_mfma<25>(kind, p, nfine, ops, nif, Uc, Uf, nrm, rc, rf, s);
    default: break;
    }
    const int ipb = 256 / nm;
    const size_t base = sizeof(double) * (size_t)ipb * ((size_t)(1 + nfine) * nm + 2 * (size_t)nm);
    const size_t opb = sizeof(double) * (size_t)(3 + 2 * nfine) * nm * nm;
    const int64_t need = (nif + ipb - 1) / ipb;
    const unsigned grid = (unsigned)(need < 4096 ? need : 4096);
    if (base + opb <= 64 * 1024)
        hipLaunchKernelGGL(mortar_kernel<true>, dim3(grid), dim3(256), base + opb, s, kind, p, nm,
                           nfine, ops, nif, Uc, Uf, nrm, rc, rf);
    else
        hipLaunchKernelGGL(mortar_kernel<false>, dim3(grid), dim3(256), base, s, kind, p, nm,
                           nfine, ops, nif, Uc, Uf, nrm, rc, rf);
    return hipGetLastError();
}

hipError_t launch_mortar_conservation(int nm, int nfine, int64_t nif, const double *rc,
                                     const double *rf, const double *wc, const double *wf,
                                     double *out, hipStream_t s)
{
    if (nif <= 0)
        return hipSuccess;
    const int64_t b = (nif + 255) / 256;
    hipLaunchKernelGGL(mortar_conservation_kernel, dim3((unsigned)(b < 65536 ? b : 65536)),
                       dim3(256), 0, s, nm, nfine, nif, rc, rf, wc, wf, out);
    return hipGetLastError();
}

hipError_t launch_gather_faces(int n, int64_t nelem, int64_t elem_stride, int64_t nfaces,
                               const int64_t *elems, const int8_t *faces, const double *field,
                               double *out, hipStream_t s)
{
    const int64_t total = nfaces * n * n;
    if (total <= 0)
        return hipSuccess;
    const dim3 g((unsigned)((total + 255) / 256)), b(256);
#define DRF_GATHER(NN)                                                                         \
    hipLaunchKernelGGL(gather_faces_kernel<NN>, g, b, 0, s, nelem, elem_stride, nfaces, elems,  \
                       faces, field, out);                                                     \
    break
    switch (n) {
    case 2: DRF_GATHER(2);
    case 3: DRF_GATHER(3);
    case 4: DRF_GATHER(4);
    case 5: DRF_GATHER(5);
    case 6: DRF_GATHER(6);
    case 7: DRF_GATHER(7);
    case 8: DRF_GATHER(8);
    case 9: DRF_GATHER(9);
    default: return hipErrorInvalidValue;
    }
#undef DRF_GATHER
    return hipGetLastError();
}

} // namespace drf
