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
// run persistent workgroups over the interfaces of one connection type: the
// type's projection and lift operators (built once on the host) are staged in
// LDS when they fit, and every lane owns one mortar DOF of one of the
// workgroup's interfaces.
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
    const double vn = p.v[0] * n[0] + p.v[1] * n[1] + p.v[2] * n[2];
    if (kind == DRF_FLUX_UPWIND)
        return vn * ((vn >= 0) ? a : b);
    double FnL = 0.0, FnR = 0.0;
    for (int d = 0; d < 3; ++d) {
        FnL += (p.v[d] * a) * n[d];
        FnR += (p.v[d] * b) * n[d];
    }
    return 0.5 * (FnL + FnR) - 0.5 * p.v[3] * (b - a);
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

// Face traces of a nodal field: out[k][q] = field[elems[k] * es + node(face[k], q)].
__global__ void __launch_bounds__(256)
    gather_faces_kernel(int n, int64_t nelem, int64_t es, int64_t nfaces,
                        const int64_t *__restrict__ elems, const int8_t *__restrict__ faces,
                        const double *__restrict__ field, double *__restrict__ out)
{
    const int nf = n * n;
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

hipError_t launch_gather_faces(int n, int64_t nelem, int64_t elem_stride, int64_t nfaces,
                               const int64_t *elems, const int8_t *faces, const double *field,
                               double *out, hipStream_t s)
{
    const int64_t total = nfaces * n * n;
    if (total <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(gather_faces_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       n, nelem, elem_stride, nfaces, elems, faces, field, out);
    return hipGetLastError();
}

} // namespace drf
