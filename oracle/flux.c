/*
 * ORACLE (test infrastructure only) -- A14-A16, A18: numerical fluxes,
 * conforming interfaces and 2:1 mortars.
 *
 * Follows include/flux/numerical_flux.hpp:57-397 (LaxFriedrichsFlux,
 * CentralFlux, UpwindFlux, ShallowWaterHLLCFlux, LocalLaxFriedrichs),
 * src/flux/numerical_flux.cpp:13-200 (ShallowWaterRoeFlux and the factory's
 * shallow-water Lax-Friedrichs), the 5-variable Lax-Friedrichs lambda of
 * src/physics/primitive_equations.cpp:452-492, ConformingInterface
 * (src/dg/mortar.cpp:297-343), MortarSpace (:13-258), the LGL face extraction
 * of build_face_interpolation (src/dg/basis_hexahedron.cpp:431-522) and
 * LagrangeBasis1D::evaluate (include/dg/basis_hexahedron.hpp:316-348).
 *
 * Pinned by (tests/test_oracle_flux.py): consistency F*(U,U,n) = F(U).n and
 * the conservation identity F*(UL,UR,n) + F*(UR,UL,-n) = 0 for LF/HLLC/Roe
 * with n = (0.6, 0.8, 0) (tests/unit/test_numerical_flux.cpp:17-178), upwind
 * direction, dry states without NaN, a textbook (Toro) HLL/Roe evaluation in
 * numpy on random states; mortar sizes, mass matrix symmetric positive
 * definite, P = I for a single full sub-face, exact projection of
 * polynomials of degree <= N, and flux conservation of a constant flux
 * (verify_mortar_conservation, src/dg/mortar.cpp:386-413)
 * (tests/unit/test_mortar.cpp).  The inverse of the mortar mass matrix uses
 * Gauss-Jordan with partial pivoting, not Eigen's LU: entries agree to a few
 * ulp ("parity unpinned" at the ulp level).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* NumericalFluxFactory::shallow_water_lax_friedrichs, numerical_flux.cpp:134-181
 * with LaxFriedrichsFlux::flux, numerical_flux.hpp:222-239 */
static void sw_lf(const double *L, const double *R, const double *n, double g, double *out)
{
    double FL[3][3], FR[3][3]; /* [dir][var] */
    const double *U[2] = {L, R};
    for (int s = 0; s < 2; ++s) {
        double(*F)[3] = s ? FR : FL;
        double h = U[s][0];
        const double hu = U[s][1], hv = U[s][2];
        if (h < 1e-10)
            h = 1e-10;
        const double u = hu / h, v = hv / h;
        F[0][0] = hu;
        F[0][1] = hu * u + 0.5 * g * h * h;
        F[0][2] = hu * v;
        F[1][0] = hv;
        F[1][1] = hv * u;
        F[1][2] = hv * v + 0.5 * g * h * h;
        F[2][0] = F[2][1] = F[2][2] = 0.0;
    }
    double FnL[3], FnR[3];
    for (int i = 0; i < 3; ++i) {
        FnL[i] = 0.0;
        FnR[i] = 0.0;
        for (int d = 0; d < 3; ++d) {
            FnL[i] += FL[d][i] * n[d];
            FnR[i] += FR[d][i] * n[d];
        }
    }
    const double hL = dmax(L[0], 1e-10), hR = dmax(R[0], 1e-10);
    const double uL = L[1] / hL, vL = L[2] / hL, uR = R[1] / hR, vR = R[2] / hR;
    const double cL = sqrt(g * hL), cR = sqrt(g * hR);
    const double unL = uL * n[0] + vL * n[1], unR = uR * n[0] + vR * n[1];
    const double lam = dmax(fabs(unL) + cL, fabs(unR) + cR);
    for (int i = 0; i < 3; ++i)
        out[i] = 0.5 * (FnL[i] + FnR[i]) - 0.5 * lam * (R[i] - L[i]);
}

/* central flux with the shallow-water physical flux (CentralFlux,
 * numerical_flux.hpp:241-259) */
static void sw_central(const double *L, const double *R, const double *n, double g, double *out)
{
    double tmp[3];
    sw_lf(L, L, n, g, tmp); /* lambda term vanishes for identical states */
    double FnL[3], FnR[3];
    memcpy(FnL, tmp, sizeof(tmp));
    sw_lf(R, R, n, g, tmp);
    memcpy(FnR, tmp, sizeof(tmp));
    for (int i = 0; i < 3; ++i)
        out[i] = 0.5 * (FnL[i] + FnR[i]);
}

static void hllc_phys(double h, double u, double v, double g, double nx, double ny, double *F)
{
    const double un = u * nx + v * ny;
    F[0] = h * un;
    F[1] = h * u * un + 0.5 * g * h * h * nx;
    F[2] = h * v * un + 0.5 * g * h * h * ny;
}

/* ShallowWaterHLLCFlux::flux + wave_speeds, numerical_flux.hpp:272-397 */
static void sw_hllc(const double *L, const double *R, const double *n, double g, double *out)
{
    double hL = L[0], hR = R[0];
    const double huL = L[1], hvL = L[2], huR = R[1], hvR = R[2];
    if (hL < 1e-10)
        hL = 1e-10;
    if (hR < 1e-10)
        hR = 1e-10;
    const double uL = huL / hL, vL = hvL / hL, uR = huR / hR, vR = hvR / hR;
    const double nx = n[0], ny = n[1];
    const double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
    const double cL = sqrt(g * hL), cR = sqrt(g * hR);
    const double SL = dmin(unL - cL, unR - cR), SR = dmax(unL + cL, unR + cR);
    double Ss;
    const double den = hL * (SL - unL) - hR * (SR - unR);
    if (fabs(den) > 1e-10)
        Ss = (hL * unL * (SL - unL) - hR * unR * (SR - unR) + 0.5 * g * (hR * hR - hL * hL)) / den;
    else
        Ss = 0.5 * (unL + unR);
    double FL[3], FR[3];
    hllc_phys(hL, uL, vL, g, nx, ny, FL);
    hllc_phys(hR, uR, vR, g, nx, ny, FR);
    if (SL >= 0) {
        memcpy(out, FL, sizeof(FL));
    } else if (SR <= 0) {
        memcpy(out, FR, sizeof(FR));
    } else if (Ss >= 0) {
        const double hs = hL * (SL - unL) / (SL - Ss);
        const double Us[3] = {hs, hs * (uL + (Ss - unL) * nx), hs * (vL + (Ss - unL) * ny)};
        const double Uv[3] = {hL, huL, hvL};
        for (int i = 0; i < 3; ++i)
            out[i] = FL[i] + SL * (Us[i] - Uv[i]);
    } else {
        const double hs = hR * (SR - unR) / (SR - Ss);
        const double Us[3] = {hs, hs * (uR + (Ss - unR) * nx), hs * (vR + (Ss - unR) * ny)};
        const double Uv[3] = {hR, huR, hvR};
        for (int i = 0; i < 3; ++i)
            out[i] = FR[i] + SR * (Us[i] - Uv[i]);
    }
}

/* ShallowWaterRoeFlux::flux, numerical_flux.cpp:13-128 */
static void sw_roe(const double *L, const double *R, const double *n, double g, double efix,
                   double *out)
{
    double hL = L[0], hR = R[0];
    if (hL < 1e-10)
        hL = 1e-10;
    if (hR < 1e-10)
        hR = 1e-10;
    const double uL = L[1] / hL, vL = L[2] / hL, uR = R[1] / hR, vR = R[2] / hR;
    const double nx = n[0], ny = n[1];
    const double unL = uL * nx + vL * ny, utL = -uL * ny + vL * nx;
    const double unR = uR * nx + vR * ny, utR = -uR * ny + vR * nx;
    const double sL = sqrt(hL), sR = sqrt(hR), den = sL + sR;
    const double hr = 0.5 * (hL + hR);
    const double ur = (sL * uL + sR * uR) / den, vr = (sL * vL + sR * vR) / den;
    const double unr = ur * nx + vr * ny, cr = sqrt(g * hr);
    double l1 = unr - cr, l2 = unr, l3 = unr + cr;
    const double delta = efix * cr;
    if (fabs(l1) < delta) {
        l1 = (l1 * l1 + delta * delta) / (2.0 * delta);
        if (l1 < 0)
            l1 = -l1;
    }
    if (fabs(l3) < delta) {
        l3 = (l3 * l3 + delta * delta) / (2.0 * delta);
        if (l3 < 0)
            l3 = -l3;
    }
    const double dh = hR - hL, dun = unR - unL, dut = utR - utL;
    const double a1 = 0.5 * (dh - hr / cr * dun), a2 = hr * dut, a3 = 0.5 * (dh + hr / cr * dun);
    const double F1L = hL * unL, F2L = hL * uL * unL + 0.5 * g * hL * hL * nx,
                 F3L = hL * vL * unL + 0.5 * g * hL * hL * ny;
    const double F1R = hR * unR, F2R = hR * uR * unR + 0.5 * g * hR * hR * nx,
                 F3R = hR * vR * unR + 0.5 * g * hR * hR * ny;
    const double A1 = fabs(l1), A2 = fabs(l2), A3 = fabs(l3);
    const double r11 = a1, r12 = a1 * (ur - cr * nx), r13 = a1 * (vr - cr * ny);
    const double r21 = 0.0, r22 = a2 * (-ny), r23 = a2 * (nx);
    const double r31 = a3, r32 = a3 * (ur + cr * nx), r33 = a3 * (vr + cr * ny);
    out[0] = 0.5 * (F1L + F1R) - 0.5 * (A1 * r11 + A2 * r21 + A3 * r31);
    out[1] = 0.5 * (F2L + F2R) - 0.5 * (A1 * r12 + A2 * r22 + A3 * r32);
    out[2] = 0.5 * (F3L + F3R) - 0.5 * (A1 * r13 + A2 * r23 + A3 * r33);
}

/* 5-variable Lax-Friedrichs of apply_interface_fluxes, primitive_equations.cpp:452-492 */
static void pe_lf5(const double *L, const double *R, const double *n, double g, double *out)
{
    const double HL = L[2] + 10.0, HR = R[2] + 10.0;
    const double cmax = sqrt(g * dmax(HL, HR));
    double FL[5], FR[5];
    const double *U[2] = {L, R};
    for (int s = 0; s < 2; ++s) {
        double *F = s ? FR : FL;
        const double H = U[s][2] + 10.0;
        const double u = H > 1e-10 ? U[s][0] / H : 0.0;
        const double v = H > 1e-10 ? U[s][1] / H : 0.0;
        const double un = u * n[0] + v * n[1];
        F[2] = U[s][0] * n[0] + U[s][1] * n[1];
        F[0] = U[s][0] * un + 0.5 * g * H * H * n[0];
        F[1] = U[s][1] * un + 0.5 * g * H * H * n[1];
        F[3] = U[s][3] * un;
        F[4] = U[s][4] * un;
    }
    for (int i = 0; i < 5; ++i)
        out[i] = 0.5 * (FL[i] + FR[i]) - 0.5 * cmax * (R[i] - L[i]);
}

/* scalar fluxes used through the interface components */
static double scalar_flux(int kind, double a, double b, const double *n, const double *p)
{
    const double vn = p[0] * n[0] + p[1] * n[1] + p[2] * n[2]; /* Vec3::dot */
    if (kind == ORC_FLUX_UPWIND) /* UpwindFlux, numerical_flux.hpp:261-268 */
        return vn * ((vn >= 0) ? a : b);
    /* ORC_FLUX_LF_ADV: LaxFriedrichsFlux with F(U) = v U and speed p[3]
     * (numerical_flux.hpp:222-239): 0.5 (vn a + vn b) - 0.5 lambda (b - a) */
    double FnL = 0.0, FnR = 0.0;
    for (int d = 0; d < 3; ++d) {
        FnL += (p[d] * a) * n[d];
        FnR += (p[d] * b) * n[d];
    }
    return 0.5 * (FnL + FnR) - 0.5 * p[3] * (b - a);
}

int orc_flux(int kind, const double *UL, const double *UR, const double *n, const double *p,
             double *out)
{
    switch (kind) {
    case ORC_FLUX_SW_LF: sw_lf(UL, UR, n, p[0], out); return 3;
    case ORC_FLUX_SW_HLLC: sw_hllc(UL, UR, n, p[0], out); return 3;
    case ORC_FLUX_SW_ROE: sw_roe(UL, UR, n, p[0], p[1], out); return 3;
    case ORC_FLUX_SW_CENTRAL: sw_central(UL, UR, n, p[0], out); return 3;
    case ORC_FLUX_PE_LF5: pe_lf5(UL, UR, n, p[0], out); return 5;
    case ORC_FLUX_UPWIND:
    case ORC_FLUX_LF_ADV: out[0] = scalar_flux(kind, UL[0], UR[0], n, p); return 1;
    default: return -1;
    }
}

/* ---------------- LGL face extraction (A18) ---------------- */
/* build_face_interpolation, basis_hexahedron.cpp:431-522: face DOF
 * it + n jt <- volume node with the normal index at its boundary value and
 * the two tangent indices (it, jt) in tangent-axis order. */
void orc_face_map(int order, int face, int *vol_index)
{
    const int n = order + 1;
    const int nax = face / 2, pos = face % 2;
    const int t1 = (nax == 0) ? 1 : 0, t2 = (nax == 2) ? 1 : 2;
    for (int jt = 0; jt < n; ++jt)
        for (int it = 0; it < n; ++it) {
            int idx[3];
            idx[nax] = pos ? n - 1 : 0;
            idx[t1] = it;
            idx[t2] = jt;
            vol_index[it + n * jt] = idx[0] + n * (idx[1] + n * idx[2]);
        }
}

/* ---------------- ConformingInterface (A15) ---------------- */
/* compute_flux, mortar.cpp:320-343, with FaceQuadrature(face, N, LGL): N
 * points per direction (GaussQuadrature1D uses `order` points,
 * quadrature_3d.cpp:13-31), so only the first N^2 face values are used. */
int orc_conforming_flux(int order, int kind, const double *p, const double *UL,
                        const double *UR, const double *n, double *rhs_l, double *rhs_r)
{
    const int nf = (order + 1) * (order + 1);
    int nq1 = order < 2 ? 2 : order;
    double x[ORC_MAX_N + 1], w[ORC_MAX_N + 1];
    if (orc_lgl_nodes(nq1, x, w))
        return -1;
    for (int i = 0; i < nf; ++i)
        rhs_l[i] = rhs_r[i] = 0.0;
    for (int j = 0; j < nq1; ++j)
        for (int i = 0; i < nq1; ++i) {
            const int q = i + nq1 * j;
            const double wq = w[i] * w[j];
            const double F = scalar_flux(kind, UL[q], UR[q], n, p);
            rhs_l[q] += wq * F;
            rhs_r[q] -= wq * F;
        }
    return nq1 * nq1;
}

/* ---------------- MortarSpace (A16) ---------------- */
static int nfine_of(int type)
{
    switch (type) {
    case ORC_CONN_SAME: return 1;
    case ORC_CONN_FINE2X1:
    case ORC_CONN_FINE1X2: return 2;
    case ORC_CONN_FINE3_2P1:
    case ORC_CONN_FINE3_1P2: return 3;
    case ORC_CONN_FINE2X2: return 4;
    default: return 0;
    }
}

/* compute_subface_bounds, mortar.cpp:27-78: {t1min, t2min, t1max, t2max} */
static void subface_bounds(int type, int f, double *bb)
{
    static const double same[1][4] = {{-1, -1, 1, 1}};
    static const double f21[2][4] = {{-1, -1, 0, 1}, {0, -1, 1, 1}};
    static const double f12[2][4] = {{-1, -1, 1, 0}, {-1, 0, 1, 1}};
    static const double f22[4][4] = {{-1, -1, 0, 0}, {0, -1, 1, 0}, {-1, 0, 0, 1}, {0, 0, 1, 1}};
    static const double f3a[3][4] = {{-1, -1, 0, 0}, {0, -1, 1, 0}, {-1, 0, 1, 1}};
    static const double f3b[3][4] = {{-1, -1, 0, 0}, {-1, 0, 0, 1}, {0, -1, 1, 1}};
    const double *s = same[0];
    switch (type) {
    case ORC_CONN_FINE2X1: s = f21[f]; break;
    case ORC_CONN_FINE1X2: s = f12[f]; break;
    case ORC_CONN_FINE2X2: s = f22[f]; break;
    case ORC_CONN_FINE3_2P1: s = f3a[f]; break;
    case ORC_CONN_FINE3_1P2: s = f3b[f]; break;
    default: break;
    }
    memcpy(bb, s, 4 * sizeof(double));
}

/* LagrangeBasis1D::evaluate, include/dg/basis_hexahedron.hpp:316-348 */
static void lagrange_eval(int n, const double *x, const double *bw, double xi, double *phi)
{
    double tmp[ORC_MAX_N], sum = 0.0;
    int at = -1;
    for (int j = 0; j < n; ++j) {
        const double d = xi - x[j];
        if (fabs(d) < 1e-15) {
            at = j;
            break;
        }
        tmp[j] = bw[j] / d;
        sum += tmp[j];
    }
    if (at >= 0) {
        for (int j = 0; j < n; ++j)
            phi[j] = 0.0;
        phi[at] = 1.0;
    } else {
        for (int j = 0; j < n; ++j)
            phi[j] = tmp[j] / sum;
    }
}

/* Gauss-Jordan inverse with partial pivoting (stands in for Eigen's inverse) */
static int invert(int m, const double *A, double *Ai)
{
    double *a = (double *)malloc(sizeof(double) * m * m);
    memcpy(a, A, sizeof(double) * m * m);
    for (int i = 0; i < m * m; ++i)
        Ai[i] = 0.0;
    for (int i = 0; i < m; ++i)
        Ai[i * m + i] = 1.0;
    for (int c = 0; c < m; ++c) {
        int piv = c;
        for (int r = c + 1; r < m; ++r)
            if (fabs(a[r * m + c]) > fabs(a[piv * m + c]))
                piv = r;
        if (a[piv * m + c] == 0.0) {
            free(a);
            return -1;
        }
        if (piv != c)
            for (int k = 0; k < m; ++k) {
                double t = a[c * m + k];
                a[c * m + k] = a[piv * m + k];
                a[piv * m + k] = t;
                t = Ai[c * m + k];
                Ai[c * m + k] = Ai[piv * m + k];
                Ai[piv * m + k] = t;
            }
        const double d = a[c * m + c];
        for (int k = 0; k < m; ++k) {
            a[c * m + k] /= d;
            Ai[c * m + k] /= d;
        }
        for (int r = 0; r < m; ++r) {
            if (r == c)
                continue;
            const double fct = a[r * m + c];
            if (fct == 0.0)
                continue;
            for (int k = 0; k < m; ++k) {
                a[r * m + k] -= fct * a[c * m + k];
                Ai[r * m + k] -= fct * Ai[c * m + k];
            }
        }
    }
    free(a);
    return 0;
}

/* MortarSpace ctor + build_projection_operators, mortar.cpp:13-25, 80-210.
 * Outputs (row-major): M, Minv (nm x nm), Pc, Lc (nm x nf / nf x nm) and per
 * fine face Pf[f], Lf[f].  nm = nf = (N+1)^2.  Returns the number of fine
 * faces. */
int orc_mortar_operators(int order, int type, double *M, double *Minv, double *Pc, double *Lc,
                         double *Pf, double *Lf)
{
    const int np = order + 1, nm = np * np, nf = nm;
    const int nfine = nfine_of(type);
    const int nq1 = 2 * order + 1;
    double gx[2 * ORC_MAX_N + 1], gw[2 * ORC_MAX_N + 1], lx[ORC_MAX_N], lw[ORC_MAX_N],
        bw[ORC_MAX_N];
    orc_gl_nodes(nq1, gx, gw);
    orc_lgl_nodes(np, lx, lw);
    orc_barycentric_weights(np, lx, bw);
    const int nq = nq1 * nq1;
    double *Phi = (double *)malloc(sizeof(double) * nq * nm);
    double *qx = (double *)malloc(sizeof(double) * nq * 3);
    for (int j = 0; j < nq1; ++j)
        for (int i = 0; i < nq1; ++i) {
            const int q = i + nq1 * j; /* GaussQuadrature2D ordering, quadrature_3d.cpp:60-75 */
            qx[3 * q] = gx[i];
            qx[3 * q + 1] = gx[j];
            qx[3 * q + 2] = gw[i] * gw[j];
            double p1[ORC_MAX_N], p2[ORC_MAX_N];
            lagrange_eval(np, lx, bw, gx[i], p1);
            lagrange_eval(np, lx, bw, gx[j], p2);
            for (int b = 0; b < np; ++b)
                for (int a = 0; a < np; ++a)
                    Phi[q * nm + a + np * b] = p1[a] * p2[b];
        }
    for (int i = 0; i < nm * nm; ++i)
        M[i] = 0.0;
    for (int q = 0; q < nq; ++q)
        for (int i = 0; i < nm; ++i)
            for (int j = 0; j < nm; ++j)
                M[i * nm + j] += qx[3 * q + 2] * Phi[q * nm + i] * Phi[q * nm + j];
    if (invert(nm, M, Minv)) {
        free(Phi);
        free(qx);
        return -1;
    }
    double *B = (double *)malloc(sizeof(double) * nm * nf);
    /* coarse: B_c(m, ic + np jc) += w Phi(q,m) psi(ic) psi(jc) */
    for (int i = 0; i < nm * nf; ++i)
        B[i] = 0.0;
    for (int q = 0; q < nq; ++q) {
        double p1[ORC_MAX_N], p2[ORC_MAX_N];
        lagrange_eval(np, lx, bw, qx[3 * q], p1);
        lagrange_eval(np, lx, bw, qx[3 * q + 1], p2);
        for (int m = 0; m < nm; ++m)
            for (int jc = 0; jc < np; ++jc)
                for (int ic = 0; ic < np; ++ic)
                    B[m * nf + ic + np * jc] += qx[3 * q + 2] * Phi[q * nm + m] * p1[ic] * p2[jc];
    }
    for (int r = 0; r < nm; ++r)
        for (int c = 0; c < nf; ++c) {
            double s = 0.0;
            for (int k = 0; k < nm; ++k)
                s += Minv[r * nm + k] * B[k * nf + c];
            Pc[r * nf + c] = s;
            Lc[c * nm + r] = B[r * nf + c];
        }
    for (int f = 0; f < nfine; ++f) {
        double bb[4];
        subface_bounds(type, f, bb);
        for (int i = 0; i < nm * nf; ++i)
            B[i] = 0.0;
        for (int q = 0; q < nq; ++q) {
            const double x0 = qx[3 * q], x1 = qx[3 * q + 1];
            if (!(x0 >= bb[0] && x0 <= bb[2] && x1 >= bb[1] && x1 <= bb[3]))
                continue; /* inclusive bounds, mortar.cpp:184-185 */
            const double f0 = 2.0 * (x0 - bb[0]) / (bb[2] - bb[0]) - 1.0;
            const double f1 = 2.0 * (x1 - bb[1]) / (bb[3] - bb[1]) - 1.0;
            double p1[ORC_MAX_N], p2[ORC_MAX_N];
            lagrange_eval(np, lx, bw, f0, p1);
            lagrange_eval(np, lx, bw, f1, p2);
            for (int m = 0; m < nm; ++m)
                for (int jf = 0; jf < np; ++jf)
                    for (int iff = 0; iff < np; ++iff)
                        B[m * nf + iff + np * jf] +=
                            qx[3 * q + 2] * Phi[q * nm + m] * p1[iff] * p2[jf];
        }
        double *P = Pf + (size_t)f * nm * nf, *L = Lf + (size_t)f * nf * nm;
        for (int r = 0; r < nm; ++r)
            for (int c = 0; c < nf; ++c) {
                double s = 0.0;
                for (int k = 0; k < nm; ++k)
                    s += Minv[r * nm + k] * B[k * nf + c];
                P[r * nf + c] = s;
                L[c * nm + r] = B[r * nf + c];
            }
    }
    free(B);
    free(Phi);
    free(qx);
    return nfine;
}

/* MortarSpace::compute_mortar_flux, mortar.cpp:212-258 (scalar flux) */
int orc_mortar_flux(int order, int type, int kind, const double *p, const double *Uc,
                    const double *Uf, const double *n, double *rhs_c, double *rhs_f)
{
    const int np = order + 1, nm = np * np, nf = nm;
    const int nfine = nfine_of(type);
    double *M = (double *)malloc(sizeof(double) * (3 * nm * nm + 2 * nm * nf * (1 + 4)));
    double *Minv = M + nm * nm, *Pc = Minv + nm * nm, *Lc = Pc + nm * nf, *Pf = Lc + nf * nm;
    double *Lf = Pf + 4 * nm * nf;
    if (orc_mortar_operators(order, type, M, Minv, Pc, Lc, Pf, Lf) < 0) {
        free(M);
        return -1;
    }
    double ucm[ORC_MAX_N * ORC_MAX_N], ufm[ORC_MAX_N * ORC_MAX_N], F[ORC_MAX_N * ORC_MAX_N],
        MF[ORC_MAX_N * ORC_MAX_N];
    for (int r = 0; r < nm; ++r) {
        double s = 0.0;
        for (int c = 0; c < nf; ++c)
            s += Pc[r * nf + c] * Uc[c];
        ucm[r] = s;
        ufm[r] = 0.0;
    }
    for (int f = 0; f < nfine; ++f)
        for (int r = 0; r < nm; ++r) {
            double s = 0.0;
            for (int c = 0; c < nf; ++c)
                s += Pf[(size_t)f * nm * nf + r * nf + c] * Uf[f * nf + c];
            ufm[r] += s;
        }
    for (int d = 0; d < nm; ++d)
        F[d] = scalar_flux(kind, ucm[d], ufm[d], n, p);
    for (int r = 0; r < nm; ++r) {
        double s = 0.0;
        for (int c = 0; c < nm; ++c)
            s += Minv[r * nm + c] * F[c];
        MF[r] = s;
    }
    for (int r = 0; r < nf; ++r) {
        double s = 0.0;
        for (int c = 0; c < nm; ++c)
            s += Lc[r * nm + c] * MF[c];
        rhs_c[r] = s;
    }
    for (int f = 0; f < nfine; ++f)
        for (int r = 0; r < nf; ++r) {
            double s = 0.0;
            for (int c = 0; c < nm; ++c)
                s += -Lf[(size_t)f * nf * nm + r * nm + c] * MF[c];
            rhs_f[f * nf + r] = s;
        }
    free(M);
    return nfine;
}

/* verify_mortar_conservation, src/dg/mortar.cpp:386-413: |sum_i wc_i rc_i +
 * sum_f sum_i wf_fi rf_fi| / (|coarse integral| + |fine integral|), 0 when
 * that norm is below 1e-15 */
double orc_verify_mortar_conservation(int nm, int nfine, const double *rc, const double *rf,
                                      const double *wc, const double *wf)
{
    double ic = 0.0, ifn = 0.0;
    for (int i = 0; i < nm; ++i)
        ic += wc[i] * rc[i];
    for (int f = 0; f < nfine; ++f)
        for (int i = 0; i < nm; ++i)
            ifn += wf[f * nm + i] * rf[f * nm + i];
    const double total = ic + ifn, norm = fabs(ic) + fabs(ifn);
    if (norm < 1e-15)
        return 0.0;
    return fabs(total) / norm;
}
