// C ABI of the face components: pointwise numerical fluxes, face-trace
// gathers, conforming interfaces and 2:1 mortars.  Host code; kernels are in
// interface.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "drf_internal.hpp"
#include "interface.hpp"

namespace drf {
namespace {

// LagrangeBasis1D::evaluate (include/dg/basis_hexahedron.hpp:316-348):
// barycentric form, exact nodal value within 1e-15 of a node.
void lagrange_eval(int n, const double *x, const double *bw, double xi, double *phi)
{
    double tmp[kMaxN], sum = 0.0;
    for (int j = 0; j < n; ++j) {
        const double d = xi - x[j];
        if (std::fabs(d) < 1e-15) {
            for (int k = 0; k < n; ++k)
                phi[k] = (k == j) ? 1.0 : 0.0;
            return;
        }
        tmp[j] = bw[j] / d;
        sum += tmp[j];
    }
    for (int j = 0; j < n; ++j)
        phi[j] = tmp[j] / sum;
}

// compute_subface_bounds (src/dg/mortar.cpp:27-78): {t1min, t2min, t1max,
// t2max} of fine face f in the coarse face's reference square.
void subface_bounds(int type, int f, double *bb)
{
    const double lo[2] = {-1.0, 0.0}, hi[2] = {0.0, 1.0};
    double a0 = -1, b0 = -1, a1 = 1, b1 = 1;
    switch (type) {
    case DRF_CONN_FINE_2X1: a0 = lo[f]; a1 = hi[f]; break;
    case DRF_CONN_FINE_1X2: b0 = lo[f]; b1 = hi[f]; break;
    case DRF_CONN_FINE_2X2:
        a0 = lo[f & 1]; a1 = hi[f & 1];
        b0 = lo[f >> 1]; b1 = hi[f >> 1];
        break;
    case DRF_CONN_FINE_3_2PLUS1: // two quarters below, one half above
        if (f < 2) {
            a0 = lo[f]; a1 = hi[f]; b0 = -1; b1 = 0;
        } else {
            b0 = 0; b1 = 1;
        }
        break;
    case DRF_CONN_FINE_3_1PLUS2: // two quarters on the left, one half right
        if (f < 2) {
            a0 = -1; a1 = 0; b0 = lo[f]; b1 = hi[f];
        } else {
            a0 = 0; a1 = 1;
        }
        break;
    default: break;
    }
    bb[0] = a0;
    bb[1] = b0;
    bb[2] = a1;
    bb[3] = b1;
}

// Inverse of a symmetric positive definite matrix by Cholesky, M = L L^T,
// column j of M^-1 from L L^T x = e_j.
bool spd_inverse(int m, const double *A, double *Ai)
{
    std::vector<double> L((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
        double d = A[j * m + j];
        for (int k = 0; k < j; ++k)
            d -= L[j * m + k] * L[j * m + k];
        if (!(d > 0.0))
            return false;
        const double ljj = std::sqrt(d);
        L[j * m + j] = ljj;
        for (int i = j + 1; i < m; ++i) {
            double s = A[i * m + j];
            for (int k = 0; k < j; ++k)
                s -= L[i * m + k] * L[j * m + k];
            L[i * m + j] = s / ljj;
        }
    }
    std::vector<double> y(m);
    for (int c = 0; c < m; ++c) {
        for (int i = 0; i < m; ++i) {
            double s = (i == c) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k)
                s -= L[i * m + k] * y[k];
            y[i] = s / L[i * m + i];
        }
        for (int i = m - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < m; ++k)
                s -= L[k * m + i] * Ai[k * m + c];
            Ai[i * m + c] = s / L[i * m + i];
        }
    }
    return true;
}

} // namespace

// MortarSpace::build_projection_operators (src/dg/mortar.cpp:80-210): mortar
// space = the face's own LGL basis of degree N, (2N+1)^2-point Gauss rule
// on the coarse face, M = sum_q w Phi Phi^T, B_c and B_f the mixed mass
// matrices against the coarse / fine-face bases (fine points: those inside
// the sub-face, bounds inclusive), P = M^-1 B, L = B^T.
int build_mortar_operators(int order, int type, double *M, double *Minv, double *Pc, double *Lc,
                           double *Pf, double *Lf)
{
    const int np = order + 1, nm = np * np;
    const int nfine = mortar_num_fine(type);
    if (order < 1 || np > kMaxN || nfine == 0)
        return -1;
    const int nq1 = 2 * order + 1, nq = nq1 * nq1;
    double gx[2 * kMaxN], gw[2 * kMaxN], lx[kMaxN], lw[kMaxN], bw[kMaxN];
    if (gl_nodes(nq1, gx, gw) || lgl_nodes(np, lx, lw))
        return -1;
    for (int j = 0; j < np; ++j) {
        double p = 1.0;
        for (int k = 0; k < np; ++k)
            if (k != j)
                p *= (lx[j] - lx[k]);
        bw[j] = 1.0 / p;
    }
    // mortar basis at the quadrature points, point q = i + nq1 j
    std::vector<double> Phi((size_t)nq * nm), wq(nq);
    std::vector<double> p1((size_t)nq1 * np);
    for (int i = 0; i < nq1; ++i)
        lagrange_eval(np, lx, bw, gx[i], &p1[(size_t)i * np]);
    for (int j = 0; j < nq1; ++j)
        for (int i = 0; i < nq1; ++i) {
            const int q = i + nq1 * j;
            wq[q] = gw[i] * gw[j];
            for (int b = 0; b < np; ++b)
                for (int a = 0; a < np; ++a)
                    Phi[(size_t)q * nm + a + np * b] = p1[(size_t)i * np + a] * p1[(size_t)j * np + b];
        }
    for (int i = 0; i < nm * nm; ++i)
        M[i] = 0.0;
    for (int q = 0; q < nq; ++q)
        for (int i = 0; i < nm; ++i)
            for (int j = 0; j < nm; ++j)
                M[i * nm + j] += wq[q] * Phi[(size_t)q * nm + i] * Phi[(size_t)q * nm + j];
    if (!spd_inverse(nm, M, Minv))
        return -1;

    std::vector<double> B((size_t)nm * nm);
    // bb == nullptr: the coarse face itself, basis evaluated at the points
    auto project = [&](const double *bb, double *P, double *L) {
        std::fill(B.begin(), B.end(), 0.0);
        for (int j = 0; j < nq1; ++j)
            for (int i = 0; i < nq1; ++i) {
                const double x0 = gx[i], x1 = gx[j];
                double f0 = x0, f1 = x1;
                if (bb) {
                    if (!(x0 >= bb[0] && x0 <= bb[2] && x1 >= bb[1] && x1 <= bb[3]))
                        continue;
                    f0 = 2.0 * (x0 - bb[0]) / (bb[2] - bb[0]) - 1.0;
                    f1 = 2.0 * (x1 - bb[1]) / (bb[3] - bb[1]) - 1.0;
                }
                double s1[kMaxN], s2[kMaxN];
                lagrange_eval(np, lx, bw, f0, s1);
                lagrange_eval(np, lx, bw, f1, s2);
                const int q = i + nq1 * j;
                for (int m = 0; m < nm; ++m)
                    for (int jf = 0; jf < np; ++jf)
                        for (int iff = 0; iff < np; ++iff)
                            B[(size_t)m * nm + iff + np * jf] +=
                                wq[q] * Phi[(size_t)q * nm + m] * s1[iff] * s2[jf];
            }
        for (int r = 0; r < nm; ++r)
            for (int c = 0; c < nm; ++c) {
                double s = 0.0;
                for (int k = 0; k < nm; ++k)
                    s += Minv[r * nm + k] * B[(size_t)k * nm + c];
                P[r * nm + c] = s;
                L[c * nm + r] = B[(size_t)r * nm + c];
            }
    };
    project(nullptr, Pc, Lc);
    for (int f = 0; f < nfine; ++f) {
        double bb[4];
        subface_bounds(type, f, bb);
        project(bb, Pf + (size_t)f * nm * nm, Lf + (size_t)f * nm * nm);
    }
    return nfine;
}

} // namespace drf

using namespace drf;

struct drf_mortar {
    int order, type, nm, nfine, device;
    std::vector<double> host; // M | Minv | Pc | Lc | Pf[nfine] | Lf[nfine]
    double *d_ops;            // Minv | Pc | Lc | Pf[nfine] | Lf[nfine]
};

namespace {

#define DRF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond))                                                                         \
            return ::drf::fail((code), (msg));                                               \
    } while (0)

bool flux_params(int kind, const double *params, FluxParams &p)
{
    std::memset(&p, 0, sizeof(p));
    int np = 0;
    switch (kind) {
    case DRF_FLUX_SW_LF:
    case DRF_FLUX_SW_HLLC:
    case DRF_FLUX_SW_CENTRAL:
    case DRF_FLUX_PE_LF5: np = 1; break;
    case DRF_FLUX_SW_ROE: np = 2; break;
    case DRF_FLUX_UPWIND: np = 3; break;
    case DRF_FLUX_LF_ADV: np = 4; break;
    default: return false;
    }
    if (!params)
        return false;
    for (int i = 0; i < np; ++i)
        p.v[i] = params[i];
    return true;
}

bool scalar_kind(int kind) { return kind == DRF_FLUX_UPWIND || kind == DRF_FLUX_LF_ADV; }

} // namespace

extern "C" {

drf_status drf_flux_points(int kind, int64_t npts, const void *UL, const void *UR,
                           const void *normals, const double *params, void *out, void *stream)
{
    FluxParams p;
    DRF_REQUIRE(flux_params(kind, params, p), DRF_ERR_ARG,
                "drf_flux_points: unknown flux kind or missing params");
    DRF_REQUIRE(npts >= 0 && npts < ((int64_t)1 << 39), DRF_ERR_ARG,
                "drf_flux_points: npts out of range");
    if (npts == 0)
        return DRF_OK;
    DRF_REQUIRE(UL && UR && normals && out, DRF_ERR_ARG, "drf_flux_points: null pointer");
    DRF_HIP_CHECK(launch_flux_points(kind, npts, (const double *)UL, (const double *)UR,
                                     (const double *)normals, p, (double *)out,
                                     (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_face_map(int order, int face, int32_t *vol_index)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_face_map: order out of range");
    DRF_REQUIRE(face >= 0 && face < 6, DRF_ERR_RANGE, "drf_face_map: face must be 0..5");
    DRF_REQUIRE(vol_index, DRF_ERR_ARG, "drf_face_map: null output");
    const int n = order + 1;
    for (int jt = 0; jt < n; ++jt)
        for (int it = 0; it < n; ++it)
            vol_index[it + n * jt] = face_node(n, face, it, jt);
    return DRF_OK;
}

drf_status drf_gather_face_traces(int order, int64_t n_elem, int64_t elem_stride, int64_t nfaces,
                                  const void *elems, const void *faces, const void *field,
                                  void *out, void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_gather_face_traces: order out of range");
    const int n = order + 1;
    DRF_REQUIRE(n_elem >= 0 && nfaces >= 0 && elem_stride >= (int64_t)n * n * n, DRF_ERR_ARG,
                "drf_gather_face_traces: bad sizes (elem_stride must be >= (N+1)^3)");
    DRF_REQUIRE(nfaces * n * n < ((int64_t)1 << 39), DRF_ERR_ARG,
                "drf_gather_face_traces: too many faces");
    if (nfaces == 0)
        return DRF_OK;
    DRF_REQUIRE(elems && faces && field && out, DRF_ERR_ARG,
                "drf_gather_face_traces: null pointer");
    DRF_HIP_CHECK(launch_gather_faces(n, n_elem, elem_stride, nfaces, (const int64_t *)elems,
                                      (const int8_t *)faces, (const double *)field,
                                      (double *)out, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_conforming_flux_batch(int order, int kind, const double *params, int64_t nfaces,
                                     const void *UL, const void *UR, const void *normals,
                                     void *rhs_l, void *rhs_r, void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_conforming_flux_batch: order out of range");
    DRF_REQUIRE(scalar_kind(kind), DRF_ERR_ARG,
                "drf_conforming_flux_batch: kind must be DRF_FLUX_UPWIND or DRF_FLUX_LF_ADV");
    FluxParams p;
    DRF_REQUIRE(flux_params(kind, params, p), DRF_ERR_ARG,
                "drf_conforming_flux_batch: missing params");
    DRF_REQUIRE(nfaces >= 0 && nfaces < ((int64_t)1 << 31), DRF_ERR_ARG,
                "drf_conforming_flux_batch: nfaces out of range");
    if (nfaces == 0)
        return DRF_OK;
    DRF_REQUIRE(UL && UR && normals && rhs_l && rhs_r, DRF_ERR_ARG,
                "drf_conforming_flux_batch: null pointer");
    FaceRule rule;
    std::memset(&rule, 0, sizeof(rule));
    rule.nq1 = order < 2 ? 2 : order; // GaussQuadrature1D(order, LGL)
    double x[kMaxN];
    DRF_REQUIRE(lgl_nodes(rule.nq1, x, rule.w) == 0, DRF_ERR_UNSUPPORTED,
                "drf_conforming_flux_batch: LGL rule failed");
    const int nf = (order + 1) * (order + 1);
    DRF_HIP_CHECK(launch_conforming(kind, p, nf, rule, nfaces, (const double *)UL,
                                    (const double *)UR, (const double *)normals, (double *)rhs_l,
                                    (double *)rhs_r, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_mortar_create(int order, int type, int device, drf_mortar **out)
{
    DRF_REQUIRE(out, DRF_ERR_ARG, "drf_mortar_create: null output");
    *out = nullptr;
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_mortar_create: order out of range");
    const int nfine = mortar_num_fine(type);
    DRF_REQUIRE(nfine > 0, DRF_ERR_ARG, "drf_mortar_create: type must be a face connection type");
    const int nm = (order + 1) * (order + 1);
    const size_t mm = (size_t)nm * nm;
    std::vector<double> host((4 + 2 * (size_t)nfine) * mm);
    double *M = host.data(), *Minv = M + mm, *Pc = Minv + mm, *Lc = Pc + mm, *Pf = Lc + mm,
           *Lf = Pf + nfine * mm;
    DRF_REQUIRE(build_mortar_operators(order, type, M, Minv, Pc, Lc, Pf, Lf) == nfine,
                DRF_ERR_ARG, "drf_mortar_create: mortar mass matrix is singular");
    double *d = nullptr;
    if (device < 0) { // operators on the host only
        *out = new drf_mortar{order, type, nm, nfine, device, std::move(host), nullptr};
        return DRF_OK;
    }
    DRF_HIP_CHECK(hipSetDevice(device));
    const size_t dev_bytes = sizeof(double) * (3 + 2 * (size_t)nfine) * mm;
    DRF_HIP_CHECK(hipMalloc(&d, dev_bytes));
    hipError_t e = hipMemcpy(d, Minv, dev_bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(DRF_ERR_HIP, std::string("drf_mortar_create: ") + hipGetErrorString(e));
    }
    *out = new drf_mortar{order, type, nm, nfine, device, std::move(host), d};
    return DRF_OK;
}

void drf_mortar_destroy(drf_mortar *m)
{
    if (!m)
        return;
    if (m->d_ops) {
        (void)hipSetDevice(m->device);
        (void)hipFree(m->d_ops);
    }
    delete m;
}

int drf_mortar_num_fine(const drf_mortar *m) { return m ? m->nfine : 0; }
int drf_mortar_num_dofs(const drf_mortar *m) { return m ? m->nm : 0; }

drf_status drf_mortar_operators(const drf_mortar *m, double *M, double *Minv, double *Pc,
                                double *Lc, double *Pf, double *Lf)
{
    DRF_REQUIRE(m, DRF_ERR_ARG, "drf_mortar_operators: null mortar");
    const size_t mm = (size_t)m->nm * m->nm, nf = (size_t)m->nfine;
    const double *h = m->host.data();
    double *dst[6] = {M, Minv, Pc, Lc, Pf, Lf};
    const size_t off[6] = {0, mm, 2 * mm, 3 * mm, 4 * mm, (4 + nf) * mm};
    const size_t len[6] = {mm, mm, mm, mm, nf * mm, nf * mm};
    for (int i = 0; i < 6; ++i)
        if (dst[i])
            std::memcpy(dst[i], h + off[i], sizeof(double) * len[i]);
    return DRF_OK;
}

drf_status drf_mortar_flux_batch(drf_mortar *m, int kind, const double *params,
                                 int64_t n_interfaces, const void *Uc, const void *Uf,
                                 const void *normals, void *rhs_c, void *rhs_f, void *stream)
{
    DRF_REQUIRE(m, DRF_ERR_ARG, "drf_mortar_flux_batch: null mortar");
    DRF_REQUIRE(scalar_kind(kind), DRF_ERR_ARG,
                "drf_mortar_flux_batch: kind must be DRF_FLUX_UPWIND or DRF_FLUX_LF_ADV");
    FluxParams p;
    DRF_REQUIRE(flux_params(kind, params, p), DRF_ERR_ARG, "drf_mortar_flux_batch: missing params");
    DRF_REQUIRE(n_interfaces >= 0, DRF_ERR_ARG, "drf_mortar_flux_batch: negative count");
    if (n_interfaces == 0)
        return DRF_OK;
    DRF_REQUIRE(Uc && Uf && normals && rhs_c && rhs_f, DRF_ERR_ARG,
                "drf_mortar_flux_batch: null pointer");
    DRF_REQUIRE(m->d_ops, DRF_ERR_STATE,
                "drf_mortar_flux_batch: mortar was created without a device (device < 0)");
    DRF_HIP_CHECK(hipSetDevice(m->device));
    DRF_HIP_CHECK(launch_mortar(kind, p, m->nm, m->nfine, m->d_ops, n_interfaces,
                                (const double *)Uc, (const double *)Uf, (const double *)normals,
                                (double *)rhs_c, (double *)rhs_f, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_verify_mortar_conservation(const drf_mortar *m, const double *rhs_coarse,
                                          const double *rhs_fine, const double *w_coarse,
                                          const double *w_fine, double *rel_error)
{
    DRF_REQUIRE(m && rhs_coarse && rhs_fine && w_coarse && w_fine && rel_error, DRF_ERR_ARG,
                "drf_verify_mortar_conservation: null argument");
    double ic = 0.0, ifn = 0.0;
    for (int i = 0; i < m->nm; ++i)
        ic += w_coarse[i] * rhs_coarse[i];
    for (int f = 0; f < m->nfine; ++f)
        for (int i = 0; i < m->nm; ++i)
            ifn += w_fine[f * m->nm + i] * rhs_fine[f * m->nm + i];
    const double total = ic + ifn, norm = std::fabs(ic) + std::fabs(ifn);
    *rel_error = (norm < 1e-15) ? 0.0 : std::fabs(total) / norm;
    return DRF_OK;
}

drf_status drf_mortar_conservation_dev(const drf_mortar *m, int64_t n_interfaces,
                                       const void *rhs_c, const void *rhs_f, const void *w_coarse,
                                       const void *w_fine, void *rel_error, void *stream)
{
    DRF_REQUIRE(m, DRF_ERR_ARG, "drf_mortar_conservation_dev: null mortar");
    DRF_REQUIRE(n_interfaces >= 0, DRF_ERR_ARG, "drf_mortar_conservation_dev: negative count");
    if (n_interfaces == 0)
        return DRF_OK;
    DRF_REQUIRE(rhs_c && rhs_f && w_coarse && w_fine && rel_error, DRF_ERR_ARG,
                "drf_mortar_conservation_dev: null pointer");
    DRF_HIP_CHECK(launch_mortar_conservation(m->nm, m->nfine, n_interfaces, (const double *)rhs_c,
                                             (const double *)rhs_f, (const double *)w_coarse,
                                             (const double *)w_fine, (double *)rel_error,
                                             (hipStream_t)stream));
    return DRF_OK;
}

} // extern "C"
