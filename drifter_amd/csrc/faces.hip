// Face term of the DG right-hand side on an adaptive octree: conforming
// interfaces and 2:1 mortars, each fused from the state's face traces to the
// tendencies (gather -> numerical flux -> lift -> inverse mass matrix).
//
// Replaces the face loop of DG3DIntegration::compute_rhs
// (src/dg/operators_3d.cpp:255-275) with compute_interface_flux_conforming
// (:320-343, through ConformingInterface::compute_flux, src/dg/mortar.cpp:
// 320-343) and compute_interface_flux_nonconforming (:345-391, through
// MortarSpace::compute_mortar_flux, src/dg/mortar.cpp:212-258), and the
// inverse mass matrix of :277-281 (DG3DElementOperator::mass_inv :52-55, the
// diagonal 1/(w_i w_j w_k) of src/dg/basis_hexahedron.cpp:393-420) applied to
// what the faces add.
//
// Every launch covers the interfaces of one normal axis and one kind.  An
// element's two faces of one axis hold disjoint nodes (index 0 and N along
// the axis), and every (element, face) belongs to at most one interface (the
// plan checks it), so within a launch no two lanes update the same tendency:
// the lifts are plain read-add-writes.  Launches of different axes are
// ordered by the stream.
#include "drf_internal.hpp"
#include "interface.hpp"

// No FMA contraction: the fluxes, operator products and lifts round as the
// reference's sequence of operations.
#pragma clang fp contract(off)

namespace drf {
namespace {

template <class T> __device__ __forceinline__ int64_t node_off(const FaceArgs<T> &a, int n,
                                                               int64_t e, int d)
{
    const int i = d % n, L = d / n;
    return elem_offset(e, a.es, a.bs, a.bb) + (int64_t)L * a.ls + (int64_t)i * a.is;
}

// 1 / (w_i w_j w_k) of volume node d, as build_mass_matrices forms it
template <class T> __device__ __forceinline__ T inv_mass(const FaceArgs<T> &a, int n, int d)
{
    const int i = d % n, j = (d / n) % n, k = d / (n * n);
    const T w = a.w[i] * a.w[j] * a.w[k];
    return T(1) / w;
}

// ConformingInterface on the interfaces list[4 c .. 4 c + 3] = (left element,
// left face, right element, right face).  LIFT: one lane per (interface,
// quadrature point) -- the face rule has nq1 = N points per direction
// (FaceQuadrature(face, order), quadrature_3d.cpp:13-31), so only face DOFs
// q < nq1^2 carry a flux and the others add nothing -- looping over the five
// fields.  Otherwise one lane per (interface, face DOF), writing rhs_l and
// rhs_r as ConformingInterface::compute_flux returns them, [c][field][nh].
template <int n, class T, bool LIFT>
__global__ void __launch_bounds__(256)
    conf_face_kernel(int64_t nif, const int64_t *__restrict__ list, FaceArgs<T> a,
                     T *__restrict__ out_l, T *__restrict__ out_r)
{
    constexpr int nh = n * n;
    const int nq1 = a.nq1, nq = nq1 * nq1;
    const int per = LIFT ? nq : nh;
    const int64_t total = nif * per;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = t / per;
        const int q = (int)(t - c * per);
        const int64_t el = list[4 * c], er = list[4 * c + 2];
        const int fl = (int)list[4 * c + 1], fr = (int)list[4 * c + 3];
        const int dl = face_node(n, fl, q % n, q / n), dr = face_node(n, fr, q % n, q / n);
        const int64_t ol = node_off(a, n, el, dl), orr = node_off(a, n, er, dr);
        T nrm[3] = {T(0), T(0), T(0)};
        nrm[fl >> 1] = (fl & 1) ? T(1) : T(-1);
        const bool on = q < nq;
        const T wq = on ? a.rw[q % nq1] * a.rw[q / nq1] : T(0);
        const T ml = inv_mass(a, n, dl), mr = inv_mass(a, n, dr);
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            T rl = T(0), rr = T(0);
            if (on) {
                const T F = scalar_flux_t<T>(a.kind, a.U[f * a.fs + ol], a.U[f * a.fs + orr], nrm,
                                             a.p);
                rl = T(0) + wq * F;
                rr = T(0) - wq * F;
            }
            if (LIFT) {
                a.dU[f * a.fs + ol] += ml * rl;
                a.dU[f * a.fs + orr] += mr * rr;
            } else {
                out_l[(c * 5 + f) * nh + q] = rl;
                out_r[(c * 5 + f) * nh + q] = rr;
            }
        }
    }
}

// The conforming part of the face term, element by element: every interface
// seen from each of its two elements, so that a thread adds to its own
// element's nodes only.  eface[6 e + f] = (neighbour << 4) | (neighbour's face
// << 1) | side for a conforming face (side 0: this element is the
// interface's left, whose face gives the normal), -1 otherwise; a conforming
// neighbour always meets the face with its opposite face f ^ 1.  One thread
// per (element, field, xi-line): it reads and rewrites its line of U and dU
// (contiguous across the workgroup in the LINES layout, each element's block
// once) and reads the neighbours' nodes on the shared faces.
// Per node the contributions of the x, y and z faces are summed (in that
// order, the order of the interface lists) before M^-1 is applied
// (operators_3d.cpp:277-281).  Each contribution is evaluated with the
// interface's own left/right order, as the interface kernel adds it.
// Addresses: a wave-uniform base per (field, xi-plane) and a 32-bit per-lane
// byte offset (element and line), which the host checks fits.
struct FaceNb {
    uint32_t off; // byte offset of the neighbour's line, plane-relative
    int side;
    bool on;
};

// r = (side 0) 0 + wq F(u_own, u_nb, n_f) or (side 1) 0 - wq F(u_nb, u_own, n_(f^1))
template <class T>
__device__ __forceinline__ T face_contrib(const FaceArgs<T> &a, int f, int side, T own, T other,
                                          T wq)
{
    const int fl = side ? (f ^ 1) : f;
    T nrm[3] = {T(0), T(0), T(0)};
    nrm[fl >> 1] = (fl & 1) ? T(1) : T(-1);
    if (side == 0)
        return T(0) + wq * scalar_flux_t<T>(a.kind, own, other, nrm, a.p);
    return T(0) - wq * scalar_flux_t<T>(a.kind, other, own, nrm, a.p);
}

template <class T> __device__ __forceinline__ T ldb(const T *base, uint32_t off)
{
    return *(const T *)((const char *)base + off);
}

template <int n, class T>
__global__ void __launch_bounds__(256)
    face_elem_kernel(int64_t E, const int64_t *__restrict__ eface, FaceArgs<T> a)
{
    // one thread per (element, field, xi-line): every load of a thread is
    // issued before its arithmetic, and a workgroup's lanes for one field
    // and xi-plane read one contiguous run
    constexpr int n2 = n * n, n3 = n2 * n, PER = 5 * n2, EPB = 256 / PER, NT = EPB * PER;
    // face-rule weights by face DOF (0 past the N-point rule) and M^-1 by node
    __shared__ T sWq[n2], sMinv[n3];
    const int tid = threadIdx.x;
    const int nq1 = a.nq1;
    for (int q = tid; q < n2; q += blockDim.x)
        sWq[q] = (q < nq1 * nq1) ? a.rw[q % nq1] * a.rw[q / nq1] : T(0);
    for (int d = tid; d < n3; d += blockDim.x)
        sMinv[d] = inv_mass(a, n, d);
    __syncthreads();
    if (tid >= NT)
        return;
    const int le = tid / PER, r = tid - le * PER, fld = r / n2, L = r - fld * n2;
    const int j = L % n, k = L / n;
    const int nq = nq1 * nq1;
    const int64_t nbatch = (E + EPB - 1) / EPB;
    const T *__restrict__ Uf = a.U + fld * a.fs;
    T *__restrict__ dUf = a.dU + fld * a.fs;
    auto line_off = [&](int64_t el, int line) {
        return (uint32_t)((elem_offset(el, a.es, a.bs, a.bb) + (int64_t)line * a.ls) *
                          (int64_t)sizeof(T));
    };
    auto entry = [&](int64_t v, int line) {
        FaceNb r;
        r.on = v >= 0;
        r.side = (int)(v & 1);
        r.off = r.on ? line_off(v >> 4, line) : 0u;
        return r;
    };
    // the faces this line lies on: both x-faces (its end nodes), a y-face
    // when j is 0 or N, a z-face when k is 0 or N (every node of the line);
    // the neighbour's line across a y-face has j = N - j, across a z-face
    // k = N - k, across an x-face it is the same line
    const int fy = (j == 0) ? 2 : (j == n - 1) ? 3 : -1;
    const int fz = (k == 0) ? 4 : (k == n - 1) ? 5 : -1;
    const int Ly = (n - 1 - j) + n * k, Lz = j + n * (n - 1 - k);
    // XCD-aware order: workgroup w runs on XCD w mod 8, and XCD x takes the
    // x-th eighth of the batches in order, so that neighbouring elements (a
    // few Morton positions apart) are read behind one L2 (grid: a multiple
    // of 8 workgroups)
    const int64_t G = gridDim.x;
    const int64_t lw = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8
                                    : (int64_t)blockIdx.x;
    for (int64_t b = lw; b < nbatch; b += G) {
        const int64_t e = b * EPB + le;
        if (e >= E)
            continue;
        const FaceNb x0 = entry(eface[6 * e], L), x1 = entry(eface[6 * e + 1], L);
        const FaceNb y = entry(fy >= 0 ? eface[6 * e + fy] : -1, Ly);
        const FaceNb z = entry(fz >= 0 ? eface[6 * e + fz] : -1, Lz);
        const uint32_t own = line_off(e, L);
        T u[n], du[n], ux0 = T(0), ux1 = T(0), uy[n], uz[n];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            u[i] = ldb(Uf + i * a.is, own);
            du[i] = ldb(dUf + i * a.is, own);
            uy[i] = y.on ? ldb(Uf + i * a.is, y.off) : T(0);
            uz[i] = z.on ? ldb(Uf + i * a.is, z.off) : T(0);
        }
        if (x0.on)
            ux0 = ldb(Uf + (n - 1) * a.is, x0.off);
        if (x1.on)
            ux1 = ldb(Uf, x1.off);
        T acc[n];
#pragma unroll
        for (int i = 0; i < n; ++i)
            acc[i] = T(0);
        // x-faces: node i = 0 (face 0) or N (face 1), face DOF (j, k) = L;
        // the neighbour's node is at the other end of the same line
        if (L < nq) {
            const T wq = sWq[L];
            if (x0.on)
                acc[0] += face_contrib<T>(a, 0, x0.side, u[0], ux0, wq);
            if (x1.on)
                acc[n - 1] += face_contrib<T>(a, 1, x1.side, u[n - 1], ux1, wq);
        }
        // y-face: face DOF (i, k); z-face: face DOF (i, j); the same plane i
        if (y.on) {
#pragma unroll
            for (int i = 0; i < n; ++i)
                if (i + n * k < nq)
                    acc[i] += face_contrib<T>(a, fy, y.side, u[i], uy[i], sWq[i + n * k]);
        }
        if (z.on) {
#pragma unroll
            for (int i = 0; i < n; ++i)
                if (i + n * j < nq)
                    acc[i] += face_contrib<T>(a, fz, z.side, u[i], uz[i], sWq[i + n * j]);
        }
#pragma unroll
        for (int i = 0; i < n; ++i)
            *(T *)((char *)(dUf + i * a.is) + own) = du[i] + sMinv[i + n * L] * acc[i];
    }
}

// MortarSpace::compute_mortar_flux on the interfaces list[6 m .. 6 m + 5] =
// (coarse element, coarse face, fine elements of sub-faces 0..NF-1), the fine
// faces being the coarse face's opposite.  ops = [Minv | Pc | Lc | Pf[NF] |
// Lf[NF]] (nm x nm each) of the group's connection type, staged in LDS once
// per persistent workgroup.  One lane per (interface, field, mortar DOF).
// LIFT adds M^-1 I^T of rhs_c and rhs_f to the tendencies; otherwise rhs_c
// is written [m][field][nm] and rhs_f [m][field][NF][nm].
template <int n, int NF, class T, bool LIFT>
__global__ void __launch_bounds__(256)
    mortar_face_kernel(int64_t nif, const int64_t *__restrict__ list,
                       const T *__restrict__ ops, FaceArgs<T> a, T *__restrict__ out_c,
                       T *__restrict__ out_f)
{
    constexpr int nm = n * n, MM = nm * nm, NMAT = 3 + 2 * NF;
    constexpr int ipb = 256 / nm;                  // (interface, field) units per pass
    constexpr int per = (1 + NF) * nm + 2 * nm;    // Uc | Uf[NF] | F | M^-1 F
    __shared__ T sops[NMAT * MM];
    __shared__ T sbuf[ipb * per];
    for (int i = threadIdx.x; i < NMAT * MM; i += blockDim.x)
        sops[i] = ops[i];
    const T *Minv = sops, *Pc = sops + MM, *Lc = sops + 2 * MM, *Pf = sops + 3 * MM,
            *Lf = sops + (3 + NF) * MM;
    const int g = threadIdx.x / nm, r = threadIdx.x - g * nm;
    const bool lane_on = g < ipb;
    T *su = sbuf + (lane_on ? g : 0) * per;
    T *sF = su + (1 + NF) * nm;
    const int it = r % n, jt = r / n;
    const int64_t units = nif * 5;
    for (int64_t u0 = (int64_t)blockIdx.x * ipb; u0 < units; u0 += (int64_t)gridDim.x * ipb) {
        const int64_t u = u0 + g;
        const bool act = lane_on && u < units;
        const int64_t m = act ? u / 5 : 0;
        const int f = act ? (int)(u - 5 * m) : 0;
        const int fc = act ? (int)list[6 * m + 1] : 0, ff = fc ^ 1;
        const int dc = face_node(n, fc, it, jt), df = face_node(n, ff, it, jt);
        const int64_t ec = act ? list[6 * m] : 0;
        int64_t oc = 0, of[NF];
        __syncthreads(); // operators staged; the previous units are done with sbuf
        if (act) {
            oc = f * a.fs + node_off(a, n, ec, dc);
            su[r] = a.U[oc];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                of[k] = f * a.fs + node_off(a, n, list[6 * m + 2 + k], df);
                su[nm + k * nm + r] = a.U[of[k]];
            }
        }
        __syncthreads();
        if (act) {
            T nrm[3] = {T(0), T(0), T(0)};
            nrm[fc >> 1] = (fc & 1) ? T(1) : T(-1);
            T uc = T(0);
            for (int c = 0; c < nm; ++c)
                uc += Pc[r * nm + c] * su[c];
            T uf = T(0);
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                T s = T(0);
                for (int c = 0; c < nm; ++c)
                    s += Pf[k * MM + r * nm + c] * su[nm + k * nm + c];
                uf += s;
            }
            sF[r] = scalar_flux_t<T>(a.kind, uc, uf, nrm, a.p);
        }
        __syncthreads();
        if (act) {
            T mf = T(0);
            for (int c = 0; c < nm; ++c)
                mf += Minv[r * nm + c] * sF[c];
            sF[nm + r] = mf;
        }
        __syncthreads();
        if (act) {
            T rc = T(0);
            for (int c = 0; c < nm; ++c)
                rc += Lc[r * nm + c] * sF[nm + c];
            if (LIFT)
                a.dU[oc] += inv_mass(a, n, dc) * rc;
            else
                out_c[(m * 5 + f) * nm + r] = rc;
            const T mfi = inv_mass(a, n, df);
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                T v = T(0);
                for (int c = 0; c < nm; ++c)
                    v += -Lf[k * MM + r * nm + c] * sF[nm + c];
                if (LIFT)
                    a.dU[of[k]] += mfi * v;
                else
                    out_f[((m * 5 + f) * NF + k) * nm + r] = v;
            }
        }
    }
}

unsigned grid_for(int64_t work, int per_block, unsigned cap)
{
    const int64_t b = (work + per_block - 1) / per_block;
    return (unsigned)(b < (int64_t)cap ? (b < 1 ? 1 : b) : cap);
}

template <int n, class T>
hipError_t conf_n(bool lift, int64_t nif, const int64_t *list, const FaceArgs<T> &a, T *ol, T *orr,
                  hipStream_t s)
{
    const int per = lift ? a.nq1 * a.nq1 : n * n;
    const dim3 g(grid_for(nif * per, 256, 8192)), b(256);
    if (lift)
        hipLaunchKernelGGL((conf_face_kernel<n, T, true>), g, b, 0, s, nif, list, a, ol, orr);
    else
        hipLaunchKernelGGL((conf_face_kernel<n, T, false>), g, b, 0, s, nif, list, a, ol, orr);
    return hipGetLastError();
}

template <int n, class T>
hipError_t elem_n(int64_t E, const int64_t *eface, const FaceArgs<T> &a, hipStream_t s)
{
    constexpr int EPB = 256 / (5 * n * n);
    const unsigned g8 = (grid_for(E, EPB, 1u << 20) + 7u) & ~7u; // a multiple of 8
    const dim3 g(g8), b(256);
    hipLaunchKernelGGL((face_elem_kernel<n, T>), g, b, 0, s, E, eface, a);
    return hipGetLastError();
}

template <int n, int NF, class T>
hipError_t mortar_nf(bool lift, int64_t nif, const int64_t *list, const T *ops,
                     const FaceArgs<T> &a, T *oc, T *of, hipStream_t s)
{
    constexpr int ipb = 256 / (n * n);
    const dim3 g(grid_for(nif * 5, ipb, 2048)), b(256);
    if (lift)
        hipLaunchKernelGGL((mortar_face_kernel<n, NF, T, true>), g, b, 0, s, nif, list, ops, a, oc,
                           of);
    else
        hipLaunchKernelGGL((mortar_face_kernel<n, NF, T, false>), g, b, 0, s, nif, list, ops, a,
                           oc, of);
    return hipGetLastError();
}

template <int n, class T>
hipError_t mortar_n(int nfine, bool lift, int64_t nif, const int64_t *list, const T *ops,
                    const FaceArgs<T> &a, T *oc, T *of, hipStream_t s)
{
    switch (nfine) {
    case 2: return mortar_nf<n, 2, T>(lift, nif, list, ops, a, oc, of, s);
    case 4: return mortar_nf<n, 4, T>(lift, nif, list, ops, a, oc, of, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace

template <class T>
hipError_t launch_face_conforming(int n, bool lift, int64_t nif, const int64_t *list,
                                  const FaceArgs<T> &a, T *out_l, T *out_r, hipStream_t s)
{
    if (nif <= 0)
        return hipSuccess;
    switch (n) {
    case 2: return conf_n<2, T>(lift, nif, list, a, out_l, out_r, s);
    case 3: return conf_n<3, T>(lift, nif, list, a, out_l, out_r, s);
    case 4: return conf_n<4, T>(lift, nif, list, a, out_l, out_r, s);
    case 5: return conf_n<5, T>(lift, nif, list, a, out_l, out_r, s);
    default: return hipErrorInvalidValue;
    }
}

template <class T>
hipError_t launch_face_elements(int n, int64_t E, const int64_t *eface, const FaceArgs<T> &a,
                                hipStream_t s)
{
    if (E <= 0)
        return hipSuccess;
    switch (n) {
    case 2: return elem_n<2, T>(E, eface, a, s);
    case 3: return elem_n<3, T>(E, eface, a, s);
    case 4: return elem_n<4, T>(E, eface, a, s);
    case 5: return elem_n<5, T>(E, eface, a, s);
    default: return hipErrorInvalidValue;
    }
}

template <class T>
hipError_t launch_face_mortar(int n, int nfine, bool lift, int64_t nif, const int64_t *list,
                              const T *ops, const FaceArgs<T> &a, T *out_c, T *out_f,
                              hipStream_t s)
{
    if (nif <= 0)
        return hipSuccess;
    switch (n) {
    case 2: return mortar_n<2, T>(nfine, lift, nif, list, ops, a, out_c, out_f, s);
    case 3: return mortar_n<3, T>(nfine, lift, nif, list, ops, a, out_c, out_f, s);
    case 4: return mortar_n<4, T>(nfine, lift, nif, list, ops, a, out_c, out_f, s);
    case 5: return mortar_n<5, T>(nfine, lift, nif, list, ops, a, out_c, out_f, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_face_conforming<double>(int, bool, int64_t, const int64_t *,
                                                   const FaceArgs<double> &, double *, double *,
                                                   hipStream_t);
template hipError_t launch_face_conforming<float>(int, bool, int64_t, const int64_t *,
                                                  const FaceArgs<float> &, float *, float *,
                                                  hipStream_t);
template hipError_t launch_face_elements<double>(int, int64_t, const int64_t *,
                                                 const FaceArgs<double> &, hipStream_t);
template hipError_t launch_face_elements<float>(int, int64_t, const int64_t *,
                                                const FaceArgs<float> &, hipStream_t);
template hipError_t launch_face_mortar<double>(int, int, bool, int64_t, const int64_t *,
                                               const double *, const FaceArgs<double> &, double *,
                                               double *, hipStream_t);
template hipError_t launch_face_mortar<float>(int, int, bool, int64_t, const int64_t *,
                                              const float *, const FaceArgs<float> &, float *,
                                              float *, hipStream_t);

} // namespace drf
