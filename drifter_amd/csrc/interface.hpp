// Face fluxes, conforming interfaces, mortars and face traces: declarations
// shared by interface.hip (kernels) and interface_api.cpp (host + C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "drf_internal.hpp"
#include "drifter_hip/drf.h"

namespace drf {

constexpr int kMaxFaceQ = 9; // 1D points of the conforming face rule (<= N + 1)

// Flux parameters, by value in kernel arguments (meaning per kind in drf.h).
struct FluxParams {
    double v[4];
};

// 1D LGL weights of the conforming face rule, by value.
struct FaceRule {
    double w[kMaxFaceQ];
    int nq1;
};

__host__ __device__ inline int flux_num_vars(int kind)
{
    switch (kind) {
    case DRF_FLUX_SW_LF:
    case DRF_FLUX_SW_HLLC:
    case DRF_FLUX_SW_ROE:
    case DRF_FLUX_SW_CENTRAL: return 3;
    case DRF_FLUX_PE_LF5: return 5;
    case DRF_FLUX_UPWIND:
    case DRF_FLUX_LF_ADV: return 1;
    default: return 0;
    }
}

// Scalar numerical flux carried through ConformingInterface / MortarSpace
// (U_L(1), U_R(1): src/dg/mortar.cpp:236-242, 330-336): UpwindFlux
// (include/flux/numerical_flux.hpp:92-107) or the LaxFriedrichsFlux of the
// advective flux F(U) = v U (:57-75).  v = {vx, vy, vz, lambda}.  No FMA
// contraction, so every product and sum rounds as the reference's sequence.
template <class T>
__host__ __device__ inline T scalar_flux_t(int kind, T a, T b, const T *n, const T *v)
{
#pragma clang fp contract(off)
    const T vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    if (kind == DRF_FLUX_UPWIND)
        return vn * ((vn >= T(0)) ? a : b);
    T FnL = T(0), FnR = T(0);
    for (int d = 0; d < 3; ++d) {
        FnL += (v[d] * a) * n[d];
        FnR += (v[d] * b) * n[d];
    }
    return T(0.5) * (FnL + FnR) - T(0.5) * v[3] * (b - a);
}

// Volume node of face DOF it + n jt on face `face` (0..5 = -xi, +xi, -eta,
// +eta, -zeta, +zeta): the normal index sits at its boundary value and the
// two tangent indices run over the remaining axes in increasing axis order
// (LGL face extraction of build_face_interpolation,
// src/dg/basis_hexahedron.cpp:431-522).
__host__ __device__ inline int face_node(int n, int face, int it, int jt)
{
    const int nax = face >> 1;
    const int b = (face & 1) ? n - 1 : 0;
    const int t1 = (nax == 0) ? 1 : 0, t2 = (nax == 2) ? 1 : 2;
    int idx[3];
    idx[nax] = b;
    idx[t1] = it;
    idx[t2] = jt;
    return idx[0] + n * (idx[1] + n * idx[2]);
}

// Arguments of the face-term kernels (faces.hip), by value: the state U and
// tendencies dU in one strided layout (node d = i + n L of element e, field
// f at f fs + e es + L ls + i is), the 1D LGL weights of the inverse mass
// matrix, the conforming face rule and the scalar flux.
template <class T> struct FaceArgs {
    const T *U;
    T *dU;
    int64_t fs, es, ls, is;
    int64_t bs, bb; // blocked layout: block stride and elements per block (0: unblocked)
    T w[kMaxN];     // 1D LGL weights of the element basis
    T rw[kMaxFaceQ]; // 1D weights of the conforming face rule (nq1 points)
    int nq1;
    int kind;        // DRF_FLUX_UPWIND or DRF_FLUX_LF_ADV
    T p[4];          // vx, vy, vz, lambda
};

template <class T>
hipError_t launch_face_conforming(int n, bool lift, int64_t nif, const int64_t *list,
                                  const FaceArgs<T> &a, T *out_l, T *out_r, hipStream_t s);
// The conforming face term element by element (eface: [E][6] packed
// neighbour, neighbour face and side, -1 where the face is not conforming).
template <class T>
hipError_t launch_face_elements(int n, int64_t E, const int64_t *eface, const FaceArgs<T> &a,
                                hipStream_t s);
template <class T>
hipError_t launch_face_mortar(int n, int nfine, bool lift, int64_t nif, const int64_t *list,
                              const T *ops, const FaceArgs<T> &a, T *out_c, T *out_f,
                              hipStream_t s);

// Number of fine faces of a connection type (0 for boundary / invalid).
inline int mortar_num_fine(int type)
{
    switch (type) {
    case DRF_CONN_SAME_LEVEL: return 1;
    case DRF_CONN_FINE_2X1:
    case DRF_CONN_FINE_1X2: return 2;
    case DRF_CONN_FINE_3_2PLUS1:
    case DRF_CONN_FINE_3_1PLUS2: return 3;
    case DRF_CONN_FINE_2X2: return 4;
    default: return 0;
    }
}

// Host construction of the mortar operators (interface_api.cpp).  Outputs
// row-major nm x nm matrices, nm = (order+1)^2: M, Minv, Pc, Lc and nfine
// each of Pf, Lf.  Returns nfine, or -1 on failure.
int build_mortar_operators(int order, int type, double *M, double *Minv, double *Pc, double *Lc,
                           double *Pf, double *Lf);

// Gauss-Legendre nodes/weights (basis.cpp), n <= 2 kMaxN.
int gl_nodes(int n, double *x, double *w);

hipError_t launch_flux_points(int kind, int64_t npts, const double *UL, const double *UR,
                              const double *nrm, const FluxParams &p, double *out, hipStream_t s);
hipError_t launch_conforming(int kind, const FluxParams &p, int nf, const FaceRule &rule,
                             int64_t nfaces, const double *UL, const double *UR,
                             const double *nrm, double *rl, double *rr, hipStream_t s);
hipError_t launch_mortar_conservation(int nm, int nfine, int64_t nif, const double *rc,
                                     const double *rf, const double *wc, const double *wf,
                                     double *out, hipStream_t s);
hipError_t launch_mortar(int kind, const FluxParams &p, int nm, int nfine, const double *ops,
                         int64_t nif, const double *Uc, const double *Uf, const double *nrm,
                         double *rc, double *rf, hipStream_t s);
hipError_t launch_gather_faces(int n, int64_t nelem, int64_t elem_stride, int64_t nfaces,
                               const int64_t *elems, const int8_t *faces, const double *field,
                               double *out, hipStream_t s);

} // namespace drf
