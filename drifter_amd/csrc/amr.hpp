// AMR solution transfer (SolutionProjection, AdaptiveMeshRefinement::
// transfer_solution): declarations shared by amr.hip (kernels) and
// amr_api.cpp (host + C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "drf_internal.hpp"

namespace drf {

// How an output element finds its source element and its sub-box of the
// source's reference cube [-1, 1]^3.
enum class BoxMode : int {
    Transfer = 0, // transfer_solution: copies of the non-refined, then 8 XYZ children per parent
    Children = 1, // project_to_children: n_child children per parent, mask given
    Boxes = 2,    // explicit (source, box) per output element
};

// LGL nodes of the basis and the 1D interpolation matrices of the three
// standard sub-intervals of [-1, 1] (0 whole, 1 lower half, 2 upper half):
// M[b][p][i] = L_i(lo_b + (x_p + 1) (hi_b - lo_b) / 2), by value.
struct NodeSet {
    double x[kMaxN];
    double M[3][kMaxN * kMaxN];
};

// project_from_children's per-axis tables: Q[a][p][i] = L_i(local coordinate
// of parent node p in its child along axis a); half[a][p] = that child's index
// along a (0 if the axis is not refined).
struct FromChildrenOps {
    double Q[3][kMaxN * kMaxN];
    int8_t half[3][kMaxN];
};

// The fields one launch transfers: field f reads element e at U[f] + e es[f]
// and writes output element o at out[f] + o os[f].  Boxes mode takes up to
// kMaxBoxFields fields; Transfer and Children take one.
constexpr int kMaxBoxFields = 8;
struct BoxFields {
    const double *U[kMaxBoxFields];
    int64_t es[kMaxBoxFields];
    double *out[kMaxBoxFields];
    int64_t os[kMaxBoxFields];
    int nf;
};
inline BoxFields one_field(const double *U, int64_t es, double *out, int64_t os)
{
    BoxFields bf{};
    bf.U[0] = U;
    bf.es[0] = es;
    bf.out[0] = out;
    bf.os[0] = os;
    bf.nf = 1;
    return bf;
}

// Transfer: n_out = n_keep + 8 n_parent, parents idx[n_parent]; Children:
// every one of the n_src elements is a parent, n_out = n_src nchild(mask);
// Boxes: idx[n_out] sources and boxes[6 n_out].
hipError_t launch_interp_boxes(int n, BoxMode mode, int mask, int64_t n_out, int64_t n_src,
                               const BoxFields &bf, int64_t n_keep, int64_t n_parent,
                               const int64_t *idx, const double *boxes, const NodeSet &nodes,
                               hipStream_t s);
hipError_t launch_from_children(int n, int mask, int64_t n_parent, const double *children,
                                int64_t cs, const FromChildrenOps &ops, double *U, int64_t es,
                                hipStream_t s);

} // namespace drf
