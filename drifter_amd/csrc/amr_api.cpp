// C ABI of the AMR solution transfer: SolutionProjection::project_to_children
// / project_from_children (src/amr/refinement.cpp:217-342),
// AdaptiveMeshRefinement::transfer_solution (:533-567), and the transfer onto
// a refined mesh's own leaf order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "amr.hpp"
#include "drf_internal.hpp"

using namespace drf;

namespace {

#define DRF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond))                                                                         \
            return ::drf::fail((code), (msg));                                               \
    } while (0)

// L_i(lo + (x_p + 1) (hi - lo) / 2) in the reference's arithmetic
// (project_to_children :252-263 with interpolate_3d :350-362).
double lagrange_at(const double *x, int n, int p, int i, double lo, double hi)
{
#pragma clang fp contract(off)
    const double xi = lo + 0.5 * (x[p] + 1.0) * (hi - lo);
    double L = 1.0;
    for (int j = 0; j < n; ++j)
        if (j != i)
            L *= (xi - x[j]) / (x[i] - x[j]);
    return L;
}

bool node_set(int n, NodeSet &ns)
{
    std::memset(&ns, 0, sizeof(ns));
    double w[kMaxN];
    if (lgl_nodes(n, ns.x, w))
        return false;
    const double lo[3] = {-1.0, -1.0, 0.0}, hi[3] = {1.0, 0.0, 1.0};
    for (int b = 0; b < 3; ++b)
        for (int p = 0; p < n; ++p)
            for (int i = 0; i < n; ++i)
                ns.M[b][p * kMaxN + i] = lagrange_at(ns.x, n, p, i, lo[b], hi[b]);
    return true;
}

int num_children(int mask) { return 1 << (((mask & 1) != 0) + ((mask & 2) != 0) + ((mask & 4) != 0)); }

// project_from_children's tables (refinement.cpp:296-336): the child half
// holding each parent node and the Lagrange weights at its child coordinate,
// in the reference's arithmetic.
bool from_children_ops(int n, int mask, FromChildrenOps &ops)
{
    NodeSet ns;
    if (!node_set(n, ns))
        return false;
    std::memset(&ops, 0, sizeof(ops));
    for (int a = 0; a < 3; ++a) {
        const bool ref = (mask >> a) & 1;
        for (int p = 0; p < n; ++p) {
            const double xp = ns.x[p];
            const int h = ref ? (xp >= 0 ? 1 : 0) : 0;
            const double lo = ref ? (-1.0 + h) : -1.0, hi = ref ? (double)h : 1.0;
            const double xc = std::clamp(2.0 * (xp - lo) / (hi - lo) - 1.0, -1.0, 1.0);
            ops.half[a][p] = (int8_t)h;
            for (int i = 0; i < n; ++i) {
                double L = 1.0;
                for (int j = 0; j < n; ++j)
                    if (j != i)
                        L *= (xc - ns.x[j]) / (ns.x[i] - ns.x[j]);
                ops.Q[a][p * n + i] = L;
            }
        }
    }
    return true;
}

} // namespace

extern "C" {

drf_status drf_project_to_children_dev(int order, int mask, int64_t n_parent, const void *U,
                                       int64_t elem_stride, void *children, int64_t child_stride,
                                       void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_project_to_children_dev: order out of range");
    DRF_REQUIRE(mask >= 1 && mask <= 7, DRF_ERR_ARG,
                "drf_project_to_children_dev: mask must be 1..7 (RefineMask bits)");
    const int n = order + 1, nd = n * n * n;
    DRF_REQUIRE(n_parent >= 0 && elem_stride >= nd && child_stride >= nd, DRF_ERR_ARG,
                "drf_project_to_children_dev: bad sizes (strides must be >= (N+1)^3)");
    if (n_parent == 0)
        return DRF_OK;
    DRF_REQUIRE(U && children, DRF_ERR_ARG, "drf_project_to_children_dev: null pointer");
    NodeSet ns;
    DRF_REQUIRE(node_set(n, ns), DRF_ERR_ARG, "drf_project_to_children_dev: LGL nodes failed");
    DRF_HIP_CHECK(launch_interp_boxes(
        n, BoxMode::Children, mask, n_parent * num_children(mask), n_parent,
        one_field((const double *)U, elem_stride, (double *)children, child_stride), 0, 0, nullptr,
        nullptr, ns, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_project_from_children_dev(int order, int mask, int64_t n_parent,
                                         const void *children, int64_t child_stride, void *U,
                                         int64_t elem_stride, void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_project_from_children_dev: order out of range");
    DRF_REQUIRE(mask >= 1 && mask <= 7, DRF_ERR_ARG,
                "drf_project_from_children_dev: mask must be 1..7 (RefineMask bits)");
    const int n = order + 1, nd = n * n * n;
    DRF_REQUIRE(n_parent >= 0 && elem_stride >= nd && child_stride >= nd, DRF_ERR_ARG,
                "drf_project_from_children_dev: bad sizes (strides must be >= (N+1)^3)");
    if (n_parent == 0)
        return DRF_OK;
    DRF_REQUIRE(U && children, DRF_ERR_ARG, "drf_project_from_children_dev: null pointer");
    FromChildrenOps ops;
    DRF_REQUIRE(from_children_ops(n, mask, ops), DRF_ERR_ARG,
                "drf_project_from_children_dev: LGL nodes failed");
    DRF_HIP_CHECK(launch_from_children(n, mask, n_parent, (const double *)children, child_stride,
                                       ops, (double *)U, elem_stride, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_transfer_solution_dev(int order, int64_t n_elem, const void *U, int64_t elem_stride,
                                     int64_t n_refine, const void *parents, void *U_new,
                                     int64_t new_stride, void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_transfer_solution_dev: order out of range");
    const int n = order + 1, nd = n * n * n;
    DRF_REQUIRE(n_elem >= 0 && n_refine >= 0 && n_refine <= n_elem && elem_stride >= nd &&
                    new_stride >= nd,
                DRF_ERR_ARG, "drf_transfer_solution_dev: bad sizes");
    const int64_t n_keep = n_elem - n_refine, n_out = n_keep + 8 * n_refine;
    if (n_out == 0)
        return DRF_OK;
    DRF_REQUIRE(U && U_new && (n_refine == 0 || parents), DRF_ERR_ARG,
                "drf_transfer_solution_dev: null pointer");
    NodeSet ns;
    DRF_REQUIRE(node_set(n, ns), DRF_ERR_ARG, "drf_transfer_solution_dev: LGL nodes failed");
    DRF_HIP_CHECK(launch_interp_boxes(
        n, BoxMode::Transfer, 7, n_out, n_elem,
        one_field((const double *)U, elem_stride, (double *)U_new, new_stride), n_keep, n_refine,
        (const int64_t *)parents, nullptr, ns, (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_transfer_solution(int order, int64_t n_elem, const double *U, int64_t n_refine,
                                 const int64_t *parents, double *U_new, int device)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_transfer_solution: order out of range");
    DRF_REQUIRE(n_elem >= 0 && n_refine >= 0 && n_refine <= n_elem, DRF_ERR_ARG,
                "drf_transfer_solution: bad sizes");
    const int64_t nd = (int64_t)(order + 1) * (order + 1) * (order + 1);
    const int64_t n_out = n_elem + 7 * n_refine;
    if (n_out == 0)
        return DRF_OK;
    DRF_REQUIRE(U && U_new && (n_refine == 0 || parents), DRF_ERR_ARG,
                "drf_transfer_solution: null pointer");
    for (int64_t r = 0; r < n_refine; ++r)
        DRF_REQUIRE(parents[r] >= 0 && parents[r] < n_elem && (r == 0 || parents[r - 1] < parents[r]),
                    DRF_ERR_ARG,
                    "drf_transfer_solution: parents must be strictly ascending element indices");
    DRF_HIP_CHECK(hipSetDevice(device));
    // one allocation per call: the adapt cycle is rare next to the RHS
    struct Bufs {
        void *p = nullptr;
        ~Bufs() { (void)hipFree(p); }
    } b;
    const size_t bu = sizeof(double) * n_elem * nd, bo = sizeof(double) * n_out * nd;
    DRF_HIP_CHECK(hipMalloc(&b.p, bu + bo + sizeof(int64_t) * (n_refine + 1)));
    double *dU = (double *)b.p, *dN = (double *)((char *)b.p + bu);
    int64_t *dP = (int64_t *)((char *)b.p + bu + bo);
    DRF_HIP_CHECK(hipMemcpy(dU, U, bu, hipMemcpyHostToDevice));
    if (n_refine)
        DRF_HIP_CHECK(hipMemcpy(dP, parents, sizeof(int64_t) * n_refine, hipMemcpyHostToDevice));
    drf_status st = drf_transfer_solution_dev(order, n_elem, dU, nd, n_refine, dP, dN, nd, nullptr);
    if (st)
        return st;
    DRF_HIP_CHECK(hipMemcpy(U_new, dN, bo, hipMemcpyDeviceToHost));
    return DRF_OK;
}

drf_status drf_interpolate_boxes_fields_dev(int order, int64_t n_out, int64_t n_elem,
                                            const void *src, const void *boxes, int n_fields,
                                            const void *const *U, const int64_t *elem_stride,
                                            void *const *U_new, const int64_t *new_stride,
                                            void *stream)
{
    DRF_REQUIRE(order >= 1 && order + 1 <= kMaxN, DRF_ERR_UNSUPPORTED,
                "drf_interpolate_boxes_dev: order out of range");
    DRF_REQUIRE(n_fields >= 1 && n_fields <= kMaxBoxFields, DRF_ERR_ARG,
                "drf_interpolate_boxes_fields_dev: 1 to " + std::to_string(kMaxBoxFields) +
                    " fields per call");
    DRF_REQUIRE(U && elem_stride && U_new && new_stride, DRF_ERR_ARG,
                "drf_interpolate_boxes_fields_dev: null field array");
    const int n = order + 1, nd = n * n * n;
    DRF_REQUIRE(n_out >= 0 && n_elem >= 0, DRF_ERR_ARG, "drf_interpolate_boxes_dev: bad sizes");
    BoxFields bf{};
    bf.nf = n_fields;
    for (int f = 0; f < n_fields; ++f) {
        DRF_REQUIRE(elem_stride[f] >= nd && new_stride[f] >= nd, DRF_ERR_ARG,
                    "drf_interpolate_boxes_dev: bad sizes (strides must be >= (N+1)^3)");
        DRF_REQUIRE(n_out == 0 || (U_new[f] && (n_elem == 0 || U[f])), DRF_ERR_ARG,
                    "drf_interpolate_boxes_dev: null pointer");
        bf.U[f] = (const double *)U[f];
        bf.es[f] = elem_stride[f];
        bf.out[f] = (double *)U_new[f];
        bf.os[f] = new_stride[f];
    }
    if (n_out == 0)
        return DRF_OK;
    DRF_REQUIRE(src && boxes, DRF_ERR_ARG, "drf_interpolate_boxes_dev: null pointer");
    NodeSet ns;
    DRF_REQUIRE(node_set(n, ns), DRF_ERR_ARG, "drf_interpolate_boxes_dev: LGL nodes failed");
    DRF_HIP_CHECK(launch_interp_boxes(n, BoxMode::Boxes, 7, n_out, n_elem, bf, 0, 0,
                                      (const int64_t *)src, (const double *)boxes, ns,
                                      (hipStream_t)stream));
    return DRF_OK;
}

drf_status drf_interpolate_boxes_dev(int order, int64_t n_out, int64_t n_elem, const void *src,
                                     const void *boxes, const void *U, int64_t elem_stride,
                                     void *U_new, int64_t new_stride, void *stream)
{
    return drf_interpolate_boxes_fields_dev(order, n_out, n_elem, src, boxes, 1, &U, &elem_stride,
                                            &U_new, &new_stride, stream);
}

drf_status drf_refine_transfer_map(int64_t n_old, const double *old_bounds, int64_t n_new,
                                   const double *new_bounds, int64_t *src, double *boxes)
{
    DRF_REQUIRE(n_old >= 0 && n_new >= 0, DRF_ERR_ARG, "drf_refine_transfer_map: negative count");
    DRF_REQUIRE(n_new == 0 || (old_bounds && new_bounds && src && boxes), DRF_ERR_ARG,
                "drf_refine_transfer_map: null pointer");
    // Leaves are numbered depth first, so the leaves inside an old leaf form
    // one contiguous run at its position: walk both lists once.
    int64_t e = 0;
    for (int64_t k = 0; k < n_new; ++k) {
        const double *nb = new_bounds + 6 * k;
        double c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = 0.5 * (nb[2 * a] + nb[2 * a + 1]);
        auto inside = [&](int64_t o) {
            const double *ob = old_bounds + 6 * o;
            for (int a = 0; a < 3; ++a)
                if (!(c[a] > ob[2 * a] && c[a] < ob[2 * a + 1]))
                    return false;
            return true;
        };
        while (e < n_old && !inside(e))
            ++e;
        DRF_REQUIRE(e < n_old, DRF_ERR_ARG,
                    "drf_refine_transfer_map: the new mesh is not a refinement of the old one "
                    "in leaf order");
        const double *ob = old_bounds + 6 * e;
        src[k] = e;
        for (int a = 0; a < 3; ++a) {
            const double w = ob[2 * a + 1] - ob[2 * a];
            // Sub-box bounds are dyadic (-1, -1/2, 0, ...): snap the rounding
            // of the bounds arithmetic away, so an unrefined axis is exactly
            // [-1, 1] (an exact copy) and one level of XYZ refinement gives
            // exactly the boxes of project_to_children.
            auto snap = [](double v) {
                const double r = std::nearbyint(v * 1048576.0) / 1048576.0;
                return std::fabs(r - v) < 1e-12 ? r : v;
            };
            boxes[6 * k + 2 * a] = snap(2.0 * (nb[2 * a] - ob[2 * a]) / w - 1.0);
            boxes[6 * k + 2 * a + 1] = snap(2.0 * (nb[2 * a + 1] - ob[2 * a]) / w - 1.0);
        }
    }
    return DRF_OK;
}

} // extern "C"
