// Face plans: the interface lists of an octree mesh grouped for the face-term
// kernels (faces.hip), and their C ABI.
//
// Mirrors the face loop of DG3DIntegration::compute_rhs
// (src/dg/operators_3d.cpp:255-281): conforming faces as
// compute_interface_flux_conforming (:320-343) would take them, and the 2:1
// mortars that AdaptiveMeshRefinement::update_mortar_interfaces registers
// (src/amr/refinement.cpp:575-596: every non-SameLevel, non-Boundary
// connection seen from its coarse side) as compute_interface_flux_nonconforming
// (:345-391) takes them.  The interface lists are sorted by normal axis, kind
// and level: one persistent launch per (axis, kind) walks its level groups in
// order.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "drf_internal.hpp"
#include "interface.hpp"

namespace drf {
// the library's own mortar operators (interface_api.cpp)
int build_mortar_operators(int order, int type, double *M, double *Minv, double *Pc, double *Lc,
                           double *Pf, double *Lf);
}

using namespace drf;

struct drf_face_plan {
    int order = 0, n = 0, nh = 0, precision = DRF_PREC_FP64, device = 0;
    int64_t E = 0;
    int kind = 0;
    double params[4] = {0, 0, 0, 0};
    // groups in launch order: axis, kind (DRF_CONN_SAME_LEVEL for conforming
    // faces, else the mortar's connection type), level, first, count
    std::vector<std::array<int64_t, 5>> groups;
    std::vector<int64_t> conf; // [nc][4]: left element, left face, right element, right face
    std::vector<int64_t> mort; // [nm][6]: coarse element, coarse face, fine elements by sub-face
    std::vector<int> mort_type;
    // [E][6]: (neighbour << 4) | (neighbour's face << 1) | side of every
    // conforming face (side 0: the element is the interface's left), else -1
    std::vector<int64_t> eface;
    int64_t *d_conf = nullptr, *d_mort = nullptr, *d_eface = nullptr;
    void *d_wq = nullptr;   // face-rule weight by face DOF (plan precision), for the fused phase
    // conforming interfaces: DRF_FACES_BY_INTERFACE, _BY_ELEMENT or _IN_VOLUME
    int conf_mode = DRF_FACES_IN_VOLUME;
    void *d_ops[7] = {};       // per connection type: Minv | Pc | Lc | Pf[] | Lf[] (context precision)
    drf_ctx *attached = nullptr;
};

namespace {

#define DRF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond))                                                                         \
            return ::drf::fail((code), (msg));                                               \
    } while (0)

// tangent axes of a face in increasing axis order (get_face_tangent_axes,
// include/dg/face_connection.hpp:157-171)
void tangents(int face, int &t1, int &t2)
{
    const int a = face >> 1;
    t1 = (a == 0) ? 1 : 0;
    t2 = (a == 2) ? 1 : 2;
}

bool close_to(double a, double b, double scale) { return std::fabs(a - b) <= 1e-9 * scale; }

// Sub-face of the coarse face (MortarSpace::compute_subface_bounds,
// src/dg/mortar.cpp:27-78: t1 halves first) that a fine leaf covers, or -1
// when it covers no sub-face of the type exactly.
int subface_of(const double *cb, const double *fb, int face, int type)
{
    int t1, t2;
    tangents(face, t1, t2);
    const bool s1 = type == DRF_CONN_FINE_2X1 || type == DRF_CONN_FINE_2X2;
    const bool s2 = type == DRF_CONN_FINE_1X2 || type == DRF_CONN_FINE_2X2;
    int half[2] = {0, 0};
    const int ax[2] = {t1, t2};
    const bool split[2] = {s1, s2};
    for (int k = 0; k < 2; ++k) {
        const double lo = cb[2 * ax[k]], hi = cb[2 * ax[k] + 1], sc = hi - lo;
        const double flo = fb[2 * ax[k]], fhi = fb[2 * ax[k] + 1];
        if (!split[k]) {
            if (!close_to(flo, lo, sc) || !close_to(fhi, hi, sc))
                return -1;
            continue;
        }
        const double mid = lo + 0.5 * sc;
        if (close_to(flo, lo, sc) && close_to(fhi, mid, sc))
            half[k] = 0;
        else if (close_to(flo, mid, sc) && close_to(fhi, hi, sc))
            half[k] = 1;
        else
            return -1;
    }
    if (type == DRF_CONN_FINE_2X1)
        return half[0];
    if (type == DRF_CONN_FINE_1X2)
        return half[1];
    return half[0] + 2 * half[1];
}

template <class T> drf_status upload(const std::vector<T> &h, T **d)
{
    *d = nullptr;
    if (h.empty())
        return DRF_OK;
    DRF_HIP_CHECK(hipMalloc((void **)d, sizeof(T) * h.size()));
    DRF_HIP_CHECK(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return DRF_OK;
}

void free_plan(drf_face_plan *p)
{
    if (!p)
        return;
    if (p->attached && p->attached->faces == p)
        p->attached->faces = nullptr;
    (void)hipFree(p->d_conf);
    (void)hipFree(p->d_mort);
    (void)hipFree(p->d_eface);
    (void)hipFree(p->d_wq);
    for (void *o : p->d_ops)
        (void)hipFree(o);
    delete p;
}

template <class T> void fill_args(const drf_face_plan *p, FaceArgs<T> &a, const void *U, void *dU,
                                  int layout)
{
    std::memset(&a, 0, sizeof(a));
    a.U = (const T *)U;
    a.dU = (T *)dU;
    layout_strides(layout, p->E, p->n, 5, a.fs, a.es, a.ls, a.is, a.bs, a.bb);
    double x[kMaxN], w[kMaxN];
    lgl_nodes(p->n, x, w);
    for (int i = 0; i < p->n; ++i)
        a.w[i] = (T)w[i];
    a.nq1 = p->order < 2 ? 2 : p->order; // GaussQuadrature1D(order, LGL), quadrature_3d.cpp:13-31
    double rx[kMaxN], rw[kMaxN];
    lgl_nodes(a.nq1, rx, rw);
    for (int i = 0; i < a.nq1; ++i)
        a.rw[i] = (T)rw[i];
    a.kind = p->kind;
    for (int i = 0; i < 4; ++i)
        a.p[i] = (T)p->params[i];
}

// The element-ordered conforming launch (standalone or fused) addresses the
// state through 32-bit per-lane byte offsets: the element span must fit.
template <class T> bool element_offsets_fit(const drf_face_plan *p, const FaceArgs<T> &a)
{
    const int64_t span =
        (elem_offset(p->E - 1, a.es, a.bs, a.bb) + (int64_t)(p->nh - 1) * a.ls) * (int64_t)sizeof(T);
    return span < ((int64_t)1 << 32);
}

// One group (or, with group < 0, every group with LIFT) of a plan; with
// skip_conforming the conforming groups are left out (the volume kernel has
// added them).
template <class T>
drf_status run_groups(const drf_face_plan *p, int group, const void *U, void *dU, int layout,
                      void *out_a, void *out_b, hipStream_t s, bool skip_conforming = false)
{
    FaceArgs<T> a;
    fill_args<T>(p, a, U, dU, layout);
    const bool lift = group < 0;
    size_t g0 = lift ? 0 : (size_t)group, g1 = lift ? p->groups.size() : (size_t)group + 1;
    // The whole face term: every conforming interface in one element-ordered
    // launch (each element's block read and rewritten once; when the state's
    // element span exceeds 4 GiB the interface-ordered launches below take
    // the conforming groups), then the mortar groups.
    const bool by_element = lift && p->conf_mode != DRF_FACES_BY_INTERFACE &&
                            !p->conf.empty() && element_offsets_fit<T>(p, a);
    if (by_element && !skip_conforming) {
        const hipError_t e = launch_face_elements<T>(p->n, p->E, p->d_eface, a, s);
        if (e != hipSuccess)
            return fail(DRF_ERR_HIP, std::string("face kernel launch: ") + hipGetErrorString(e));
    }
    // consecutive groups of one (axis, kind) are one launch
    while (g0 < g1) {
        size_t ge = g0 + 1;
        if (lift)
            while (ge < g1 && p->groups[ge][0] == p->groups[g0][0] &&
                   p->groups[ge][1] == p->groups[g0][1])
                ++ge;
        const int kind = (int)p->groups[g0][1];
        if ((by_element || skip_conforming) && kind == DRF_CONN_SAME_LEVEL) {
            g0 = ge;
            continue;
        }
        const int64_t first = p->groups[g0][3];
        int64_t cnt = 0;
        for (size_t g = g0; g < ge; ++g)
            cnt += p->groups[g][4];
        hipError_t e;
        if (kind == DRF_CONN_SAME_LEVEL)
            e = launch_face_conforming<T>(p->n, lift, cnt, p->d_conf + 4 * first, a, (T *)out_a,
                                          (T *)out_b, s);
        else
            e = launch_face_mortar<T>(p->n, mortar_num_fine(kind), lift, cnt,
                                      p->d_mort + 6 * first, (const T *)p->d_ops[kind], a,
                                      (T *)out_a, (T *)out_b, s);
        if (e != hipSuccess)
            return fail(DRF_ERR_HIP, std::string("face kernel launch: ") + hipGetErrorString(e));
        g0 = ge;
    }
    return DRF_OK;
}

} // namespace

namespace drf {

// dU += the plan's face term (every group), on stream s; used by the
// right-hand-side entry points of a context with an attached plan.
drf_status face_term(const drf_face_plan *p, const void *U, void *dU, int layout, hipStream_t s)
{
    if (p->groups.empty())
        return DRF_OK;
    if (p->precision == DRF_PREC_FP64)
        return run_groups<double>(p, -1, U, dU, layout, nullptr, nullptr, s);
    return run_groups<float>(p, -1, U, dU, layout, nullptr, nullptr, s);
}

// The fused form: the conforming part for the volume kernel (false when the
// plan keeps it out: another mode, no conforming interface, or offsets that
// do not fit), and the mortar groups on their own.
template <class T> bool volume_faces(const drf_face_plan *p, int layout, VolFaces<T> *out)
{
    // the volume kernel's face phase addresses unblocked states: a blocked
    // state takes the element-ordered launch after the volume kernel
    if (p->conf_mode != DRF_FACES_IN_VOLUME || p->conf.empty() || !p->d_wq ||
        layout == DRF_LAYOUT_BLOCKS)
        return false;
    FaceArgs<T> a;
    fill_args<T>(p, a, nullptr, nullptr, layout);
    if (!element_offsets_fit<T>(p, a))
        return false;
    out->eface = p->d_eface;
    out->wq = (const T *)p->d_wq;
    for (int i = 0; i < 4; ++i)
        out->p[i] = a.p[i];
    out->kind = p->kind;
    out->nq = a.nq1 * a.nq1;
    return true;
}
template bool volume_faces<double>(const drf_face_plan *, int, VolFaces<double> *);
template bool volume_faces<float>(const drf_face_plan *, int, VolFaces<float> *);

drf_status face_term_mortars(const drf_face_plan *p, const void *U, void *dU, int layout,
                             hipStream_t s)
{
    if (p->groups.empty())
        return DRF_OK;
    if (p->precision == DRF_PREC_FP64)
        return run_groups<double>(p, -1, U, dU, layout, nullptr, nullptr, s, true);
    return run_groups<float>(p, -1, U, dU, layout, nullptr, nullptr, s, true);
}

} // namespace drf

extern "C" {

drf_status drf_face_plan_create(const drf_ctx *ctx, int64_t n_elem, const double *bounds,
                                const int32_t *levels, const uint8_t *type, const int8_t *count,
                                const int64_t *elems, const int8_t *faces, int kind,
                                const double *params, drf_face_plan **out)
{
    DRF_REQUIRE(out, DRF_ERR_ARG, "drf_face_plan_create: null output");
    *out = nullptr;
    DRF_REQUIRE(ctx && ctx->E > 0, DRF_ERR_STATE,
                "drf_face_plan_create: context not initialized");
    DRF_REQUIRE(n_elem == ctx->E, DRF_ERR_ARG,
                "drf_face_plan_create: element count differs from the context's");
    DRF_REQUIRE(ctx->n >= 2 && ctx->n <= 5, DRF_ERR_UNSUPPORTED,
                "drf_face_plan_create: face kernels are built for orders 1-4");
    DRF_REQUIRE(bounds && levels && type && count && elems && faces, DRF_ERR_ARG,
                "drf_face_plan_create: null argument");
    DRF_REQUIRE(kind == DRF_FLUX_UPWIND || kind == DRF_FLUX_LF_ADV, DRF_ERR_ARG,
                "drf_face_plan_create: kind must be DRF_FLUX_UPWIND or DRF_FLUX_LF_ADV");
    DRF_REQUIRE(params, DRF_ERR_ARG, "drf_face_plan_create: null params");
    const int64_t E = n_elem;
    auto lvl = [&](int64_t e) {
        return std::max(levels[3 * e], std::max(levels[3 * e + 1], levels[3 * e + 2]));
    };
    // (axis, kind, level, element, face, index into the kind's list)
    struct Item {
        int axis, kind, level;
        int64_t e;
        int f;
        std::array<int64_t, 4> other;
    };
    std::vector<Item> items;
    std::vector<char> used((size_t)E * 6, 0);
    auto take = [&](int64_t e, int f) {
        char &u = used[(size_t)(6 * e + f)];
        if (u)
            return false;
        u = 1;
        return true;
    };
    for (int64_t e = 0; e < E; ++e)
        for (int f = 0; f < 6; ++f) {
            const int64_t k = 6 * e + f;
            const int t = type[k];
            if (t == DRF_CONN_BOUNDARY)
                continue;
            if (t == DRF_CONN_SAME_LEVEL) {
                DRF_REQUIRE(count[k] == 1, DRF_ERR_ARG,
                            "drf_face_plan_create: SameLevel connection without one neighbour");
                const int64_t nb = elems[4 * k];
                DRF_REQUIRE(nb >= 0 && nb < E && faces[4 * k] == (f ^ 1), DRF_ERR_ARG,
                            "drf_face_plan_create: invalid neighbour");
                const int64_t kb = 6 * nb + (f ^ 1);
                const bool back = type[kb] == DRF_CONN_SAME_LEVEL && count[kb] == 1 &&
                                  elems[4 * kb] == e;
                // a reciprocal pair once, from its lower element; a face
                // against a coarser leaf adds nothing of its own
                if (!back || nb < e)
                    continue;
                DRF_REQUIRE(take(e, f) && take(nb, f ^ 1), DRF_ERR_ARG,
                            "drf_face_plan_create: a face belongs to two interfaces");
                items.push_back({f >> 1, DRF_CONN_SAME_LEVEL, lvl(e), e, f, {nb, f ^ 1, -1, -1}});
                continue;
            }
            DRF_REQUIRE(t == DRF_CONN_FINE_2X1 || t == DRF_CONN_FINE_1X2 || t == DRF_CONN_FINE_2X2,
                        DRF_ERR_UNSUPPORTED,
                        "drf_face_plan_create: connection type without a 2:1 mortar (1:3 faces)");
            const int nf = mortar_num_fine(t);
            DRF_REQUIRE(count[k] == nf, DRF_ERR_UNSUPPORTED,
                        "drf_face_plan_create: element " + std::to_string(e) + " face " +
                            std::to_string(f) + ": fine face count does not match its type");
            std::array<int64_t, 4> fine = {-1, -1, -1, -1};
            for (int s = 0; s < nf; ++s) {
                const int64_t fe = elems[4 * k + s];
                DRF_REQUIRE(fe >= 0 && fe < E && faces[4 * k + s] == (f ^ 1), DRF_ERR_ARG,
                            "drf_face_plan_create: invalid fine neighbour");
                const int sub = subface_of(bounds + 6 * e, bounds + 6 * fe, f, t);
                DRF_REQUIRE(sub >= 0 && fine[sub] < 0, DRF_ERR_UNSUPPORTED,
                            "drf_face_plan_create: fine faces of element " + std::to_string(e) +
                                " face " + std::to_string(f) + " do not tile it");
                fine[sub] = fe;
            }
            DRF_REQUIRE(take(e, f), DRF_ERR_ARG,
                        "drf_face_plan_create: a face belongs to two interfaces");
            for (int s = 0; s < nf; ++s)
                DRF_REQUIRE(take(fine[s], f ^ 1), DRF_ERR_ARG,
                            "drf_face_plan_create: a face belongs to two interfaces");
            items.push_back({f >> 1, t, lvl(e), e, f, fine});
        }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        if (a.axis != b.axis)
            return a.axis < b.axis;
        if (a.kind != b.kind)
            return a.kind < b.kind;
        return a.level < b.level;
    });
    drf_face_plan *p = new drf_face_plan();
    p->order = ctx->order;
    p->n = ctx->n;
    p->nh = ctx->nh;
    p->precision = ctx->precision;
    p->device = ctx->device;
    p->E = E;
    p->kind = kind;
    for (int i = 0; i < (kind == DRF_FLUX_UPWIND ? 3 : 4); ++i)
        p->params[i] = params[i];
    for (const Item &it : items) {
        const bool conf = it.kind == DRF_CONN_SAME_LEVEL;
        const int64_t idx = conf ? (int64_t)p->conf.size() / 4 : (int64_t)p->mort.size() / 6;
        if (p->groups.empty() || p->groups.back()[0] != it.axis || p->groups.back()[1] != it.kind ||
            p->groups.back()[2] != it.level)
            p->groups.push_back({it.axis, it.kind, it.level, idx, 0});
        ++p->groups.back()[4];
        if (conf) {
            p->conf.insert(p->conf.end(), {it.e, (int64_t)it.f, it.other[0], it.other[1]});
        } else {
            p->mort.insert(p->mort.end(), {it.e, (int64_t)it.f, it.other[0], it.other[1],
                                           it.other[2], it.other[3]});
            p->mort_type.push_back(it.kind);
        }
    }
    if (!p->conf.empty()) {
        p->eface.assign((size_t)E * 6, -1);
        for (size_t c = 0; c < p->conf.size(); c += 4) {
            const int64_t el = p->conf[c], fl = p->conf[c + 1], er = p->conf[c + 2],
                          fr = p->conf[c + 3];
            p->eface[(size_t)(6 * el + fl)] = (er << 4) | (fr << 1);
            p->eface[(size_t)(6 * er + fr)] = (el << 4) | (fl << 1) | 1;
        }
    }
    drf_status st = hipSetDevice(ctx->device) == hipSuccess
                        ? DRF_OK
                        : fail(DRF_ERR_HIP, "drf_face_plan_create: hipSetDevice failed");
    if (!st)
        st = upload(p->conf, &p->d_conf);
    if (!st)
        st = upload(p->eface, &p->d_eface);
    // face-rule weight by face DOF, as the conforming kernels form it, in the
    // plan's precision (the fused volume phase reads it)
    if (!st && !p->conf.empty()) {
        const int nq1 = p->order < 2 ? 2 : p->order, nh = p->nh;
        double rx[kMaxN], rw[kMaxN];
        lgl_nodes(nq1, rx, rw);
        if (p->precision == DRF_PREC_FP64) {
            std::vector<double> wq(nh, 0.0);
            for (int q = 0; q < nq1 * nq1 && q < nh; ++q)
                wq[q] = rw[q % nq1] * rw[q / nq1];
            st = upload(wq, (double **)&p->d_wq);
        } else {
            std::vector<float> wq(nh, 0.0f);
            for (int q = 0; q < nq1 * nq1 && q < nh; ++q)
                wq[q] = (float)rw[q % nq1] * (float)rw[q / nq1];
            st = upload(wq, (float **)&p->d_wq);
        }
    }
    if (!st)
        st = upload(p->mort, &p->d_mort);
    // mortar operators per connection type present, in the context's precision
    for (int t = DRF_CONN_FINE_2X1; t <= DRF_CONN_FINE_2X2 && !st; ++t) {
        if (std::find(p->mort_type.begin(), p->mort_type.end(), t) == p->mort_type.end())
            continue;
        const int nf = mortar_num_fine(t), nm = p->nh, mm = nm * nm;
        std::vector<double> M(mm), Minv(mm), Pc(mm), Lc(mm), Pf((size_t)nf * mm),
            Lf((size_t)nf * mm);
        if (build_mortar_operators(p->order, t, M.data(), Minv.data(), Pc.data(), Lc.data(),
                                   Pf.data(), Lf.data()) != nf) {
            st = fail(DRF_ERR_ARG, "drf_face_plan_create: mortar mass matrix is singular");
            break;
        }
        std::vector<double> ops;
        for (auto *v : {&Minv, &Pc, &Lc, &Pf, &Lf})
            ops.insert(ops.end(), v->begin(), v->end());
        if (p->precision == DRF_PREC_FP64) {
            st = upload(ops, (double **)&p->d_ops[t]);
        } else {
            std::vector<float> o32(ops.begin(), ops.end());
            st = upload(o32, (float **)&p->d_ops[t]);
        }
    }
    if (st) {
        free_plan(p);
        return st;
    }
    *out = p;
    return DRF_OK;
}

void drf_face_plan_destroy(drf_face_plan *plan) { free_plan(plan); }

drf_status drf_face_plan_set_conforming_mode(drf_face_plan *p, int mode)
{
    DRF_REQUIRE(p, DRF_ERR_ARG, "drf_face_plan_set_conforming_mode: null plan");
    DRF_REQUIRE(mode == DRF_FACES_BY_INTERFACE || mode == DRF_FACES_BY_ELEMENT ||
                    mode == DRF_FACES_IN_VOLUME,
                DRF_ERR_ARG, "drf_face_plan_set_conforming_mode: unknown mode");
    p->conf_mode = mode;
    return DRF_OK;
}

drf_status drf_face_plan_info(const drf_face_plan *p, int *n_groups, int64_t *n_conforming,
                              int64_t *n_mortar)
{
    DRF_REQUIRE(p, DRF_ERR_ARG, "drf_face_plan_info: null plan");
    if (n_groups)
        *n_groups = (int)p->groups.size();
    if (n_conforming)
        *n_conforming = (int64_t)p->conf.size() / 4;
    if (n_mortar)
        *n_mortar = (int64_t)p->mort.size() / 6;
    return DRF_OK;
}

drf_status drf_face_plan_groups(const drf_face_plan *p, int64_t *desc)
{
    DRF_REQUIRE(p && desc, DRF_ERR_ARG, "drf_face_plan_groups: null argument");
    for (size_t g = 0; g < p->groups.size(); ++g)
        for (int i = 0; i < 5; ++i)
            desc[5 * g + i] = p->groups[g][i];
    return DRF_OK;
}

drf_status drf_face_plan_lists(const drf_face_plan *p, int64_t *conforming, int64_t *mortar)
{
    DRF_REQUIRE(p, DRF_ERR_ARG, "drf_face_plan_lists: null plan");
    if (conforming && !p->conf.empty())
        std::memcpy(conforming, p->conf.data(), sizeof(int64_t) * p->conf.size());
    if (mortar && !p->mort.empty())
        std::memcpy(mortar, p->mort.data(), sizeof(int64_t) * p->mort.size());
    return DRF_OK;
}

drf_status drf_face_term_dev(drf_face_plan *p, const void *U, void *dU, int layout, void *stream)
{
    DRF_REQUIRE(p, DRF_ERR_ARG, "drf_face_term_dev: null plan");
    DRF_REQUIRE(layout == DRF_LAYOUT_PACKED || layout == DRF_LAYOUT_SOA ||
                    layout == DRF_LAYOUT_LINES || layout == DRF_LAYOUT_BLOCKS,
                DRF_ERR_ARG, "drf_face_term_dev: unknown layout");
    DRF_REQUIRE(U && dU, DRF_ERR_ARG, "drf_face_term_dev: null state");
    DRF_HIP_CHECK(hipSetDevice(p->device));
    return face_term(p, U, dU, layout, (hipStream_t)stream);
}

drf_status drf_face_group_fluxes_dev(drf_face_plan *p, int group, const void *U, int layout,
                                     void *out_a, void *out_b, void *stream)
{
    DRF_REQUIRE(p, DRF_ERR_ARG, "drf_face_group_fluxes_dev: null plan");
    DRF_REQUIRE(group >= 0 && group < (int)p->groups.size(), DRF_ERR_RANGE,
                "drf_face_group_fluxes_dev: group out of range");
    DRF_REQUIRE(layout == DRF_LAYOUT_PACKED || layout == DRF_LAYOUT_SOA ||
                    layout == DRF_LAYOUT_LINES || layout == DRF_LAYOUT_BLOCKS,
                DRF_ERR_ARG, "drf_face_group_fluxes_dev: unknown layout");
    DRF_REQUIRE(U && out_a && out_b, DRF_ERR_ARG, "drf_face_group_fluxes_dev: null pointer");
    DRF_HIP_CHECK(hipSetDevice(p->device));
    if (p->precision == DRF_PREC_FP64)
        return run_groups<double>(p, group, U, nullptr, layout, out_a, out_b,
                                  (hipStream_t)stream);
    return run_groups<float>(p, group, U, nullptr, layout, out_a, out_b, (hipStream_t)stream);
}

drf_status drf_set_face_plan(drf_ctx *ctx, drf_face_plan *p)
{
    DRF_REQUIRE(ctx, DRF_ERR_ARG, "drf_set_face_plan: null context");
    if (p) {
        DRF_REQUIRE(p->E == ctx->E && p->order == ctx->order && p->precision == ctx->precision &&
                        p->device == ctx->device,
                    DRF_ERR_ARG, "drf_set_face_plan: plan built for another context");
        DRF_REQUIRE(!p->attached || p->attached == ctx, DRF_ERR_STATE,
                    "drf_set_face_plan: plan attached to another context");
    }
    if (ctx->faces && ctx->faces != p)
        ctx->faces->attached = nullptr;
    ctx->faces = p;
    if (p)
        p->attached = ctx;
    return DRF_OK;
}

} // extern "C"
