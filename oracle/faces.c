/*
 * ORACLE (test infrastructure only) -- the face term of the DG right-hand
 * side on an octree mesh, and the interface lists it runs over.
 *
 * orc_face_term follows the face loop of DG3DIntegration::compute_rhs
 * (src/dg/operators_3d.cpp:255-281) with compute_interface_flux_conforming
 * (:320-343) and compute_interface_flux_nonconforming (:345-391): face traces
 * by the LGL face extraction (orc_face_map), the flux by
 * ConformingInterface::compute_flux (orc_conforming_flux) or
 * MortarSpace::compute_mortar_flux (orc_mortar_flux), the lift I_f^T of each
 * face vector added to its element's face-contribution vector, and the
 * inverse LGL mass matrix (DG3DElementOperator::mass_inv :52-55, diagonal
 * 1/(w_i w_j w_k), src/dg/basis_hexahedron.cpp:393-420) applied to that sum
 * before it is added to dU.  The normal is the outward normal of the
 * interface's left (conforming) or coarse (mortar) face, as :372-374 forms it.
 * The same scalar flux acts on each of the five fields of the packed state.
 *
 * orc_face_lists states the rule that selects the interfaces: a SameLevel
 * connection whose neighbour points back is one conforming interface (taken
 * from its lower element); every non-SameLevel, non-Boundary connection is a
 * mortar from its coarse side (update_mortar_interfaces,
 * src/amr/refinement.cpp:575-596), its fine elements put on the sub-faces of
 * MortarSpace::compute_subface_bounds (src/dg/mortar.cpp:27-78) by their
 * bounds; the lists are ordered by (normal axis, kind, level = largest
 * directional level of the owning element, element, face).
 *
 * Pinned by (tests/test_oracle_faces.py): the lists against the geometry of
 * refined meshes of every 2:1 type (conforming faces coincide, fine faces
 * tile their coarse face on the right sub-faces, every interior face in
 * exactly one interface, the group order); the face term in closed form for
 * piecewise-constant states with the upwind flux (conforming faces: M^-1 of
 * w_q v.n u on the N-point rule's face DOFs; a mortar's coarse side: M^-1 of
 * v.n u, since L_c M^-1 F = F there); the interface seen from either side
 * giving the same term; the cancellation of the mass-weighted conforming
 * contributions over a mesh; a planted wrong normal sign and a left-out
 * inverse mass each failing the closed form.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int lvl_of(const int32_t *levels, int64_t e)
{
    int a = levels[3 * e], b = levels[3 * e + 1], c = levels[3 * e + 2];
    int m = a > b ? a : b;
    return m > c ? m : c;
}

static int nfine(int type)
{
    switch (type) {
    case ORC_CONN_FINE2X1:
    case ORC_CONN_FINE1X2: return 2;
    case ORC_CONN_FINE2X2: return 4;
    default: return 0;
    }
}

static int near(double a, double b, double s) { return fabs(a - b) <= 1e-9 * s; }

/* sub-face of the coarse face covered by a fine leaf, -1 if none */
static int subface(const double *cb, const double *fb, int face, int type)
{
    const int ax = face / 2;
    const int t[2] = {ax == 0 ? 1 : 0, ax == 2 ? 1 : 2};
    const int split[2] = {type == ORC_CONN_FINE2X1 || type == ORC_CONN_FINE2X2,
                          type == ORC_CONN_FINE1X2 || type == ORC_CONN_FINE2X2};
    int half[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
        const double lo = cb[2 * t[k]], hi = cb[2 * t[k] + 1], s = hi - lo;
        const double flo = fb[2 * t[k]], fhi = fb[2 * t[k] + 1];
        if (!split[k]) {
            if (!near(flo, lo, s) || !near(fhi, hi, s))
                return -1;
        } else if (near(flo, lo, s) && near(fhi, lo + 0.5 * s, s)) {
            half[k] = 0;
        } else if (near(flo, lo + 0.5 * s, s) && near(fhi, hi, s)) {
            half[k] = 1;
        } else {
            return -1;
        }
    }
    if (type == ORC_CONN_FINE2X1)
        return half[0];
    if (type == ORC_CONN_FINE1X2)
        return half[1];
    return half[0] + 2 * half[1];
}

typedef struct {
    int axis, kind, level, face;
    int64_t e, other[4];
} item;

static int cmp_item(const void *pa, const void *pb)
{
    const item *a = (const item *)pa, *b = (const item *)pb;
    if (a->axis != b->axis)
        return a->axis < b->axis ? -1 : 1;
    if (a->kind != b->kind)
        return a->kind < b->kind ? -1 : 1;
    if (a->level != b->level)
        return a->level < b->level ? -1 : 1;
    if (a->e != b->e)
        return a->e < b->e ? -1 : 1;
    return a->face < b->face ? -1 : (a->face > b->face);
}

int orc_face_lists(int64_t E, const double *bounds, const int32_t *levels, const uint8_t *type,
                   const int8_t *count, const int64_t *elems, int64_t *conf, int64_t *nconf,
                   int64_t *mort, int64_t *nmort, int64_t *groups, int64_t *ngroups)
{
    item *it = (item *)malloc(sizeof(item) * (size_t)(6 * E + 1));
    int64_t n = 0;
    for (int64_t e = 0; e < E; ++e)
        for (int f = 0; f < 6; ++f) {
            const int64_t k = 6 * e + f;
            if (type[k] == ORC_CONN_BOUNDARY)
                continue;
            if (type[k] == ORC_CONN_SAME) {
                const int64_t nb = elems[4 * k];
                const int64_t kb = 6 * nb + (f ^ 1);
                if (!(type[kb] == ORC_CONN_SAME && count[kb] == 1 && elems[4 * kb] == e) || nb < e)
                    continue;
                item x = {f / 2, ORC_CONN_SAME, lvl_of(levels, e), f, e, {nb, f ^ 1, -1, -1}};
                it[n++] = x;
                continue;
            }
            const int nf = nfine(type[k]);
            if (!nf || count[k] != nf) {
                free(it);
                return -1;
            }
            item x = {f / 2, type[k], lvl_of(levels, e), f, e, {-1, -1, -1, -1}};
            for (int s = 0; s < nf; ++s) {
                const int64_t fe = elems[4 * k + s];
                const int sub = subface(bounds + 6 * e, bounds + 6 * fe, f, type[k]);
                if (sub < 0 || x.other[sub] >= 0) {
                    free(it);
                    return -1;
                }
                x.other[sub] = fe;
            }
            it[n++] = x;
        }
    qsort(it, (size_t)n, sizeof(item), cmp_item);
    int64_t nc = 0, nm = 0, ng = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int conf_kind = it[i].kind == ORC_CONN_SAME;
        if (i == 0 || it[i].axis != it[i - 1].axis || it[i].kind != it[i - 1].kind ||
            it[i].level != it[i - 1].level) {
            groups[5 * ng] = it[i].axis;
            groups[5 * ng + 1] = it[i].kind;
            groups[5 * ng + 2] = it[i].level;
            groups[5 * ng + 3] = conf_kind ? nc : nm;
            groups[5 * ng + 4] = 0;
            ++ng;
        }
        ++groups[5 * (ng - 1) + 4];
        if (conf_kind) {
            conf[4 * nc] = it[i].e;
            conf[4 * nc + 1] = it[i].face;
            conf[4 * nc + 2] = it[i].other[0];
            conf[4 * nc + 3] = it[i].other[1];
            ++nc;
        } else {
            mort[6 * nm] = it[i].e;
            mort[6 * nm + 1] = it[i].face;
            for (int s = 0; s < 4; ++s)
                mort[6 * nm + 2 + s] = it[i].other[s];
            ++nm;
        }
    }
    *nconf = nc;
    *nmort = nm;
    *ngroups = ng;
    free(it);
    return 0;
}

int orc_face_term(int order, int64_t E, int64_t nconf, const int64_t *conf, int64_t nmort,
                  const int64_t *mort, const int *mort_type, int kind, const double *p,
                  const double *U, double *dU)
{
    const int n = order + 1, nd = n * n * n, nh = n * n;
    double x[ORC_MAX_N], w[ORC_MAX_N];
    if (orc_lgl_nodes(n, x, w))
        return -1;
    double *acc = (double *)calloc((size_t)E * 5 * nd, sizeof(double));
    int fmap[6][ORC_MAX_N * ORC_MAX_N];
    for (int f = 0; f < 6; ++f)
        orc_face_map(order, f, fmap[f]);
    double a[ORC_MAX_N * ORC_MAX_N], b[4][ORC_MAX_N * ORC_MAX_N];
    double ra[ORC_MAX_N * ORC_MAX_N], rb[4 * ORC_MAX_N * ORC_MAX_N];
    for (int64_t c = 0; c < nconf; ++c) {
        const int64_t el = conf[4 * c], er = conf[4 * c + 2];
        const int fl = (int)conf[4 * c + 1], fr = (int)conf[4 * c + 3];
        double nrm[3] = {0.0, 0.0, 0.0};
        nrm[fl / 2] = (fl % 2) ? 1.0 : -1.0;
        for (int fld = 0; fld < 5; ++fld) {
            const double *ul = U + (el * 5 + fld) * nd, *ur = U + (er * 5 + fld) * nd;
            for (int q = 0; q < nh; ++q) {
                a[q] = ul[fmap[fl][q]];
                b[0][q] = ur[fmap[fr][q]];
            }
            if (orc_conforming_flux(order, kind, p, a, b[0], nrm, ra, rb) < 0) {
                free(acc);
                return -1;
            }
            double *al = acc + (el * 5 + fld) * nd, *ar = acc + (er * 5 + fld) * nd;
            for (int q = 0; q < nh; ++q) {
                al[fmap[fl][q]] += ra[q];
                ar[fmap[fr][q]] += rb[q];
            }
        }
    }
    for (int64_t m = 0; m < nmort; ++m) {
        const int64_t ec = mort[6 * m];
        const int fc = (int)mort[6 * m + 1], ff = fc ^ 1;
        const int nf = nfine(mort_type[m]);
        double nrm[3] = {0.0, 0.0, 0.0};
        nrm[fc / 2] = (fc % 2) ? 1.0 : -1.0;
        for (int fld = 0; fld < 5; ++fld) {
            const double *uc = U + (ec * 5 + fld) * nd;
            for (int q = 0; q < nh; ++q)
                a[q] = uc[fmap[fc][q]];
            double uf[4 * ORC_MAX_N * ORC_MAX_N];
            for (int s = 0; s < nf; ++s) {
                const double *u = U + (mort[6 * m + 2 + s] * 5 + fld) * nd;
                for (int q = 0; q < nh; ++q)
                    uf[s * nh + q] = u[fmap[ff][q]];
            }
            if (orc_mortar_flux(order, mort_type[m], kind, p, a, uf, nrm, ra, rb) < 0) {
                free(acc);
                return -1;
            }
            double *ac = acc + (ec * 5 + fld) * nd;
            for (int q = 0; q < nh; ++q)
                ac[fmap[fc][q]] += ra[q];
            for (int s = 0; s < nf; ++s) {
                double *af = acc + (mort[6 * m + 2 + s] * 5 + fld) * nd;
                for (int q = 0; q < nh; ++q)
                    af[fmap[ff][q]] += rb[s * nh + q];
            }
        }
    }
    /* rhs = M^-1 rhs (operators_3d.cpp:277-281) on the face contributions */
    for (int64_t e = 0; e < E; ++e)
        for (int fld = 0; fld < 5; ++fld)
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j)
                    for (int i = 0; i < n; ++i) {
                        const int d = i + n * (j + n * k);
                        const double wl = w[i] * w[j] * w[k];
                        const double minv = 1.0 / wl;
                        const int64_t o = (e * 5 + fld) * nd + d;
                        dU[o] = dU[o] + minv * acc[o];
                    }
    free(acc);
    return 0;
}
