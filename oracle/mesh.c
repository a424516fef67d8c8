/*
 * ORACLE (test infrastructure only) -- A17: directional octree, 2:1 balance
 * and face connectivity.
 *
 * Follows src/mesh/octree_adapter.cpp: build_uniform :23-84, balance
 * :138-182, refine :184-203, get_neighbor :259-359, build_face_connections
 * :369-389, rebuild_leaf_list / collect_leaves :445-468, create_children
 * :477-521, find_neighbor_same_or_coarser :523-580, find_fine_neighbors
 * :582-639; include/mesh/morton.hpp:64-143 (MortonUtil::refine and the
 * X/Y/Z masks :33-35); include/dg/face_connection.hpp:110-152
 * (get_connection_type) and ElementBounds::contains
 * (include/mesh/octree_adapter.hpp:39-42, inclusive).
 *
 * Note on the reference's behaviour that this restates exactly:
 * find_neighbor_same_or_coarser descends to the LEAF that contains the centre
 * of the would-be same-size neighbour (first child in creation order on ties),
 * so a face whose neighbour region is refined is reported as SameLevel with
 * that one fine leaf; the "refined neighbour" branch (:344-356) is reached
 * only when the descent finds no child containing the point.
 *
 * orc_mesh_connections_2to1 (not the reference's): every face lists the
 * leaves across it that share a piece of it, found by a descent over the
 * tree's boxes; the face plans' 2:1 mortars come from it.
 *
 * Pinned by (tests/test_oracle_mesh.py): on random anisotropic refine +
 * balance meshes, every FaceConnection against a linear scan (the leaf
 * holding the probe point with the smallest depth-first number, for the
 * reference's connectivity; the leaves sharing a piece of the face, for the
 * 2:1 one); the reference's count tests
 * (tests/unit/test_amr.cpp, tests/integration/test_mesh.cpp: 2x2x2 = 8 and
 * boundary/interior mix of 2x2x1), brute-force geometric adjacency on uniform
 * meshes (every interior face's neighbour shares the face and points back),
 * 2:1 balance of every leaf pair sharing a face after refine, minimality of
 * balance (every split it made has a node across a face whose level gap
 * exceeds one on exactly the split's axes), leaf order =
 * depth-first order with children iz -> iy -> ix, and Morton codes against an
 * independent bit-interleaving in Python.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t morton;
    int lev[3];
    double b[6]; /* xmin xmax ymin ymax zmin zmax */
    int mask;    /* bit0 X, bit1 Y, bit2 Z */
    int nchild;
    int64_t child[8];
    int64_t parent;
    int64_t leaf;
} onode;

struct orc_mesh {
    double dom[6];
    onode *nodes;
    int64_t nn, cap;
    int64_t *leaves;
    int64_t nl, lcap;
};

static const uint64_t XM = 0x1249249249249249ULL, YM = 0x2492492492492492ULL,
                      ZM = 0x4924924924924924ULL;

/* MortonUtil::child_x/y/z and refine, include/mesh/morton.hpp:104-143 */
static uint64_t child_axis(uint64_t code, uint64_t mask, int bit, uint64_t bitval)
{
    uint64_t bits = code & mask;
    bits = (bits << 3) | (bit ? bitval : 0ULL);
    code &= ~mask;
    code |= (bits & mask);
    return code;
}
static uint64_t morton_refine(uint64_t code, int ix, int iy, int iz, int rx, int ry, int rz)
{
    if (rx)
        code = child_axis(code, XM, ix, 0x1ULL);
    if (ry)
        code = child_axis(code, YM, iy, 0x2ULL);
    if (rz)
        code = child_axis(code, ZM, iz, 0x4ULL);
    return code;
}

static int64_t new_node(orc_mesh *m)
{
    if (m->nn == m->cap) {
        m->cap = m->cap ? 2 * m->cap : 1024;
        m->nodes = (onode *)realloc(m->nodes, sizeof(onode) * m->cap);
    }
    memset(&m->nodes[m->nn], 0, sizeof(onode));
    m->nodes[m->nn].parent = -1;
    m->nodes[m->nn].leaf = -1;
    return m->nn++;
}

/* create_children :477-521 */
static void create_children(orc_mesh *m, int64_t id)
{
    if (m->nodes[id].nchild)
        return;
    const int mask = m->nodes[id].mask;
    const int rx = mask & 1, ry = (mask >> 1) & 1, rz = (mask >> 2) & 1;
    const int nx = rx ? 2 : 1, ny = ry ? 2 : 1, nz = rz ? 2 : 1;
    double b[6];
    memcpy(b, m->nodes[id].b, sizeof(b));
    const double dx = (b[1] - b[0]) / nx, dy = (b[3] - b[2]) / ny, dz = (b[5] - b[4]) / nz;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int64_t c = new_node(m);
                onode *P = &m->nodes[id], *C = &m->nodes[c];
                C->morton = morton_refine(P->morton, ix, iy, iz, rx, ry, rz);
                C->lev[0] = P->lev[0] + (rx ? 1 : 0);
                C->lev[1] = P->lev[1] + (ry ? 1 : 0);
                C->lev[2] = P->lev[2] + (rz ? 1 : 0);
                C->b[0] = b[0] + ix * dx;
                C->b[1] = b[0] + (ix + 1) * dx;
                C->b[2] = b[2] + iy * dy;
                C->b[3] = b[2] + (iy + 1) * dy;
                C->b[4] = b[4] + iz * dz;
                C->b[5] = b[4] + (iz + 1) * dz;
                C->parent = id;
                P->child[P->nchild++] = c;
            }
}

static void collect(orc_mesh *m, int64_t id)
{
    if (!m->nodes[id].nchild) {
        if (m->nl == m->lcap) {
            m->lcap = m->lcap ? 2 * m->lcap : 1024;
            m->leaves = (int64_t *)realloc(m->leaves, sizeof(int64_t) * m->lcap);
        }
        m->leaves[m->nl++] = id;
        return;
    }
    for (int c = 0; c < m->nodes[id].nchild; ++c)
        collect(m, m->nodes[id].child[c]);
}

/* rebuild_leaf_list :445-458 */
static void rebuild(orc_mesh *m)
{
    m->nl = 0;
    collect(m, 0);
    for (int64_t i = 0; i < m->nl; ++i)
        m->nodes[m->leaves[i]].leaf = i;
}

static int contains(const double *b, const double *p)
{
    return p[0] >= b[0] && p[0] <= b[1] && p[1] >= b[2] && p[1] <= b[3] && p[2] >= b[4] &&
           p[2] <= b[5];
}

/* find_neighbor_same_or_coarser :523-580 */
static int64_t neighbor_same_or_coarser(const orc_mesh *m, int64_t id, int face)
{
    const double *b = m->nodes[id].b;
    double p[3];
    switch (face) {
    case 0: p[0] = b[0] - 0.5 * (b[1] - b[0]); p[1] = 0.5 * (b[2] + b[3]); p[2] = 0.5 * (b[4] + b[5]); break;
    case 1: p[0] = b[1] + 0.5 * (b[1] - b[0]); p[1] = 0.5 * (b[2] + b[3]); p[2] = 0.5 * (b[4] + b[5]); break;
    case 2: p[0] = 0.5 * (b[0] + b[1]); p[1] = b[2] - 0.5 * (b[3] - b[2]); p[2] = 0.5 * (b[4] + b[5]); break;
    case 3: p[0] = 0.5 * (b[0] + b[1]); p[1] = b[3] + 0.5 * (b[3] - b[2]); p[2] = 0.5 * (b[4] + b[5]); break;
    case 4: p[0] = 0.5 * (b[0] + b[1]); p[1] = 0.5 * (b[2] + b[3]); p[2] = b[4] - 0.5 * (b[5] - b[4]); break;
    default: p[0] = 0.5 * (b[0] + b[1]); p[1] = 0.5 * (b[2] + b[3]); p[2] = b[5] + 0.5 * (b[5] - b[4]); break;
    }
    if (!contains(m->dom, p))
        return -1;
    int64_t cur = 0;
    while (m->nodes[cur].nchild) {
        int found = 0;
        for (int c = 0; c < m->nodes[cur].nchild; ++c) {
            const int64_t ch = m->nodes[cur].child[c];
            if (contains(m->nodes[ch].b, p)) {
                cur = ch;
                found = 1;
                break;
            }
        }
        if (!found)
            break;
    }
    return cur;
}

/* tangent axes, include/dg/face_connection.hpp:157-171 */
static void tangents(int face, int *t1, int *t2)
{
    if (face <= 1) {
        *t1 = 1;
        *t2 = 2;
    } else if (face <= 3) {
        *t1 = 0;
        *t2 = 2;
    } else {
        *t1 = 0;
        *t2 = 1;
    }
}

/* get_connection_type, include/dg/face_connection.hpp:110-152 */
static int connection_type(const int *coarse, const int *fine, int face)
{
    int t1, t2;
    tangents(face, &t1, &t2);
    const int d1 = fine[t1] - coarse[t1], d2 = fine[t2] - coarse[t2];
    if (d1 == 0 && d2 == 0)
        return ORC_CONN_SAME;
    if (d1 == 1 && d2 == 0)
        return ORC_CONN_FINE2X1;
    if (d1 == 0 && d2 == 1)
        return ORC_CONN_FINE1X2;
    return ORC_CONN_FINE2X2;
}

/* find_fine_neighbors :582-639 (BFS from the coarse neighbour) */
static int fine_neighbors(const orc_mesh *m, int64_t id, int face, int64_t *out, int maxout)
{
    const int64_t coarse = neighbor_same_or_coarser(m, id, face);
    if (coarse < 0)
        return 0;
    int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (m->nn + 1));
    int64_t qh = 0, qt = 0;
    int cnt = 0;
    q[qt++] = coarse;
    const double *b = m->nodes[id].b;
    const double tol = 1e-12;
    while (qh < qt) {
        const int64_t x = q[qh++];
        const onode *N = &m->nodes[x];
        if (!N->nchild) {
            int sh = 0;
            switch (face) {
            case 0: sh = fabs(N->b[1] - b[0]) < tol; break;
            case 1: sh = fabs(N->b[0] - b[1]) < tol; break;
            case 2: sh = fabs(N->b[3] - b[2]) < tol; break;
            case 3: sh = fabs(N->b[2] - b[3]) < tol; break;
            case 4: sh = fabs(N->b[5] - b[4]) < tol; break;
            default: sh = fabs(N->b[4] - b[5]) < tol; break;
            }
            if (sh) {
                if (cnt < maxout)
                    out[cnt] = x;
                ++cnt;
            }
        } else {
            for (int c = 0; c < N->nchild; ++c)
                q[qt++] = N->child[c];
        }
    }
    free(q);
    return cnt;
}

orc_mesh *orc_mesh_create(double xmin, double xmax, double ymin, double ymax, double zmin,
                          double zmax)
{
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    m->dom[0] = xmin;
    m->dom[1] = xmax;
    m->dom[2] = ymin;
    m->dom[3] = ymax;
    m->dom[4] = zmin;
    m->dom[5] = zmax;
    return m;
}

void orc_mesh_destroy(orc_mesh *m)
{
    if (!m)
        return;
    free(m->nodes);
    free(m->leaves);
    free(m);
}

/* build_uniform :23-84 (breadth-first construction, leaf order depth-first) */
void orc_mesh_build_uniform(orc_mesh *m, int nx, int ny, int nz)
{
    int lx = 0, ly = 0, lz = 0, c;
    for (c = 1; c < nx; c *= 2)
        ++lx;
    for (c = 1; c < ny; c *= 2)
        ++ly;
    for (c = 1; c < nz; c *= 2)
        ++lz;
    m->nn = 0;
    const int64_t root = new_node(m);
    memcpy(m->nodes[root].b, m->dom, sizeof(m->dom));
    m->nodes[root].mask = 7;
    int64_t *q = NULL;
    int64_t qh = 0, qt = 0, qc = 0;
#define PUSH(v)                                                                              \
    do {                                                                                     \
        if (qt == qc) {                                                                      \
            qc = qc ? 2 * qc : 1024;                                                         \
            q = (int64_t *)realloc(q, sizeof(int64_t) * qc);                                 \
        }                                                                                    \
        q[qt++] = (v);                                                                       \
    } while (0)
    PUSH(root);
    while (qh < qt) {
        const int64_t id = q[qh++];
        const int nxn = m->nodes[id].lev[0] < lx, nyn = m->nodes[id].lev[1] < ly,
                  nzn = m->nodes[id].lev[2] < lz;
        if (!nxn && !nyn && !nzn)
            continue;
        m->nodes[id].mask = (nxn ? 1 : 0) | (nyn ? 2 : 0) | (nzn ? 4 : 0);
        create_children(m, id);
        for (int k = 0; k < m->nodes[id].nchild; ++k)
            PUSH(m->nodes[id].child[k]);
    }
#undef PUSH
    free(q);
    rebuild(m);
}

/* balance :138-182 */
void orc_mesh_balance(orc_mesh *m)
{
    int changed = 1;
    while (changed) {
        changed = 0;
        const int64_t nl = m->nl;
        int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (nl ? nl : 1));
        memcpy(cur, m->leaves, sizeof(int64_t) * nl);
        for (int64_t li = 0; li < nl; ++li) {
            const int64_t id = cur[li];
            for (int face = 0; face < 6; ++face) {
                const int64_t nb = neighbor_same_or_coarser(m, id, face);
                if (nb < 0)
                    continue;
                int needed = 0;
                if (m->nodes[id].lev[0] - m->nodes[nb].lev[0] > 1)
                    needed |= 1;
                if (m->nodes[id].lev[1] - m->nodes[nb].lev[1] > 1)
                    needed |= 2;
                if (m->nodes[id].lev[2] - m->nodes[nb].lev[2] > 1)
                    needed |= 4;
                if (needed && !m->nodes[nb].nchild) {
                    m->nodes[nb].mask = needed;
                    create_children(m, nb);
                    changed = 1;
                }
            }
        }
        free(cur);
        if (changed)
            rebuild(m);
    }
}

/* refine :184-203 */
void orc_mesh_refine(orc_mesh *m, int64_t count, const int64_t *elems, const int *masks)
{
    for (int64_t i = 0; i < count; ++i) {
        const int64_t e = elems[i];
        if (e < 0 || e >= m->nl)
            continue;
        const int64_t id = m->leaves[e];
        if (m->nodes[id].nchild)
            continue;
        m->nodes[id].mask = masks[i];
        create_children(m, id);
    }
    rebuild(m);
    orc_mesh_balance(m);
}

int64_t orc_mesh_num_elements(const orc_mesh *m) { return m->nl; }

void orc_mesh_elements(const orc_mesh *m, double *bounds, int *levels, uint64_t *morton)
{
    for (int64_t i = 0; i < m->nl; ++i) {
        const onode *N = &m->nodes[m->leaves[i]];
        if (bounds)
            memcpy(bounds + 6 * i, N->b, sizeof(N->b));
        if (levels)
            for (int a = 0; a < 3; ++a)
                levels[3 * i + a] = N->lev[a];
        if (morton)
            morton[i] = N->morton;
    }
}

int64_t orc_mesh_num_nodes(const orc_mesh *m) { return m->nn; }

/* every node of the tree in creation order (ids are stable: nodes are only
 * appended), leaves and interior nodes alike; mask is the split an interior
 * node was given, leaf its element number or -1 */
void orc_mesh_nodes(const orc_mesh *m, double *bounds, int *levels, int *mask, int *nchild,
                    int64_t *leaf)
{
    for (int64_t i = 0; i < m->nn; ++i) {
        const onode *N = &m->nodes[i];
        memcpy(bounds + 6 * i, N->b, sizeof(N->b));
        for (int a = 0; a < 3; ++a)
            levels[3 * i + a] = N->lev[a];
        mask[i] = N->mask;
        nchild[i] = N->nchild;
        leaf[i] = N->nchild ? -1 : N->leaf;
    }
}

/* The 2:1 mortar connectivity's neighbours: every leaf on the other side of
 * face f of leaf id whose face overlaps it with positive area, in depth-first
 * order (recursion over the children in creation order). */
static void adjacent_rec(const orc_mesh *m, int64_t cur, int64_t id, int f, int64_t *out, int maxout,
                         int *cnt)
{
    const double *b = m->nodes[id].b, *x = m->nodes[cur].b;
    const int a = f / 2;
    const double plane = (f % 2) ? b[2 * a + 1] : b[2 * a];
    const double tol = 1e-12 * (fabs(plane) + (b[2 * a + 1] - b[2 * a]));
    if (cur == id)
        return;
    if (f % 2) {
        if (!(x[2 * a] <= plane + tol && x[2 * a + 1] > plane + tol))
            return;
    } else if (!(x[2 * a + 1] >= plane - tol && x[2 * a] < plane - tol)) {
        return;
    }
    for (int t = 0; t < 3; ++t) {
        if (t == a)
            continue;
        const double lo = x[2 * t] > b[2 * t] ? x[2 * t] : b[2 * t];
        const double hi = x[2 * t + 1] < b[2 * t + 1] ? x[2 * t + 1] : b[2 * t + 1];
        if (!(hi - lo > 1e-12 * (b[2 * t + 1] - b[2 * t])))
            return;
    }
    if (!m->nodes[cur].nchild) {
        const double side = (f % 2) ? x[2 * a] : x[2 * a + 1];
        if (fabs(side - plane) <= tol) {
            if (*cnt < maxout)
                out[*cnt] = cur;
            ++*cnt;
        }
        return;
    }
    for (int c = 0; c < m->nodes[cur].nchild; ++c)
        adjacent_rec(m, m->nodes[cur].child[c], id, f, out, maxout, cnt);
}

static int adjacent(const orc_mesh *m, int64_t id, int f, int64_t *out, int maxout)
{
    int cnt = 0;
    adjacent_rec(m, 0, id, f, out, maxout, &cnt);
    return cnt;
}

/* get_neighbor :259-359 and build_face_connections :369-389 */
static void connections(const orc_mesh *m, int at_level, uint8_t *type, int8_t *count,
                        int64_t *elems, int8_t *faces, int8_t *subs)
{
    for (int64_t e = 0; e < m->nl; ++e) {
        const int64_t id = m->leaves[e];
        const double *b = m->nodes[id].b;
        for (int f = 0; f < 6; ++f) {
            const int64_t k = e * 6 + f;
            type[k] = ORC_CONN_BOUNDARY;
            count[k] = 0;
            for (int s = 0; s < 4; ++s) {
                elems[4 * k + s] = -1;
                faces[4 * k + s] = -1;
                subs[4 * k + s] = -1;
            }
            int at = 0;
            switch (f) {
            case 0: at = b[0] <= m->dom[0] + 1e-12; break;
            case 1: at = b[1] >= m->dom[1] - 1e-12; break;
            case 2: at = b[2] <= m->dom[2] + 1e-12; break;
            case 3: at = b[3] >= m->dom[3] - 1e-12; break;
            case 4: at = b[4] <= m->dom[4] + 1e-12; break;
            default: at = b[5] >= m->dom[5] - 1e-12; break;
            }
            if (at)
                continue;
            const int opp = (f % 2 == 0) ? f + 1 : f - 1;
            if (at_level) {
                int64_t adj[64];
                const int cnt = adjacent(m, id, f, adj, 64);
                if (!cnt)
                    continue;
                const onode *x = &m->nodes[adj[0]];
                type[k] = (uint8_t)(cnt == 1 ? ORC_CONN_SAME
                                             : connection_type(m->nodes[id].lev, x->lev, f));
                count[k] = (int8_t)(cnt < 4 ? cnt : 4);
                for (int s = 0; s < cnt && s < 4; ++s) {
                    elems[4 * k + s] = m->nodes[adj[s]].leaf;
                    faces[4 * k + s] = (int8_t)opp;
                    subs[4 * k + s] = (int8_t)s;
                }
                continue;
            }
            const int64_t nb = neighbor_same_or_coarser(m, id, f);
            if (nb < 0)
                continue;
            if (!m->nodes[nb].nchild) {
                type[k] = ORC_CONN_SAME; /* :331-338: SameLevel either way */
                count[k] = 1;
                elems[4 * k] = m->nodes[nb].leaf;
                faces[4 * k] = (int8_t)opp;
                subs[4 * k] = 0;
            } else {
                int64_t fine[64];
                const int cnt = fine_neighbors(m, id, f, fine, 64);
                /* fine_neighbors[0] is read unconditionally by the reference (:349) */
                type[k] = (uint8_t)connection_type(m->nodes[id].lev, m->nodes[fine[0]].lev, f);
                count[k] = (int8_t)(cnt < 4 ? cnt : 4);
                for (int s = 0; s < cnt && s < 4; ++s) {
                    elems[4 * k + s] = m->nodes[fine[s]].leaf;
                    faces[4 * k + s] = (int8_t)opp;
                    subs[4 * k + s] = (int8_t)s;
                }
            }
        }
    }
}

void orc_mesh_connections(const orc_mesh *m, uint8_t *type, int8_t *count, int64_t *elems,
                          int8_t *faces, int8_t *subs)
{
    connections(m, 0, type, count, elems, faces, subs);
}

/* The 2:1 mortar connectivity: every face lists the leaves across it that
 * share a piece of it (depth-first order); one leaf is SameLevel, several are
 * typed by get_connection_type from the first, as get_neighbor's
 * refined-neighbour branch (:344-356) would type them. */
void orc_mesh_connections_2to1(const orc_mesh *m, uint8_t *type, int8_t *count, int64_t *elems,
                               int8_t *faces, int8_t *subs)
{
    connections(m, 1, type, count, elems, faces, subs);
}
