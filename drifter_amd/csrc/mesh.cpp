// Directional octree mesh, 2:1 balance and face connectivity (host side).
//
// Replaces OctreeAdapter (include/mesh/octree_adapter.hpp:103-273):
// build_uniform (src/mesh/octree_adapter.cpp:23-84), refine (:184-203),
// balance (:138-182), element_bounds (:234-239) and build_face_connections
// (:369-389 via get_neighbor :259-359).  Element numbering, bounds arithmetic,
// tie-breaks and connection types are the reference's, so the flat index maps
// handed to the face kernels are bit-identical to what the reference would
// build:
//   - children are created iz -> iy -> ix with bounds min + i * (max - min)/c;
//   - elements are numbered in depth-first leaf order;
//   - a face's neighbour is the leaf containing the centre of the would-be
//     same-size neighbour (inclusive bounds, first child wins on ties),
//     reported as SameLevel; only when that descent stops at an interior node
//     are the fine leaves sharing the face listed in breadth-first order;
//   - a face is a boundary face when it lies within 1e-12 of the domain.
// Storage differs from the reference's pointer tree: nodes live in one
// vector with contiguous children, and traversals use explicit stacks/queues.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "drf_internal.hpp"

struct drf_mesh {
    struct Node {
        double b[6];      // xmin xmax ymin ymax zmin zmax
        int lev[3];       // per-axis refinement level
        uint64_t morton;  // directional Morton code
        int mask;         // refinement mask used for the children (bit0 X, bit1 Y, bit2 Z)
        int64_t child0;   // first child (children are contiguous), -1 for a leaf
        int nchild;
        int64_t leaf;     // element number, -1 if not a leaf
        int64_t parent;   // -1 for the root
    };
    double dom[6];
    std::vector<Node> nodes;
    std::vector<int64_t> leaves;
};

namespace {

using Node = drf_mesh::Node;

constexpr uint64_t kXMask = 0x1249249249249249ULL; // Morton3D masks, include/mesh/morton.hpp:33-35
constexpr uint64_t kYMask = 0x2492492492492492ULL;
constexpr uint64_t kZMask = 0x4924924924924924ULL;

// MortonUtil::child_{x,y,z} / refine, include/mesh/morton.hpp:104-143
uint64_t morton_child(uint64_t code, uint64_t mask, uint64_t bitval, bool set)
{
    const uint64_t bits = ((code & mask) << 3) | (set ? bitval : 0ULL);
    return (code & ~mask) | (bits & mask);
}

void create_children(drf_mesh &m, int64_t id)
{
    if (m.nodes[id].nchild)
        return;
    const int mask = m.nodes[id].mask;
    const bool rx = mask & 1, ry = mask & 2, rz = mask & 4;
    const int cx = rx ? 2 : 1, cy = ry ? 2 : 1, cz = rz ? 2 : 1;
    const Node p = m.nodes[id];
    const double dx = (p.b[1] - p.b[0]) / cx, dy = (p.b[3] - p.b[2]) / cy,
                 dz = (p.b[5] - p.b[4]) / cz;
    const int64_t first = (int64_t)m.nodes.size();
    for (int iz = 0; iz < cz; ++iz)
        for (int iy = 0; iy < cy; ++iy)
            for (int ix = 0; ix < cx; ++ix) {
                Node c{};
                uint64_t code = p.morton;
                if (rx)
                    code = morton_child(code, kXMask, 0x1, ix);
                if (ry)
                    code = morton_child(code, kYMask, 0x2, iy);
                if (rz)
                    code = morton_child(code, kZMask, 0x4, iz);
                c.morton = code;
                c.lev[0] = p.lev[0] + (rx ? 1 : 0);
                c.lev[1] = p.lev[1] + (ry ? 1 : 0);
                c.lev[2] = p.lev[2] + (rz ? 1 : 0);
                c.b[0] = p.b[0] + ix * dx;
                c.b[1] = p.b[0] + (ix + 1) * dx;
                c.b[2] = p.b[2] + iy * dy;
                c.b[3] = p.b[2] + (iy + 1) * dy;
                c.b[4] = p.b[4] + iz * dz;
                c.b[5] = p.b[4] + (iz + 1) * dz;
                c.child0 = -1;
                c.leaf = -1;
                c.parent = id;
                m.nodes.push_back(c);
            }
    m.nodes[id].child0 = first;
    m.nodes[id].nchild = cx * cy * cz;
}

void number_leaves(drf_mesh &m)
{
    for (auto &nd : m.nodes)
        nd.leaf = -1;
    m.leaves.clear();
    std::vector<int64_t> stack{0};
    while (!stack.empty()) {
        const int64_t id = stack.back();
        stack.pop_back();
        const Node &nd = m.nodes[id];
        if (!nd.nchild) {
            m.nodes[id].leaf = (int64_t)m.leaves.size();
            m.leaves.push_back(id);
            continue;
        }
        for (int c = nd.nchild - 1; c >= 0; --c) // depth first, children in creation order
            stack.push_back(nd.child0 + c);
    }
}

// Renumbers after splits: a split leaf's place in the depth-first order is
// taken by its subtree's leaves, so the previous leaf list with each split
// leaf expanded depth first is the new order (the same numbering as
// number_leaves, without walking the unsplit part of the tree).
void renumber_leaves(drf_mesh &m)
{
    std::vector<int64_t> old;
    old.swap(m.leaves);
    m.leaves.reserve(m.nodes.size());
    std::vector<int64_t> stack;
    for (int64_t id0 : old) {
        if (!m.nodes[id0].nchild) {
            m.nodes[id0].leaf = (int64_t)m.leaves.size();
            m.leaves.push_back(id0);
            continue;
        }
        stack.assign(1, id0);
        while (!stack.empty()) {
            const int64_t id = stack.back();
            stack.pop_back();
            Node &nd = m.nodes[id];
            if (!nd.nchild) {
                nd.leaf = (int64_t)m.leaves.size();
                m.leaves.push_back(id);
                continue;
            }
            nd.leaf = -1;
            for (int c = nd.nchild - 1; c >= 0; --c)
                stack.push_back(nd.child0 + c);
        }
    }
}

bool inside(const double *b, const double *p)
{
    return p[0] >= b[0] && p[0] <= b[1] && p[1] >= b[2] && p[1] <= b[3] && p[2] >= b[4] &&
           p[2] <= b[5];
}

// centre of the same-size box across face f (find_neighbor_same_or_coarser :530-558)
void across(const double *b, int f, double *p)
{
    p[0] = 0.5 * (b[0] + b[1]);
    p[1] = 0.5 * (b[2] + b[3]);
    p[2] = 0.5 * (b[4] + b[5]);
    const int a = f / 2;
    const double lo = b[2 * a], hi = b[2 * a + 1];
    p[a] = (f % 2) ? hi + 0.5 * (hi - lo) : lo - 0.5 * (hi - lo);
}

// The deepest node containing the centre p of the same-size box across
// face f of node id (the first child containing p at every level, as the
// reference descends from the root), reached from the nearest ancestor of id
// that contains p instead of from the root.  p lies half a
// cell of id's levels off every grid line of those levels, so it is strictly
// inside that ancestor and every node above it: the root's descent passes
// through the ancestor, and below it both descents take the same children.
int64_t neighbour_node(const drf_mesh &m, int64_t id, int f)
{
    double p[3];
    across(m.nodes[id].b, f, p);
    if (!inside(m.dom, p))
        return -1;
    int64_t cur = m.nodes[id].parent;
    while (cur > 0 && !inside(m.nodes[cur].b, p))
        cur = m.nodes[cur].parent;
    if (cur < 0)
        cur = 0;
    while (m.nodes[cur].nchild) {
        // A child's bounds on one axis depend only on its index on that axis
        // (create_children), so the children containing p are a product of
        // per-axis index sets, and the first of them in creation order takes
        // on each axis the first interval containing p: read it off the
        // first child's bounds, then test the one child so found.
        const Node &nd = m.nodes[cur];
        const Node &c0 = m.nodes[nd.child0];
        const int cx = (nd.mask & 1) ? 2 : 1, cy = (nd.mask & 2) ? 2 : 1;
        const int ix = (cx == 2 && p[0] > c0.b[1]) ? 1 : 0;
        const int iy = (cy == 2 && p[1] > c0.b[3]) ? 1 : 0;
        const int iz = ((nd.mask & 4) && p[2] > c0.b[5]) ? 1 : 0;
        const int64_t next = nd.child0 + (iz * cy + iy) * cx + ix;
        if (!inside(m.nodes[next].b, p))
            break;
        cur = next;
    }
    return cur;
}

// Every leaf across face f of node id that shares a piece of the face of
// positive area, in depth-first (element) order: a descent into the
// children that reach across the face plane and overlap the face on both
// tangent axes (the 2:1 mortar connectivity, not the reference's).
void adjacent_leaves(const drf_mesh &m, int64_t id, int f, std::vector<int64_t> &out,
                     std::vector<int64_t> &stack)
{
    out.clear();
    const double *b = m.nodes[id].b;
    const int a = f / 2, t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
    const double plane = (f % 2) ? b[2 * a + 1] : b[2 * a];
    const double tol = 1e-12 * (std::fabs(plane) + (b[2 * a + 1] - b[2 * a]));
    auto reaches = [&](const double *x) {
        if (f % 2) // the neighbour lies above the plane and touches it from there
            return x[2 * a] <= plane + tol && x[2 * a + 1] > plane + tol;
        return x[2 * a + 1] >= plane - tol && x[2 * a] < plane - tol;
    };
    auto overlaps = [&](const double *x, int t) {
        const double lo = std::max(x[2 * t], b[2 * t]), hi = std::min(x[2 * t + 1], b[2 * t + 1]);
        return hi - lo > 1e-12 * (b[2 * t + 1] - b[2 * t]);
    };
    // Every such leaf meets the interior of the same-size box across the
    // face next to the face, so it lies in any node containing that box
    // (the leaves partition each node): the descent starts at the nearest
    // ancestor that does (the root at the latest) and lists what a descent
    // from the root would, in the same order.  A bound that rounding moves
    // off the box only sends the search one ancestor further up.
    double box[6];
    std::memcpy(box, b, sizeof(box));
    box[2 * a] = (f % 2) ? plane : 2 * plane - b[2 * a + 1];
    box[2 * a + 1] = (f % 2) ? 2 * plane - b[2 * a] : plane;
    int64_t top = m.nodes[id].parent;
    while (top > 0) {
        const double *t = m.nodes[top].b;
        if (t[0] <= box[0] && t[1] >= box[1] && t[2] <= box[2] && t[3] >= box[3] &&
            t[4] <= box[4] && t[5] >= box[5])
            break;
        top = m.nodes[top].parent;
    }
    stack.assign(1, top > 0 ? top : 0);
    while (!stack.empty()) {
        const int64_t cur = stack.back();
        stack.pop_back();
        const Node &x = m.nodes[cur];
        if (cur == id || !reaches(x.b) || !overlaps(x.b, t1) || !overlaps(x.b, t2))
            continue;
        if (!x.nchild) {
            const double side = (f % 2) ? x.b[2 * a] : x.b[2 * a + 1];
            if (std::fabs(side - plane) <= tol)
                out.push_back(cur);
            continue;
        }
        for (int c = x.nchild - 1; c >= 0; --c)
            stack.push_back(x.child0 + c);
    }
}

// fn(i0, i1) over [0, n) in contiguous chunks on up to 16 threads (one chunk
// below 4,096 items); fn must only read shared state and write its own range.
template <class F> void parallel_chunks(int64_t n, F fn)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const int64_t nt = std::min<int64_t>({n / 4096 + 1, hw ? (int64_t)hw : 1, 16});
    if (nt <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int64_t t = 1; t < nt; ++t) {
        const int64_t i0 = std::min(n, t * chunk), i1 = std::min(n, (t + 1) * chunk);
        try {
            pool.emplace_back([=] { fn(i0, i1); });
        } catch (...) { // no thread available: this chunk runs here
            fn(i0, i1);
        }
    }
    fn((int64_t)0, std::min(n, chunk));
    for (auto &th : pool)
        th.join();
}

// Does leaf id have a face neighbour more than one level coarser on some axis?
bool needs_split(const drf_mesh &m, int64_t id)
{
    const Node &nd = m.nodes[id];
    const Node *par = nd.parent >= 0 ? &m.nodes[nd.parent] : nullptr;
    for (int f = 0; f < 6; ++f) {
        // a face between two children of the parent has a sibling of the
        // same size across it (or that sibling's finer leaves), never a
        // coarser leaf; the lower child's lower bound is the parent's
        const int a = f / 2;
        if (par && (par->mask >> a & 1)) {
            const bool upper_child = nd.b[2 * a] != par->b[2 * a];
            if ((f % 2 == 1) != upper_child)
                continue;
        }
        const int64_t nb = neighbour_node(m, id, f);
        if (nb < 0 || m.nodes[nb].nchild)
            continue;
        for (int ax = 0; ax < 3; ++ax)
            if (nd.lev[ax] - m.nodes[nb].lev[ax] > 1)
                return true;
    }
    return false;
}

// 2:1 balance, pass by pass over the leaves in leaf order, splitting a
// neighbour more than one level coarser on the axes that need it.  A leaf
// with no such neighbour gets none later (splits only make neighbours finer),
// so a pass checks, in leaf order, only the leaves that can still need a
// split: after the first pass, those the previous pass created and those
// whose check split a neighbour; and of these only the ones that a parallel
// filter against the tree at the start of the pass flags.  The splits and
// their order are those of rescanning every leaf in every pass.
//
// first_new: nodes from this index on were created since the tree was last
// balanced (0: check every leaf).  Splits only make a leaf's neighbours
// finer, so a leaf of a balanced tree keeps needing none: the first pass
// checks the new leaves only, and finds the splits (in leaf order) of a
// full scan.
void balance(drf_mesh &m, int64_t first_new = 0)
{
    bool changed = true;
    std::vector<int64_t> snapshot;
    if (first_new > 0) {
        for (int64_t id : m.leaves)
            if (id >= first_new)
                snapshot.push_back(id);
    } else {
        snapshot = m.leaves;
    }
    while (changed) {
        changed = false;
        const int64_t first_new = (int64_t)m.nodes.size();
        std::vector<char> flag(snapshot.size());
        parallel_chunks((int64_t)snapshot.size(), [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i)
                flag[i] = needs_split(m, snapshot[i]);
        });
        std::vector<int64_t> work;
        for (size_t i = 0; i < snapshot.size(); ++i)
            if (flag[i])
                work.push_back(snapshot[i]);
        std::vector<int64_t> split_by; // leaves whose check split a neighbour
        for (int64_t id : work) {
            bool split = false;
            for (int f = 0; f < 6; ++f) {
                const int64_t nb = neighbour_node(m, id, f);
                if (nb < 0)
                    continue;
                int need = 0;
                for (int a = 0; a < 3; ++a)
                    if (m.nodes[id].lev[a] - m.nodes[nb].lev[a] > 1)
                        need |= 1 << a;
                if (need && !m.nodes[nb].nchild) {
                    m.nodes[nb].mask = need;
                    create_children(m, nb);
                    changed = split = true;
                }
            }
            if (split)
                split_by.push_back(id);
        }
        if (changed) {
            // next pass: the leaves this pass created, and the leaves that
            // split a neighbour (one split narrows a level gap by one only,
            // so they may still need another), in leaf order
            renumber_leaves(m);
            std::vector<char> again(m.nodes.size(), 0);
            for (int64_t id : split_by)
                again[id] = 1;
            snapshot.clear();
            for (int64_t id : m.leaves)
                if (id >= first_new || again[id])
                    snapshot.push_back(id);
        }
    }
}

int conn_type(const int *coarse, const int *fine, int f)
{
    // get_connection_type, include/dg/face_connection.hpp:110-152
    const int t1 = (f <= 1) ? 1 : 0, t2 = (f <= 3) ? 2 : 1;
    const int d1 = fine[t1] - coarse[t1], d2 = fine[t2] - coarse[t2];
    if (d1 == 0 && d2 == 0)
        return 0; // SameLevel
    if (d1 == 1 && d2 == 0)
        return 1; // Fine2x1
    if (d1 == 0 && d2 == 1)
        return 2; // Fine1x2
    return 5;     // Fine2x2 (also the 1:3 fallback)
}

} // namespace

extern "C" {

drf_status drf_mesh_create(double xmin, double xmax, double ymin, double ymax, double zmin,
                           double zmax, drf_mesh **out)
{
    if (!out)
        return drf::fail(DRF_ERR_ARG, "null output pointer");
    if (!(xmax > xmin && ymax > ymin && zmax > zmin))
        return drf::fail(DRF_ERR_ARG, "empty domain");
    drf_mesh *m = new drf_mesh();
    const double d[6] = {xmin, xmax, ymin, ymax, zmin, zmax};
    std::memcpy(m->dom, d, sizeof(d));
    Node root{};
    std::memcpy(root.b, d, sizeof(d));
    root.child0 = -1;
    root.parent = -1;
    root.mask = 7;
    m->nodes.push_back(root);
    number_leaves(*m);
    *out = m;
    return DRF_OK;
}

void drf_mesh_destroy(drf_mesh *m) { delete m; }

drf_status drf_mesh_build_uniform(drf_mesh *m, int nx, int ny, int nz)
{
    if (!m || nx < 1 || ny < 1 || nz < 1)
        return drf::fail(DRF_ERR_ARG, "invalid mesh or element counts");
    int lv[3] = {0, 0, 0};
    const int cnt[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a)
        for (int c = 1; c < cnt[a]; c *= 2)
            ++lv[a];
    Node root = m->nodes[0];
    root.child0 = -1;
    root.parent = -1;
    root.nchild = 0;
    root.mask = 7;
    root.lev[0] = root.lev[1] = root.lev[2] = 0;
    root.morton = 0;
    m->nodes.assign(1, root);
    // breadth-first construction (:49-81)
    std::vector<int64_t> queue{0};
    for (size_t qh = 0; qh < queue.size(); ++qh) {
        const int64_t id = queue[qh];
        int mask = 0;
        for (int a = 0; a < 3; ++a)
            if (m->nodes[id].lev[a] < lv[a])
                mask |= 1 << a;
        if (!mask)
            continue;
        m->nodes[id].mask = mask;
        create_children(*m, id);
        for (int c = 0; c < m->nodes[id].nchild; ++c)
            queue.push_back(m->nodes[id].child0 + c);
    }
    number_leaves(*m);
    return DRF_OK;
}

drf_status drf_mesh_refine(drf_mesh *m, int64_t count, const int64_t *elems, const uint8_t *masks)
{
    if (!m || (count > 0 && (!elems || !masks)))
        return drf::fail(DRF_ERR_ARG, "null argument");
    // room for every child up front: growing the node array child by child
    // copied it several times over (a third of the refine at 110,000 leaves)
    if (count > 0)
        m->nodes.reserve(m->nodes.size() +
                         8 * std::min<size_t>((size_t)count, m->leaves.size()));
    // every mesh this API hands out is 2:1 balanced (build_uniform, refine
    // and balance leave it so): the balance below starts from the new leaves
    const int64_t first_new = (int64_t)m->nodes.size();
    for (int64_t i = 0; i < count; ++i) {
        const int64_t e = elems[i];
        if (e < 0 || e >= (int64_t)m->leaves.size())
            continue; // the reference skips bad indices (:190-191)
        const int64_t id = m->leaves[e];
        if (m->nodes[id].nchild)
            continue;
        m->nodes[id].mask = masks[i] & 7;
        create_children(*m, id);
    }
    renumber_leaves(*m);
    balance(*m, first_new);
    return DRF_OK;
}

drf_status drf_mesh_balance(drf_mesh *m)
{
    if (!m)
        return drf::fail(DRF_ERR_ARG, "null mesh");
    balance(*m);
    return DRF_OK;
}

int64_t drf_mesh_num_elements(const drf_mesh *m) { return m ? (int64_t)m->leaves.size() : 0; }

drf_status drf_mesh_elements(const drf_mesh *m, double *bounds, int32_t *levels, uint64_t *morton)
{
    if (!m)
        return drf::fail(DRF_ERR_ARG, "null mesh");
    for (size_t i = 0; i < m->leaves.size(); ++i) {
        const Node &nd = m->nodes[m->leaves[i]];
        if (bounds)
            std::memcpy(bounds + 6 * i, nd.b, sizeof(nd.b));
        if (levels)
            for (int a = 0; a < 3; ++a)
                levels[3 * i + a] = nd.lev[a];
        if (morton)
            morton[i] = nd.morton;
    }
    return DRF_OK;
}

drf_status drf_mesh_element_bounds(const drf_mesh *m, int64_t elem, double *b6)
{
    if (!m || !b6)
        return drf::fail(DRF_ERR_ARG, "null argument");
    if (elem < 0 || elem >= (int64_t)m->leaves.size())
        return drf::fail(DRF_ERR_RANGE, "Invalid element index"); // std::out_of_range, :235-237
    std::memcpy(b6, m->nodes[m->leaves[elem]].b, 6 * sizeof(double));
    return DRF_OK;
}

} // extern "C"

namespace {

drf_status face_connections(const drf_mesh *m, bool at_level, uint8_t *type, int8_t *count,
                            int64_t *elems, int8_t *faces, int8_t *subs)
{
    if (!m || !type || !count || !elems || !faces || !subs)
        return drf::fail(DRF_ERR_ARG, "null argument");
    const int64_t E = (int64_t)m->leaves.size();
    // elements are independent: each writes its own six faces
    parallel_chunks(E, [&](int64_t e0, int64_t e1) {
    std::vector<int64_t> adj, stack, fine; // reused by the chunk's queries
    for (int64_t e = e0; e < e1; ++e) {
        const int64_t id = m->leaves[e];
        const Node &nd = m->nodes[id];
        for (int f = 0; f < 6; ++f) {
            const int64_t k = 6 * e + f;
            type[k] = 6; // Boundary
            count[k] = 0;
            for (int s = 0; s < 4; ++s) {
                elems[4 * k + s] = -1;
                faces[4 * k + s] = -1;
                subs[4 * k + s] = -1;
            }
            const int a = f / 2;
            const bool on_boundary = (f % 2) ? nd.b[2 * a + 1] >= m->dom[2 * a + 1] - 1e-12
                                             : nd.b[2 * a] <= m->dom[2 * a] + 1e-12;
            if (on_boundary)
                continue;
            const int8_t opp = (int8_t)((f % 2 == 0) ? f + 1 : f - 1);
            if (at_level) {
                adjacent_leaves(*m, id, f, adj, stack);
                if (adj.empty())
                    continue;
                const Node &x = m->nodes[adj[0]];
                const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
                const bool covers = adj.size() == 1 && x.b[2 * t1] <= nd.b[2 * t1] &&
                                    x.b[2 * t1 + 1] >= nd.b[2 * t1 + 1] &&
                                    x.b[2 * t2] <= nd.b[2 * t2] &&
                                    x.b[2 * t2 + 1] >= nd.b[2 * t2 + 1];
                type[k] = covers ? 0 : (uint8_t)conn_type(nd.lev, x.lev, f);
                count[k] = (int8_t)std::min<size_t>(adj.size(), 4);
                for (size_t s = 0; s < adj.size() && s < 4; ++s) {
                    elems[4 * k + s] = m->nodes[adj[s]].leaf;
                    faces[4 * k + s] = opp;
                    subs[4 * k + s] = (int8_t)s;
                }
                continue;
            }
            const int64_t nb = neighbour_node(*m, id, f);
            if (nb < 0)
                continue;
            if (!m->nodes[nb].nchild) {
                type[k] = 0; // SameLevel, also when the neighbour is coarser (:331-338)
                count[k] = 1;
                elems[4 * k] = m->nodes[nb].leaf;
                faces[4 * k] = opp;
                subs[4 * k] = 0;
                continue;
            }
            // fine leaves sharing the face, breadth-first from the node (:582-639)
            std::vector<int64_t> &q = stack;
            q.assign(1, nb);
            fine.clear();
            const double plane = (f % 2) ? nd.b[2 * a + 1] : nd.b[2 * a];
            for (size_t qh = 0; qh < q.size(); ++qh) {
                const Node &x = m->nodes[q[qh]];
                if (!x.nchild) {
                    const double side = (f % 2) ? x.b[2 * a] : x.b[2 * a + 1];
                    if (std::fabs(side - plane) < 1e-12)
                        fine.push_back(q[qh]);
                } else {
                    for (int c = 0; c < x.nchild; ++c)
                        q.push_back(x.child0 + c);
                }
            }
            if (fine.empty())
                continue;
            type[k] = (uint8_t)conn_type(nd.lev, m->nodes[fine[0]].lev, f);
            count[k] = (int8_t)std::min<size_t>(fine.size(), 4);
            for (size_t s = 0; s < fine.size() && s < 4; ++s) {
                elems[4 * k + s] = m->nodes[fine[s]].leaf;
                faces[4 * k + s] = opp;
                subs[4 * k + s] = (int8_t)s;
            }
        }
    }
    });
    return DRF_OK;
}

} // namespace

extern "C" {

drf_status drf_mesh_face_connections(const drf_mesh *m, uint8_t *type, int8_t *count,
                                     int64_t *elems, int8_t *faces, int8_t *subs)
{
    return face_connections(m, false, type, count, elems, faces, subs);
}

drf_status drf_mesh_face_connections_2to1(const drf_mesh *m, uint8_t *type, int8_t *count,
                                          int64_t *elems, int8_t *faces, int8_t *subs)
{
    return face_connections(m, true, type, count, elems, faces, subs);
}

} // extern "C"
