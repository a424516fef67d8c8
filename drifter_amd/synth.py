"""Synthetic meshes, bathymetry and ocean states for tests and benchmarks.

The reference ships no seeded generator or data for its DG path, so these
definitions are this project's own:

* Uniform octree meshes follow OctreeAdapter::build_uniform
  (src/mesh/octree_adapter.cpp:23-84): the root is split per axis until the
  requested counts are reached, children are created iz -> iy -> ix
  (create_children :494-520, bounds b.min + i * (b.max - b.min) / 2), and the
  element number is the depth-first leaf order (collect_leaves :460-468).
* Node positions: x = centre + 0.5 * xi_i * size, as the reference's driver
  test does (tests/integration/test_simulation_full.cpp:54-64).
* State: for every field q in (u, v, eta, T, S) a sum of 8 Fourier modes
  a_m cos(2 pi (kx x/Lx + ky y/Ly + kz s) + phi_m), s = (zeta + 1)/2, with
  integers kx, ky in 0..3, kz in 0..1 (0 for eta), a_m ~ U(-1,1)/m and
  phi_m ~ U(0, 2 pi) drawn from splitmix64 seeded with 20261202 + q, then
  normalised to max |F| = 1.  u, v = 0.1 F; eta = 0.05 F; T = 15 + 5 F;
  S = 32.5 + 2.5 F; H = eta + h and the conserved state is H * (u, v, T, S).
"""
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea & Flood 2014): integer-exact on every host."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def uniform_mesh(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax):
    """Element bounds (E x 6: xmin, xmax, ymin, ymax, zmin, zmax) of
    OctreeAdapter::build_uniform, in leaf (element) order."""

    def levels(cnt):
        lv, c = 0, 1
        while c < cnt:
            c *= 2
            lv += 1
        return lv

    lx, ly, lz = levels(nx), levels(ny), levels(nz)
    out = []

    def rec(b, lev):
        need = (lev[0] < lx, lev[1] < ly, lev[2] < lz)
        if not any(need):
            out.append(b)
            return
        cx, cy, cz = (2 if need[0] else 1), (2 if need[1] else 1), (2 if need[2] else 1)
        dx = (b[1] - b[0]) / cx
        dy = (b[3] - b[2]) / cy
        dz = (b[5] - b[4]) / cz
        nl = (lev[0] + need[0], lev[1] + need[1], lev[2] + need[2])
        for iz in range(cz):
            for iy in range(cy):
                for ix in range(cx):
                    rec((b[0] + ix * dx, b[0] + (ix + 1) * dx,
                         b[2] + iy * dy, b[2] + (iy + 1) * dy,
                         b[4] + iz * dz, b[4] + (iz + 1) * dz), nl)

    rec((xmin, xmax, ymin, ymax, zmin, zmax), (0, 0, 0))
    return np.array(out, dtype=np.float64)


def lgl_nodes(order):
    """LGL nodes of the product library (bit-identical to the reference's)."""
    lib = _lib.load()
    n = order + 1
    x = np.zeros(n)
    w = np.zeros(n)
    D = np.zeros(n * n)
    _lib.check(lib.drf_basis(order, _lib.dptr(x), _lib.dptr(w), _lib.dptr(D)))
    return x, w, D.reshape(n, n)


def node_coords(bounds, nodes):
    """x, y, zeta-fraction s at every DOF: arrays of shape (E, ndof), d = i + n(j + n k)."""
    n = len(nodes)
    E = bounds.shape[0]
    cx = 0.5 * (bounds[:, 0] + bounds[:, 1])
    cy = 0.5 * (bounds[:, 2] + bounds[:, 3])
    sx = bounds[:, 1] - bounds[:, 0]
    sy = bounds[:, 3] - bounds[:, 2]
    x1 = cx[:, None] + 0.5 * nodes[None, :] * sx[:, None]  # (E, n) along i
    y1 = cy[:, None] + 0.5 * nodes[None, :] * sy[:, None]  # (E, n) along j
    s1 = 0.5 * (nodes + 1.0)                                # along k
    x = np.broadcast_to(x1[:, None, None, :], (E, n, n, n)).reshape(E, n ** 3)
    y = np.broadcast_to(y1[:, None, :, None], (E, n, n, n)).reshape(E, n ** 3)
    s = np.broadcast_to(s1[None, :, None, None], (E, n, n, n)).reshape(E, n ** 3)
    return np.ascontiguousarray(x), np.ascontiguousarray(y), np.ascontiguousarray(s)


def fourier_modes(q, seed_base=20261202):
    """The 8 modes (kx, ky, kz, a, phi) of field q, drawn from splitmix64."""
    rng = SplitMix64(seed_base + q)
    modes = []
    for m in range(1, 9):
        kx = rng.next() % 4
        ky = rng.next() % 4
        kz = rng.next() % 2
        if q == 2:  # eta is a surface field
            kz = 0
        a = (2.0 * rng.uniform() - 1.0) / m
        phi = 2.0 * math.pi * rng.uniform()
        modes.append((kx, ky, kz, a, phi))
    return modes


def fourier_sum(q, x, y, s, Lx, Ly, seed_base=20261202):
    """Un-normalised sum of field q's 8 modes at the points (x, y, s)."""
    F = np.zeros_like(x)
    for kx, ky, kz, a, phi in fourier_modes(q, seed_base):
        F += a * np.cos(2.0 * math.pi * (kx * x / Lx + ky * y / Ly + kz * s) + phi)
    return F


def fourier_field(q, x, y, s, Lx, Ly, seed_base=20261202):
    F = fourier_sum(q, x, y, s, Lx, Ly, seed_base)
    return F / np.max(np.abs(F))


def bathymetry(kind, x, y, Lx, Ly):
    if kind == "seamount":  # flat 50 m with a 30 m Gaussian seamount of sigma 2 km
        r2 = (x - 0.5 * Lx) ** 2 + (y - 0.5 * Ly) ** 2
        return 50.0 - 30.0 * np.exp(-r2 / (2.0 * 2000.0 ** 2))
    if kind == "shelf":  # tanh shelf from 20 m to 200 m
        return 110.0 + 90.0 * np.tanh((x - 0.5 * Lx) / (Lx / 20.0))
    if kind == "coast":  # 20 m at the sinusoidal coastline, 200 m offshore
        yc = 0.8 * Ly + 0.05 * Ly * np.sin(2.0 * np.pi * x / (0.25 * Lx))
        return 20.0 + 180.0 * np.tanh(np.maximum(yc - y, 0.0) / (0.2 * Ly))
    if kind == "flat":
        return np.full_like(x, 50.0)
    raise ValueError(kind)


class Case:
    """A seeded mesh + bathymetry + state.  Arrays are element-major."""

    def __init__(self, order, bounds, h, y, U, Lx, Ly):
        self.order = order
        self.n = order + 1
        self.nd = self.n ** 3
        self.bounds = bounds
        self.E = bounds.shape[0]
        self.h = h          # (E*nd,)
        self.y = y          # (E*nd,)
        self.U = U          # packed (E*5*nd,)
        self.Lx, self.Ly = Lx, Ly

    def full_state(self):
        """[e][Hu|Hv|eta|HT|HS|H|u|v|T|S|omega|rho][d] with derived fields from
        PrimitiveState::update_derived (omega = rho = 0)."""
        E, nd = self.E, self.nd
        P = self.U.reshape(E, 5, nd)
        H = P[:, 2] + self.h.reshape(E, nd)
        Hinv = np.where(H > 1e-10, 1.0 / np.where(H > 1e-10, H, 1.0), 0.0)
        F = np.zeros((E, 12, nd))
        F[:, :5] = P
        F[:, 5] = H
        F[:, 6] = P[:, 0] * Hinv
        F[:, 7] = P[:, 1] * Hinv
        F[:, 8] = P[:, 3] * Hinv
        F[:, 9] = P[:, 4] * Hinv
        return np.ascontiguousarray(F.reshape(-1))


def make_case(nx, ny, nz, order, L=10000.0, bathy="seamount", seed_base=20261202):
    bounds = uniform_mesh(nx, ny, nz, 0.0, L, 0.0, L, -1.0, 0.0)
    return make_case_on(bounds, order, L, bathy, seed_base)


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def make_case_on(bounds, order, L, bathy="seamount", seed_base=20261202, threads=None):
    """Seeded state on any element list (bounds (E, 6), domain [0, L]^2 x [-1, 0]).

    Every value is a function of its own node, so the elements are evaluated
    in chunks on a thread pool (numpy releases the interpreter lock inside its
    loops); the fields are normalised by their maxima over the whole mesh in a
    second pass.  The result does not depend on the chunking."""
    nodes, _, _ = lgl_nodes(order)
    bounds = np.ascontiguousarray(bounds, dtype=np.float64)
    E = bounds.shape[0]
    nd = (order + 1) ** 3
    h = np.empty((E, nd))
    y = np.empty((E, nd))
    F = np.empty((5, E, nd))
    U = np.empty((E, 5, nd))
    chunk = 4096
    starts = list(range(0, E, chunk))
    fmax = np.zeros((len(starts), 5))

    def pass1(c):
        e0 = starts[c]
        e1 = min(E, e0 + chunk)
        xc, yc, sc = node_coords(bounds[e0:e1], nodes)
        h[e0:e1] = bathymetry(bathy, xc, yc, L, L)
        y[e0:e1] = yc
        for q in range(5):
            F[q, e0:e1] = fourier_sum(q, xc, yc, sc, L, L, seed_base)
            fmax[c, q] = np.max(np.abs(F[q, e0:e1]))

    def pass2(c, M):
        e0 = starts[c]
        e1 = min(E, e0 + chunk)
        u = 0.1 * (F[0, e0:e1] / M[0])
        v = 0.1 * (F[1, e0:e1] / M[1])
        eta = 0.05 * (F[2, e0:e1] / M[2])
        T = 15.0 + 5.0 * (F[3, e0:e1] / M[3])
        S = 32.5 + 2.5 * (F[4, e0:e1] / M[4])
        H = eta + h[e0:e1]
        U[e0:e1, 0] = H * u
        U[e0:e1, 1] = H * v
        U[e0:e1, 2] = eta
        U[e0:e1, 3] = H * T
        U[e0:e1, 4] = H * S

    nt = threads or _threads()
    with ThreadPoolExecutor(nt) as ex:
        list(ex.map(pass1, range(len(starts))))
        M = fmax.max(axis=0)
        list(ex.map(lambda c: pass2(c, M), range(len(starts))))
    del F
    return Case(order, bounds, h.reshape(-1), y.reshape(-1), U.reshape(-1), L, L)


# Named workloads
def config1():
    """8x8x4 elements, N=3, 10 km square, Gaussian seamount."""
    return make_case(8, 8, 4, 3, 10000.0, "seamount")


def config2():
    """64x64x8 elements, N=4, 100 km square, tanh shelf 20 -> 200 m."""
    return make_case(64, 64, 8, 4, 100000.0, "shelf")


def config4():
    """128x128x16 elements, N=4 (262,144 elements, 32,768,000 DOF per field),
    100 km square, the tanh shelf 20 -> 200 m of the 64x64x8 case."""
    return make_case(128, 128, 16, 4, 100000.0, "shelf")


def config3(band=7000.0):
    """Coastline-refined octree: base 32x32x4, N=3, 100 km square, three
    rounds of XY refinement (with 2:1 balance) of the leaves within `band`
    of y = 0.8 L + 0.05 L sin(8 pi x / L); 39,280 elements at the default.
    Returns (case, mesh)."""
    from .mesh import coastline_mesh

    m = coastline_mesh(band=band)
    bounds, _, _ = m.elements()
    return make_case_on(bounds, 3, 100000.0, "coast"), m


def to_lines(U, E, n):
    """packed [e][f][i + n L] -> LINES [f][i][e][L]"""
    nd = n ** 3
    P = U.reshape(E, 5, n * n, n)  # e, f, L, i
    return np.ascontiguousarray(P.transpose(1, 3, 0, 2).reshape(-1))


def from_lines(V, E, n):
    P = V.reshape(5, n, E, n * n)  # f, i, e, L
    return np.ascontiguousarray(P.transpose(2, 0, 3, 1).reshape(-1))


def to_blocks(U, E, n, B):
    """packed [e][f][i + n L] -> blocked [e // B][f][i][e % B][L], the last
    block padded with zeros to B elements"""
    nb = -(-E // B)
    P = np.zeros((nb * B, 5, n * n, n))
    P[:E] = U.reshape(E, 5, n * n, n)  # e, f, L, i
    return np.ascontiguousarray(P.reshape(nb, B, 5, n * n, n).transpose(0, 2, 4, 1, 3).reshape(-1))


def from_blocks(V, E, n, B):
    nb = -(-E // B)
    P = V.reshape(nb, 5, n, B, n * n)  # block, f, i, e, L
    return np.ascontiguousarray(P.transpose(0, 3, 1, 4, 2).reshape(nb * B, -1)[:E].reshape(-1))


def to_soa(U, E, nd):
    return np.ascontiguousarray(U.reshape(E, 5, nd).transpose(1, 0, 2).reshape(-1))


def from_soa(V, E, nd):
    return np.ascontiguousarray(V.reshape(5, E, nd).transpose(1, 0, 2).reshape(-1))
