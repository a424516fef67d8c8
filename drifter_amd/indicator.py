"""Learned refinement indicator on the coastline workload.

The MLP (drifter_amd.mlp) maps per-element orthonormal-Legendre modal
coefficients of the free surface eta and the temperature T (order 3: 2 x 64
inputs, mean modes zeroed) to y = log10(max(v_eta, v_T, 1e-3)), where v is
the gradient error estimator's normalised indicator of each field
(GradientErrorEstimator, src/amr/refinement.cpp:44-93).  Decisions: refine
when y > log10(0.1), coarsen when y < log10(0.01) -- the estimator's default
thresholds (include/amr/refinement.hpp:64-65) applied to both fields at once
(refine if either field asks for it, coarsen only if both allow it).

Feature scaling mirrors the estimator's normalisation: the coefficients of
field q in element e are multiplied by h_e / R_q, h_e = cbrt(element
volume) and R_q the estimator's un-normalised maximum for that field.

All arithmetic runs in the HIP library (modal features, estimator, MLP
forward/backward/Adam, classification); torch provides device buffers and
the synthetic seeded fields (the same Fourier construction as
drifter_amd.synth, evaluated on the device).
"""
import math

import numpy as np

from . import adapt, mlp, synth

REFINE_LEVEL = math.log10(0.1)
COARSEN_LEVEL = math.log10(0.01)
FLOOR = 1e-3
ORDER = 3


class CoastlineFields:
    """Node coordinates of the coastline mesh on the device, and seeded eta
    and T fields evaluated there."""

    def __init__(self, mesh, device="cuda"):
        import torch

        bounds, levels, _ = mesh.elements()
        self.E = bounds.shape[0]
        self.nd = (ORDER + 1) ** 3
        nodes, _, _ = synth.lgl_nodes(ORDER)
        x, y, s = synth.node_coords(bounds, nodes)
        self.x = torch.from_numpy(x).to(device)
        self.y = torch.from_numpy(y).to(device)
        self.s = torch.from_numpy(s).to(device)
        self.h_char = torch.from_numpy(adapt.element_char_length(bounds)).to(device)
        self.levels = torch.from_numpy(levels.astype(np.int32)).to(device)
        self.L = 100000.0
        self.device = device

    def fourier(self, q, seed_base):
        """synth.fourier_field on the device"""
        import torch

        rng = synth.SplitMix64(seed_base + q)
        F = torch.zeros_like(self.x)
        for m in range(1, 9):
            kx = rng.next() % 4
            ky = rng.next() % 4
            kz = rng.next() % 2
            if q == 2:
                kz = 0
            a = (2.0 * rng.uniform() - 1.0) / m
            phi = 2.0 * math.pi * rng.uniform()
            F += a * torch.cos(2.0 * math.pi * (kx * self.x / self.L + ky * self.y / self.L
                                                + kz * self.s) + phi)
        return F / F.abs().max()

    def sample(self, seed):
        """(eta, T) nodal fields (E, nd) of the seeded state"""
        eta = 0.05 * self.fourier(2, seed)
        T = 15.0 + 5.0 * self.fourier(3, seed)
        return eta.contiguous(), T.contiguous()


def build_rows(fields, seed, estimator=None):
    """Features X (E, 128) BF16 bits, target y (E,) FP32 and the
    estimator's combined decision action (E,) uint8 for one seeded state."""
    eta, T = fields.sample(seed)
    return rows_from_fields(fields, eta, T, estimator)


def build_dataset(fields, seeds):
    import torch

    est = adapt.GradientErrorEstimator(ORDER)
    parts = [build_rows(fields, s, est) for s in seeds]
    return (torch.cat([p[0] for p in parts]), torch.cat([p[1] for p in parts]),
            torch.cat([p[2] for p in parts]))


def train(model, X, y, epochs=4, batch=8192, lr=2e-3, seed=0):
    """Adam over shuffled minibatches with a cosine-decayed learning rate;
    returns the last epoch's mean loss."""
    import torch

    n = X.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device=X.device)
    total = 0.0
    for ep in range(epochs):
        lr_ep = lr * 0.5 * (1.0 + math.cos(math.pi * ep / max(epochs, 1)))
        perm = torch.randperm(n, generator=g).to(X.device)
        Xs, ys = X[perm].contiguous(), y[perm].contiguous()
        acc = torch.zeros(1, dtype=torch.float64, device=X.device)
        nb = 0
        for b0 in range(0, n - batch + 1, batch):
            model.train_step(Xs[b0:b0 + batch], ys[b0:b0 + batch], lr_ep, loss=loss)
            acc += loss
            nb += 1
        total = float(acc.cpu()[0]) / max(nb, 1)
    return total


def misclassification(model, X, action):
    pred = model.forward(X)
    got = mlp.classify(pred, REFINE_LEVEL, COARSEN_LEVEL)
    return float((got != action).float().mean().cpu())


class SteppedCoastline:
    """Coastline states that the model has stepped: the seeded Fourier state
    (u, v, eta, T, S as in drifter_amd.synth, evaluated on the device) on the
    coastline mesh with its bathymetry, advanced by fused RK4 steps of the
    right-hand side at the CFL time step of a reference state."""

    def __init__(self, fields, h_nodal, solver, dt):
        self.fields, self.h, self.solver, self.dt = fields, h_nodal, solver, dt
        self.n = ORDER + 1

    def state(self, seed):
        """The seed's state in the line layout [field][i][e][L], on the device."""
        import torch

        f = self.fields
        u = 0.1 * f.fourier(0, seed)
        v = 0.1 * f.fourier(1, seed)
        eta = 0.05 * f.fourier(2, seed)
        T = 15.0 + 5.0 * f.fourier(3, seed)
        S = 32.5 + 2.5 * f.fourier(4, seed)
        H = eta + self.h
        n, E = self.n, f.E
        P = torch.stack([H * u, H * v, eta, H * T, H * S], dim=1)  # (E, 5, nd)
        return P.reshape(E, 5, n * n, n).permute(1, 3, 0, 2).contiguous().reshape(-1)

    def eta_T(self, V):
        n, E, nd = self.n, self.fields.E, self.fields.nd
        P = V.view(5, n, E, n * n).permute(2, 0, 3, 1).reshape(E, 5, nd)
        H = P[:, 2] + self.h
        return P[:, 2].contiguous(), (P[:, 3] / H).contiguous()

    def step(self, V, k):
        from . import _lib

        for _ in range(k):
            self.solver.rk_step_dev(_lib.RK_RK4, 0.0, self.dt, V, _lib.LAYOUT_LINES, fused=True)


def rows_from_fields(fields, eta, T, estimator=None):
    """Features X (E, 128) BF16 bits, target y (E,) FP32 and the estimator's
    combined decision (E,) uint8 of nodal (eta, T) fields (E, nd)."""
    import torch

    est = estimator or adapt.GradientErrorEstimator(ORDER)
    E, nd = fields.E, fields.nd
    X = torch.zeros((E, 2 * nd), dtype=torch.int16, device=eta.device)
    vs = []
    for col, f in ((0, eta), (nd, T)):
        r = est.estimate(f, fields.h_char)
        vs.append(r["value"])
        inv = 1.0 / r["max_value"]  # device scalar: no host round trip
        adapt.modal_features_bf16(ORDER, f, E, X, col, scale=1.0,
                                  elem_scale=(fields.h_char * inv).contiguous())
    v = torch.maximum(vs[0], vs[1]).clamp_min(FLOOR)
    y = torch.log10(v).to(torch.float32)
    return X, y, mlp.classify(y, REFINE_LEVEL, COARSEN_LEVEL)
