"""Pins the oracle's indicator MLP (A21): BF16 rounding, the FP64 gradient
against central finite differences and against torch autograd on the CPU
(library use only), and the quantised mode's consistency."""
import numpy as np
import pytest


def _params(rng, scale=0.3):
    return rng.normal(scale=scale, size=12481)


def test_bf16_round_matches_bit_formula(orc):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=1000) * 10.0 ** rng.uniform(-5, 5, 1000),
                        [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.5 - 2 ** -8, 0.0]])
    lib = orc.load()
    got = np.array([lib.orc_bf16_round(v) for v in x])
    b = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    exp = (((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    np.testing.assert_array_equal(got, exp.astype(np.float64))
    # ties to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert lib.orc_bf16_round(1.0 + 2 ** -8) == 1.0
    assert lib.orc_bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def test_fp64_gradient_central_differences(orc):
    rng = np.random.default_rng(1)
    p = _params(rng)
    X = rng.normal(size=(7, 128))
    y = rng.normal(size=7)
    loss, _, g = orc.mlp_loss_grad(p, X, y, quantize=False)
    idx = np.concatenate([rng.choice(8192, 12, replace=False), 8192 + rng.choice(64, 4),
                          8256 + rng.choice(4096, 8), 12352 + rng.choice(64, 4),
                          12416 + rng.choice(64, 4), [12480]])
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.mlp_loss_grad(pp, X, y, False, False)[0]
              - orc.mlp_loss_grad(pm, X, y, False, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (abs(g[i]) + 1e-3), (i, fd, g[i])


def test_fp64_gradient_matches_torch_autograd(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    p = _params(rng)
    X = rng.normal(size=(11, 128))
    y = rng.normal(size=11)
    loss, pred, g = orc.mlp_loss_grad(p, X, y, quantize=False)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1 = P[:8192].view(64, 128)
    b1 = P[8192:8256]
    W2 = P[8256:12352].view(64, 64)
    b2 = P[12352:12416]
    W3 = P[12416:12480]
    b3 = P[12480]
    gelu = torch.nn.GELU(approximate="tanh")
    Xt = torch.tensor(X)
    out = gelu(gelu(Xt @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    L = ((out - torch.tensor(y)) ** 2).mean()
    L.backward()
    np.testing.assert_allclose(pred, out.detach().numpy(), rtol=1e-13, atol=1e-13)
    assert abs(loss - L.item()) < 1e-12 * abs(loss)
    np.testing.assert_allclose(g, P.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_quantised_mode_rounds_operands(orc):
    """with every weight and input already BF16-representable and tiny
    activations' rounding effect, quantised and FP64 modes agree closely;
    a non-representable weight changes the quantised result"""
    rng = np.random.default_rng(3)
    p = orc.load().orc_bf16_round
    params = np.array([p(v) for v in _params(rng)])
    X = np.array([[p(v) for v in row] for row in rng.normal(size=(5, 128))])
    y = rng.normal(size=5)
    lq, pq, gq = orc.mlp_loss_grad(params, X, y, quantize=True)
    lf, pf, gf = orc.mlp_loss_grad(params, X, y, quantize=False)
    np.testing.assert_allclose(pq, pf, rtol=2e-2, atol=2e-2)  # only activations rounded
    assert np.linalg.norm(gq - gf) < 5e-2 * np.linalg.norm(gf)
    params2 = params.copy()
    params2[0] += 1e-5  # below BF16 resolution of a ~0.3 weight: no effect when quantised
    assert orc.mlp_loss_grad(params2, X, y, True, False)[0] == lq
