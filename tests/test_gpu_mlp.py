"""GPU refinement-indicator MLP (BF16 MFMA forward/backward, FP32
accumulation) against the oracle's quantised model (A21).

Bars: forward within 5e-3 of the output scale (activations rounded to BF16
on both sides can fall on opposite sides of a rounding tie: FP32 vs FP64
sums), median difference below 1e-5 of the scale; gradients within 1e-2
relative L2 per parameter block; the same call twice gives identical bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (0, 8192), "b1": (8192, 8256), "W2": (8256, 12352), "b2": (12352, 12416),
          "W3": (12416, 12480), "b3": (12480, 12481)}


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _setup(torch, seed, n, scale_x=1.0):
    from drifter_amd import mlp

    rng = np.random.default_rng(seed)
    m = mlp.IndicatorMLP(seed=seed)
    p = m.get_params().astype(np.float64)
    p[8192:8256] = rng.normal(scale=0.1, size=64)  # non-zero biases
    p[12352:12416] = rng.normal(scale=0.1, size=64)
    p[12480] = 0.2
    m.set_params(p.astype(np.float32))
    p = m.get_params().astype(np.float64)
    xb = mlp.bf16_bits(rng.normal(scale=scale_x, size=(n, 128)))
    X = torch.from_numpy(xb.view(np.int16)).cuda()
    Xf = mlp.bf16_value(xb).astype(np.float64)
    y = rng.normal(size=n).astype(np.float32)
    return m, p, X, Xf, y


@pytest.mark.parametrize("n", [1, 63, 1000])
def test_forward_matches_oracle(torch, orc, n):
    print(f"[mlp forward] n={n}", flush=True)
    m, p, X, Xf, _ = _setup(torch, 10 + n, n)
    got = m.forward(X).cpu().numpy().astype(np.float64)
    _, ref, _ = orc.mlp_loss_grad(p, Xf, np.zeros(n), quantize=True, want_grad=False)
    scale = max(1.0, np.abs(ref).max())
    d = np.abs(got - ref)
    assert d.max() <= 5e-3 * scale, d.max()
    assert np.median(d) <= 1e-5 * scale, np.median(d)


@pytest.mark.parametrize("n", [1, 33, 300])
def test_gradient_matches_oracle(torch, orc, n):
    print(f"[mlp grad] n={n}", flush=True)
    m, p, X, Xf, y = _setup(torch, 20 + n, n)
    loss, grad = m.loss_grad(X, torch.from_numpy(y).cuda())
    loss = float(loss.cpu()[0])
    grad = grad.cpu().numpy().astype(np.float64)
    rl, _, rg = orc.mlp_loss_grad(p, Xf, y.astype(np.float64), quantize=True)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    for name, (a, b) in BLOCKS.items():
        den = np.linalg.norm(rg[a:b])
        err = np.linalg.norm(grad[a:b] - rg[a:b]) / max(den, 1e-30)
        print(f"  {name}: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-2, (name, err)


def test_gradient_deterministic(torch):
    m, p, X, Xf, y = _setup(torch, 5, 5000)
    yt = torch.from_numpy(y).cuda()
    l1, g1 = m.loss_grad(X, yt)
    l2, g2 = m.loss_grad(X, yt)
    assert torch.equal(g1, g2) and torch.equal(l1, l2)


def test_training_fits_a_teacher(torch, orc):
    """Adam on a target produced by another network of the same shape"""
    from drifter_amd import mlp

    rng = np.random.default_rng(9)
    n = 4096
    teacher = rng.normal(scale=0.15, size=12481)
    xb = mlp.bf16_bits(rng.normal(size=(n, 128)))
    Xf = mlp.bf16_value(xb).astype(np.float64)
    _, y, _ = orc.mlp_loss_grad(teacher, Xf, np.zeros(n), quantize=False, want_grad=False)
    X = torch.from_numpy(xb.view(np.int16)).cuda()
    yt = torch.from_numpy(y.astype(np.float32)).cuda()
    m = mlp.IndicatorMLP(seed=3)
    l0 = float(m.loss_grad(X, yt)[0].cpu()[0])
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    for it in range(400):
        m.train_step(X, yt, 3e-3, loss=loss)
    l1 = float(loss.cpu()[0])
    print(f"  teacher fit: loss {l0:.4f} -> {l1:.5f}", flush=True)
    assert l1 < 0.05 * l0


def test_params_roundtrip_and_classify(torch):
    from drifter_amd import mlp

    m = mlp.IndicatorMLP(seed=4)
    p = np.random.default_rng(0).normal(size=12481).astype(np.float32)
    m.set_params(p)
    np.testing.assert_array_equal(m.get_params(), p)
    pred = torch.tensor([-3.0, -2.0, -1.5, -1.0, -0.5, 0.0], device="cuda")
    act = mlp.classify(pred, -1.0, -2.0).cpu().numpy()
    np.testing.assert_array_equal(act, [2, 0, 0, 0, 1, 1])
