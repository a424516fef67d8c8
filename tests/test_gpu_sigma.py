"""GPU sigma-coordinate metrics (SigmaMetrics::update,
include/mesh/sigma_coordinate.hpp:340-354) against the oracle, bit-exact
(the kernel does not contract products into FMAs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("steady", [False, True])
def test_sigma_metrics_bit_exact(orc, steady):
    import torch

    from drifter_amd.sigma import sigma_metrics

    rng = np.random.default_rng(3)
    n = 100003
    sig = rng.uniform(-1, 0, n)
    eta = rng.normal(scale=0.5, size=n)
    h = rng.uniform(0, 200, n)
    h[:50] = -eta[:50] + rng.uniform(-1e-9, 1e-9, 50)  # dry and nearly dry columns
    g = [rng.normal(scale=1e-3, size=n) for _ in range(4)]
    et = None if steady else rng.normal(scale=1e-4, size=n)
    d = lambda a: None if a is None else torch.from_numpy(a).cuda()
    got = [t.cpu().numpy() for t in sigma_metrics(d(sig), d(eta), d(h), *[d(x) for x in g],
                                                  deta_dt=d(et))]
    ref = orc.sigma_metrics(sig, eta, h, *g, et)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    if steady:
        assert np.all(got[4] == 0)
