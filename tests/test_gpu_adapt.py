"""GPU parity of the refinement-side kernels against the oracle:
GradientErrorEstimator::estimate + mark_elements (src/amr/refinement.cpp:44-161,
488-520) bit-exact (values, masks, flags, actions), and the modal
coefficients (FP64 within 1e-13 of the largest coefficient; BF16 features
within one BF16 ulp of the oracle's coefficients rounded the same way)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _run_estimator(torch, order, field, bounds, levels, **kw):
    from drifter_amd import adapt

    est = adapt.GradientErrorEstimator(order, **kw)
    h = adapt.element_char_length(bounds)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = est.estimate(d(field), d(h), levels=d(levels.astype(np.int32)),
                       elem_stride=field.shape[1])
    return {k: v.cpu().numpy() for k, v in out.items()}


def _check(orc, order, field, bounds, levels, got, **kw):
    rt, ct, sf = kw.get("refine_threshold", 0.1), kw.get("coarsen_threshold", 0.01), \
        kw.get("smoothness_factor", 0.5)
    nd = (order + 1) ** 3
    v, m, r, c = orc.estimate(order, field[:, :nd], bounds, rt, ct, sf)
    # un-normalised maximum: (max - min) * cbrt(volume), the same operations
    from drifter_amd import adapt

    f = field[:, :nd]
    raw = (f.max(axis=1) - f.min(axis=1)) * adapt.element_char_length(bounds)
    assert got["max_value"][0] == raw.max()
    np.testing.assert_array_equal(got["value"], v)
    np.testing.assert_array_equal(got["mask"], m)
    np.testing.assert_array_equal(got["refine"], r)
    np.testing.assert_array_equal(got["coarsen"], c)
    a, mo = orc.mark_elements(r, c, m, levels, kw.get("max_level", 8), kw.get("min_level", 0))
    np.testing.assert_array_equal(got["action"], a)
    np.testing.assert_array_equal(got["mask_out"], mo)
    return r.sum(), c.sum()


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 8])
def test_estimator_random_fields(torch, orc, order):
    print(f"[estimator] order={order}", flush=True)
    rng = np.random.default_rng(order)
    n = order + 1
    nd = n ** 3
    E = 1000
    lo = rng.uniform(0, 10, size=(E, 3))
    sz = 2.0 ** -rng.integers(0, 4, size=(E, 3))
    bounds = np.stack([lo[:, 0], lo[:, 0] + sz[:, 0], lo[:, 1], lo[:, 1] + sz[:, 1],
                       lo[:, 2], lo[:, 2] + sz[:, 2]], axis=1)
    x = np.linspace(-1, 1, n)
    amp = 10.0 ** rng.uniform(-4, 0, size=(E, 3))
    F = (amp[:, 0, None, None, None] * np.sin(3 * x)[None, None, None, :]
         + amp[:, 1, None, None, None] * x[None, None, :, None] ** 2
         + amp[:, 2, None, None, None] * np.cos(2 * x)[None, :, None, None])
    field = np.zeros((E, nd + 3))  # padded element stride
    field[:, :nd] = F.reshape(E, nd) + rng.normal(size=(E, 1))
    levels = rng.integers(0, 9, size=(E, 3))
    got = _run_estimator(torch, order, field, bounds, levels)
    nr, nc = _check(orc, order, field, bounds, levels, got)
    assert nr > 0 and nc > 0


def test_estimator_thresholds_and_levels(torch, orc):
    rng = np.random.default_rng(5)
    order, E = 3, 300
    field = rng.normal(size=(E, 64)) * (10.0 ** rng.uniform(-3, 0, size=(E, 1)))
    bounds = np.tile([0, 1, 0, 1, 0, 0.5], (E, 1)).astype(float)
    levels = rng.integers(0, 4, size=(E, 3))
    kw = dict(refine_threshold=0.3, coarsen_threshold=0.05, smoothness_factor=0.8, max_level=3,
              min_level=1)
    got = _run_estimator(torch, order, field, bounds, levels, **kw)
    _check(orc, order, field, bounds, levels, got, **kw)


def test_estimator_constant_and_single(torch, orc):
    bounds = np.tile([0, 1, 0, 1, 0, 1], (8, 1)).astype(float)
    levels = np.zeros((8, 3), int)
    got = _run_estimator(torch, 2, np.ones((8, 27)), bounds, levels)
    _check(orc, 2, np.ones((8, 27)), bounds, levels, got)
    assert np.all(got["value"] == 0) and not got["refine"].any()
    f1 = np.arange(27, dtype=float)[None, :]
    got = _run_estimator(torch, 2, f1, bounds[:1], levels[:1])
    _check(orc, 2, f1, bounds[:1], levels[:1], got)
    assert got["mask"][0] == 4 and got["action"][0] == 1


def test_estimator_coastline_mesh(torch, orc):
    """39,280-element coastline octree, N=3: estimator on eta and on T"""
    from drifter_amd import synth

    case, mesh = synth.config3()
    print(f"[estimator] coastline E={case.E}", flush=True)
    _, levels, _ = mesh.elements()
    U = case.U.reshape(case.E, 5, case.nd)
    for q in (2, 3):
        field = np.ascontiguousarray(U[:, q])
        got = _run_estimator(torch, 3, field, case.bounds, levels)
        nr, nc = _check(orc, 3, field, case.bounds, levels, got)
        print(f"  field {q}: refine {nr}, coarsen {nc}", flush=True)


def _bf16_rne(x):
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def _bf16_to_f32(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6, 8])
def test_modal_coeffs(torch, orc, order):
    from drifter_amd import adapt

    print(f"[modal] order={order}", flush=True)
    n = order + 1
    nd = n ** 3
    E = 333
    rng = np.random.default_rng(order)
    u = rng.normal(size=(E, nd + 5))
    d = torch.from_numpy(u).cuda()
    coef = adapt.modal_coeffs(order, d, E, elem_stride=nd + 5).cpu().numpy()
    ref = orc.modal_coeffs(order, u[:, :nd])
    np.testing.assert_allclose(coef, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    # BF16 features into columns [7, 7 + nd) of a wider row, per-element scale
    feat = torch.zeros((E, nd + 16), dtype=torch.int16, device="cuda")
    es = rng.uniform(0.5, 2.0, E)
    adapt.modal_features_bf16(order, d, E, feat, 7, scale=0.25,
                              elem_scale=torch.from_numpy(es).cuda(), elem_stride=nd + 5)
    fb = feat.cpu().numpy().view(np.uint16)
    exp = _bf16_rne(ref * (0.25 * es)[:, None])
    exp[:, 0] = 0  # mean mode zeroed
    got = fb[:, 7:7 + nd]
    assert np.all(fb[:, :7] == 0) and np.all(fb[:, 7 + nd:] == 0)
    diff = np.abs(got.astype(np.int64) - exp.astype(np.int64))
    assert diff.max() <= 1 and (diff == 0).mean() > 0.999
    sref = ref * (0.25 * es)[:, None]
    sref[:, 0] = 0.0
    np.testing.assert_allclose(_bf16_to_f32(got), sref, rtol=2 ** -8, atol=1e-30)
