"""GPU parity of the face components against the FP64 oracle:
pointwise numerical fluxes (NumericalFlux::flux, include/flux/numerical_flux.hpp),
face-trace gathers (src/dg/basis_hexahedron.cpp:431-522),
ConformingInterface::compute_flux (src/dg/mortar.cpp:320-343) and
MortarSpace::compute_mortar_flux (src/dg/mortar.cpp:212-258).

Bars: gathers bit-exact; fluxes and conforming interfaces relative 1e-13 per
value (the kernels do not contract products into FMAs, so they are expected
to round like the oracle); mortars relative 1e-12 of the interface's largest
value (the inverse mass matrix is factored differently, Cholesky here and
Gauss-Jordan in the oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _unit_normals(rng, n, planar=True):
    v = rng.normal(size=(n, 3))
    if planar:
        v[:, 2] = 0.0
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _states(kind, rng, n):
    if kind in (0, 1, 2, 3):
        U = np.stack([rng.uniform(0.5, 5.0, n), rng.normal(size=n), rng.normal(size=n)])
        U[0, :3] = [0.0, 1e-12, 3e-11]           # dry and nearly dry cells
        U[1, 3], U[0, 3] = 40.0, 1.0              # supersonic
        return U
    if kind == 4:
        return np.stack([rng.normal(size=n), rng.normal(size=n), rng.uniform(-1, 1, n),
                         rng.uniform(100, 200, n), rng.uniform(300, 350, n)])
    return rng.normal(size=(1, n))


PARAMS = {0: [9.81], 1: [9.81], 2: [9.81, 0.1], 3: [9.81], 4: [9.81],
          5: [0.7, -0.4, 0.2], 6: [0.7, -0.4, 0.2, 1.3]}


@pytest.mark.parametrize("kind", range(7))
def test_flux_points(torch, orc, kind):
    from drifter_amd import interface as itf

    print(f"[flux points] kind={kind}", flush=True)
    rng = np.random.default_rng(100 + kind)
    npts = 777
    UL, UR = _states(kind, rng, npts), _states(kind, rng, npts)
    nrm = _unit_normals(rng, npts, planar=(kind != 5 and kind != 6))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    got = itf.flux_points(kind, d(UL), d(UR), d(nrm.T), PARAMS[kind]).cpu().numpy()
    torch.cuda.synchronize()
    ref = np.stack([orc.flux(kind, UL[:, i], UR[:, i], nrm[i], PARAMS[kind])
                    for i in range(npts)], axis=1)
    assert np.all(np.isfinite(got))
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    print(f"  bit-identical values: {np.mean(got == ref):.4f}", flush=True)


def test_flux_points_empty_and_single(torch, orc):
    from drifter_amd import interface as itf

    e = torch.zeros((3, 0), dtype=torch.float64, device="cuda")
    assert itf.flux_points(0, e, e, torch.zeros((3, 0), dtype=torch.float64,
                                                 device="cuda"), [9.81]).shape == (3, 0)
    UL = np.array([[2.0], [0.3], [-0.1]])
    UR = np.array([[1.5], [-0.2], [0.4]])
    n = np.array([[0.6], [0.8], [0.0]])
    d = lambda a: torch.from_numpy(a).cuda()
    got = itf.flux_points(1, d(UL), d(UR), d(n), [9.81]).cpu().numpy()[:, 0]
    np.testing.assert_allclose(got, orc.flux(1, UL[:, 0], UR[:, 0], n[:, 0], [9.81]), rtol=1e-14)


@pytest.mark.parametrize("order", [1, 3, 4, 8])
def test_gather_face_traces_bit_exact(torch, orc, order):
    from drifter_amd import interface as itf

    print(f"[gather] order={order}", flush=True)
    n = order + 1
    nd = n ** 3
    E = 37
    rng = np.random.default_rng(order)
    U = rng.normal(size=(E, 5, nd))
    el = np.repeat(np.arange(E), 6).astype(np.int64)
    fa = np.tile(np.arange(6), E).astype(np.int8)
    el[5], fa[7] = E, 6                       # invalid element / face -> NaN
    Ud = torch.from_numpy(U).cuda()
    got = itf.gather_face_traces(order, Ud.reshape(-1)[2 * nd:], torch.from_numpy(el).cuda(),
                                 torch.from_numpy(fa).cuda(), E, elem_stride=5 * nd).cpu().numpy()
    for k in range(len(el)):
        if k in (5, 7):
            assert np.all(np.isnan(got[k]))
            continue
        np.testing.assert_array_equal(got[k], U[el[k], 2][orc.face_map(order, int(fa[k]))])


@pytest.mark.parametrize("kind", [5, 6])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_conforming_batch(torch, orc, order, kind):
    from drifter_amd import interface as itf

    print(f"[conforming] order={order} kind={kind}", flush=True)
    nf = (order + 1) ** 2
    nfaces = 50
    rng = np.random.default_rng(order * 10 + kind)
    UL, UR = rng.normal(size=(nfaces, nf)), rng.normal(size=(nfaces, nf))
    nrm = _unit_normals(rng, nfaces, planar=False)
    ci = itf.ConformingInterface(order)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rl, rr = (t.cpu().numpy() for t in ci.compute_flux(kind, PARAMS[kind], d(UL), d(UR), d(nrm)))
    for k in range(nfaces):
        el, er = orc.conforming_flux(order, kind, PARAMS[kind], UL[k], UR[k], nrm[k])
        np.testing.assert_allclose(rl[k], el, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(rr[k], er, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("ctype", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_mortar_batch(torch, orc, order, ctype):
    from drifter_amd import interface as itf

    print(f"[mortar] order={order} type={ctype}", flush=True)
    ms = itf.MortarSpace(order, ctype, device=0)
    nm, nfine = ms.num_dofs(), ms.num_fine_elements()
    n_if = 40
    rng = np.random.default_rng(order * 10 + ctype)
    Uc = rng.normal(size=(n_if, nm))
    Uf = rng.normal(size=(n_if, nfine, nm))
    nrm = _unit_normals(rng, n_if, planar=False)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    for kind in (5, 6):
        rc, rf = (t.cpu().numpy() for t in ms.compute_mortar_flux(kind, PARAMS[kind], d(Uc),
                                                                   d(Uf), d(nrm)))
        for k in range(n_if):
            ec, ef = orc.mortar_flux(order, ctype, kind, PARAMS[kind], Uc[k], Uf[k], nrm[k])
            s = max(np.abs(ec).max(), np.abs(ef).max())
            np.testing.assert_allclose(rc[k], ec, rtol=0, atol=1e-12 * s)
            np.testing.assert_allclose(rf[k], ef, rtol=0, atol=1e-12 * s)


def test_mortar_large_batch_stride(torch, orc):
    """more interfaces than workgroups: the grid-stride loop covers them all"""
    from drifter_amd import interface as itf

    order, ctype, kind = 3, 5, 6
    ms = itf.MortarSpace(order, ctype, device=0)
    nm, nfine = ms.num_dofs(), ms.num_fine_elements()
    n_if = 9000
    rng = np.random.default_rng(7)
    Uc, Uf = rng.normal(size=(n_if, nm)), rng.normal(size=(n_if, nfine, nm))
    nrm = _unit_normals(rng, n_if, planar=False)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rc, rf = (t.cpu().numpy() for t in ms.compute_mortar_flux(kind, PARAMS[kind], d(Uc), d(Uf),
                                                               d(nrm)))
    for k in list(range(0, n_if, 997)) + [n_if - 1]:
        ec, ef = orc.mortar_flux(order, ctype, kind, PARAMS[kind], Uc[k], Uf[k], nrm[k])
        s = max(np.abs(ec).max(), np.abs(ef).max())
        np.testing.assert_allclose(rc[k], ec, rtol=0, atol=1e-12 * s)
        np.testing.assert_allclose(rf[k], ef, rtol=0, atol=1e-12 * s)


@pytest.mark.parametrize("n_if", [0, 1, 15, 16, 17])
@pytest.mark.parametrize("order", [3, 4])
def test_mortar_small_counts(torch, orc, order, n_if):
    """counts around one wavefront's 16 interfaces (and none), with one
    (order 3) and two (order 4) 16-row operator tiles"""
    from drifter_amd import interface as itf

    ctype, kind = 5, 6
    ms = itf.MortarSpace(order, ctype, device=0)
    nm, nfine = ms.num_dofs(), ms.num_fine_elements()
    rng = np.random.default_rng(n_if + 3 + 100 * order)
    Uc, Uf = rng.normal(size=(n_if, nm)), rng.normal(size=(n_if, nfine, nm))
    nrm = _unit_normals(rng, max(n_if, 1), planar=False)[:n_if]
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    print(f"[mortar] order {order} small count {n_if}", flush=True)
    rc, rf = (t.cpu().numpy() for t in ms.compute_mortar_flux(kind, PARAMS[kind], d(Uc), d(Uf),
                                                               d(nrm)))
    assert rc.shape == (n_if, nm) and rf.shape == (n_if, nfine, nm)
    for k in range(n_if):
        ec, ef = orc.mortar_flux(order, ctype, kind, PARAMS[kind], Uc[k], Uf[k], nrm[k])
        s = max(np.abs(ec).max(), np.abs(ef).max())
        np.testing.assert_allclose(rc[k], ec, rtol=0, atol=1e-12 * s)
        np.testing.assert_allclose(rf[k], ef, rtol=0, atol=1e-12 * s)


@pytest.mark.parametrize("order,ctype", [(2, 1), (4, 5)])
def test_mortar_matrix_core_grid_stride(torch, orc, order, ctype):
    """more interfaces than one pass of the persistent grid (one workgroup
    per resident slot, 16 interfaces per wavefront: at most 256 CUs x 64
    wavefronts x 16 = 262,144 interfaces per pass)"""
    from drifter_amd import interface as itf

    kind = 5
    ms = itf.MortarSpace(order, ctype, device=0)
    nm, nfine = ms.num_dofs(), ms.num_fine_elements()
    n_if = 262144 + 4099
    g = torch.Generator(device="cuda").manual_seed(5)
    Uc = torch.randn((n_if, nm), dtype=torch.float64, device="cuda", generator=g)
    Uf = torch.randn((n_if, nfine, nm), dtype=torch.float64, device="cuda", generator=g)
    nrm = torch.zeros((n_if, 3), dtype=torch.float64, device="cuda")
    nrm[:, 0], nrm[:, 1] = 0.6, 0.8
    print(f"[mortar] grid stride n_if={n_if}", flush=True)
    rc, rf = ms.compute_mortar_flux(kind, PARAMS[kind], Uc, Uf, nrm)
    torch.cuda.synchronize()
    ks = [0, 65535, 65536, 81919, 81920, 262143, 262144, n_if - 1]
    Ucn, Ufn, nn = Uc[ks].cpu().numpy(), Uf[ks].cpu().numpy(), nrm[ks].cpu().numpy()
    rcn, rfn = rc[ks].cpu().numpy(), rf[ks].cpu().numpy()
    for i, k in enumerate(ks):
        ec, ef = orc.mortar_flux(order, ctype, kind, PARAMS[kind], Ucn[i], Ufn[i], nn[i])
        s = max(np.abs(ec).max(), np.abs(ef).max())
        np.testing.assert_allclose(rcn[i], ec, rtol=0, atol=1e-12 * s)
        np.testing.assert_allclose(rfn[i], ef, rtol=0, atol=1e-12 * s)


@pytest.mark.parametrize("ctype", [0, 2, 4, 5])
def test_mortar_conservation_batch_bit_exact(torch, orc, ctype):
    """batched verify_mortar_conservation on the mortar flux's own output,
    bit-exact against the oracle per interface; empty batch is a no-op"""
    from drifter_amd import interface as itf

    order, kind = 3, 6
    print(f"[mortar conservation] type={ctype}", flush=True)
    ms = itf.MortarSpace(order, ctype, device=0)
    nm, nfine = ms.num_dofs(), ms.num_fine_elements()
    n_if = 300
    rng = np.random.default_rng(40 + ctype)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rc, rf = ms.compute_mortar_flux(kind, PARAMS[kind], d(rng.normal(size=(n_if, nm))),
                                    d(rng.normal(size=(n_if, nfine, nm))),
                                    d(_unit_normals(rng, n_if, planar=False)))
    rc[0].zero_()
    rf[0].zero_()
    wc, wf = rng.uniform(0, 1, nm), rng.uniform(0, 1, (nfine, nm))
    got = ms.verify_conservation_batch(rc, rf, d(wc), d(wf)).cpu().numpy()
    rcn, rfn = rc.cpu().numpy(), rf.cpu().numpy()
    exp = np.array([orc.verify_mortar_conservation(rcn[k], rfn[k], wc, wf)
                    for k in range(n_if)])
    np.testing.assert_array_equal(got, exp)
    assert got[0] == 0.0
    empty = ms.verify_conservation_batch(rc[:0], rf[:0], d(wc), d(wf))
    assert empty.numel() == 0
