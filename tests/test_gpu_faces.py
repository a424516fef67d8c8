"""The batched adaptive DG workload on the coastline octree (base 32x32x4,
N=3, three XY refinement rounds near y = 0.8 L + 0.05 L sin(8 pi x / L),
39,280 elements) and the face term on the uniform 128x128x16, N=4 mesh.

  - connectivity: the product's octree against the oracle's, bit-exact, for
    the reference's connectivity (OctreeAdapter::build_face_connections,
    src/mesh/octree_adapter.cpp:369-389) and the 2:1 one;
  - face plans: interface lists and groups bit-exact against the oracle's;
  - every group's face vectors against ConformingInterface::compute_flux /
    MortarSpace::compute_mortar_flux (src/dg/mortar.cpp:212-343) per
    interface (conforming bit-exact, mortars within 1e-12 of the largest
    value), and the mortar conservation check (verify_mortar_conservation,
    :386-413) of every mortar against the oracle's;
  - the face term (DG3DIntegration::compute_rhs face loop and M^-1,
    src/dg/operators_3d.cpp:255-281) against the oracle: FP64 <= 1e-12 and
    FP32 <= 1e-5 relative L2 per field;
  - the right-hand side with the face term attached, and an un-fused RK4
    step with it, against the oracle's volume RHS plus face term (<= 1e-12);
  - mode splitting with 30 barotropic substeps on the coastline mesh, as a
    HIP graph replayed bit-identically to eager steps, against the oracle
    on sampled elements (<= 1e-12).
"""
import numpy as np
import pytest

from drifter_amd import _lib, faces, synth
from drifter_amd.mesh import OctreeAdapter
from drifter_amd.solver import CoriolisParameter, ModeSplittingSolver, PrimitiveEquationsSolver
from sampling import rel_l2_fields, rows, sample_elements

pytestmark = pytest.mark.gpu

PARAMS = (0.3, -0.2, 0.1, 0.6)  # LF_ADV: v = (0.3, -0.2, 0.1), lambda = 0.6
TOL64 = 1e-12
TOL32 = 1e-5


@pytest.fixture(scope="module")
def coast():
    print("[faces] building the coastline case", flush=True)
    case, mesh = synth.config3()
    return case, mesh


def _orc_coast(orc):
    L = 100000.0
    o = orc.Mesh(0.0, L, 0.0, L, -1.0, 0.0)
    o.build_uniform(32, 32, 4)
    for _ in range(3):
        b, _, _ = o.elements()
        cx = 0.5 * (b[:, 0] + b[:, 1])
        cy = 0.5 * (b[:, 2] + b[:, 3])
        yc = 0.8 * L + 0.05 * L * np.sin(2.0 * np.pi * cx / (0.25 * L))
        sel = np.nonzero(np.abs(cy - yc) < 7000.0)[0]
        o.refine(sel, np.full(sel.size, 3))
    return o


def _mort_types(groups, nm):
    t = np.zeros(nm, dtype=np.int32)
    for g in groups:
        if g[1] != 0:
            t[g[3]:g[3] + g[4]] = g[1]
    return t


def _solver(case, prec=_lib.PREC_FP64):
    s = PrimitiveEquationsSolver(case.order, prec)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    return s


def test_coastline_connectivity_and_lists_bit_exact(orc, coast):
    case, mesh = coast
    o = _orc_coast(orc)
    b, lv, mo = mesh.elements()
    ob, olv, omo = o.elements()
    assert b.shape[0] == 39280
    assert np.array_equal(b, ob) and np.array_equal(lv, olv) and np.array_equal(mo, omo)
    s = _solver(case)
    for two in (False, True):
        conn = mesh.build_face_connections(two_to_one=two)
        oc = o.connections(two_to_one=two)
        for x, y in zip(conn, oc):
            assert np.array_equal(x, y)
        plan = faces.FacePlan(s, b, lv, conn, kind=faces.LF_ADV, params=PARAMS)
        cf, mt = plan.lists()
        ocf, omt, ogr = orc.face_lists(b, lv, oc)
        assert np.array_equal(cf, ocf) and np.array_equal(mt, omt)
        assert np.array_equal(plan.groups(), ogr)
        ng, nc, nm = plan.info()
        print(f"[faces] two_to_one={two}: {ng} groups, {nc} conforming, {nm} mortars", flush=True)
        if two:
            assert nm == 3012 and (ogr[:, 1] == 1).any()
        else:
            assert nm == 0  # the reference's search never reports a refined neighbour
        plan.close()
    s.close()


def test_group_fluxes_per_interface(orc, coast):
    import torch

    case, mesh = coast
    s = _solver(case)
    plan = faces.FacePlan.from_mesh(s, mesh, two_to_one=True, kind=faces.LF_ADV, params=PARAMS)
    E, n, nd, nh = case.E, case.n, case.nd, case.n ** 2
    U = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda()
    P = case.U.reshape(E, 5, nd)
    cf, mt = plan.lists()
    groups = plan.groups()
    fmap = [orc.face_map(case.order, f) for f in range(6)]
    rng = np.random.default_rng(5)
    nmort_checked = 0
    for gi, g in enumerate(groups):
        a, bb = plan.group_fluxes(gi, U)
        a, bb = a.cpu().numpy(), bb.cpu().numpy()
        first, cnt, kind = int(g[3]), int(g[4]), int(g[1])
        pick = rng.choice(cnt, min(cnt, 60), replace=False) if kind == 0 else np.arange(cnt)
        for i in pick:
            for fld in range(5):
                if kind == 0:
                    el, fl, er, fr = cf[first + i]
                    nrm = np.zeros(3)
                    nrm[fl // 2] = 1.0 if fl % 2 else -1.0
                    rl, rr = orc.conforming_flux(case.order, faces.LF_ADV, PARAMS,
                                                 P[el, fld][fmap[fl]], P[er, fld][fmap[fr]], nrm)
                    assert np.array_equal(a[i, fld], rl) and np.array_equal(bb[i, fld], rr)
                else:
                    ec, fc = mt[first + i, 0], int(mt[first + i, 1])
                    fine = [x for x in mt[first + i, 2:] if x >= 0]
                    nrm = np.zeros(3)
                    nrm[fc // 2] = 1.0 if fc % 2 else -1.0
                    Uf = np.stack([P[x, fld][fmap[fc ^ 1]] for x in fine])
                    rc, rf = orc.mortar_flux(case.order, kind, faces.LF_ADV, PARAMS,
                                             P[ec, fld][fmap[fc]], Uf, nrm)
                    sc = max(np.abs(rc).max(), np.abs(rf).max(), 1e-300)
                    assert np.abs(a[i, fld] - rc).max() <= 1e-12 * sc
                    assert np.abs(bb[i, fld] - rf).max() <= 1e-12 * sc
                    nmort_checked += 1
        if kind != 0:
            # the conservation check of every (mortar, field) on the device
            # against the oracle's on the device's own face vectors
            # (LGL face weights on the coarse face, scaled by the sub-face's
            # share of the area on the fine faces)
            m = faces_mortar(case.order, kind)
            nfine = bb.shape[2]
            _, w1 = orc.lgl(n)
            wc = np.outer(w1, w1).reshape(-1)
            wf = np.tile(wc / nfine, (nfine, 1))
            A = a.reshape(-1, nh)
            B = bb.reshape(-1, nfine, nh)
            rel = m.verify_conservation_batch(
                torch.from_numpy(A).cuda(), torch.from_numpy(np.ascontiguousarray(B)).cuda(),
                torch.from_numpy(wc).cuda(), torch.from_numpy(wf).cuda()).cpu().numpy()
            for k in range(0, rel.size, 7):
                assert rel[k] == orc.verify_mortar_conservation(A[k], B[k], wc, wf)
    assert nmort_checked >= 3012 * 5
    plan.close()
    s.close()


def faces_mortar(order, kind):
    from drifter_amd import interface as itf

    return itf.MortarSpace(order, kind)


@pytest.mark.parametrize("mode", [faces.IN_VOLUME, faces.BY_INTERFACE])
@pytest.mark.parametrize("two", [True, False])
def test_face_term_fp64_against_oracle(orc, coast, two, mode):
    import torch

    case, mesh = coast
    s = _solver(case)
    plan = faces.FacePlan.from_mesh(s, mesh, two_to_one=two, kind=faces.LF_ADV, params=PARAMS)
    plan.set_conforming_mode(mode)
    E, n = case.E, case.n
    U = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda()
    dU = torch.zeros_like(U)
    plan.face_term(U, dU)
    got = synth.from_lines(dU.cpu().numpy(), E, n)
    cf, mt = plan.lists()
    want = orc.face_term(case.order, E, cf, mt, _mort_types(plan.groups(), mt.shape[0]),
                         faces.LF_ADV, PARAMS, case.U)
    errs = rel_l2_fields(got, want, case.nd)
    print(f"[faces] FP64 face term (two_to_one={two}, mode={mode}) rel L2 "
          f"{['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= TOL64, errs
    plan.close()
    s.close()


@pytest.mark.parametrize("mode", [faces.IN_VOLUME, faces.BY_INTERFACE])
def test_face_term_fp32_against_oracle(orc, coast, mode):
    import torch

    case, mesh = coast
    s = _solver(case, _lib.PREC_MIXED)
    plan = faces.FacePlan.from_mesh(s, mesh, two_to_one=True, kind=faces.LF_ADV, params=PARAMS)
    plan.set_conforming_mode(mode)
    E, n = case.E, case.n
    U = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda().float()
    dU = torch.zeros_like(U)
    plan.face_term(U, dU)
    got = synth.from_lines(dU.double().cpu().numpy(), E, n)
    cf, mt = plan.lists()
    want = orc.face_term(case.order, E, cf, mt, _mort_types(plan.groups(), mt.shape[0]),
                         faces.LF_ADV, PARAMS, case.U)
    errs = rel_l2_fields(got, want, case.nd)
    print(f"[faces] FP32 face term rel L2 {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= TOL32, errs
    plan.close()
    s.close()


@pytest.mark.parametrize("kind", [faces.UPWIND, faces.LF_ADV])
@pytest.mark.parametrize("layout", [_lib.LAYOUT_LINES, _lib.LAYOUT_SOA, _lib.LAYOUT_PACKED])
def test_face_term_conforming_orders_agree(coast, layout, kind):
    """The element-ordered conforming launch (each interface evaluated from
    both of its elements, the x, y and z contributions of a node summed
    before M^-1) and the interface-ordered launches give the same face term
    to rounding in every state layout: <= 1e-13 relative L2 per field of
    what they add to a non-zero dU, on the same set of nodes."""
    import torch

    case, mesh = coast
    s = _solver(case)
    params = PARAMS if kind == faces.LF_ADV else PARAMS[:3]
    plan = faces.FacePlan.from_mesh(s, mesh, two_to_one=True, kind=kind, params=params)
    E, n, nd = case.E, case.n, case.nd
    if layout == _lib.LAYOUT_LINES:
        host = synth.to_lines(case.U, E, n)
    elif layout == _lib.LAYOUT_SOA:
        host = synth.to_soa(case.U, E, nd)
    else:
        host = np.ascontiguousarray(case.U)
    U = torch.from_numpy(host).cuda()
    base = torch.from_numpy(np.sin(np.arange(host.size) * 0.37)).cuda()
    got = {}
    for mode in (faces.BY_ELEMENT, faces.BY_INTERFACE):
        plan.set_conforming_mode(mode)
        dU = base.clone()
        plan.face_term(U, dU, layout)
        torch.cuda.synchronize()
        got[mode] = dU.cpu().numpy()
    a, b = got[faces.BY_ELEMENT], got[faces.BY_INTERFACE]
    if layout == _lib.LAYOUT_LINES:
        a, b = synth.from_lines(a, E, n), synth.from_lines(b, E, n)
    elif layout == _lib.LAYOUT_SOA:
        a, b = synth.from_soa(a, E, nd), synth.from_soa(b, E, nd)
    ref = np.ascontiguousarray(np.sin(np.arange(host.size) * 0.37))
    if layout == _lib.LAYOUT_LINES:
        ref = synth.from_lines(ref, E, n)
    elif layout == _lib.LAYOUT_SOA:
        ref = synth.from_soa(ref, E, nd)
    errs = rel_l2_fields(a - ref, b - ref, nd)
    print(f"[faces] element vs interface order (layout {layout}, kind {kind}): "
          f"{['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-13, errs
    assert np.array_equal(a == ref, b == ref)  # the same nodes are touched
    plan.close()
    s.close()


@pytest.mark.parametrize("prec", [_lib.PREC_FP64, _lib.PREC_MIXED])
def test_face_kernels_on_blocked_state_bit_identical(coast, prec):
    """The coastline plan's face term (element by element and interface by
    interface), every group's face vectors and the right-hand side with the
    plan attached, on the blocked state, give the bits of the line layout."""
    import torch

    case, mesh = coast
    s = _solver(case, prec)
    plan = faces.FacePlan.from_mesh(s, mesh, two_to_one=True, kind=faces.LF_ADV, params=PARAMS)
    E, n = case.E, case.n
    B = _lib.load().drf_layout_block_elements(case.order)
    dt = torch.float64 if prec == _lib.PREC_FP64 else torch.float32
    Ul = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda().to(dt)
    Ub = torch.from_numpy(synth.to_blocks(case.U, E, n, B)).cuda().to(dt)

    def same(a, b):
        return np.array_equal(synth.from_lines(a.double().cpu().numpy(), E, n),
                              synth.from_blocks(b.double().cpu().numpy(), E, n, B))

    for mode in (faces.IN_VOLUME, faces.BY_INTERFACE):
        plan.set_conforming_mode(mode)
        Rl, Rb = torch.zeros_like(Ul), torch.zeros_like(Ub)
        plan.face_term(Ul, Rl, _lib.LAYOUT_LINES)
        plan.face_term(Ub, Rb, _lib.LAYOUT_BLOCKS)
        torch.cuda.synchronize()
        assert same(Rl, Rb), mode
    for g in range(len(plan.groups())):
        al, bl = plan.group_fluxes(g, Ul, _lib.LAYOUT_LINES)
        ab, bb = plan.group_fluxes(g, Ub, _lib.LAYOUT_BLOCKS)
        torch.cuda.synchronize()
        assert torch.equal(al, ab) and torch.equal(bl, bb), g
    plan.set_conforming_mode(faces.IN_VOLUME)
    plan.attach()
    Rl, Rb = torch.empty_like(Ul), torch.full_like(Ub, float("nan"))
    s.compute_rhs_dev(Ul, Rl, _lib.LAYOUT_LINES)
    s.compute_rhs_dev(Ub, Rb, _lib.LAYOUT_BLOCKS)
    torch.cuda.synchronize()
    assert same(Rl, Rb)
    plan.detach()
    plan.close()
    s.close()


def _rhs_with_faces(orc, case, U, cf, mt, mtypes):
    f = np.full(case.h.size, 1e-4)
    r = orc.rhs(case.order, U, case.h, f)
    return orc.face_term(case.order, case.E, cf, mt, mtypes, faces.LF_ADV, PARAMS, U, r)


@pytest.mark.parametrize("prec", [_lib.PREC_FP64, _lib.PREC_MIXED])
@pytest.mark.parametrize("layout", [_lib.LAYOUT_LINES, _lib.LAYOUT_SOA])
def test_rhs_faces_in_volume_kernel_bit_identical(coast, prec, layout):
    """The conforming face term added in the volume kernel's last phase gives
    the same bits as the volume kernel followed by the element-ordered face
    launch (the same arithmetic in the same order), at every workgroup size,
    FP64 and mixed; the mortars run as their own launches in both."""
    import torch

    case, mesh = coast
    s = _solver(case, prec)
    plan = faces.FacePlan.from_mesh(s, mesh, two_to_one=True, kind=faces.LF_ADV, params=PARAMS)
    E, n, nd = case.E, case.n, case.nd
    host = synth.to_lines(case.U, E, n) if layout == _lib.LAYOUT_LINES else \
        synth.to_soa(case.U, E, nd)
    dt = torch.float64 if prec == _lib.PREC_FP64 else torch.float32
    U = torch.from_numpy(host).cuda().to(dt)
    plan.attach()
    plan.set_conforming_mode(faces.BY_ELEMENT)
    ref = torch.empty_like(U)
    s.compute_rhs_dev(U, ref, layout)
    plan.set_conforming_mode(faces.IN_VOLUME)
    out = torch.empty_like(U)
    for lines in (128, 256, 512):
        s.set_launch(0, True, lines)
        out.fill_(float("nan"))
        s.compute_rhs_dev(U, out, layout)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), lines
    plan.detach()
    plan.close()
    s.close()


def test_rhs_and_unfused_rk4_with_faces(orc, coast):
    import torch

    case, mesh = coast
    s = _solver(case)
    plan = faces.FacePlan.from_mesh(s, mesh, two_to_one=True, kind=faces.LF_ADV, params=PARAMS)
    cf, mt = plan.lists()
    mtypes = _mort_types(plan.groups(), mt.shape[0])
    E, n = case.E, case.n
    plan.attach()
    U = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda()
    R = torch.empty_like(U)
    s.compute_rhs_dev(U, R, _lib.LAYOUT_LINES)
    want = _rhs_with_faces(orc, case, case.U, cf, mt, mtypes)
    errs = rel_l2_fields(synth.from_lines(R.cpu().numpy(), E, n), want, case.nd)
    print(f"[faces] RHS + faces rel L2 {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= TOL64, errs
    # the host-pointer RHSFunction drop-in adds the same term
    errs = rel_l2_fields(s.compute_rhs(case.U), want, case.nd)
    assert max(errs) <= TOL64, errs
    # fused stages carry no face term: refused
    with pytest.raises(_lib.DrfError) as ei:
        s.rk_step_dev(_lib.RK_RK4, 0.0, 1e-3, U, _lib.LAYOUT_LINES, fused=True)
    assert ei.value.status == _lib.DRF_ERR_UNSUPPORTED
    # one un-fused RK4 step (RK4Classic, src/solver/time_stepper.cpp:86-119)
    dt = 1e-3
    s.rk_step_dev(_lib.RK_RK4, 0.0, dt, U, _lib.LAYOUT_LINES, fused=False)
    f = lambda x: _rhs_with_faces(orc, case, x, cf, mt, mtypes)  # noqa: E731
    U0 = case.U
    k1 = f(U0)
    k2 = f(U0 + 0.5 * dt * k1)
    k3 = f(U0 + 0.5 * dt * k2)
    k4 = f(U0 + dt * k3)
    ref = U0 + dt / 6.0 * (((k1 + 2.0 * k2) + 2.0 * k3) + k4)
    errs = rel_l2_fields(synth.from_lines(U.cpu().numpy(), E, n), ref, case.nd)
    print(f"[faces] un-fused RK4 step with faces rel L2 {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= TOL64, errs
    plan.detach()
    s.close()
    plan.close()


def test_mode_split_30_substeps_graph_on_coastline(orc, coast):
    import torch

    case, _ = coast
    E, nd = case.E, case.nd
    ms = ModeSplittingSolver(case.order, subcycles=30)
    ms.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    full = case.full_state()
    eager = torch.from_numpy(full.copy()).cuda()
    replay = eager.clone()
    dt = 2e-3
    g = ms.capture_steps_dev(dt, replay, n_steps=3)
    for rep in range(2):
        for _ in range(3):
            ms.step_dev(dt, eager)
        g.launch()
        torch.cuda.synchronize()
        assert np.isfinite(replay.cpu().numpy()).all()
        assert torch.equal(eager, replay), f"replay {rep} differs"
    g.close()
    # the mode split is element-local: the sampled elements alone give the
    # oracle's values for them
    idx = sample_elements(E, 300, seed=8)
    F = rows(full, idx, 12 * nd)
    h = rows(case.h, idx, nd)
    f = np.full(h.size, 1e-4)
    for _ in range(6):
        F = orc.mode_split_step(case.order, dt, 30, F, h, f)
    got = rows(eager.cpu().numpy(), idx, 12 * nd).reshape(-1, 12, nd)
    ref = F.reshape(-1, 12, nd)
    for fld in range(12):
        den = np.linalg.norm(ref[:, fld])
        err = np.linalg.norm(got[:, fld] - ref[:, fld]) / (den if den > 0 else 1.0)
        assert err <= TOL64, (fld, err)


@pytest.fixture(scope="module")
def c4():
    print("[faces] building the 128x128x16 N=4 case", flush=True)
    case = synth.config4()
    m = OctreeAdapter(0.0, 100000.0, 0.0, 100000.0, -1.0, 0.0)
    m.build_uniform(128, 128, 16)
    b, lv, _ = m.elements()
    assert np.array_equal(b, case.bounds)
    return case, m, b, lv


@pytest.mark.parametrize("prec", [_lib.PREC_FP64, _lib.PREC_MIXED])
def test_face_term_128x128x16_on_sampled_elements(orc, c4, prec):
    """Every face of the uniform mesh is conforming (three groups, one per
    axis); the face term of an element depends on it and its face
    neighbours, so the oracle run over the interfaces that touch the sampled
    elements gives their exact values."""
    import torch

    case, m, b, lv = c4
    s = _solver(case, prec)
    plan = faces.FacePlan(s, b, lv, m.build_face_connections(two_to_one=True),
                          kind=faces.LF_ADV, params=PARAMS)
    ng, nc, nm = plan.info()
    assert (ng, nm) == (3, 0)
    E, n, nd = case.E, case.n, case.nd
    dt = torch.float64 if prec == _lib.PREC_FP64 else torch.float32
    U = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda().to(dt)
    dU = torch.zeros_like(U)
    plan.face_term(U, dU)
    idx = sample_elements(E, 300, seed=9)
    from sampling import lines_rows

    got = lines_rows(dU, idx, E, n)
    cf, _ = plan.lists()
    touch = np.isin(cf[:, 0], idx) | np.isin(cf[:, 2], idx)
    want = orc.face_term(case.order, E, cf[touch], np.zeros((0, 6), np.int64), [],
                         faces.LF_ADV, PARAMS, case.U)
    want = rows(want, idx, 5 * nd)
    errs = rel_l2_fields(got, want, nd)
    tol = TOL64 if prec == _lib.PREC_FP64 else TOL32
    print(f"[faces] 128x128x16 face term prec={prec} rel L2 {['%.2e' % e for e in errs]}",
          flush=True)
    assert max(errs) <= tol, errs
    plan.close()
    s.close()
