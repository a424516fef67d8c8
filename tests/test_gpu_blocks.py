"""The blocked state layout [block][field][xi][element in block][line]
gives the same bits as the line layout in every path that takes a layout:
the right-hand side at every order, precision and workgroup size, the fused
and un-fused Runge-Kutta steps, the CFL wave speed (also through the
entry point that allocates nothing), the autotuner, and the right-hand side
with a face plan attached.  A mesh of
13 x 7 x 5 = 455 elements leaves the last block partly filled at every order
(blocks of 128, 56, 32, 20, 3, 2, 2 and 1 elements at orders 1 to 8)."""
import numpy as np
import pytest

from drifter_amd import _lib, faces, synth
from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

pytestmark = pytest.mark.gpu


def _pair(case, prec=_lib.PREC_FP64):
    """(solver, LINES state, BLOCKS state) of one case on the device"""
    import torch

    s = PrimitiveEquationsSolver(case.order, prec)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 2e-11, 0.0),
                 y_positions=case.y)
    B = _lib.load().drf_layout_block_elements(case.order)
    dt = torch.float64 if prec == _lib.PREC_FP64 else torch.float32
    Ul = torch.from_numpy(synth.to_lines(case.U, case.E, case.n)).cuda().to(dt)
    Ub = torch.from_numpy(synth.to_blocks(case.U, case.E, case.n, B)).cuda().to(dt)
    assert Ub.numel() == s.state_len(_lib.LAYOUT_BLOCKS)
    return s, Ul, Ub, B


def _same(lines, blocks, case, B):
    a = synth.from_lines(lines.double().cpu().numpy(), case.E, case.n)
    b = synth.from_blocks(blocks.double().cpu().numpy(), case.E, case.n, B)
    return np.array_equal(a, b)


@pytest.mark.parametrize("order", range(1, 9))
@pytest.mark.parametrize("prec", [_lib.PREC_FP64, _lib.PREC_MIXED])
def test_rhs_blocks_bit_identical_to_lines(order, prec):
    import torch

    case = synth.make_case(13, 7, 5, order, 10000.0, "seamount")
    print(f"[blocks] rhs order={order} prec={prec} E={case.E}", flush=True)
    s, Ul, Ub, B = _pair(case, prec)
    Rl = torch.empty_like(Ul)
    s.compute_rhs_dev(Ul, Rl, _lib.LAYOUT_LINES)
    Rb = torch.full_like(Ub, float("nan"))
    for lines in ((128, 256, 512) if order <= 4 else (128,)):
        for wpc in (0, 1, 3):
            s.set_launch(wpc, True, lines)
            Rb.fill_(float("nan"))
            s.compute_rhs_dev(Ub, Rb, _lib.LAYOUT_BLOCKS)
            torch.cuda.synchronize()
            assert _same(Rl, Rb, case, B), (lines, wpc)
    s.close()


@pytest.mark.parametrize("scheme", [_lib.RK_RK4, _lib.RK_SSPRK3, _lib.RK_LSRK45, _lib.RK_DP45])
@pytest.mark.parametrize("fused", [True, False])
def test_rk_steps_blocks_bit_identical_to_lines(scheme, fused):
    import torch

    case = synth.make_case(13, 7, 5, 3, 10000.0, "seamount")
    print(f"[blocks] rk scheme={scheme} fused={fused}", flush=True)
    s, Ul, Ub, B = _pair(case)
    c = s.max_wave_speed_dev(Ul, _lib.LAYOUT_LINES)
    dt = _lib.load().drf_compute_dt(c, 2.0 / case.n, 0.5, 1e-6, 3600.0)  # the driver's CFL step
    for _ in range(3):
        s.rk_step_dev(scheme, 0.0, dt, Ul, _lib.LAYOUT_LINES, fused=fused)
        s.rk_step_dev(scheme, 0.0, dt, Ub, _lib.LAYOUT_BLOCKS, fused=fused)
    torch.cuda.synchronize()
    assert torch.isfinite(Ul).all()
    assert _same(Ul, Ub, case, B)
    s.close()


@pytest.mark.parametrize("order", [1, 3, 4, 7])
def test_wave_speed_blocks_bit_identical_to_lines(order):
    import torch

    case = synth.make_case(13, 7, 5, order, 10000.0, "seamount")
    s, Ul, Ub, _ = _pair(case)
    ref = s.max_wave_speed_dev(Ul, _lib.LAYOUT_LINES)
    assert s.max_wave_speed_dev(Ub, _lib.LAYOUT_BLOCKS) == ref
    c = torch.zeros(1, dtype=torch.float64, device="cuda")
    s.max_wave_speed_noalloc_dev(Ub, c, _lib.LAYOUT_BLOCKS)
    torch.cuda.synchronize()
    assert float(c.cpu()[0]) == ref
    s.close()


def test_autotune_on_blocks_keeps_results():
    import torch

    case = synth.make_case(13, 7, 5, 4, 10000.0, "seamount")
    s, Ul, Ub, B = _pair(case)
    Rl = torch.empty_like(Ul)
    s.compute_rhs_dev(Ul, Rl, _lib.LAYOUT_LINES)
    Rb = torch.empty_like(Ub)
    wpc, remap, lines, us = s.autotune(Ub, Rb, _lib.LAYOUT_BLOCKS, reps=2)
    assert 1 <= wpc <= 8 and lines in (128, 256, 512) and us > 0
    Rb.fill_(float("nan"))
    s.compute_rhs_dev(Ub, Rb, _lib.LAYOUT_BLOCKS)
    torch.cuda.synchronize()
    assert _same(Rl, Rb, case, B)
    s.close()


def test_rhs_with_face_plan_blocks_bit_identical_to_lines():
    """With a face plan attached the blocked state takes the element-ordered
    face launch after the volume kernel, the line-layout state the fused face
    phase: the same arithmetic, the same bits."""
    import torch

    from drifter_amd.mesh import OctreeAdapter

    L = 40000.0
    m = OctreeAdapter(0.0, L, 0.0, L, -1.0, 0.0)
    m.build_uniform(4, 4, 2)
    bounds, _, _ = m.elements()
    case = synth.make_case_on(bounds, 3, L, "seamount")
    for prec in (_lib.PREC_FP64, _lib.PREC_MIXED):
        s, Ul, Ub, B = _pair(case, prec)
        plan = faces.FacePlan.from_mesh(s, m, two_to_one=False, kind=faces.UPWIND)
        plan.attach()
        Rl, Rb = torch.empty_like(Ul), torch.full_like(Ub, float("nan"))
        s.compute_rhs_dev(Ul, Rl, _lib.LAYOUT_LINES)
        s.compute_rhs_dev(Ub, Rb, _lib.LAYOUT_BLOCKS)
        torch.cuda.synchronize()
        assert _same(Rl, Rb, case, B), prec
        plan.detach()
        plan.close()
        s.close()


def test_adaptive_steps_and_graph_on_blocks():
    """SimulationDriver::step on the device (CFL dt from the wave speed, fused
    RK4) on the blocked state gives the line layout's state, dt and time;
    a captured graph of those steps replays them exactly, with eager work
    between the replays."""
    import torch

    case = synth.make_case(13, 7, 5, 3, 10000.0, "seamount")
    s, Ul, Ub, B = _pair(case)
    z = lambda: torch.zeros(1, dtype=torch.float64, device="cuda")  # noqa: E731
    dtl, tl, dtb, tb = z(), z(), z(), z()
    for _ in range(3):
        s.rk_step_adaptive_dev(_lib.RK_RK4, Ul, dtl, tl, cfl=0.5, layout=_lib.LAYOUT_LINES)
        s.rk_step_adaptive_dev(_lib.RK_RK4, Ub, dtb, tb, cfl=0.5, layout=_lib.LAYOUT_BLOCKS)
    torch.cuda.synchronize()
    assert _same(Ul, Ub, case, B)
    assert dtl.item() == dtb.item() > 0 and tl.item() == tb.item()
    Ug = Ub.clone()
    dtg, tg = dtb.clone(), tb.clone()
    g = s.capture_adaptive_steps(_lib.RK_RK4, Ug, dtg, tg, n_steps=2, cfl=0.5,
                                 layout=_lib.LAYOUT_BLOCKS)
    for rep in range(3):
        for _ in range(2):
            s.rk_step_adaptive_dev(_lib.RK_RK4, Ub, dtb, tb, cfl=0.5, layout=_lib.LAYOUT_BLOCKS)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.isfinite(torch.randn(1000, device="cuda").sum())
        assert torch.equal(Ug, Ub), rep
        assert tg.item() == tb.item() and dtg.item() == dtb.item(), rep
    g.close()
    s.close()


@pytest.mark.parametrize("order", [1, 4, 8])
def test_mesh_smaller_than_one_block(order):
    """A mesh of 8 elements (the uniform mesh rounds 7 x 1 x 1 up to a power
    of two): one partly filled block at orders 1 and 4, whose blocks hold 128
    and 20 elements; eight full blocks of one element at order 8."""
    import torch

    case = synth.make_case(7, 1, 1, order, 10000.0, "seamount")
    E = case.E
    s, Ul, Ub, B = _pair(case)
    Rl, Rb = torch.empty_like(Ul), torch.full_like(Ub, float("nan"))
    s.compute_rhs_dev(Ul, Rl, _lib.LAYOUT_LINES)
    s.compute_rhs_dev(Ub, Rb, _lib.LAYOUT_BLOCKS)
    s.rk_step_dev(_lib.RK_RK4, 0.0, 1e-3, Ul, _lib.LAYOUT_LINES, fused=True)
    s.rk_step_dev(_lib.RK_RK4, 0.0, 1e-3, Ub, _lib.LAYOUT_BLOCKS, fused=True)
    torch.cuda.synchronize()
    assert _same(Rl, Rb, case, B) and _same(Ul, Ub, case, B)
    # the slots past the last element keep what the caller put there
    if B > E:
        n = case.n
        assert torch.isnan(Rb.view(1, 5, n, B, n * n)[:, :, :, E:, :]).all()
        assert not torch.isnan(Rb.view(1, 5, n, B, n * n)[:, :, :, :E, :]).any()
    assert s.max_wave_speed_dev(Ul, _lib.LAYOUT_LINES) == \
        s.max_wave_speed_dev(Ub, _lib.LAYOUT_BLOCKS)
    s.close()
