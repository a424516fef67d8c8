"""GPU parity of the time steppers, CFL reduction and mode splitting against
the FP64 oracle.

RK steps: TimeStepper::step for ForwardEuler, RK2Midpoint, SSPRK3, RK4Classic,
LSRK45 and RK45DormandPrince (src/solver/time_stepper.cpp:12-201), fused into
the RHS epilogue and un-fused, relative L2 <= 1e-12 per field.
Wave speed: SimulationDriver::compute_max_wave_speed, bit-exact.
Mode split: ModeSplittingSolver::step (src/physics/mode_splitting.cpp:434-453)
driven like SimulationDriver::step (dt from the CFL controller), 1 and 10
steps with 20 substeps, relative L2 <= 1e-12 per field of the full state.
"""
import numpy as np
import pytest

from drifter_amd import _lib, synth
from drifter_amd.solver import (CoriolisParameter, ModeSplittingSolver,
                                PrimitiveEquationsSolver)

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_l2(a, b, E, nf, nd):
    A = a.reshape(E, nf, nd)
    B = b.reshape(E, nf, nd)
    out = []
    for f in range(nf):
        den = np.linalg.norm(B[:, f])
        num = np.linalg.norm(A[:, f] - B[:, f])
        out.append(num / den if den > 0 else num)
    return out


def _driver_dt(orc, case):
    H, u, v, _, _ = orc.derived(case.order, case.U, case.h)
    c = orc.max_wave_speed(case.order, H, u, v)
    return orc.compute_dt(c, 2.0 / (case.order + 1), 0.5, 1e-6, 3600.0), c


@pytest.mark.parametrize("scheme", range(6))
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("layout", [_lib.LAYOUT_SOA, _lib.LAYOUT_LINES])
def test_rk_step_matches_oracle(orc, scheme, fused, layout):
    import torch

    case = synth.config1()
    dt, _ = _driver_dt(orc, case)
    print(f"[rk] scheme={scheme} fused={fused} layout={layout} dt={dt:.4e}", flush=True)
    ref = orc.rk_step(scheme, 3, 0.0, dt, case.U, case.h, np.full(case.h.size, 1e-4))
    s = PrimitiveEquationsSolver(3)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    if layout == _lib.LAYOUT_SOA:
        U = torch.from_numpy(synth.to_soa(case.U, case.E, case.nd)).cuda()
    else:
        U = torch.from_numpy(synth.to_lines(case.U, case.E, case.n)).cuda()
    s.rk_step_dev(scheme, 0.0, dt, U, layout, fused=fused)
    torch.cuda.synchronize()
    if layout == _lib.LAYOUT_SOA:
        got = synth.from_soa(U.cpu().numpy(), case.E, case.nd)
    else:
        got = synth.from_lines(U.cpu().numpy(), case.E, case.n)
    errs = rel_l2(got, ref, case.E, 5, case.nd)
    print(f"  rel L2: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= TOL
    s.close()


def test_host_rk4_matches_oracle(orc):
    case = synth.make_case(4, 4, 2, 4, 100000.0, "shelf")
    dt, _ = _driver_dt(orc, case)
    s = PrimitiveEquationsSolver(4)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    got = s.rk_step(_lib.RK_RK4, 0.0, dt, case.U)
    ref = orc.rk_step(3, 4, 0.0, dt, case.U, case.h, np.full(case.h.size, 1e-4))
    assert max(rel_l2(got, ref, case.E, 5, case.nd)) <= TOL
    s.close()


def test_wave_speed_bit_exact(orc):
    import torch

    for case in (synth.config1(), synth.make_case(16, 16, 2, 4, 100000.0, "shelf")):
        H, u, v, _, _ = orc.derived(case.order, case.U, case.h)
        ref = orc.max_wave_speed(case.order, H, u, v)
        s = PrimitiveEquationsSolver(case.order)
        s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
        assert s.max_wave_speed(case.U) == ref
        Ud = torch.from_numpy(synth.to_soa(case.U, case.E, case.nd)).cuda()
        assert s.max_wave_speed_dev(Ud, _lib.LAYOUT_SOA) == ref
        Ul = torch.from_numpy(synth.to_lines(case.U, case.E, case.n)).cuda()
        assert s.max_wave_speed_dev(Ul, _lib.LAYOUT_LINES) == ref
        Hd, ud, vd = (torch.from_numpy(a).cuda() for a in (H, u, v))
        assert s.max_wave_speed_derived_dev(Hd, ud, vd, case.nd) == ref
        s.close()


@pytest.mark.parametrize("order", range(1, 9))
def test_wave_speed_every_order_and_ragged_batches(orc, order):
    """one thread per grid line, n^2 lines per element: element counts that
    leave the last batch of a block part-filled, every order"""
    import torch

    case = synth.make_case(5, 3, 3, order, 100000.0, "shelf")  # 45 elements
    H, u, v, _, _ = orc.derived(case.order, case.U, case.h)
    ref = orc.max_wave_speed(case.order, H, u, v)
    s = PrimitiveEquationsSolver(case.order)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    print(f"[wave speed] order={order} E={case.E}", flush=True)
    Ud = torch.from_numpy(synth.to_soa(case.U, case.E, case.nd)).cuda()
    assert s.max_wave_speed_dev(Ud, _lib.LAYOUT_SOA) == ref
    Ul = torch.from_numpy(synth.to_lines(case.U, case.E, case.n)).cuda()
    assert s.max_wave_speed_dev(Ul, _lib.LAYOUT_LINES) == ref
    s.close()


def test_wave_speed_grid_stride(orc):
    """64x64x8 at N = 4: more element batches (3,277) than blocks (2,048)"""
    import torch

    case = synth.config2()
    H, u, v, _, _ = orc.derived(case.order, case.U, case.h)
    ref = orc.max_wave_speed(case.order, H, u, v)
    s = PrimitiveEquationsSolver(4)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    Ul = torch.from_numpy(synth.to_lines(case.U, case.E, case.n)).cuda()
    print(f"[wave speed] grid stride E={case.E}", flush=True)
    assert s.max_wave_speed_dev(Ul, _lib.LAYOUT_LINES) == ref
    s.close()


@pytest.mark.parametrize("order,mesh", [(3, (8, 8, 4)), (4, (8, 8, 2))])
def test_mode_split_matches_oracle(orc, order, mesh):
    import torch

    case = synth.make_case(*mesh, order, 100000.0 if order == 4 else 10000.0,
                           "shelf" if order == 4 else "seamount")
    E, nd = case.E, case.nd
    f = np.full(E * nd, 1e-4)
    ms = ModeSplittingSolver(order, subcycles=20)
    ms.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    ref = case.full_state()
    dev = torch.from_numpy(ref.copy()).cuda()
    for step in range(10):
        F = ref.reshape(E, 12, nd)
        c = orc.max_wave_speed(order, np.ascontiguousarray(F[:, 5]),
                               np.ascontiguousarray(F[:, 6]), np.ascontiguousarray(F[:, 7]))
        dt = orc.compute_dt(c, 2.0 / (order + 1), 0.5, 1e-6, 3600.0)
        ref = orc.mode_split_step(order, dt, 20, ref, case.h, f)
        ms.step_dev(dt, dev)
        torch.cuda.synchronize()
        got = dev.cpu().numpy()
        errs = rel_l2(got, ref, E, 12, nd)
        print(f"[mode split] N={order} step {step + 1}: max rel L2 {max(errs):.2e}", flush=True)
        assert max(errs) <= TOL, errs
    # the host-pointer entry point gives the same result as the device one
    one = ms.step(1e-3, case.full_state())
    ref1 = orc.mode_split_step(order, 1e-3, 20, case.full_state(), case.h, f)
    assert max(rel_l2(one, ref1, E, 12, nd)) <= TOL
