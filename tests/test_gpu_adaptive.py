"""Device-resident adaptive stepping: SimulationDriver::step
(src/solver/simulation_driver.cpp:108-131) with the CFL time step computed on
the GPU (compute_max_wave_speed :153-168, AdaptiveTimeController::compute_dt
src/solver/time_stepper.cpp:261-268) and read by the fused stage kernels.
It must equal, bit for bit, the host loop SimulationDriver::step runs: wave speed to the
host, compute_dt there, a fused step with that dt; and a captured graph of
such steps must equal the eager steps over several replays.  One step is
also checked against the oracle driven the same way."""
import numpy as np
import pytest

from drifter_amd import _lib, synth
from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

pytestmark = pytest.mark.gpu

CFL, DT_MIN, DT_MAX = 0.5, 1e-6, 3600.0


def _setup(layout, precision=_lib.PREC_FP64):
    import torch

    case = synth.config1()
    s = PrimitiveEquationsSolver(3, precision)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    if layout == _lib.LAYOUT_SOA:
        U = torch.from_numpy(synth.to_soa(case.U, case.E, case.nd)).cuda()
    else:
        U = torch.from_numpy(synth.to_lines(case.U, case.E, case.n)).cuda()
    return case, s, U


def _host_loop(s, scheme, U, n, layout, t0=0.0, dt_max=DT_MAX):
    lib = _lib.load()
    t, dt = t0, None
    for _ in range(n):
        c = s.max_wave_speed_dev(U, layout)
        dt = lib.drf_compute_dt(c, 2.0 / (s.order + 1), CFL, DT_MIN, dt_max)
        s.rk_step_dev(scheme, t, dt, U, layout, fused=True)
        t += dt
    return t, dt


@pytest.mark.parametrize("scheme", [_lib.RK_EULER, _lib.RK_RK2, _lib.RK_SSPRK3, _lib.RK_RK4,
                                    _lib.RK_LSRK45])
@pytest.mark.parametrize("layout", [_lib.LAYOUT_SOA, _lib.LAYOUT_LINES])
def test_device_steps_equal_the_host_loop(scheme, layout):
    import torch

    case, s, U = _setup(layout)
    print(f"[adaptive] scheme={scheme} layout={layout}", flush=True)
    Uh = U.clone()
    t_ref, dt_ref = _host_loop(s, scheme, Uh, 4, layout)
    dt_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    t_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    for _ in range(4):
        s.rk_step_adaptive_dev(scheme, U, dt_dev, t_dev, cfl=CFL, dt_min=DT_MIN, dt_max=DT_MAX,
                               layout=layout)
    torch.cuda.synchronize()
    assert torch.equal(U, Uh)
    assert dt_dev.item() == dt_ref and t_dev.item() == t_ref


def test_dt_clamps_and_time_is_optional():
    import torch

    layout = _lib.LAYOUT_SOA
    case, s, U = _setup(layout)
    print("[adaptive] clamp to dt_max, no time output", flush=True)
    Uh = U.clone()
    _, dt_ref = _host_loop(s, _lib.RK_RK4, Uh, 2, layout, dt_max=1e-5)
    assert dt_ref == 1e-5  # the CFL step is larger: clamped
    dt_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    for _ in range(2):
        s.rk_step_adaptive_dev(_lib.RK_RK4, U, dt_dev, None, cfl=CFL, dt_min=DT_MIN, dt_max=1e-5,
                               layout=layout)
    torch.cuda.synchronize()
    assert torch.equal(U, Uh) and dt_dev.item() == 1e-5


def test_graph_of_adaptive_steps_replays_exactly():
    import torch

    layout = _lib.LAYOUT_LINES
    case, s, U = _setup(layout)
    print("[adaptive] graph of 3 steps x 3 replays", flush=True)
    Ue = U.clone()
    dte = torch.zeros(1, dtype=torch.float64, device="cuda")
    te = torch.zeros(1, dtype=torch.float64, device="cuda")
    dtg = torch.zeros(1, dtype=torch.float64, device="cuda")
    tg = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = s.capture_adaptive_steps(_lib.RK_SSPRK3, U, dtg, tg, n_steps=3, cfl=CFL, dt_min=DT_MIN,
                                 dt_max=DT_MAX, layout=layout)
    for rep in range(3):
        for _ in range(3):
            s.rk_step_adaptive_dev(_lib.RK_SSPRK3, Ue, dte, te, cfl=CFL, dt_min=DT_MIN,
                                   dt_max=DT_MAX, layout=layout)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        scratch = torch.randn(1000, device="cuda").sum()  # eager work between replays
        assert torch.isfinite(scratch)
        assert torch.equal(U, Ue), rep
        assert tg.item() == te.item() and dtg.item() == dte.item(), rep


def test_one_adaptive_step_matches_the_oracle(orc):
    import torch

    layout = _lib.LAYOUT_SOA
    case, s, U = _setup(layout)
    H, u, v, _, _ = orc.derived(case.order, case.U, case.h)
    c = orc.max_wave_speed(case.order, H, u, v)
    dt = orc.compute_dt(c, 2.0 / (case.order + 1), CFL, DT_MIN, DT_MAX)
    ref = orc.rk_step(_lib.RK_RK4, 3, 0.0, dt, case.U, case.h, np.full(case.h.size, 1e-4))
    dt_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    print("[adaptive] one RK4 step vs oracle", flush=True)
    s.rk_step_adaptive_dev(_lib.RK_RK4, U, dt_dev, None, cfl=CFL, dt_min=DT_MIN, dt_max=DT_MAX,
                           layout=layout)
    torch.cuda.synchronize()
    assert dt_dev.item() == dt
    got = synth.from_soa(U.cpu().numpy(), case.E, case.nd)
    A = got.reshape(case.E, 5, case.nd)
    B = ref.reshape(case.E, 5, case.nd)
    for f in range(5):
        assert np.linalg.norm(A[:, f] - B[:, f]) <= 1e-12 * np.linalg.norm(B[:, f]), f


def test_unsupported_configurations_are_rejected():
    import torch

    case, s, U = _setup(_lib.LAYOUT_SOA)
    dt_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    print("[adaptive] rejections", flush=True)
    with pytest.raises(_lib.DrfError):
        s.rk_step_adaptive_dev(_lib.RK_DP45, U, dt_dev)
    with pytest.raises(_lib.DrfError):
        s.rk_step_adaptive_dev(_lib.RK_RK4, U, None)
    _, s32, U32 = _setup(_lib.LAYOUT_SOA, _lib.PREC_MIXED)
    with pytest.raises(_lib.DrfError):
        s32.rk_step_adaptive_dev(_lib.RK_RK4, U32.float(), dt_dev)
