"""drf_max_wave_speed_noalloc_dev (the wave-speed maximum with its partial
maxima in the library's file-scope scratch array) against
drf_max_wave_speed_dev (SimulationDriver::compute_max_wave_speed,
src/solver/simulation_driver.cpp:153-168) on the same states: bit-identical,
on two calls in a row, the second on a smaller mesh, so the second call
reads only partials it wrote itself."""
import numpy as np
import pytest

from drifter_amd import _lib, synth
from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

pytestmark = pytest.mark.gpu


def _solver_and_state(case, E, layout):
    import torch

    nd = case.nd
    s = PrimitiveEquationsSolver(case.order)
    s.initialize(E, case.h[:E * nd], coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    U = case.U[:E * 5 * nd]
    if layout == _lib.LAYOUT_LINES:
        U = synth.to_lines(U, E, case.n)
    elif layout == _lib.LAYOUT_SOA:
        U = synth.to_soa(U, E, nd)
    return s, torch.from_numpy(np.ascontiguousarray(U)).cuda()


@pytest.mark.parametrize("layout", [_lib.LAYOUT_LINES, _lib.LAYOUT_PACKED])
def test_noalloc_wave_speed_matches_the_allocating_entry_point(layout):
    import torch

    big = synth.config2()  # 32,768 elements: 2,048 partial maxima, the scratch's size
    for E in (big.E, 1000):  # the second call smaller
        print(f"[no-alloc] E={E} layout={layout}", flush=True)
        s, U = _solver_and_state(big, E, layout)
        want = s.max_wave_speed_dev(U, layout)
        out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
        s.max_wave_speed_noalloc_dev(U, out, layout)
        torch.cuda.synchronize()
        got = float(out.item())
        assert got == want and want > 0.0
        s.close()


def test_noalloc_wave_speed_rejects_a_mixed_context():
    import torch

    case = synth.config1()
    s = PrimitiveEquationsSolver(case.order, _lib.PREC_MIXED)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    U = torch.zeros(case.E * 5 * case.nd, dtype=torch.float32, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.DrfError) as ei:
        s.max_wave_speed_noalloc_dev(U, out, _lib.LAYOUT_LINES)
    assert ei.value.status == _lib.DRF_ERR_UNSUPPORTED
    s.close()
