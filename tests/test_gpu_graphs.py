"""HIP graphs of repeated steps (drf_graph_rk, drf_graph_mode_split): every
replay must give exactly the bits of the same steps launched eagerly.  Each
graph is replayed several times, with the program's own eager work (RHS
evaluations, torch copies, synchronisations) between replays."""
import numpy as np
import pytest

from drifter_amd import _lib, synth
from drifter_amd.solver import CoriolisParameter, ModeSplittingSolver, PrimitiveEquationsSolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme,fused", [(_lib.RK_RK4, True), (_lib.RK_SSPRK3, True),
                                          (_lib.RK_LSRK45, True), (_lib.RK_DP45, False)])
def test_rk_graph_replays_bitwise(scheme, fused):
    import torch

    print(f"[graph rk] scheme={scheme} fused={fused}", flush=True)
    case = synth.config1()
    E, n = case.E, case.n
    s = PrimitiveEquationsSolver(case.order)
    s.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    U0 = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda()
    eager, replay = U0.clone(), U0.clone()
    g = s.capture_rk_steps(scheme, 1e-4, replay, _lib.LAYOUT_LINES, fused, n_steps=3)
    scratch = torch.empty_like(U0)
    for rep in range(3):
        for _ in range(3):
            s.rk_step_dev(scheme, 0.0, 1e-4, eager, _lib.LAYOUT_LINES, fused)
        g.launch()
        s.compute_rhs_dev(eager, scratch, _lib.LAYOUT_LINES)  # eager work between replays
        scratch.copy_(eager)
        torch.cuda.synchronize()
        assert torch.equal(eager, replay), f"replay {rep} differs"
    g.close()


def test_mode_split_graph_replays_bitwise():
    import torch

    case = synth.config1()
    E = case.E
    ms = ModeSplittingSolver(case.order, subcycles=20)
    ms.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    full = case.full_state()
    eager = torch.from_numpy(full.copy()).cuda()
    replay = eager.clone()
    g = ms.capture_steps_dev(2e-3, replay, n_steps=4)
    for rep in range(3):
        for _ in range(4):
            ms.step_dev(2e-3, eager)
        g.launch()
        torch.cuda.synchronize()
        tmp = replay.cpu().numpy()  # host round trip between replays
        assert np.isfinite(tmp).all()
        assert torch.equal(eager, replay), f"replay {rep} differs"
    g.close()


def test_graph_arguments_rejected():
    import ctypes as C

    lib = _lib.load()
    h = C.c_void_p()
    assert lib.drf_graph_rk(None, 3, 1e-3, None, 2, 1, 1, C.byref(h)) == _lib.DRF_ERR_ARG
