"""GPU parity of the volume right-hand side against the FP64 oracle
(PrimitiveEquationsSolver::compute_rhs, src/physics/primitive_equations.cpp:428-442).

Bar: relative L2 <= 1e-12 per field (Hu, Hv, eta, HT, HS).
"""
import numpy as np
import pytest

from drifter_amd import _lib, synth
from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_l2_fields(a, b, E, nd):
    A = a.reshape(E, 5, nd)
    B = b.reshape(E, 5, nd)
    out = []
    for f in range(5):
        den = np.linalg.norm(B[:, f])
        out.append(np.linalg.norm(A[:, f] - B[:, f]) / (den if den > 0 else 1.0))
    return out


def _check(case, f_nodes, coriolis, orc, label):
    print(f"[rhs parity] {label}: E={case.E} N={case.order}", flush=True)
    s = PrimitiveEquationsSolver(case.order)
    s.initialize(case.E, case.h, coriolis=coriolis, y_positions=case.y)
    got = s.compute_rhs(case.U)
    ref = orc.rhs(case.order, case.U, case.h, f_nodes)
    errs = rel_l2_fields(got, ref, case.E, case.nd)
    print(f"  rel L2 per field: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= TOL, errs
    s.close()
    return got


def test_config1_seamount_n3(orc):
    case = synth.config1()
    _check(case, np.full(case.h.size, 1e-4), CoriolisParameter(1e-4, 0.0, 0.0), orc, "8x8x4 N=3")


def test_shelf_n4_with_beta_plane(orc):
    case = synth.make_case(16, 16, 4, 4, 100000.0, "shelf")
    cor = CoriolisParameter(1e-4, 2e-11, 0.0)
    _check(case, cor(case.y), cor, orc, "16x16x4 N=4 beta-plane")


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 8])
def test_all_orders(orc, order):
    case = synth.make_case(4, 2, 2, order, 10000.0, "seamount", seed_base=77 + order)
    cor = CoriolisParameter(1.2e-4, 1e-11, 3000.0)
    _check(case, cor(case.y), cor, orc, f"4x2x2 N={order}")


def test_ragged_last_block(orc):
    # element counts that do not fill the last workgroup
    for nx, ny, nz in ((1, 1, 1), (3, 1, 1), (7, 1, 1)):
        case = synth.make_case(nx, ny, nz, 4, 5000.0, "flat")
        _check(case, np.full(case.h.size, 1e-4), CoriolisParameter(1e-4, 0.0, 0.0), orc,
               f"ragged {nx}x{ny}x{nz}")


def test_layouts_agree_bitwise_and_full_config2_sampled(orc):
    """64x64x8, N=4 (32,768 elements) on device in the packed, SoA and LINES layouts;
    every element is checked against the oracle on a sample (the RHS is
    element-local, so a sample of elements is a self-contained input)."""
    import torch

    case = synth.config2()
    E, nd = case.E, case.nd
    print(f"[rhs parity] config2 E={E}", flush=True)
    s = PrimitiveEquationsSolver(4)
    s.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    dev = torch.device("cuda", 0)
    Up = torch.from_numpy(case.U).to(dev)
    Us = torch.from_numpy(synth.to_soa(case.U, E, nd)).to(dev)
    Ul = torch.from_numpy(synth.to_lines(case.U, E, 5)).to(dev)
    Rp = torch.empty_like(Up)
    Rs = torch.empty_like(Us)
    Rl = torch.empty_like(Ul)
    s.compute_rhs_dev(Up, Rp, _lib.LAYOUT_PACKED)
    s.compute_rhs_dev(Us, Rs, _lib.LAYOUT_SOA)
    s.compute_rhs_dev(Ul, Rl, _lib.LAYOUT_LINES)
    torch.cuda.synchronize()
    rp = Rp.cpu().numpy()
    rs = synth.from_soa(Rs.cpu().numpy(), E, nd)
    rl = synth.from_lines(Rl.cpu().numpy(), E, 5)
    assert np.array_equal(rp, rs)
    assert np.array_equal(rp, rl)
    assert np.all(np.isfinite(rp))
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(E, 400, replace=False))
    idx = np.unique(np.r_[0, idx, E - 1])
    Usub = np.ascontiguousarray(case.U.reshape(E, -1)[idx].reshape(-1))
    hsub = np.ascontiguousarray(case.h.reshape(E, nd)[idx].reshape(-1))
    ref = orc.rhs(4, Usub, hsub, np.full(hsub.size, 1e-4))
    got = np.ascontiguousarray(rp.reshape(E, -1)[idx].reshape(-1))
    errs = rel_l2_fields(got, ref, idx.size, nd)
    print(f"  sampled rel L2: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= TOL
    s.close()


def test_error_before_initialize():
    s = PrimitiveEquationsSolver(3)
    with pytest.raises(_lib.DrfError) as ei:
        s.compute_rhs(np.zeros(5 * 64))
    assert ei.value.status == _lib.DRF_ERR_STATE
    s.close()


@pytest.mark.parametrize("order", [3, 4, 1])
@pytest.mark.parametrize("prec", [_lib.PREC_FP64, _lib.PREC_MIXED])
def test_launch_shape_does_not_change_results(prec, order):
    """every element is computed the same way whatever the grid: all launch
    shapes (workgroup sizes, workgroups per CU, batch orders), and the
    autotuner's choice, give identical bits; 13 x 9 x 4 elements at orders 1
    and 4 leave partly filled last batches at every workgroup size"""
    import torch

    case = synth.config1() if order == 3 else synth.make_case(13, 9, 4, order, 10000.0,
                                                              "seamount")
    E, n = case.E, case.n
    s = PrimitiveEquationsSolver(case.order, prec)
    s.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    dt = torch.float64 if prec == _lib.PREC_FP64 else torch.float32
    U = torch.from_numpy(synth.to_lines(case.U, E, n)).cuda().to(dt)
    ref = torch.empty_like(U)
    s.set_launch(0, True)
    s.compute_rhs_dev(U, ref, _lib.LAYOUT_LINES)
    out = torch.empty_like(U)
    for lines in (128, 256, 512):
        for wpc in (1, 2, 3, 8):
            for remap in (False, True):
                s.set_launch(wpc, remap, lines)
                out.zero_()
                s.compute_rhs_dev(U, out, _lib.LAYOUT_LINES)
                torch.cuda.synchronize()
                assert torch.equal(out, ref), (lines, wpc, remap)
    wpc, remap, lines, us = s.autotune(U, out, _lib.LAYOUT_LINES, reps=2)
    assert 1 <= wpc <= 8 and lines in (128, 256, 512) and us > 0
    assert s.get_launch() == (wpc, remap, lines)
    s.compute_rhs_dev(U, out, _lib.LAYOUT_LINES)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    s.close()


def test_wrappers_reject_wrong_buffers():
    """Host and device buffers are checked before the library is called:
    a short, mistyped or host-resident state raises ValueError."""
    import torch

    case = synth.config1()
    s = PrimitiveEquationsSolver(case.order)
    s.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    with pytest.raises(ValueError):
        s.compute_rhs(case.U[:-1])
    with pytest.raises(ValueError):
        s.rk_step(_lib.RK_RK4, 0.0, 1e-4, case.U[:-5])
    with pytest.raises(ValueError):
        s.max_wave_speed(case.U[:10])
    U = torch.from_numpy(case.U).cuda()
    R = torch.empty_like(U)
    with pytest.raises(ValueError):
        s.compute_rhs_dev(U[:-1], R)
    with pytest.raises(ValueError):
        s.compute_rhs_dev(U, R[:100])
    with pytest.raises(ValueError):
        s.compute_rhs_dev(U.float(), R)
    with pytest.raises(ValueError):
        s.compute_rhs_dev(U.cpu(), R)
    with pytest.raises(ValueError):
        s.compute_rhs_dev(U.view(5, -1).t(), R)
    with pytest.raises(ValueError):
        s.rk_step_dev(_lib.RK_RK4, 0.0, 1e-4, U[:-1])
    dt = torch.zeros(1, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        s.rk_step_adaptive_dev(_lib.RK_RK4, U, dt)
    s.close()
    m = PrimitiveEquationsSolver(case.order, _lib.PREC_MIXED)
    m.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    with pytest.raises(ValueError):  # the mixed state is float32
        m.compute_rhs_dev(U, R, _lib.LAYOUT_PACKED)
    m.close()
