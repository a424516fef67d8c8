"""Benchmark: DG right-hand side throughput on one MI355X.

Workload (one "step" = one evaluation of the full right-hand side):
uniform 128 x 128 x 16 octree, order N = 4 (262,144 elements, 125 LGL nodes
per element, 32,768,000 DOF per field), 100 km square, tanh shelf
20 -> 200 m, seeded Fourier state; FP64 storage and arithmetic (the parity
path), device-resident structure-of-arrays state, Coriolis f0 = 1e-4,
beta = 0.

value = GDOF/s = E * (N+1)^3 * 5 / t_RHS / 1e9 with inputs already in HBM
(each launch reads 1.44 GB and writes 1.31 GB, far beyond the 256 MB
last-level cache), timed with HIP events on the launch stream.
Extras: the mixed-precision (FP32 volume) kernel on the same mesh with its
error against the FP64 result per RHS and after 1 and 100 RK4 steps at the
driver's CFL time step, fused / un-fused RK4 steps, the 64 x 64 x 8 mesh,
the coastline octree, the face, transfer, sigma, adaptation and indicator
components, and the oracle's FP64 CPU baseline on a bounded sample of
elements (1 thread and the host's threads).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launches(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps  # ms per call


def time_rotating(solver, U, R, nset, steps, warmup, stream, layout=None):
    """ms per compute_rhs_dev launch cycling over nset input/output sets
    (U, R and nset - 1 copies), so that each launch reads its state from HBM
    rather than from the last-level cache."""
    import torch

    from drifter_amd import _lib

    Us = [U] + [U.clone() for _ in range(nset - 1)]
    Rs = [R] + [torch.empty_like(R) for _ in range(nset - 1)]
    it = [0]

    def one():
        k = it[0] % nset
        it[0] += 1
        solver.compute_rhs_dev(Us[k], Rs[k], _lib.LAYOUT_LINES if layout is None else layout,
                               stream)

    ms = time_launches(one, steps, warmup, stream)
    del Us, Rs
    return ms


def rel_l2_fields(a, b, E, nd):
    """Relative L2 difference per field; None where it is not finite (a
    diverged state)."""
    import numpy as np

    A = a.reshape(5, -1)
    B = b.reshape(5, -1)
    out = []
    for f in range(5):
        with np.errstate(all="ignore"):
            r = float(np.linalg.norm(A[f] - B[f]) / np.linalg.norm(B[f]))
        out.append(r if np.isfinite(r) else None)
    return out


def head_layout(args):
    from drifter_amd import _lib

    return _lib.LAYOUT_BLOCKS if args.layout == "blocks" else _lib.LAYOUT_LINES


def state_in(args, case, dev):
    """The case's FP64 state on the device in the headline layout."""
    import torch

    from drifter_amd import _lib, synth

    n = case.order + 1
    if args.layout == "blocks":
        B = _lib.load().drf_layout_block_elements(case.order)
        host = synth.to_blocks(case.U, case.E, n, B)
    else:
        host = synth.to_lines(case.U, case.E, n)
    return torch.from_numpy(host).to(dev)


def per_field(x, args, case):
    """(5, m) field-major view of a host state in the headline layout (the
    blocked layout's padding included: zero in the inputs, never written)."""
    from drifter_amd import _lib

    if args.layout == "blocks":
        blk = 5 * (case.order + 1) ** 3 * _lib.load().drf_layout_block_elements(case.order)
        return x.reshape(-1, 5, blk // 5).transpose(1, 0, 2).reshape(5, -1)
    return x.reshape(5, -1)


def _fmt(x):
    return None if x is None else float("%.3e" % x)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads this process may use, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(case, label, target_s=10.0):
    """The oracle's FP64 right-hand side (dense Kronecker mat-vecs, the
    reference algorithm) on a bounded prefix of the mesh's elements: once on 1
    thread (the reference as built: CMake forces OpenMP off) and once with the
    elements split over the host threads (as the reference's element-loop
    pragma would, were OpenMP on).  The element-local RHS makes every chunk
    an independent call; ctypes releases the interpreter lock during each."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.load()
    nd = case.nd
    f = np.full(case.h.size, 1e-4)

    def run(e0, e1):
        oracle.rhs(case.order, case.U[e0 * 5 * nd:e1 * 5 * nd], case.h[e0 * nd:e1 * nd],
                   f[e0 * nd:e1 * nd])

    probe = 256
    t = time.perf_counter()
    run(0, probe)
    per = (time.perf_counter() - t) / probe
    ns = int(min(case.E, max(probe, target_s / max(per, 1e-9))))
    t = time.perf_counter()
    run(0, ns)
    dt1 = time.perf_counter() - t
    nt = cpu_threads()
    nm = int(min(case.E, max(nt * probe, nt * target_s / max(per, 1e-9))))
    cuts = np.linspace(0, nm, nt + 1).astype(int)
    with ThreadPoolExecutor(nt) as ex:
        t = time.perf_counter()
        list(ex.map(lambda k: run(cuts[k], cuts[k + 1]), range(nt)))
        dtm = time.perf_counter() - t
    model = cpu_model()
    return {
        "value": ns * nd * 5 / dt1 / 1e9,
        "unit": "GDOF/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{ns} of {case.E} elements of the {label} mesh, one RHS ({dt1:.1f} s), "
                  "dense Kronecker mat-vecs as the reference computes them, 1 thread (the "
                  "reference builds with OpenMP off)",
        "cpu_model": model,
        "threaded": {
            "value": nm * nd * 5 / dtm / 1e9,
            "cores": nt,
            "sample": f"{nm} elements over {nt} threads ({dtm:.1f} s)",
        },
    }


def copy_probe(steps, warmup):
    """Plain device copies with the RHS's field sizes (64x64x8 elements,
    125 nodes, FP64): (a) 5 fields copied, 1 read + 1 write per value (the
    only copy-engine kernel in this run: every other buffer is filled, not
    cloned, so the per-dispatch HBM counters of that kernel are of one size,
    163,840,000 B read and written, and give the scale of FETCH_SIZE /
    WRITE_SIZE); (b) 5 fields plus a broadcast 6th field added into 5
    outputs, the volume kernel's 6-in/5-out pattern, on one buffer set and
    (c, as a subtraction, a kernel of its own) cycling over four sets so that
    every launch reads from HBM: the copy ceiling."""
    import torch

    nn = 64 * 64 * 8 * 125
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    srcs = [torch.full((6 * nn,), 1.0 + k, dtype=torch.float64, device=dev) for k in range(4)]
    dsts = [torch.zeros(5 * nn, dtype=torch.float64, device=dev) for _ in range(4)]
    src, dst = srcs[0], dsts[0]
    a = time_launches(lambda: dst.copy_(src[: 5 * nn]), steps, warmup, stream)
    b = time_launches(lambda: torch.add(src[: 5 * nn].view(5, nn), src[5 * nn:].view(1, nn),
                                        out=dst.view(5, nn)), steps, warmup, stream)
    it = [0]

    def sub_rot():
        k = it[0] % 4
        it[0] += 1
        torch.sub(srcs[k][: 5 * nn].view(5, nn), srcs[k][5 * nn:].view(1, nn),
                  out=dsts[k].view(5, nn))

    c = time_launches(sub_rot, steps, warmup, stream)
    ba, bb = 10 * nn * 8, 11 * nn * 8
    print(json.dumps({
        "sub_6_in_5_out_from_hbm": {"us": round(c * 1e3, 2), "bytes": bb,
                                    "gb_s": round(bb / (c * 1e-3) / 1e9, 1),
                                    "sets": 4},
        "copy_5_fields": {"us": round(a * 1e3, 2), "bytes": ba,
                          "gb_s": round(ba / (a * 1e-3) / 1e9, 1),
                          "read_bytes": 5 * nn * 8, "write_bytes": 5 * nn * 8},
        "add_6_in_5_out": {"us": round(b * 1e3, 2), "bytes": bb,
                           "gb_s": round(bb / (b * 1e-3) / 1e9, 1),
                           "read_bytes": 6 * nn * 8, "write_bytes": 5 * nn * 8},
    }), flush=True)


def measure_mixed_and_steps(args, case, s64, U, r64, stream, label):
    """On the headline mesh: the mixed-precision (FP32 volume) RHS with its
    error against the FP64 result, the RK4 step fused / un-fused / mixed, and
    the mixed state's error against the FP64 state after 1, 10 and 100 RK4
    steps of the time step SimulationDriver::step takes from this state (CFL
    0.5, dx_min = 2/(N+1)) and after 100 steps of a twentieth of it, each
    beside the FP64 state started from the FP32 rounding of the initial state.
    The FP64 device path is the reference for these errors: the tests hold it
    within 1e-12 of the oracle per step."""
    import torch

    from drifter_amd import _lib
    from drifter_amd.solver import (AdaptiveTimeController, CoriolisParameter,
                                    PrimitiveEquationsSolver)

    E, nd = case.E, case.nd
    dof = E * nd
    lay = head_layout(args)
    s32 = PrimitiveEquationsSolver(4, _lib.PREC_MIXED)
    s32.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    U32 = U.to(torch.float32)
    R32 = torch.empty_like(U32)
    tune32 = s32.autotune(U32, R32, lay, reps=4, stream=stream)
    ms32 = time_launches(lambda: s32.compute_rhs_dev(U32, R32, lay, stream),
                         args.steps, args.warmup, stream)
    err32 = rel_l2_fields(per_field(R32.double().cpu().numpy(), args, case), r64, E, nd)
    del R32
    log(f"mixed RHS ({label}): {ms32 * 1e3:.1f} us, {dof * 5 / (ms32 * 1e-3) / 1e9:.1f} GDOF/s, "
        f"rel L2 vs FP64 {err32}")

    c = s64.max_wave_speed_dev(U, lay, stream)
    dt = AdaptiveTimeController(0.5, 1e-6, 3600.0).compute_dt(c, 2.0 / 5)
    Uw = U.clone()
    ms_rk_f = time_launches(lambda: s64.rk_step_dev(_lib.RK_RK4, 0.0, dt, Uw, lay,
                                                    True, stream), 10, 2, stream)
    Uw.copy_(U)
    ms_rk_u = time_launches(lambda: s64.rk_step_dev(_lib.RK_RK4, 0.0, dt, Uw, lay,
                                                    False, stream), 10, 2, stream)
    Uw32 = U32.clone()
    ms_rk_m = time_launches(lambda: s32.rk_step_dev(_lib.RK_RK4, 0.0, dt, Uw32, lay,
                                                    True, stream), 10, 2, stream)
    del Uw32
    log(f"RK4 step ({label}) FP64 fused {ms_rk_f:.3f} ms, un-fused {ms_rk_u:.3f} ms, mixed fused "
        f"{ms_rk_m:.3f} ms")
    # accuracy of the mixed state against the FP64 state: after 1, 10 and
    # 100 steps of the CFL dt, and after 100 steps of dt / 20; beside it the
    # FP64 path started from the FP32 rounding of the same initial state --
    # how far the system itself carries that 2.5e-8 perturbation (the element-
    # local reference system amplifies perturbations fast: at the CFL dt both
    # diverge within 100 steps)
    Ur = U32.double()  # the FP32-rounded initial state, in FP64
    steps_err, floor_err = {}, {}
    for label_k, step, marks in (("dt", dt, (1, 10, 100)), ("dt/20", dt / 20.0, (100,))):
        Uw.copy_(U)
        Ub = U32.clone()
        Uf = Ur.clone()
        for k in range(1, max(marks) + 1):
            s64.rk_step_dev(_lib.RK_RK4, 0.0, step, Uw, lay, True, stream)
            s32.rk_step_dev(_lib.RK_RK4, 0.0, step, Ub, lay, True, stream)
            s64.rk_step_dev(_lib.RK_RK4, 0.0, step, Uf, lay, True, stream)
            if k in marks:
                ref = per_field(Uw.cpu().numpy(), args, case)
                key = f"{k}x{label_k}"
                steps_err[key] = [_fmt(x) for x in
                                  rel_l2_fields(per_field(Ub.double().cpu().numpy(), args, case), ref, E, nd)]
                floor_err[key] = [_fmt(x) for x in rel_l2_fields(per_field(Uf.cpu().numpy(), args, case), ref, E, nd)]
        del Ub, Uf
    log(f"mixed vs FP64 rel L2 (dt {dt:.4g}): {steps_err}; FP64 from the FP32-rounded state: "
        f"{floor_err}")
    del Uw, U32, Ur
    s32.close()
    return {
        "autotune_mixed": {"wg_per_cu": tune32[0], "xcd_remap": tune32[1],
                           "wg_lines": tune32[2], "us": round(tune32[3], 2)},
        "mixed_gdof_s": round(dof * 5 / (ms32 * 1e-3) / 1e9, 3),
        "mixed_ms": round(ms32, 6),
        "mixed_vs_fp64_gdof_ratio": None,
        "mixed_hbm_gb_s": round(case.E * nd * 11 * 4 / (ms32 * 1e-3) / 1e9, 1),
        "mixed_rhs_rel_l2_vs_fp64_per_field": [_fmt(e) for e in err32],
        "rk4_dt_from_cfl": dt,
        "mixed_rel_l2_after_1_rk4_step": steps_err["1xdt"],
        "mixed_rel_l2_after_100_rk4_steps_of_dt_over_20": steps_err["100xdt/20"],
        "mixed_rel_l2_vs_fp64_by_steps": steps_err,
        "fp64_from_fp32_rounded_state_rel_l2_by_steps": floor_err,
        "rk4_step_ms_fp64_fused": round(ms_rk_f, 4),
        "rk4_step_ms_fp64_unfused": round(ms_rk_u, 4),
        "rk4_step_ms_mixed_fused": round(ms_rk_m, 4),
        "mixed_fused_vs_plain_fp64_rk4_speedup": round(ms_rk_u / ms_rk_m, 3),
    }


def measure_config2(args, stream):
    """The 64x64x8, N=4 mesh (32,768 elements, one eighth of the headline
    mesh): the FP64 and mixed RHS with the state read from HBM (launches
    cycle over input/output sets, since one 361 MB pair is partly held by the
    256 MB last-level cache), the driver's adaptive step, and mode splitting
    (10 baroclinic steps x 20 barotropic substeps)."""
    import torch

    from drifter_amd import _lib, synth
    from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

    case = synth.config2()
    E, nd = case.E, case.nd
    dof = E * nd
    out = {}
    s64 = PrimitiveEquationsSolver(4, _lib.PREC_FP64)
    s64.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    lay = head_layout(args)
    Uh = state_in(args, case, "cuda")  # the headline's layout, for the RHS timings
    R = torch.empty_like(Uh)
    tune = s64.autotune(Uh, R, lay, reps=8, stream=stream)
    ms_pair = time_launches(lambda: s64.compute_rhs_dev(Uh, R, lay, stream),
                            args.steps, args.warmup, stream)
    ms = time_rotating(s64, Uh, R, 4, args.steps, args.warmup, stream, lay)
    out["c2_fp64_us"] = round(ms * 1e3, 2)
    out["c2_fp64_one_pair_us"] = round(ms_pair * 1e3, 2)
    out["c2_fp64_gdof_s"] = round(dof * 5 / (ms * 1e-3) / 1e9, 3)
    out["c2_fp64_hbm_frac"] = round(s64.rhs_bytes() / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    out["c2_autotune_fp64"] = {"wg_per_cu": tune[0], "xcd_remap": tune[1],
                               "wg_lines": tune[2], "us": round(tune[3], 2)}
    s32 = PrimitiveEquationsSolver(4, _lib.PREC_MIXED)
    s32.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    U32 = Uh.to(torch.float32)
    R32 = torch.empty_like(U32)
    s32.autotune(U32, R32, lay, reps=8, stream=stream)
    ms32 = time_rotating(s32, U32, R32, 8, args.steps, args.warmup, stream, lay)
    out["c2_mixed_us"] = round(ms32 * 1e3, 2)
    out["c2_mixed_gdof_s"] = round(dof * 5 / (ms32 * 1e-3) / 1e9, 3)
    log(f"64x64x8 N=4: FP64 RHS {ms * 1e3:.1f} us from HBM ({ms_pair * 1e3:.1f} us on one pair), "
        f"mixed {ms32 * 1e3:.1f} us")
    s32.close()
    del U32, R32, Uh
    U = torch.from_numpy(synth.to_lines(case.U, E, 5)).to("cuda")  # the steps below: LINES

    # the driver's step with the CFL dt (SimulationDriver::step): host loop
    # (wave speed to the host, compute_dt, fused step) vs device-resident
    # (dt on the device, eager and as one graph of 20 steps)
    lib = _lib.load()
    nstep, dx = 20, 2.0 / 5
    Uw = U.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nstep):
        c = s64.max_wave_speed_dev(Uw, _lib.LAYOUT_LINES, stream)
        dtk = lib.drf_compute_dt(c, dx, 0.5, 1e-6, 3600.0)
        s64.rk_step_dev(_lib.RK_RK4, 0.0, dtk, Uw, _lib.LAYOUT_LINES, True, stream)
    torch.cuda.synchronize()
    ms_host = (time.perf_counter() - t0) * 1e3 / nstep
    Uw.copy_(U)
    dtd = torch.zeros(1, dtype=torch.float64, device=U.device)
    td = torch.zeros(1, dtype=torch.float64, device=U.device)
    ms_dev = time_launches(lambda: s64.rk_step_adaptive_dev(
        _lib.RK_RK4, Uw, dtd, td, cfl=0.5, dt_min=1e-6, dt_max=3600.0, layout=_lib.LAYOUT_LINES,
        stream=stream), nstep, 2, stream)
    g = s64.capture_adaptive_steps(_lib.RK_RK4, Uw, dtd, td, n_steps=nstep, cfl=0.5, dt_min=1e-6,
                                   dt_max=3600.0, layout=_lib.LAYOUT_LINES)
    ms_graph = time_launches(lambda: g.launch(stream), 3, 1, stream) / nstep
    g.close()
    log(f"adaptive RK4 step: host loop {ms_host:.3f} ms, device {ms_dev:.3f} ms, "
        f"graph {ms_graph:.3f} ms")
    out["c2_adaptive_rk4_step_ms_host_loop"] = round(ms_host, 4)
    out["c2_adaptive_rk4_step_ms_device"] = round(ms_dev, 4)
    out["c2_adaptive_rk4_step_ms_graph"] = round(ms_graph, 4)
    del Uw, U, R
    s64.close()
    out.update(measure_mode_split(case, stream))
    return out


def measure_coastline(args, stream):
    """RHS throughput on the coastline-refined octree (39,280 elements, N=3)."""
    import torch

    from drifter_amd import _lib, synth
    from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

    case, _ = synth.config3()
    E, nd = case.E, case.nd
    out = {}
    for name, prec, dt in (("fp64", _lib.PREC_FP64, torch.float64),
                           ("mixed", _lib.PREC_MIXED, torch.float32)):
        s = PrimitiveEquationsSolver(3, prec)
        s.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
        U = state_in(args, case, "cuda").to(dt)  # the headline's layout
        R = torch.empty_like(U)
        s.autotune(U, R, head_layout(args), reps=8, stream=stream)
        # 8 sets (0.9-1.7 GB): state from HBM every launch, as for the value
        ms = time_rotating(s, U, R, 8, args.steps, args.warmup, stream, head_layout(args))
        out[f"coastline_39280_n3_{name}_gdof_s"] = round(E * nd * 5 / (ms * 1e-3) / 1e9, 3)
        out[f"coastline_39280_n3_{name}_us"] = round(ms * 1e3, 2)
    log(f"coastline N=3 RHS: {out}")
    return out


def measure_faces_coastline(args, stream):
    """The batched adaptive DG right-hand side on the coastline octree
    (39,280 elements, N=3): the face term of the 2:1 plan (103,116 conforming
    interfaces in one element-ordered launch, 3,012 Fine2x1 mortars in 3
    per-axis launches) alone, and the volume RHS with it attached (conforming
    faces in the volume kernel's last phase, mortars after it), FP64 and
    mixed, with the state read from HBM (8 rotating sets)."""
    import torch

    from drifter_amd import _lib, faces, synth
    from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

    case, mesh = synth.config3()
    E, nd = case.E, case.nd
    b, lv, _ = mesh.elements()
    conn = mesh.build_face_connections(two_to_one=True)
    out = {}
    for name, prec, dt in (("fp64", _lib.PREC_FP64, torch.float64),
                           ("mixed", _lib.PREC_MIXED, torch.float32)):
        s = PrimitiveEquationsSolver(3, prec)
        s.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
        plan = faces.FacePlan(s, b, lv, conn, kind=faces.LF_ADV, params=(0.3, -0.2, 0.1, 0.6))
        U = torch.from_numpy(synth.to_lines(case.U, E, 4)).to("cuda").to(dt)
        R = torch.empty_like(U)
        ms_face = time_launches(lambda: plan.face_term(U, R, _lib.LAYOUT_LINES, stream), 50, 10,
                                stream)
        plan.attach()  # conforming faces in the volume kernel, launch shape tuned with them
        s.autotune(U, R, _lib.LAYOUT_LINES, reps=8, stream=stream)
        ms = time_rotating(s, U, R, 8, args.steps, args.warmup, stream)
        plan.detach()
        ng, nc, nm = plan.info()
        out[f"coastline_faces_{name}_face_term_us"] = round(ms_face * 1e3, 2)
        out[f"coastline_faces_{name}_rhs_with_faces_us"] = round(ms * 1e3, 2)
        out[f"coastline_faces_{name}_rhs_with_faces_gdof_s"] = round(E * nd * 5 / (ms * 1e-3) / 1e9, 3)
        out["coastline_face_plan_groups_conforming_mortars"] = [ng, nc, nm]
        plan.close()
        s.close()
        del U, R
    log(f"coastline RHS with faces: {out}")
    return out


def measure_faces_uniform(args, case, stream):
    """The face term on the headline mesh (every face conforming: 765,952
    interfaces) alone and the RHS with it attached (in the volume kernel's
    last phase), FP64 and mixed (FP32 volume and surface kernels), with the
    state read from HBM."""
    import torch

    from drifter_amd import _lib, faces, synth
    from drifter_amd.mesh import OctreeAdapter
    from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

    E, nd = case.E, case.nd
    m = OctreeAdapter(0.0, 100000.0, 0.0, 100000.0, -1.0, 0.0)
    m.build_uniform(128, 128, 16)
    b, lv, _ = m.elements()
    conn = m.build_face_connections(two_to_one=True)
    out = {}
    for name, prec, dt in (("fp64", _lib.PREC_FP64, torch.float64),
                           ("mixed", _lib.PREC_MIXED, torch.float32)):
        s = PrimitiveEquationsSolver(4, prec)
        s.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
        plan = faces.FacePlan(s, b, lv, conn, kind=faces.LF_ADV, params=(0.3, -0.2, 0.1, 0.6))
        U = torch.from_numpy(synth.to_lines(case.U, E, 5)).to("cuda").to(dt)
        R = torch.empty_like(U)
        ms_face = time_launches(lambda: plan.face_term(U, R, _lib.LAYOUT_LINES, stream), 20, 5,
                                stream)
        plan.attach()  # conforming faces in the volume kernel, launch shape tuned with them
        s.autotune(U, R, _lib.LAYOUT_LINES, reps=4, stream=stream)
        ms = time_launches(lambda: s.compute_rhs_dev(U, R, _lib.LAYOUT_LINES, stream), 20, 5,
                           stream)
        plan.detach()
        out[f"headline_faces_{name}_face_term_us"] = round(ms_face * 1e3, 1)
        out[f"headline_faces_{name}_rhs_with_faces_us"] = round(ms * 1e3, 1)
        out[f"headline_faces_{name}_rhs_with_faces_gdof_s"] = round(E * nd * 5 / (ms * 1e-3) / 1e9, 3)
        plan.close()
        s.close()
        del U, R
        torch.cuda.empty_cache()
    out["headline_faces_mixed_vs_fp64_rhs_with_faces"] = round(
        out["headline_faces_mixed_rhs_with_faces_gdof_s"] /
        out["headline_faces_fp64_rhs_with_faces_gdof_s"], 3)
    log(f"headline RHS with faces: {out}")
    return out


def measure_mode_split(case, stream):
    """Mode splitting on the 64x64x8, N=4 state: 10 baroclinic steps with 20
    barotropic substeps each, launched step by step and as one HIP graph."""
    import torch

    from drifter_amd.solver import CoriolisParameter, ModeSplittingSolver

    ms = ModeSplittingSolver(4, subcycles=20)
    ms.initialize(case.E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    full = torch.from_numpy(case.full_state()).to("cuda")
    dt = 1e-3
    work = full.clone()
    eager = time_launches(lambda: [ms.step_dev(dt, work, stream) for _ in range(10)], 3, 1, stream)
    work = full.clone()
    g = ms.capture_steps_dev(dt, work, 10)
    graph = time_launches(lambda: g.launch(stream), 3, 1, stream)
    g.close()
    log(f"mode split, 10 steps x 20 substeps: eager {eager:.3f} ms, graph {graph:.3f} ms")
    return {"mode_split_10x20_eager_ms": round(eager, 4), "mode_split_10x20_graph_ms": round(graph, 4)}


def measure_faces(stream):
    """Face components at coastline-mesh scale, N = 3: pointwise fluxes,
    conforming faces, 2:1 mortars (2x2 sub-faces; also at N = 4) and
    face-trace gathers."""
    import numpy as np
    import torch

    from drifter_amd import interface as itf

    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(1)
    out = {}
    npts = 1 << 22
    UL = (torch.rand((3, npts), device=dev, dtype=torch.float64, generator=g) + 0.5)
    UR = (torch.rand((3, npts), device=dev, dtype=torch.float64, generator=g) + 0.5)
    nrm = torch.zeros((3, npts), device=dev, dtype=torch.float64)
    nrm[0] = 0.6
    nrm[1] = 0.8
    ms = time_launches(lambda: itf.flux_points(itf.SW_HLLC, UL, UR, nrm, [9.81], stream=stream),
                       20, 3, stream)
    out["flux_hllc_4M_points_us"] = round(ms * 1e3, 1)
    out["flux_hllc_gb_s"] = round(npts * 8 * 12 / (ms * 1e-3) / 1e9, 1)
    nf, nfaces = 16, 200000
    A = torch.rand((nfaces, nf), device=dev, dtype=torch.float64, generator=g)
    B = torch.rand((nfaces, nf), device=dev, dtype=torch.float64, generator=g)
    fn = torch.zeros((nfaces, 3), device=dev, dtype=torch.float64)
    fn[:, 2] = 1.0
    ci = itf.ConformingInterface(3)
    ms = time_launches(lambda: ci.compute_flux(itf.UPWIND, [0.3, 0.2, 0.1], A, B, fn, stream),
                       20, 3, stream)
    out["conforming_200k_faces_n3_us"] = round(ms * 1e3, 1)
    m = itf.MortarSpace(3, itf.CONN_FINE_2X2)
    nif = 100000
    Uc = torch.rand((nif, nf), device=dev, dtype=torch.float64, generator=g)
    Uf = torch.rand((nif, 4, nf), device=dev, dtype=torch.float64, generator=g)
    mn = fn[:nif].contiguous()
    ms = time_launches(lambda: m.compute_mortar_flux(itf.LF_ADV, [0.3, 0.2, 0.1, 0.5], Uc, Uf, mn,
                                                     stream), 20, 3, stream)
    out["mortar_2x2_100k_interfaces_n3_us"] = round(ms * 1e3, 1)
    m4 = itf.MortarSpace(4, itf.CONN_FINE_2X2)
    Uc4 = torch.rand((nif, 25), device=dev, dtype=torch.float64, generator=g)
    Uf4 = torch.rand((nif, 4, 25), device=dev, dtype=torch.float64, generator=g)
    ms = time_launches(lambda: m4.compute_mortar_flux(itf.LF_ADV, [0.3, 0.2, 0.1, 0.5], Uc4, Uf4,
                                                      mn, stream), 20, 3, stream)
    out["mortar_2x2_100k_interfaces_n4_us"] = round(ms * 1e3, 1)
    E = 39280
    field = torch.rand((E, 64), device=dev, dtype=torch.float64, generator=g)
    el = torch.arange(E, device=dev, dtype=torch.int64).repeat_interleave(6)
    fa = torch.arange(6, device=dev, dtype=torch.int8).repeat(E)
    ms = time_launches(lambda: itf.gather_face_traces(3, field, el, fa, E, stream=stream), 20, 3,
                       stream)
    out["gather_235680_faces_n3_us"] = round(ms * 1e3, 1)
    log(f"faces: {out}")
    return out


def measure_amr(stream):
    """Solution transfer of one adapt cycle at the bench mesh size (32768
    elements, N = 4, one field): transfer_solution with 10% of the elements
    refined, and project_to_children (XYZ) of every element."""
    import torch

    from drifter_amd import amr

    dev = "cuda"
    order, E = 4, 32768
    nd = (order + 1) ** 3
    g = torch.Generator(device=dev).manual_seed(2)
    U = torch.rand((E, nd), device=dev, dtype=torch.float64, generator=g)
    parents = torch.arange(0, E, 10, device=dev, dtype=torch.int64)
    R = parents.numel()
    out = {}
    ms = time_launches(lambda: amr.transfer_solution(order, U, parents, stream=stream), 20, 3,
                       stream)
    keep = E - R
    moved = (2 * keep + R + 8 * R) * nd * 8 + R * 8
    out["transfer_solution_32768_elems_10pct_n4_us"] = round(ms * 1e3, 1)
    out["transfer_solution_gb_s"] = round(moved / (ms * 1e-3) / 1e9, 1)
    sp = amr.SolutionProjection(order)
    ms = time_launches(lambda: sp.project_to_children(U, amr.MASK_XYZ, stream=stream), 20, 3,
                       stream)
    out["project_to_children_32768_parents_n4_us"] = round(ms * 1e3, 1)
    out["project_to_children_gb_s"] = round(9 * E * nd * 8 / (ms * 1e-3) / 1e9, 1)
    log(f"amr transfer: {out}")
    return out


def measure_adapt_cycle(stream):
    """One adaptation cycle of the 64x64x8 N = 4 octree mesh, stage by stage
    (wall clock, each stage synchronised): the gradient estimator and its
    marking on the device, the refine list on the host, octree refine +
    balance and the face connectivity on the host (the reference's, and the
    2:1 lists the face plans are built from), the transfer map on the
    host, and the transfer of the five state fields and the bathymetry to the
    refined mesh on the device."""
    import numpy as np
    import torch

    from drifter_amd import adapt, amr, synth
    from drifter_amd.mesh import OctreeAdapter

    order, nd, L = 4, 125, 100000.0
    m = OctreeAdapter(0.0, L, 0.0, L, -1.0, 0.0)
    m.build_uniform(64, 64, 8)
    old_b, levels, _ = m.elements()
    case = synth.make_case_on(old_b, order, L, "shelf")
    E = old_b.shape[0]
    U = torch.from_numpy(case.U.reshape(E, 5 * nd)).to("cuda")
    h = torch.from_numpy(case.h.reshape(E, nd)).to("cuda")
    hc = torch.from_numpy(adapt.element_char_length(old_b)).to("cuda")
    lv = torch.from_numpy(np.ascontiguousarray(levels, dtype=np.int32)).to("cuda")
    est = adapt.GradientErrorEstimator(order, refine_threshold=0.5)
    eta = U.view(-1)[2 * nd:]  # field 2 of every element, element stride 5 nd
    est.estimate(eta, hc, lv, elem_stride=5 * nd, stream=stream)  # warm-up
    torch.cuda.synchronize()
    t = {}
    t0 = time.perf_counter()
    r = est.estimate(eta, hc, lv, elem_stride=5 * nd, stream=stream)
    torch.cuda.synchronize()
    t["estimate_and_mark_device"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    act = r["action"].cpu().numpy()
    els = np.nonzero(act == 1)[0]
    t["refine_list_to_host"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    m.refine(els, [7] * els.size)  # refines, then balances
    t["octree_refine_balance_host"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    m.build_face_connections()
    t["face_connections_host"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    m.build_face_connections(two_to_one=True)  # what the face plans are built from
    t["face_connections_2to1_host"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    new_b, _, _ = m.elements()
    src, boxes = amr.refine_transfer_map(old_b, new_b)
    src_d, boxes_d = torch.from_numpy(src).to("cuda"), torch.from_numpy(boxes).to("cuda")
    t["transfer_map_host"] = time.perf_counter() - t0
    cols = [U[:, f * nd:(f + 1) * nd] for f in range(5)] + [h]

    def transfer():  # the five state fields and the bathymetry in one launch
        return amr.interpolate_boxes_fields(order, cols, src_d, boxes_d, stream)

    def transfer_by_field():
        return [amr.interpolate_boxes(order, c, src_d, boxes_d, stream) for c in cols]

    for fn, key in ((transfer, "transfer_device"), (transfer_by_field, "transfer_by_field_device")):
        fn()  # the first call's output allocations are not the transfer's cost
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fields = fn()
        torch.cuda.synchronize()
        t[key] = time.perf_counter() - t0
        del fields
    out = {"adapt_cycle_elements_before_after": [int(E), int(new_b.shape[0])],
           "adapt_cycle_refined": int(els.size)}
    out.update({f"adapt_cycle_{k}_ms": round(v * 1e3, 3) for k, v in t.items()})
    log(f"adapt cycle: {out}")
    return out


def measure_vertical(stream):
    """sigma-coordinate metrics and operators and the omega diagnosis on the
    64x64x8 N = 4 node count (4,096,000 nodes), FP64, random fields."""
    import torch

    from drifter_amd import sigma, vertical

    E, nd = 32768, 125
    g = torch.Generator(device="cuda").manual_seed(3)
    r = lambda: torch.rand((E, nd), dtype=torch.float64, device="cuda", generator=g)
    eta, h = 0.1 * r(), 20.0 + 100.0 * r()
    dex, dey, dhx, dhy, det = (1e-3 * r() for _ in range(5))
    sg = -r()
    out = {}
    n = E * nd
    ms = time_launches(lambda: sigma.sigma_metrics(sg.view(-1), eta.view(-1), h.view(-1),
                                                   dex.view(-1), dey.view(-1), dhx.view(-1),
                                                   dhy.view(-1), stream=stream), 20, 3, stream)
    out["sigma_metrics_4096000_points_us"] = round(ms * 1e3, 1)
    ops = sigma.SigmaCoordinateOperators(4)
    ops.update_sigma_metrics(eta, h, dex, dey, dhx, dhy, det, stream=stream)
    U, dU, u, v, om, wm = (r() for _ in range(6))
    ms = time_launches(lambda: ops.horizontal_gradient(U, stream), 20, 3, stream)
    out["sigma_horizontal_gradient_32768_n4_us"] = round(ms * 1e3, 1)
    ms = time_launches(lambda: ops.material_derivative(U, dU, u, v, om, wm, stream), 20, 3,
                       stream)
    out["sigma_material_derivative_32768_n4_us"] = round(ms * 1e3, 1)
    vv = vertical.VerticalVelocityDiagnosis(4)
    ms = time_launches(lambda: vv.diagnose_omega(U, h, stream), 20, 3, stream)
    out["diagnose_omega_32768_n4_us"] = round(ms * 1e3, 1)
    out["diagnose_omega_gb_s"] = round(n * 24 / (ms * 1e-3) / 1e9, 1)
    # pressure gradient (density Jacobian and standard with the metric inputs)
    from drifter_amd import pressure

    rho = 1020.0 + 10.0 * r()
    H = h + eta
    for name, method, nbytes in (("jacobian", pressure.DENSITY_JACOBIAN, 6 * 8),
                                 ("standard", pressure.STANDARD, 10 * 8)):
        pg = pressure.PressureGradientSolver(4, method)
        ms = time_launches(lambda: pg.compute(eta, rho, H, sg, dex, dey, dhx, dhy,
                                              stream=stream), 20, 3, stream)
        out[f"pressure_gradient_{name}_32768_n4_us"] = round(ms * 1e3, 1)
        out[f"pressure_gradient_{name}_gb_s"] = round(n * nbytes / (ms * 1e-3) / 1e9, 1)
    log(f"sigma / vertical: {out}")
    return out


def measure_indicator(stream):
    """Learned indicator: training on 50 seeded coastline states (about 2M
    element/label pairs, 16 epochs), misclassification against the estimator on 5
    unseen seeds, and inference time on 196,400 elements (5 states) beside
    the plain estimator on the same two fields."""
    import torch

    from drifter_amd import _lib, adapt, indicator, mlp
    from drifter_amd.mesh import coastline_mesh

    fields = indicator.CoastlineFields(coastline_mesh(band=7000.0))
    t0 = time.perf_counter()
    Xtr, ytr, _ = indicator.build_dataset(fields, range(1000, 1050))
    Xte, yte, ate = indicator.build_dataset(fields, range(5000, 5005))
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    model = mlp.IndicatorMLP(seed=7)
    t0 = time.perf_counter()
    loss = indicator.train(model, Xtr, ytr, epochs=16, batch=2048, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    err = indicator.misclassification(model, Xte, ate)
    n = Xte.shape[0]
    est = adapt.GradientErrorEstimator(3)
    eta, T = fields.sample(5000)
    eta5, T5 = eta.repeat(5, 1).contiguous(), T.repeat(5, 1).contiguous()
    h5 = fields.h_char.repeat(5).contiguous()
    X5 = torch.zeros((n, 128), dtype=torch.int16, device="cuda")

    def infer():  # features of both fields from the nodal state, forward, decisions
        adapt.modal_features_bf16(3, eta5, n, X5, 0, elem_scale=h5, stream=stream)
        adapt.modal_features_bf16(3, T5, n, X5, 64, elem_scale=h5, stream=stream)
        return mlp.classify(model.forward(X5, stream), indicator.REFINE_LEVEL,
                            indicator.COARSEN_LEVEL, stream)

    # Timed through the bare library calls with the buffers allocated once:
    # both sides then issue a handful of calls per pass and the events
    # measure the device work rather than the interpreter's issue rate.
    from drifter_amd.solver import _ptr, _sptr

    L = _lib.load()
    pred = torch.empty(n, dtype=torch.float32, device="cuda")
    act = torch.empty(n, dtype=torch.uint8, device="cuda")
    est_out = [torch.empty(n, dtype=torch.float64, device="cuda")] + \
        [torch.empty(n, dtype=torch.uint8, device="cuda") for _ in range(3)] + \
        [torch.empty(1, dtype=torch.float64, device="cuda")]
    sp, pX5, ph5 = _sptr(stream), _ptr(X5), _ptr(h5)
    pe, pT, ppred, pact = _ptr(eta5), _ptr(T5), _ptr(pred), _ptr(act)
    pv, pm, pr, pc, pmax = (_ptr(t) for t in est_out)
    nd3 = 64

    def infer_bare():
        st = L.drf_modal_features_bf16_dev(3, n, pe, nd3, 1.0, ph5, pX5, 128, 0, sp)
        st |= L.drf_modal_features_bf16_dev(3, n, pT, nd3, 1.0, ph5, pX5, 128, 64, sp)
        st |= L.drf_mlp_forward_dev(model._m, n, pX5, ppred, sp)
        st |= L.drf_mlp_classify_dev(n, ppred, indicator.REFINE_LEVEL, indicator.COARSEN_LEVEL,
                                     pact, sp)
        assert st == 0, st

    def est_bare():
        for pf in (pe, pT):
            st = L.drf_estimate_dev(est._e, n, pf, nd3, ph5, None, pv, pm, pr, pc, None, None,
                                    pmax, sp)
            assert st == 0, st

    infer()  # through the checking wrappers once
    est.estimate(eta5, h5, stream=stream)
    torch.cuda.synchronize()
    ms_inf = time_launches(infer_bare, 200, 20, stream)
    ms_fwd = time_launches(lambda: L.drf_mlp_forward_dev(model._m, n, pX5, ppred, sp), 200, 20,
                           stream)
    ms_est = time_launches(est_bare, 200, 20, stream)
    counts = torch.bincount(ate.long(), minlength=3).cpu().tolist()
    log(f"indicator: {Xtr.shape[0]} training rows, misclassification {err:.4f} on {n} "
        f"unseen elements {counts}; features + inference {ms_inf * 1e3:.1f} us (forward "
        f"{ms_fwd * 1e3:.1f} us) vs estimator {ms_est * 1e3:.1f} us; train {t_train:.1f} s")
    return {
        "indicator_training_rows": int(Xtr.shape[0]),
        "indicator_train_s": round(t_train, 2),
        "indicator_data_s": round(t_data, 2),
        "indicator_final_train_loss": round(loss, 5),
        "indicator_misclassification_unseen": round(err, 4),
        "indicator_unseen_decisions_none_refine_coarsen": counts,
        "indicator_features_and_inference_us_196400_elements": round(ms_inf * 1e3, 1),
        "indicator_forward_us_196400_elements": round(ms_fwd * 1e3, 1),
        "estimator_two_fields_us_196400_elements": round(ms_est * 1e3, 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 warm-up launches (about 10 ms): the first hundred launches after a
    # pause run about 8 % slower than the steady state
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--launch", default=None, metavar="WPC,REMAP[,LINES]",
                    help="fixed volume-kernel launch shape (workgroups per CU, XCD order 0/1, "
                         "grid lines per workgroup 128/256/512) instead of the autotuner's "
                         "choice")
    ap.add_argument("--layout", choices=("blocks", "lines"), default="blocks",
                    help="state layout in HBM for the headline mesh: blocked xi-planes "
                         "(DRF_LAYOUT_BLOCKS, default) or whole-mesh xi-planes "
                         "(DRF_LAYOUT_LINES)")
    ap.add_argument("--rhs-only", action="store_true",
                    help="time the FP64 RHS only (no mixed-precision or RK extras)")
    ap.add_argument("--profile-mixed", action="store_true",
                    help="launch the mixed-precision RHS only, on the headline mesh (for "
                         "instruction and traffic counters of that kernel)")
    ap.add_argument("--profile-faces", action="store_true",
                    help="launch the face term only, FP64, on the headline mesh with every "
                         "face conforming (for its counters)")
    ap.add_argument("--profile-adapt", action="store_true",
                    help="run the adaptation-cycle leg only (estimator, refinement, "
                         "connectivity, transfer map and the six-field transfer), three times")
    ap.add_argument("--copy-probe", action="store_true",
                    help="time plain copies of known size instead (HBM counter calibration "
                         "and copy ceiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the element-local "
              "RHS shards by contiguous Morton element ranges with no exchange, and no "
              "multi-GPU figure is measured here.", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.steps, args.warmup)
        return

    if args.profile_adapt:
        import torch

        for _ in range(3):
            measure_adapt_cycle(torch.cuda.current_stream())
        return

    import torch

    from drifter_amd import _lib, synth
    from drifter_amd.solver import CoriolisParameter, PrimitiveEquationsSolver

    label = "128x128x16 N=4"
    log(f"building {label} case ...")
    case = synth.config4()
    E, nd = case.E, case.nd
    dof = E * nd
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()

    if args.profile_faces:
        from drifter_amd import faces
        from drifter_amd.mesh import OctreeAdapter

        m = OctreeAdapter(0.0, 100000.0, 0.0, 100000.0, -1.0, 0.0)
        m.build_uniform(128, 128, 16)
        b, lv, _ = m.elements()
        conn = m.build_face_connections(two_to_one=True)
        s64 = PrimitiveEquationsSolver(4, _lib.PREC_FP64)
        s64.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
        plan = faces.FacePlan(s64, b, lv, conn, kind=faces.LF_ADV, params=(0.3, -0.2, 0.1, 0.6))
        U = torch.from_numpy(synth.to_lines(case.U, E, 4 + 1)).to(dev)
        R = torch.zeros_like(U)
        msf = time_launches(lambda: plan.face_term(U, R, _lib.LAYOUT_LINES, stream),
                            args.steps, args.warmup, stream)
        ng, nc, nm = plan.info()
        print(json.dumps({"kernel": "face_elem_kernel<5, double>", "us": round(msf * 1e3, 2),
                          "conforming_interfaces": nc, "elements": E}), flush=True)
        plan.close()
        s64.close()
        return

    if args.profile_mixed:
        s32 = PrimitiveEquationsSolver(4, _lib.PREC_MIXED)
        s32.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
        U32 = state_in(args, case, dev).to(torch.float32)
        R32 = torch.zeros_like(U32)
        if args.launch:
            s32.set_launch(*(int(v) for v in args.launch.split(",")))
        else:
            s32.autotune(U32, R32, head_layout(args), reps=4, stream=stream)
        ms32 = time_launches(lambda: s32.compute_rhs_dev(U32, R32, head_layout(args), stream),
                             args.steps, args.warmup, stream)
        print(json.dumps({"kernel": f"volume_kernel<4, float, float, EPI_RHS, "
                                    f"{s32.get_launch()[2]}>",
                          "us": round(ms32 * 1e3, 2),
                          "gdof_s": round(dof * 5 / (ms32 * 1e-3) / 1e9, 3),
                          "launch": list(s32.get_launch())}), flush=True)
        s32.close()
        return

    s64 = PrimitiveEquationsSolver(4, _lib.PREC_FP64)
    s64.initialize(E, case.h, coriolis=CoriolisParameter(1e-4, 0.0, 0.0))
    lay = head_layout(args)
    U = state_in(args, case, dev)
    R = torch.zeros_like(U)
    # launch shape from the library's autotuner (workgroups per CU, XCD order,
    # grid lines per workgroup)
    if args.launch:
        s64.set_launch(*(int(v) for v in args.launch.split(",")))
        tune = s64.get_launch() + (float("nan"),)
        log(f"launch shape fixed: {tune[0]} workgroups/CU, xcd_remap={tune[1]}, "
            f"{tune[2]} lines per workgroup")
    else:
        tune = s64.autotune(U, R, lay, reps=4, stream=stream)
        log(f"autotune: {tune[0]} workgroups/CU, xcd_remap={tune[1]}, {tune[2]} lines per "
            f"workgroup, {tune[3]:.1f} us")
    # One input/output pair: each launch reads 1.44 GB and writes 1.31 GB, so
    # the 256 MB last-level cache holds under a tenth of it between launches
    # and the state is read from HBM (at 64x64x8 the timed launches had to
    # cycle over several sets for that).  HIP events on the launch stream.
    ms = time_launches(lambda: s64.compute_rhs_dev(U, R, lay, stream),
                       args.steps, args.warmup, stream)
    gdof = dof * 5 / (ms * 1e-3) / 1e9
    bytes_per = s64.rhs_bytes()
    achieved = bytes_per / (ms * 1e-3) / 1e9
    log(f"FP64 RHS ({label}): {ms * 1e3:.1f} us, {gdof:.1f} GDOF/s, {achieved:.0f} GB/s")

    extras = {"autotune_fp64": {"wg_per_cu": tune[0], "xcd_remap": tune[1],
                                "wg_lines": tune[2],
                                "us": round(tune[3], 2) if tune[3] == tune[3] else None}}
    if not args.rhs_only:
        r64 = per_field(R.cpu().numpy(), args, case)
        extras.update(measure_mixed_and_steps(args, case, s64, U, r64, stream, label))
        extras["mixed_vs_fp64_gdof_ratio"] = round(extras["mixed_gdof_s"] / gdof, 3)
        del r64
    del U, R
    s64.close()
    torch.cuda.empty_cache()
    if not args.rhs_only:
        extras.update(measure_faces_uniform(args, case, stream))
        extras.update(measure_config2(args, stream))
        extras.update(measure_coastline(args, stream))
        extras.update(measure_faces_coastline(args, stream))
        extras.update(measure_faces(stream))
        extras.update(measure_amr(stream))
        extras.update(measure_vertical(stream))
        extras.update(measure_adapt_cycle(stream))
        extras.update(measure_indicator(stream))

    cpu = None if args.no_cpu else cpu_baseline(case, label)
    if cpu:
        log(f"CPU baseline: {cpu['value']:.4f} GDOF/s on 1 thread, "
            f"{cpu['threaded']['value']:.4f} GDOF/s on {cpu['threaded']['cores']} threads "
            f"({cpu['cpu_model']})")

    # HBM bytes and FP64 FLOPs per launch of this kernel, from rocprofv3
    # counter runs of the same command (profiles/volume_rhs_fp64_128x128x16_n4.md,
    # and volume_rhs_fp64_lines_128x128x16_n4.md for the line layout)
    traffic, tsrc, flops = None, None, None
    tfile = os.path.join(ROOT, "profiles", "volume_rhs_fp64_%s128x128x16_n4.traffic.json"
                         % ("" if args.layout == "blocks" else "lines_"))
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        traffic = tj.get("bytes_per_launch")
        flops = tj.get("flops_per_launch")
        tsrc = ("rocprofv3 --pmc FETCH_SIZE (x2.000) / WRITE_SIZE (x1.000) and "
                "SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 of this kernel at launch shape "
                f"{tj.get('launch')}, profiles/" + os.path.basename(tfile))

    line = {
        "metric": "DG RHS GDOF/s on 1 MI355X (FP64 vs mixed @ stated L2 err); % HBM roofline",
        "value": round(gdof, 3),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "DG RHS, uniform 128x128x16 octree, N=4, tanh shelf, FP64, "
                        + ("xi-planes in blocks of 20 elements (DRF_LAYOUT_BLOCKS)"
                           if args.layout == "blocks" else
                           "whole-mesh xi-planes (DRF_LAYOUT_LINES)") + " in HBM",
            "layout": args.layout,
            "elements": E,
            "order": 4,
            "dof_per_field": dof,
            "fields": 5,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": tsrc,
            "fp64_flops_per_launch": flops,
            "fp64_tflops": round(flops / (ms * 1e-3) / 1e12, 2) if flops else None,
            "fp64_tflops_peak": FP64_PEAK_TFLOPS,
            "fp64_flop_frac": round(flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)
                              if flops else None,
            "algorithmic_bytes_per_launch": bytes_per,
            "kernel": f"volume_kernel<4, double, double, EPI_RHS, {tune[2]}>",
            "launch": {"wg_per_cu": tune[0], "xcd_remap": tune[1], "wg_lines": tune[2]},
        },
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
