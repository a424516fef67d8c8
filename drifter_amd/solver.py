"""Python mirror of the reference's operator interfaces for the DG RHS path,
running on the HIP library through the C ABI.

Names, argument meanings and error behaviour follow the reference:
  PrimitiveEquationsSolver      include/physics/primitive_equations.hpp:217-265
  CoriolisParameter             include/physics/primitive_equations.hpp:52-79
  TimeStepper factories         include/solver/time_stepper.hpp:13-33
  AdaptiveTimeController        include/solver/time_stepper.hpp:64-81
  ModeSplittingSolver           include/physics/mode_splitting.hpp:203-266
  SimulationDriver (step only)  src/solver/simulation_driver.cpp:108-168
Packed states are numpy float64 vectors [elem][Hu|Hv|eta|HT|HS][ndof]; the
`*_dev` methods take contiguous torch device tensors of the context's
precision (float64, or float32 for the mixed-precision state), checked for
size, dtype and device before the library is called.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr


def _ptr(x):
    """Device pointer of a torch tensor or an int address."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    return C.c_void_p(int(x))


def _host_state(U, n, name="U"):
    """A C-contiguous float64 copy-free view of a host array of exactly n values
    (n = 0 before initialize: the library then reports the missing mesh)."""
    a = np.ascontiguousarray(U, dtype=np.float64)
    if n and a.size != n:
        raise ValueError(f"{name} holds {a.size} values, the mesh needs {n}")
    return a


def _dev_array(t, n, dtype, name):
    """Device pointer of a contiguous torch tensor on the GPU holding at least
    n values of `dtype` (a torch dtype); anything else raises ValueError, so
    that no kernel reads or writes past a caller's buffer."""
    import torch

    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor on the GPU")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor, got one on {t.device}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < n:
        raise ValueError(f"{name} holds {t.numel()} values, needs {n}")
    return C.c_void_p(t.data_ptr())


def _sptr(s):
    """HIP stream handle: a torch stream, an int handle, or torch's current
    stream when None (so device calls are ordered after torch's copies)."""
    if s is None:
        import torch

        s = torch.cuda.current_stream()
    if hasattr(s, "cuda_stream"):
        return C.c_void_p(s.cuda_stream)
    return C.c_void_p(int(s))


class CoriolisParameter:
    """Beta-plane f = f0 + beta (y - y0) (primitive_equations.hpp:52-79)."""

    def __init__(self, f0=1e-4, beta=2e-11, y0=0.0):
        self.f0, self.beta, self.y0 = f0, beta, y0

    def __call__(self, y):
        return self.f0 + self.beta * (np.asarray(y) - self.y0)


class PrimitiveEquationsSolver:
    def __init__(self, order, precision=_lib.PREC_FP64, device=0):
        self._l = _lib.load()
        h = C.c_void_p()
        check(self._l.drf_create(order, precision, device, C.byref(h)))
        self._ctx = h
        self.order = order
        self.n = order + 1
        self.ndof = self.n ** 3
        self.precision = precision
        self.num_elements = 0

    def close(self):
        if self._ctx:
            self._l.drf_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self):
        return self._ctx

    def state_len(self, layout=None):
        """Values in one state: E * 5 * ndof in the packed, SoA and line
        layouts; whole blocks in the blocked layout (drf_state_len)."""
        if layout is None or layout != _lib.LAYOUT_BLOCKS:
            return self.num_elements * 5 * self.ndof
        return int(self._l.drf_state_len(self._ctx, layout))

    def _state_dtype(self):
        import torch

        return torch.float64 if self.precision == _lib.PREC_FP64 else torch.float32

    def _dev_state(self, t, name="U", layout=None):
        return _dev_array(t, self.state_len(layout), self._state_dtype(), name)

    def _dev_scalar(self, t, name):
        import torch

        return None if t is None else _dev_array(t, 1, torch.float64, name)

    def set_physics(self, **kw):
        p = _lib.DrfPhysics()
        self._l.drf_default_physics(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        check(self._l.drf_set_physics(self._ctx, C.byref(p)))

    def initialize(self, num_elements, bathymetry, dh_dx=None, dh_dy=None,
                   coriolis=None, y_positions=None):
        """PrimitiveEquationsSolver::initialize (primitive_equations.hpp:229-231)."""
        coriolis = coriolis or CoriolisParameter()
        h = np.ascontiguousarray(bathymetry, dtype=np.float64).reshape(-1)
        if h.size != num_elements * self.ndof:
            raise ValueError("bathymetry must hold ndof values per element")
        y = None if y_positions is None else np.ascontiguousarray(y_positions, np.float64).reshape(-1)
        check(self._l.drf_initialize(self._ctx, num_elements, dptr(h), None, None,
                                     coriolis.f0, coriolis.beta, coriolis.y0, dptr(y)))
        self.num_elements = num_elements

    def compute_rhs(self, U, t=0.0):
        """SimulationDriver::compute_rhs on the packed layout (host arrays)."""
        U = _host_state(U, self.state_len())
        out = np.empty_like(U)
        check(self._l.drf_rhs(self._ctx, t, dptr(U), dptr(out)))
        return out

    def rhs_function(self):
        """An RHSFunction: f(t, U, dUdt) filling dUdt in place."""
        def f(t, U, dUdt):
            dUdt[...] = self.compute_rhs(U, t)
        return f

    def set_launch(self, wg_per_cu=0, xcd_remap=True, wg_lines=0):
        """Volume-kernel launch shape: resident workgroups per CU (0 = the
        occupancy limit), XCD-aware batch order, grid lines per workgroup
        (0 = 128, 256, 512; orders above 4 run at 128)."""
        check(self._l.drf_set_launch(self._ctx, int(wg_per_cu), int(bool(xcd_remap)),
                                     int(wg_lines)))

    def get_launch(self):
        """(wg_per_cu, xcd_remap, wg_lines)"""
        w, r, n = C.c_int(), C.c_int(), C.c_int()
        check(self._l.drf_get_launch(self._ctx, C.byref(w), C.byref(r), C.byref(n)))
        return w.value, bool(r.value), n.value

    def autotune(self, U, R, layout=_lib.LAYOUT_LINES, reps=10, stream=None):
        """Time every launch shape on (U, R), keep the fastest; returns
        (wg_per_cu, xcd_remap, wg_lines, microseconds per launch)."""
        us = C.c_float()
        check(self._l.drf_autotune(self._ctx, self._dev_state(U, "U", layout),
                                   self._dev_state(R, "dUdt", layout), layout, reps, _sptr(stream),
                                   C.byref(us)))
        w, r, n = self.get_launch()
        return w, r, n, us.value

    def compute_rhs_dev(self, U, dUdt, layout=_lib.LAYOUT_SOA, stream=None, t=0.0):
        check(self._l.drf_rhs_dev(self._ctx, t, self._dev_state(U, "U", layout),
                                  self._dev_state(dUdt, "dUdt", layout), layout, _sptr(stream)))

    def rk_step(self, scheme, t, dt, U):
        U = np.array(_host_state(U, self.state_len()), copy=True)
        check(self._l.drf_rk_step(self._ctx, scheme, t, dt, dptr(U)))
        return U

    def rk_step_dev(self, scheme, t, dt, U, layout=_lib.LAYOUT_SOA, fused=True, stream=None):
        check(self._l.drf_rk_step_dev(self._ctx, scheme, t, dt, self._dev_state(U, "U", layout),
                                      layout, int(fused), _sptr(stream)))

    def capture_rk_steps(self, scheme, dt, U, layout=_lib.LAYOUT_SOA, fused=True, n_steps=1):
        """StepGraph of n_steps drf_rk_step_dev calls on U (in place)."""
        h = C.c_void_p()
        check(self._l.drf_graph_rk(self._ctx, scheme, dt, self._dev_state(U, "U", layout), layout,
                                   int(fused), n_steps, C.byref(h)))
        return StepGraph(h)

    def rk_step_adaptive_dev(self, scheme, U, dt_dev, time_dev=None, cfl=0.5, dt_min=1e-6,
                             dt_max=1.0, dx_min=None, layout=_lib.LAYOUT_SOA, stream=None):
        """SimulationDriver::step on the device: dt from the wave speed of U
        into dt_dev (device float64, one value), time_dev += dt, one fused
        step.  dx_min defaults to the driver's 2 / (N + 1)."""
        dx = 2.0 / (self.order + 1) if dx_min is None else dx_min
        check(self._l.drf_rk_step_adaptive_dev(self._ctx, scheme, cfl, dx, dt_min, dt_max,
                                               self._dev_state(U, "U", layout), layout,
                                               self._dev_scalar(dt_dev, "dt_dev"),
                                               self._dev_scalar(time_dev, "time_dev"),
                                               _sptr(stream)))

    def capture_adaptive_steps(self, scheme, U, dt_dev, time_dev=None, n_steps=1, cfl=0.5,
                               dt_min=1e-6, dt_max=1.0, dx_min=None, layout=_lib.LAYOUT_SOA):
        """StepGraph of n_steps rk_step_adaptive_dev calls on U (in place)."""
        dx = 2.0 / (self.order + 1) if dx_min is None else dx_min
        h = C.c_void_p()
        check(self._l.drf_graph_rk_adaptive(self._ctx, scheme, cfl, dx, dt_min, dt_max,
                                            self._dev_state(U, "U", layout), layout, n_steps,
                                            self._dev_scalar(dt_dev, "dt_dev"),
                                            self._dev_scalar(time_dev, "time_dev"), C.byref(h)))
        return StepGraph(h)

    def max_wave_speed(self, U):
        U = _host_state(U, self.state_len())
        c = C.c_double()
        check(self._l.drf_max_wave_speed(self._ctx, dptr(U), C.byref(c)))
        return c.value

    def max_wave_speed_dev(self, U, layout=_lib.LAYOUT_SOA, stream=None):
        c = C.c_double()
        check(self._l.drf_max_wave_speed_dev(self._ctx, self._dev_state(U, "U", layout), layout,
                                             C.byref(c), _sptr(stream)))
        return c.value

    def max_wave_speed_noalloc_dev(self, U, c_max, layout=_lib.LAYOUT_SOA, stream=None):
        """The wave-speed maximum of U into the device float64 c_max (one
        value), asynchronously and without allocating device memory (the
        reduction's partials live in the library's per-GPU scratch array;
        calls that overlap on one GPU race)."""
        import torch

        check(self._l.drf_max_wave_speed_noalloc_dev(
            self._ctx, self._dev_state(U, "U", layout), layout,
            _dev_array(c_max, 1, torch.float64, "c_max"), _sptr(stream)))

    def max_wave_speed_derived_dev(self, H, u, v, elem_stride, stream=None):
        import torch

        if elem_stride < self.ndof:
            raise ValueError(f"elem_stride {elem_stride} is below ndof {self.ndof}")
        n = (self.num_elements - 1) * elem_stride + self.ndof
        c = C.c_double()
        check(self._l.drf_max_wave_speed_derived_dev(
            self._ctx, _dev_array(H, n, torch.float64, "H"), _dev_array(u, n, torch.float64, "u"),
            _dev_array(v, n, torch.float64, "v"), elem_stride, C.byref(c), _sptr(stream)))
        return c.value

    def rhs_bytes(self):
        return self._l.drf_rhs_bytes(self._ctx)


class TimeStepper:
    """TimeStepper factories (include/solver/time_stepper.hpp:27-32); step()
    runs on the GPU through the solver that owns the RHS."""

    def __init__(self, scheme, order, stages):
        self.scheme, self._order, self._stages = scheme, order, stages

    def step(self, t, dt, U, solver):
        return solver.rk_step(self.scheme, t, dt, U)

    def order(self):
        return self._order

    def num_stages(self):
        return self._stages

    @staticmethod
    def forward_euler():
        return TimeStepper(_lib.RK_EULER, 1, 1)

    @staticmethod
    def rk2_midpoint():
        return TimeStepper(_lib.RK_RK2, 2, 2)

    @staticmethod
    def rk3_ssp():
        return TimeStepper(_lib.RK_SSPRK3, 3, 3)

    @staticmethod
    def rk4_classic():
        return TimeStepper(_lib.RK_RK4, 4, 4)

    @staticmethod
    def lsrk45():
        return TimeStepper(_lib.RK_LSRK45, 4, 5)

    @staticmethod
    def rk45_dormand_prince():
        return TimeStepper(_lib.RK_DP45, 5, 6)

    @staticmethod
    def rkdg(order):
        # RKDG order selection, src/solver/time_stepper.cpp:209-220
        if order <= 1:
            return TimeStepper.forward_euler()
        if order == 2:
            return TimeStepper.rk2_midpoint()
        if order == 3:
            return TimeStepper.rk3_ssp()
        return TimeStepper.rk4_classic()


class AdaptiveTimeController:
    def __init__(self, cfl_target=0.5, dt_min=1.0e-10, dt_max=1.0e6):
        self.cfl, self.dt_min, self.dt_max = cfl_target, dt_min, dt_max

    def compute_dt(self, max_wave_speed, min_element_size):
        return _lib.load().drf_compute_dt(max_wave_speed, min_element_size, self.cfl,
                                          self.dt_min, self.dt_max)


class StepGraph:
    """A HIP graph of repeated steps (drf_graph_rk / drf_graph_mode_split):
    launch() replays them on a stream."""

    def __init__(self, handle):
        self._l = _lib.load()
        self._g = handle

    def launch(self, stream=None):
        check(self._l.drf_graph_launch(self._g, _sptr(stream)))

    def close(self):
        if getattr(self, "_g", None):
            self._l.drf_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        self.close()


class ModeSplittingSolver:
    """ModeSplittingSolver (include/physics/mode_splitting.hpp:203-266)
    acting on the flat full state [e][Hu|Hv|eta|HT|HS|H|u|v|T|S|omega|rho][d]."""

    def __init__(self, order, subcycles=30, device=0, _solver=None):
        self._s = _solver or PrimitiveEquationsSolver(order, _lib.PREC_FP64, device)
        self.subcycles = subcycles
        self.order = order

    def initialize(self, num_elements, bathymetry, dh_dx=None, dh_dy=None, coriolis=None,
                   y_positions=None):
        self._s.initialize(num_elements, bathymetry, dh_dx, dh_dy, coriolis, y_positions)

    def _full_len(self):
        return self._s.num_elements * 12 * self._s.ndof

    def _dev_full(self, t):
        import torch

        return _dev_array(t, self._full_len(), torch.float64, "full_state")

    def step(self, dt_3d, full_state):
        full = np.array(_host_state(full_state, self._full_len(), "full_state"), copy=True)
        check(self._s._l.drf_mode_split_step(self._s.ctx, dt_3d, self.subcycles, dptr(full)))
        return full

    def capture_steps_dev(self, dt_3d, full_state, n_steps):
        """StepGraph of n_steps step_dev calls on the device full state."""
        h = C.c_void_p()
        check(self._s._l.drf_graph_mode_split(self._s.ctx, dt_3d, self.subcycles,
                                              self._dev_full(full_state), n_steps, C.byref(h)))
        return StepGraph(h)

    def step_dev(self, dt_3d, full_state, stream=None):
        check(self._s._l.drf_mode_split_step_dev(self._s.ctx, dt_3d, self.subcycles,
                                                 self._dev_full(full_state), _sptr(stream)))


class SimulationConfig:
    """SimulationConfig defaults (include/solver/simulation_driver.hpp:33-58)."""

    def __init__(self, **kw):
        self.t_start = 0.0
        self.cfl = 0.5
        self.dt_min = 1e-6
        self.dt_max = 3600.0
        self.use_mode_splitting = True
        self.barotropic_subcycles = 30
        self.polynomial_order = 3
        self.use_ssp_rk = True
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)


class SimulationDriver:
    """SimulationDriver::step (src/solver/simulation_driver.cpp:108-151) with
    the state kept as the flat full state on the host."""

    def __init__(self, config, device=0):
        self.config = config
        self.time = config.t_start
        self.total_steps = 0
        self.dt = 0.0
        order = config.polynomial_order
        self.order = order
        self.nd = (order + 1) ** 3
        self.dx_min = 2.0 / (order + 1)  # :40-42
        self.solver = PrimitiveEquationsSolver(order, _lib.PREC_FP64, device)
        self.ms = ModeSplittingSolver(order, config.barotropic_subcycles, _solver=self.solver)
        self.stepper = TimeStepper.rk3_ssp() if config.use_ssp_rk else TimeStepper.rk4_classic()
        self.ctl = AdaptiveTimeController(config.cfl, config.dt_min, config.dt_max)

    def initialize(self, num_elements, bathymetry, y_positions, full_state,
                   coriolis=CoriolisParameter(1e-4, 0.0, 0.0)):
        # the driver's own Coriolis parameter (simulation_driver.cpp:15)
        self.E = num_elements
        self.h = np.ascontiguousarray(bathymetry, np.float64).reshape(-1)
        self.solver.initialize(num_elements, self.h, coriolis=coriolis, y_positions=y_positions)
        self.state = np.array(full_state, dtype=np.float64, copy=True)

    def compute_max_wave_speed(self):
        F = self.state.reshape(self.E, 12, self.nd)
        lib = _lib.load()
        H = np.ascontiguousarray(F[:, 5])
        u = np.ascontiguousarray(F[:, 6])
        v = np.ascontiguousarray(F[:, 7])
        return _max_wave_speed_host(lib, self.solver, H, u, v, self.nd)

    def step(self):
        c = self.compute_max_wave_speed()
        self.dt = self.ctl.compute_dt(c, self.dx_min)
        if self.config.use_mode_splitting:
            self.state = self.ms.step(self.dt, self.state)
        else:
            F = self.state.reshape(self.E, 12, self.nd)
            U = np.ascontiguousarray(F[:, :5]).reshape(-1)
            U = self.stepper.step(self.time, self.dt, U, self.solver)
            # unpack_state (:215-235): resize zeroes omega, rho; update_derived
            P = U.reshape(self.E, 5, self.nd)
            H = P[:, 2] + self.h.reshape(self.E, self.nd)
            Hinv = np.where(H > 1e-10, 1.0 / np.where(H > 1e-10, H, 1.0), 0.0)
            F = np.zeros((self.E, 12, self.nd))
            F[:, :5] = P
            F[:, 5] = H
            F[:, 6] = P[:, 0] * Hinv
            F[:, 7] = P[:, 1] * Hinv
            F[:, 8] = P[:, 3] * Hinv
            F[:, 9] = P[:, 4] * Hinv
            self.state = np.ascontiguousarray(F.reshape(-1))
        self.time += self.dt
        self.total_steps += 1
        return self.dt


def _max_wave_speed_host(lib, solver, H, u, v, nd):
    """compute_max_wave_speed on host derived arrays via the device reduction."""
    import torch  # device memory plumbing only

    dev = torch.device("cuda", 0)
    tH = torch.from_numpy(H.reshape(-1)).to(dev)
    tu = torch.from_numpy(u.reshape(-1)).to(dev)
    tv = torch.from_numpy(v.reshape(-1)).to(dev)
    c = solver.max_wave_speed_derived_dev(tH, tu, tv, nd)
    return c
