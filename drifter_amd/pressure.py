"""Hydrostatic pressure and the pressure-gradient force in sigma coordinates
on the HIP library, mirroring PressureGradientElement / PressureGradientSolver
(include/physics/pressure_gradient.hpp:46-367; the reference header does not
compile, so the build states its algorithm: the vertical direction is the
element's zeta index, and d eta / dx is an input where the reference leaves
it at 0).

Fields are contiguous float64 device tensors, element-major (E, (N+1)^3)."""
import torch  # noqa: F401  (device tensors)

from . import _lib
from ._lib import check
from .solver import _dev_array, _sptr

STANDARD, DENSITY_JACOBIAN = 0, 1
BAROTROPIC, BAROCLINIC, FULL = 1, 2, 3


class PressureGradientSolver:
    """PressureGradientSolver (pressure_gradient.hpp:139-180): compute() is
    full_gradient per element (barotropic + baroclinic of `method`)."""

    def __init__(self, order, method=DENSITY_JACOBIAN, rho_0=1025.0, g=9.81):
        self.order, self.method, self.rho0, self.g = order, method, rho_0, g
        self.nd = (order + 1) ** 3
        self._l = _lib.load()

    def _arr(self, t, E, name, optional=False):
        if t is None and optional:
            return None
        return _dev_array(t, E * self.nd, torch.float64, name)

    def hydrostatic_pressure(self, rho, H, stream=None):
        E = rho.numel() // self.nd
        p = torch.empty_like(rho)
        check(self._l.drf_hydrostatic_pressure_dev(
            self.order, E, self.g, self._arr(rho, E, "rho"), self._arr(H, E, "H"), self.nd,
            self._arr(p, E, "p"), _sptr(stream)))
        return p

    def compute(self, eta, rho, H, sigma, deta_dx=None, deta_dy=None, dh_dx=None, dh_dy=None,
                parts=FULL, stream=None):
        E = eta.numel() // self.nd
        px = torch.empty_like(eta)
        py = torch.empty_like(eta)
        a = self._arr
        check(self._l.drf_pressure_gradient_dev(
            self.order, E, self.method, parts, self.g, self.rho0, a(eta, E, "eta"),
            a(rho, E, "rho"), a(H, E, "H"), a(sigma, E, "sigma"),
            a(deta_dx, E, "deta_dx", True), a(deta_dy, E, "deta_dy", True),
            a(dh_dx, E, "dh_dx", True), a(dh_dy, E, "dh_dy", True), self.nd,
            a(px, E, "pg_x"), a(py, E, "pg_y"), _sptr(stream)))
        return px, py
