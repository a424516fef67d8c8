"""Vertical velocity diagnosis on the GPU through the C ABI, mirroring
VerticalVelocityDiagnosis and ColumnVerticalVelocity
(include/physics/vertical_velocity.hpp, src/physics/vertical_velocity.cpp).
FP64 device tensors; not used by the right-hand side (omega = 0 there)."""
import numpy as np

from . import _lib
from ._lib import check
from .solver import _ptr, _sptr


def _same(*ts):
    for t in ts:
        if not (t.is_contiguous() and t.numel() == ts[0].numel()):
            raise ValueError("pointwise inputs must be contiguous and of equal length")


def _f64(*ts):
    import torch

    for t in ts:
        if t is not None and not (t.is_cuda and t.dtype == torch.float64):
            raise ValueError("inputs must be float64 device tensors")


class VerticalVelocityDiagnosis:
    """Batched over elements: fields are (E, (N+1)^3) device tensors (or any
    element stride, the same for every field of a call)."""

    def __init__(self, order):
        self.order = order
        self.ndof = (order + 1) ** 3
        self._l = _lib.load()

    def _es(self, *ts):
        """The element stride every field of a call shares (the C ABI takes
        one); rows must hold the element's (N+1)^3 values contiguously."""
        es = ts[0].stride(0) if ts[0].dim() == 2 else self.ndof
        for t in ts:
            if t.dim() != 2 or t.shape[1] != self.ndof or t.stride(1) != 1 or t.stride(0) != es \
                    or t.shape[0] != ts[0].shape[0]:
                raise ValueError("fields must be (E, (N+1)^3) views with one common element stride")
        return es

    def _out(self, like, es):
        """Output with the same element stride as the inputs."""
        import torch

        return torch.empty((like.shape[0], es), dtype=torch.float64,
                           device=like.device)[:, :self.ndof]

    def diagnose_omega(self, div_Hu, H, stream=None):
        import torch

        _f64(div_Hu, H)
        es = self._es(div_Hu, H)
        out = self._out(div_Hu, es)
        check(self._l.drf_diagnose_omega_dev(self.order, div_Hu.shape[0], _ptr(div_Hu), _ptr(H),
                                             es, _ptr(out), _sptr(stream)))
        return out

    def diagnose_omega_from_velocity(self, dHu_dx, dHv_dy, H, stream=None):
        import torch

        _f64(dHu_dx, dHv_dy, H)
        es = self._es(dHu_dx, dHv_dy, H)
        out = self._out(dHu_dx, es)
        check(self._l.drf_diagnose_omega_from_velocity_dev(
            self.order, dHu_dx.shape[0], _ptr(dHu_dx), _ptr(dHv_dy), _ptr(H), es, _ptr(out),
            _sptr(stream)))
        return out

    @staticmethod
    def omega_to_physical_w(omega, u, v, H, dz_dx, dz_dy, dz_dt, stream=None):
        import torch

        _f64(omega, u, v, H, dz_dx, dz_dy, dz_dt)
        _same(omega, u, v, H, dz_dx, dz_dy, dz_dt)
        w = torch.empty_like(omega)
        check(_lib.load().drf_omega_to_w_dev(omega.numel(), *[_ptr(t) for t in
                                                             (omega, u, v, H, dz_dx, dz_dy,
                                                              dz_dt, w)], _sptr(stream)))
        return w

    @staticmethod
    def omega_to_w_simple(omega, u, v, eta, h, sigma, deta_dx, deta_dy, deta_dt, stream=None):
        import torch

        _f64(omega, u, v, eta, h, sigma, deta_dx, deta_dy, deta_dt)
        _same(omega, u, v, eta, h, sigma, deta_dx, deta_dy, deta_dt)
        w = torch.empty_like(omega)
        check(_lib.load().drf_omega_to_w_simple_dev(
            omega.numel(), *[_ptr(t) for t in (omega, u, v, eta, h, sigma, deta_dx, deta_dy,
                                               deta_dt, w)], _sptr(stream)))
        return w

    @staticmethod
    def surface_omega(deta_dt, u_surf, v_surf, deta_dx, deta_dy, H):
        return _lib.load().drf_surface_omega(deta_dt, u_surf, v_surf, deta_dx, deta_dy, H)

    @staticmethod
    def bottom_omega(u_bot, v_bot, dh_dx, dh_dy):
        return 0.0  # sigma = -1 is the bottom: no flow through it by construction

    def continuity_residual(self, div_Hu, dHomega_dsigma, stream=None):
        """RMS residual per element (E,)."""
        import torch

        _f64(div_Hu, dHomega_dsigma)
        es = self._es(div_Hu, dHomega_dsigma)
        out = torch.empty(div_Hu.shape[0], dtype=torch.float64, device=div_Hu.device)
        check(self._l.drf_continuity_residual_dev(self.order, div_Hu.shape[0], _ptr(div_Hu),
                                                  _ptr(dHomega_dsigma), es, _ptr(out),
                                                  _sptr(stream)))
        return out


class ColumnVerticalVelocity:
    def __init__(self, order):
        self.order = order
        s = np.zeros(order + 1)
        check(_lib.load().drf_column_sigma_nodes(order, _lib.dptr(s)))
        self.sigma_nodes = s

    def diagnose_column(self, div_Hu, H, omega_bottom=None, stream=None):
        """div_Hu, H (ncol, order+1) contiguous device tensors; omega_bottom
        (ncol,) or None -> omega (ncol, order+1)."""
        import torch

        _f64(div_Hu, H, omega_bottom)
        n = self.order + 1
        for t in (div_Hu, H):
            if not (t.is_contiguous() and t.dim() == 2 and t.shape == div_Hu.shape and
                    t.shape[1] == n):
                raise ValueError("div_Hu and H must be contiguous (ncol, order+1)")
        if omega_bottom is not None and not (omega_bottom.is_contiguous() and
                                             omega_bottom.numel() == div_Hu.shape[0]):
            raise ValueError("omega_bottom must be contiguous (ncol,)")
        out = torch.empty_like(div_Hu)
        check(_lib.load().drf_diagnose_column_dev(self.order, div_Hu.shape[0], _ptr(div_Hu),
                                                  _ptr(H), _ptr(omega_bottom), _ptr(out),
                                                  _sptr(stream)))
        return out
