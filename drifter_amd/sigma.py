"""SigmaMetrics::update (include/mesh/sigma_coordinate.hpp:340-354) on the
GPU through the C ABI: an opt-in FP64 component, not used by the RHS."""
from . import _lib
from ._lib import check
from .solver import _ptr, _sptr


def sigma_metrics(sigma, eta, h, deta_dx, deta_dy, dh_dx, dh_dy, deta_dt=None, stream=None):
    """Device float64 tensors of equal length -> (H, Hinv, dsigma_dx,
    dsigma_dy, dsigma_dt) device tensors."""
    import torch

    ins = (sigma, eta, h, deta_dx, deta_dy, dh_dx, dh_dy) + (() if deta_dt is None else (deta_dt,))
    n = sigma.numel()
    for t in ins:
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == n):
            raise ValueError("inputs must be contiguous float64 device tensors of equal length")
    outs = [torch.empty(n, dtype=torch.float64, device=sigma.device) for _ in range(5)]
    check(_lib.load().drf_sigma_metrics_dev(
        n, *[_ptr(t) for t in ins[:7]], _ptr(deta_dt), *[_ptr(o) for o in outs], _sptr(stream)))
    return tuple(outs)


class SigmaCoordinateOperators:
    """SigmaCoordinateOperators (include/dg/operators_3d.hpp,
    src/dg/operators_3d.cpp:513-662) batched over elements: fields are
    (E, (N+1)^3) float64 device tensors sharing one element stride.  Like the
    reference object, update_sigma_metrics stores the metrics that the
    gradient operators then use."""

    def __init__(self, order):
        self.order = order
        self.ndof = (order + 1) ** 3
        self._l = _lib.load()
        self.H = self.H_inv = self.sigma = None
        self.dsigma_dx = self.dsigma_dy = self.dsigma_dt = None

    def _es(self, *ts):
        import torch

        es = ts[0].stride(0)
        for t in ts:
            if t is None:
                continue
            if not (t.is_cuda and t.dtype == torch.float64 and t.dim() == 2 and
                    t.shape == ts[0].shape and t.shape[1] == self.ndof and t.stride(1) == 1 and
                    t.stride(0) == es):
                raise ValueError("fields must be float64 (E, (N+1)^3) device views with one "
                                 "common element stride")
        return es

    def _out(self, like, es):
        import torch

        return torch.empty((like.shape[0], es), dtype=torch.float64,
                           device=like.device)[:, :self.ndof]

    def update_sigma_metrics(self, eta, h, deta_dx, deta_dy, dh_dx, dh_dy, deta_dt=None,
                             stream=None):
        es = self._es(eta, h, deta_dx, deta_dy, dh_dx, dh_dy, deta_dt)
        outs = [self._out(eta, es) for _ in range(6)]
        check(self._l.drf_sigma_ops_metrics_dev(
            self.order, eta.shape[0], _ptr(eta), _ptr(h), _ptr(deta_dx), _ptr(deta_dy),
            _ptr(dh_dx), _ptr(dh_dy), _ptr(deta_dt), es, *[_ptr(o) for o in outs],
            _sptr(stream)))
        (self.H, self.H_inv, self.sigma, self.dsigma_dx, self.dsigma_dy,
         self.dsigma_dt) = outs

    update_sigma_metrics_with_time = update_sigma_metrics

    def _need_metrics(self):
        if self.H_inv is None:
            raise RuntimeError("update_sigma_metrics first")

    def horizontal_gradient(self, U, stream=None):
        self._need_metrics()
        es = self._es(U, self.dsigma_dx, self.dsigma_dy)
        gx, gy = self._out(U, es), self._out(U, es)
        check(self._l.drf_sigma_horizontal_gradient_dev(
            self.order, U.shape[0], _ptr(U), _ptr(self.dsigma_dx), _ptr(self.dsigma_dy), es,
            _ptr(gx), _ptr(gy), _sptr(stream)))
        return gx, gy

    def vertical_gradient(self, U, stream=None):
        self._need_metrics()
        es = self._es(U, self.H_inv)
        out = self._out(U, es)
        check(self._l.drf_sigma_vertical_gradient_dev(self.order, U.shape[0], _ptr(U),
                                                      _ptr(self.H_inv), es, _ptr(out),
                                                      _sptr(stream)))
        return out

    def sigma_divergence(self, Hu, Hv, omega, stream=None):
        es = self._es(Hu, Hv, omega)
        out = self._out(Hu, es)
        check(self._l.drf_sigma_divergence_dev(self.order, Hu.shape[0], _ptr(Hu), _ptr(Hv),
                                               _ptr(omega), es, _ptr(out), _sptr(stream)))
        return out

    def ale_correction(self, U, w_mesh, stream=None):
        self._need_metrics()
        es = self._es(U, w_mesh, self.H_inv)
        out = self._out(U, es)
        check(self._l.drf_sigma_ale_correction_dev(self.order, U.shape[0], _ptr(U), _ptr(w_mesh),
                                                   _ptr(self.H_inv), es, _ptr(out),
                                                   _sptr(stream)))
        return out

    def material_derivative(self, U, dU_dt, u, v, omega, w_mesh, stream=None):
        self._need_metrics()
        es = self._es(U, dU_dt, u, v, omega, w_mesh, self.dsigma_dx, self.dsigma_dy, self.H_inv)
        out = self._out(U, es)
        check(self._l.drf_sigma_material_derivative_dev(
            self.order, U.shape[0], _ptr(U), _ptr(dU_dt), _ptr(u), _ptr(v), _ptr(omega),
            _ptr(w_mesh), _ptr(self.dsigma_dx), _ptr(self.dsigma_dy), _ptr(self.H_inv), es,
            _ptr(out), _sptr(stream)))
        return out
