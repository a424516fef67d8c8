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
