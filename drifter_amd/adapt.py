"""Python mirror of the reference's refinement-side interfaces, running on
the HIP library through the C ABI.

  GradientErrorEstimator       include/amr/refinement.hpp, src/amr/refinement.cpp:39-161
  mark_elements                src/amr/refinement.cpp:488-520 (no area of interest)
  modal coefficients           orthonormal Legendre, no reference counterpart

Device arrays are torch tensors on the GPU; errors raise DrfError.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .solver import _ptr, _sptr

NONE, REFINE, COARSEN = 0, 1, 2  # RefinementAction


def element_char_length(bounds):
    """cbrt(dx dy dz) per element from bounds (E, 6) (host numpy)."""
    b = np.ascontiguousarray(bounds, dtype=np.float64)
    out = np.zeros(b.shape[0])
    check(_lib.load().drf_element_char_length(b.shape[0], _lib.dptr(b), _lib.dptr(out)))
    return out


class GradientErrorEstimator:
    def __init__(self, order, refine_threshold=0.1, coarsen_threshold=0.01,
                 smoothness_factor=0.5, max_level=8, min_level=0, device=0):
        self._l = _lib.load()
        h = C.c_void_p()
        check(self._l.drf_estimator_create(order, refine_threshold, coarsen_threshold,
                                           smoothness_factor, max_level, min_level, device,
                                           C.byref(h)))
        self._e = h
        self.order = order
        self.device = device

    def __del__(self):
        if getattr(self, "_e", None):
            self._l.drf_estimator_destroy(self._e)
            self._e = None

    def estimate(self, field, h_char, levels=None, elem_stride=None, stream=None):
        """field: device float64 with element e at e*elem_stride; h_char (E,)
        device float64; levels (E, 3) device int32 or None.  Returns a dict of
        device tensors value, mask, refine, coarsen, max_value (the
        un-normalised maximum, one double) (+ action, mask_out)."""
        import torch

        nd = (self.order + 1) ** 3
        es = nd if elem_stride is None else elem_stride
        E = h_char.numel()
        dev = h_char.device
        out = {"value": torch.empty(E, dtype=torch.float64, device=dev),
               "max_value": torch.empty(1, dtype=torch.float64, device=dev)}
        for k in ("mask", "refine", "coarsen"):
            out[k] = torch.empty(E, dtype=torch.uint8, device=dev)
        if levels is not None:
            if levels.dtype != torch.int32 or levels.numel() != 3 * E:
                raise ValueError("levels must be an (E, 3) int32 device tensor")
            out["action"] = torch.empty(E, dtype=torch.uint8, device=dev)
            out["mask_out"] = torch.empty(E, dtype=torch.uint8, device=dev)
        if E and field.numel() < (E - 1) * es + nd:
            raise ValueError("field is too small for n_elem and elem_stride")
        check(self._l.drf_estimate_dev(
            self._e, E, _ptr(field), es, _ptr(h_char), _ptr(levels), _ptr(out["value"]),
            _ptr(out["mask"]), _ptr(out["refine"]), _ptr(out["coarsen"]), _ptr(out.get("action")),
            _ptr(out.get("mask_out")), _ptr(out["max_value"]), _sptr(stream)))
        return out


def modal_matrices(order):
    n = order + 1
    V, Vi = np.zeros((n, n)), np.zeros((n, n))
    check(_lib.load().drf_modal_matrices(order, _lib.dptr(V), _lib.dptr(Vi)))
    return V, Vi


def modal_coeffs(order, u, n_elem, elem_stride=None, stream=None):
    import torch

    nd = (order + 1) ** 3
    es = nd if elem_stride is None else elem_stride
    if n_elem and u.numel() < (n_elem - 1) * es + nd:
        raise ValueError("u is too small for n_elem and elem_stride")
    coef = torch.empty((n_elem, nd), dtype=torch.float64, device=u.device)
    check(_lib.load().drf_modal_coeffs_dev(order, n_elem, _ptr(u), es, _ptr(coef), _sptr(stream)))
    return coef


def modal_features_bf16(order, u, n_elem, features, col_offset, scale=1.0, elem_scale=None,
                        elem_stride=None, stream=None):
    """Write scale * elem_scale[e] * (modal coefficients, mean mode zeroed) as
    BF16 into features (E, row) 2-byte device tensor at col_offset."""
    nd = (order + 1) ** 3
    es = nd if elem_stride is None else elem_stride
    if features.element_size() != 2 or features.shape[0] < n_elem:
        raise ValueError("features must be a 2-byte (E, row) device tensor")
    if n_elem and u.numel() < (n_elem - 1) * es + nd:
        raise ValueError("u is too small for n_elem and elem_stride")
    if elem_scale is not None and elem_scale.numel() < n_elem:
        raise ValueError("elem_scale needs one value per element")
    check(_lib.load().drf_modal_features_bf16_dev(order, n_elem, _ptr(u), es, scale,
                                                  _ptr(elem_scale), _ptr(features),
                                                  features.shape[1], col_offset, _sptr(stream)))
    return features
