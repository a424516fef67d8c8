"""Python mirror of the reference's solution transfer on refinement, running
on the HIP library through the C ABI.

  SolutionProjection.project_to_children    src/amr/refinement.cpp:217-276
  SolutionProjection.project_from_children  src/amr/refinement.cpp:278-342
  transfer_solution                         AdaptiveMeshRefinement::transfer_solution :533-567
  transfer_to_refined_mesh                  the same transfer in the refined mesh's own
                                            leaf order, with each element's own mask
                                            (not the reference's behaviour)

Fields are one nodal value set per element, device float64 tensors of shape
(E, (N+1)^3) (or any element stride); errors raise DrfError.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .solver import _ptr, _sptr

MASK_X, MASK_Y, MASK_Z, MASK_XYZ = 1, 2, 4, 7


def num_children(mask):
    """include/mesh/refine_mask.hpp: 2^(number of refined axes)."""
    return 1 << bin(int(mask) & 7).count("1")


def _stride(t, nd):
    if t.dim() == 2:
        return t.stride(0)
    return nd


class SolutionProjection:
    """Batch form of SolutionProjection(basis): one call projects many
    elements with the same refine mask."""

    def __init__(self, order):
        self.order = order
        self.ndof = (order + 1) ** 3
        self._l = _lib.load()

    def project_to_children(self, U, mask, stream=None):
        """U (P, ndof) -> children (P, nchild, ndof), child c = cx + nx (cy + ny cz)."""
        import torch

        P = U.shape[0]
        nc = num_children(mask)
        out = torch.empty((P, nc, self.ndof), dtype=torch.float64, device=U.device)
        check(self._l.drf_project_to_children_dev(self.order, mask, P, _ptr(U),
                                                  _stride(U, self.ndof), _ptr(out), self.ndof,
                                                  _sptr(stream)))
        return out

    def project_from_children(self, children, mask, stream=None):
        """children (P, nchild, ndof) -> parents (P, ndof)."""
        import torch

        P = children.shape[0]
        assert children.shape[1] == num_children(mask) and children.is_contiguous()
        out = torch.empty((P, self.ndof), dtype=torch.float64, device=children.device)
        check(self._l.drf_project_from_children_dev(self.order, mask, P, _ptr(children),
                                                    self.ndof, _ptr(out), self.ndof,
                                                    _sptr(stream)))
        return out


def transfer_solution(order, U, parents, stream=None):
    """AdaptiveMeshRefinement::transfer_solution into an empty vector: the
    non-refined elements in order, then 8 XYZ children per parent in order.
    parents: strictly ascending element indices (host sequence or device
    int64 tensor)."""
    import torch

    nd = (order + 1) ** 3
    E = U.shape[0]
    if isinstance(parents, torch.Tensor) and parents.is_cuda:
        par = parents.to(torch.int64).contiguous()
    else:
        p = np.asarray(parents, dtype=np.int64).reshape(-1)
        if p.size and (np.any(np.diff(p) <= 0) or p[0] < 0 or p[-1] >= E):
            raise ValueError("parents must be strictly ascending element indices < E")
        par = torch.from_numpy(p).to(U.device)
    R = par.numel()
    out = torch.empty((E + 7 * R, nd), dtype=torch.float64, device=U.device)
    check(_lib.load().drf_transfer_solution_dev(order, E, _ptr(U), _stride(U, nd), R, _ptr(par),
                                                _ptr(out), nd, _sptr(stream)))
    return out


def refine_transfer_map(old_bounds, new_bounds):
    """(src, boxes) of the transfer from the old leaves to the refined mesh's
    leaves (host numpy bounds (E, 6) as OctreeAdapter.elements() returns)."""
    ob = np.ascontiguousarray(old_bounds, dtype=np.float64)
    nb = np.ascontiguousarray(new_bounds, dtype=np.float64)
    src = np.zeros(nb.shape[0], dtype=np.int64)
    boxes = np.zeros((nb.shape[0], 6))
    check(_lib.load().drf_refine_transfer_map(ob.shape[0], _lib.dptr(ob), nb.shape[0],
                                              _lib.dptr(nb), src.ctypes.data, _lib.dptr(boxes)))
    return src, boxes


def interpolate_boxes(order, U, src, boxes, stream=None):
    """Output element k = U[src[k]] evaluated on sub-box boxes[k] (device
    tensors int64 (K,), float64 (K, 6))."""
    import torch

    nd = (order + 1) ** 3
    K = src.numel()
    out = torch.empty((K, nd), dtype=torch.float64, device=U.device)
    check(_lib.load().drf_interpolate_boxes_dev(order, K, U.shape[0], _ptr(src), _ptr(boxes),
                                                _ptr(U), _stride(U, nd), _ptr(out), nd,
                                                _sptr(stream)))
    return out


def interpolate_boxes_fields(order, fields, src, boxes, stream=None):
    """interpolate_boxes for up to 8 fields of one mesh in one launch: fields
    is a list of (E, >= nd) device float64 views (rows may be strided, as the
    columns of a packed state are); returns one (K, nd) output per field,
    each bit-identical to interpolate_boxes of that field."""
    import torch

    nf = len(fields)
    if not 1 <= nf <= 8:
        raise ValueError("1 to 8 fields per call")
    nd = (order + 1) ** 3
    K = src.numel()
    n_elem = fields[0].shape[0]
    for U in fields:
        if U.shape[0] != n_elem or U.dtype != torch.float64 or not U.is_cuda:
            raise ValueError("fields must be float64 device tensors with one row per element")
        if U.dim() == 2 and (U.shape[1] < nd or U.stride(1) != 1):
            raise ValueError("each element's nd values must be contiguous")
    outs = [torch.empty((K, nd), dtype=torch.float64, device=fields[0].device) for _ in fields]
    Up = (C.c_void_p * nf)(*[_ptr(U) for U in fields])
    es = (C.c_int64 * nf)(*[_stride(U, nd) for U in fields])
    Op = (C.c_void_p * nf)(*[_ptr(o) for o in outs])
    os_ = (C.c_int64 * nf)(*([nd] * nf))
    check(_lib.load().drf_interpolate_boxes_fields_dev(order, K, n_elem, _ptr(src), _ptr(boxes),
                                                       nf, Up, es, Op, os_, _sptr(stream)))
    return outs


def transfer_to_refined_mesh(order, U, old_bounds, new_bounds, stream=None):
    """The solution on the refined mesh, in its leaf order, each new leaf
    evaluated from the old leaf that holds it (any masks, balance included)."""
    import torch

    src, boxes = refine_transfer_map(old_bounds, new_bounds)
    return interpolate_boxes(order, U, torch.from_numpy(src).to(U.device),
                             torch.from_numpy(boxes).to(U.device), stream)
