"""Refinement-indicator MLP (128 -> 64 -> 64 -> 1, GELU) on the GPU through
the C ABI: BF16 GEMM operands on MFMA with FP32 accumulation, mean-squared-
error loss, Adam.  No counterpart in the reference.

Parameter vector (FP32, 12,481): W1[64][128] | b1[64] | W2[64][64] | b2[64]
| W3[64] | b3.  Inputs are BF16 rows of 128 (uint16 bits in an int16 or
bfloat16 tensor, 16-byte aligned)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .solver import _ptr, _sptr

N_IN, N_HIDDEN, N_PARAM = 128, 64, 12481
OFFSETS = {"W1": 0, "b1": 8192, "W2": 8256, "b2": 12352, "W3": 12416, "b3": 12480}
NONE, REFINE, COARSEN = 0, 1, 2


class IndicatorMLP:
    def __init__(self, seed=1, device=0):
        self._l = _lib.load()
        h = C.c_void_p()
        check(self._l.drf_mlp_create(seed, device, C.byref(h)))
        self._m = h
        self.device = device

    def __del__(self):
        if getattr(self, "_m", None):
            self._l.drf_mlp_destroy(self._m)
            self._m = None

    def get_params(self):
        p = np.zeros(N_PARAM, dtype=np.float32)
        check(self._l.drf_mlp_get_params(self._m, p.ctypes.data_as(C.c_void_p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == N_PARAM
        check(self._l.drf_mlp_set_params(self._m, p.ctypes.data_as(C.c_void_p)))

    @staticmethod
    def _x(X):
        if X.element_size() != 2 or X.dim() != 2 or X.shape[1] != N_IN or not X.is_contiguous():
            raise ValueError("X must be a contiguous (n, 128) 2-byte device tensor")
        return X.shape[0]

    def forward(self, X, stream=None):
        import torch

        n = self._x(X)
        pred = torch.empty(n, dtype=torch.float32, device=X.device)
        check(self._l.drf_mlp_forward_dev(self._m, n, _ptr(X), _ptr(pred), _sptr(stream)))
        return pred

    def loss_grad(self, X, y, stream=None):
        import torch

        n = self._x(X)
        grad = torch.empty(N_PARAM, dtype=torch.float32, device=X.device)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        check(self._l.drf_mlp_loss_grad_dev(self._m, n, _ptr(X), _ptr(y), _ptr(grad),
                                            _ptr(loss), _sptr(stream)))
        return loss, grad

    def train_step(self, X, y, lr, loss=None, stream=None):
        n = self._x(X)
        check(self._l.drf_mlp_train_step_dev(self._m, n, _ptr(X), _ptr(y), lr, _ptr(loss),
                                             _sptr(stream)))


def classify(pred, refine_level, coarsen_level, stream=None):
    """RefinementAction per row: 1 if pred > refine_level, 2 if pred < coarsen_level."""
    import torch

    n = pred.numel()
    act = torch.empty(n, dtype=torch.uint8, device=pred.device)
    check(_lib.load().drf_mlp_classify_dev(n, _ptr(pred), refine_level, coarsen_level,
                                           _ptr(act), _sptr(stream)))
    return act


def bf16_bits(a):
    """numpy float array -> uint16 BF16 bits (round to nearest even via FP32)"""
    b = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)
