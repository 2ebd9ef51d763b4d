"""Conv-decoder warm start (kmi_nn): trained and evaluated on the GPU with BF16-MFMA kernels."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

from . import _lib as L
from .lattice import ArgumentError, MI355XBackend


class KmiNNDesc(ctypes.Structure):
    _fields_ = [("in_dim", c_int32), ("hidden", c_int32), ("latent_hw", c_int32), ("latent_c", c_int32),
                ("c1", c_int32), ("c2", c_int32), ("out_c", c_int32), ("max_batch", c_int32),
                ("seed", c_uint64)]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(POINTER(c_float))


class WarmStartNet:
    """x[B, in_dim] → y[B, 8hw, 8hw, out_c] (NHWC)."""

    def __init__(self, *, in_dim=2, hidden=256, latent_hw=16, latent_c=64, c1=64, c2=32, out_c=3,
                 max_batch=32, seed=1234, backend=None):
        self.backend = backend or MI355XBackend()
        d = KmiNNDesc()
        L.load().kmi_nn_desc_init(byref(d))
        d.in_dim, d.hidden, d.latent_hw, d.latent_c = in_dim, hidden, latent_hw, latent_c
        d.c1, d.c2, d.out_c, d.max_batch, d.seed = c1, c2, out_c, max_batch, seed
        self.desc = d
        self.out_hw = 8 * latent_hw
        h = c_void_p()
        L.check(L.load().kmi_nn_create(self.backend.ctx, byref(d), byref(h)), self.backend.ctx)
        self._h = h
        n = c_int64()
        L.check(L.load().kmi_nn_param_count(self._h, byref(n)), self.backend.ctx)
        self.nparam = n.value

    def close(self):
        if getattr(self, "_h", None):
            L.load().kmi_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        L.check(rc, self.backend.ctx)

    def _out_shape(self, B):
        return (B, self.out_hw, self.out_hw, self.desc.out_c)

    def get_params(self):
        p = np.empty(self.nparam, dtype=np.float32)
        self._chk(L.load().kmi_nn_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.nparam:
            raise ArgumentError(f"expected {self.nparam} parameters, got {p.size}")
        self._chk(L.load().kmi_nn_set_params(self._h, _fp(p)))

    def forward(self, x):
        x = _f32(x)
        B = x.shape[0]
        y = np.empty(self._out_shape(B), dtype=np.float32)
        self._chk(L.load().kmi_nn_forward(self._h, B, _fp(x), _fp(y)))
        return y

    def predict(self, x):
        """Forward in chunks of max_batch."""
        x = _f32(x)
        outs = [self.forward(x[i:i + self.desc.max_batch]) for i in range(0, len(x), self.desc.max_batch)]
        return np.concatenate(outs, axis=0)

    def loss_grad(self, x, target, want_grad=True):
        x, t = _f32(x), _f32(target)
        B = x.shape[0]
        loss = c_double()
        g = np.empty(self.nparam, dtype=np.float32) if want_grad else None
        self._chk(L.load().kmi_nn_loss_grad(self._h, B, _fp(x), _fp(t), byref(loss),
                                             _fp(g) if g is not None else None))
        return loss.value, g

    def train_step(self, x, target, lr):
        x, t = _f32(x), _f32(target)
        loss = c_double()
        self._chk(L.load().kmi_nn_train_step(self._h, x.shape[0], _fp(x), _fp(t), c_float(lr), byref(loss)))
        return loss.value

    def set_dataset(self, x, y):
        x, y = _f32(x), _f32(y)
        self._chk(L.load().kmi_nn_set_dataset(self._h, x.shape[0], _fp(x), _fp(y)))

    def train_epochs(self, epochs, batch, lr, seed=0):
        loss = c_double()
        self._chk(L.load().kmi_nn_train_epochs(self._h, int(epochs), int(batch), c_float(lr), c_uint64(seed),
                                                byref(loss)))
        return loss.value


def gemm_bf16(A, B, splits=1, backend=None, noalloc=False):
    """C = A·B (BF16 products on MFMA, FP32 accumulation); A, B are 2-D arrays (any strides).
    noalloc: kmi_gemm_bf16_noalloc, which stages everything in the library's static device
    workspace (KMI_NOALLOC_WORKSPACE_BYTES per GPU) and allocates nothing."""
    backend = backend or MI355XBackend()
    A = np.asarray(A, dtype=np.float32)
    B = np.asarray(B, dtype=np.float32)
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ArgumentError("inner dimensions differ")
    # pass the arrays' own element strides (views keep their layout)
    sa = [s // 4 for s in A.strides]
    sb = [s // 4 for s in B.strides]
    if min(sa + sb) < 0:
        A, B = np.ascontiguousarray(A), np.ascontiguousarray(B)
        sa, sb = [s // 4 for s in A.strides], [s // 4 for s in B.strides]
    C = np.empty((M, N), dtype=np.float32)
    pa = A.ctypes.data_as(POINTER(c_float))
    pb = B.ctypes.data_as(POINTER(c_float))
    fn = L.load().kmi_gemm_bf16_noalloc if noalloc else L.load().kmi_gemm_bf16
    L.check(fn(backend.ctx, M, N, K, pa, sa[0], sa[1], pb, sb[0], sb[1], int(splits), _fp(C)), backend.ctx)
    return C
