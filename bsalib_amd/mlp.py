"""ctypes binding of the device MLP (bsahip_mlp_*, include/bsahip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import _check, lib

F32 = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(F32) if a is not None else None


def _bind():
    L = lib()
    if getattr(L, "_mlp_bound", False):
        return L
    L.bsahip_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                    C.POINTER(C.c_void_p)]
    L.bsahip_mlp_nparams.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.bsahip_mlp_set_params.argtypes = [C.c_void_p, F32, F32, F32]
    L.bsahip_mlp_get_params.argtypes = [C.c_void_p, F32, F32]
    L.bsahip_mlp_forward.argtypes = [C.c_void_p, F32, C.c_int32, F32]
    L.bsahip_mlp_loss_grad.argtypes = [C.c_void_p, F32, F32, F32, C.c_int32, F32, F32]
    L.bsahip_mlp_train.argtypes = [C.c_void_p, F32, F32, F32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_float, C.c_float, C.c_uint64, F32]
    L.bsahip_mlp_destroy.argtypes = [C.c_void_p]
    L._mlp_bound = True
    return L


def layer_shapes(nin: int, width: int, depth: int):
    """[(out, in)] per linear layer, as the device parameter vector holds them."""
    return [((1 if l + 1 == depth else width), (nin if l == 0 else width)) for l in range(depth)]


class MLP:
    """nin -> width -> ... -> 1 with GELU between layers, on the device."""

    def __init__(self, nin: int, width: int = 128, depth: int = 4, seed: int = 0):
        L = _bind()
        self.nin, self.width, self.depth = nin, width, depth
        h = C.c_void_p()
        self.h = None
        _check(L.bsahip_mlp_create(nin, width, depth, C.c_uint64(seed), C.byref(h)))
        self.h = h
        n = C.c_int64(0)
        _check(L.bsahip_mlp_nparams(self.h, C.byref(n)))
        self.nparams = n.value

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().bsahip_mlp_get_params(self.h, _p(p), None))
        return p

    def set(self, params=None, mean=None, inv_scale=None):
        a, m, s = (None if v is None else _f32(v) for v in (params, mean, inv_scale))
        _check(lib().bsahip_mlp_set_params(self.h, _p(a), _p(m), _p(s)))

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        out = np.zeros(x.shape[0], dtype=np.float32)
        _check(lib().bsahip_mlp_forward(self.h, _p(x), x.shape[0], _p(out)))
        return out

    def loss_grad(self, x, y, w=None):
        x, y = _f32(x), _f32(y)
        w = np.ones_like(y) if w is None else _f32(w)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().bsahip_mlp_loss_grad(self.h, _p(x), _p(y), _p(w), x.shape[0], _p(loss), _p(g)))
        return float(loss[0]), g

    def train(self, x, y, w=None, epochs: int = 10, batch: int = 512, lr: float = 1e-3,
              weight_decay: float = 0.0, seed: int = 0):
        x, y = _f32(x), _f32(y)
        w = np.ones_like(y) if w is None else _f32(w)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(lib().bsahip_mlp_train(self.h, _p(x), _p(y), _p(w), x.shape[0], epochs, batch,
                                      C.c_float(lr), C.c_float(weight_decay), C.c_uint64(seed),
                                      _p(losses)))
        return losses

    def close(self):
        if self.h:
            lib().bsahip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass
