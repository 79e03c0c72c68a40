"""LM-seeding surrogate network on the GPU (libddm.so `ddm_mlp_*`).

A 1536 -> 256 -> 256 -> 256 -> 68 GELU MLP; BF16 weights/activations on MFMA
with FP32 accumulation, FP32 master weights and Adam, all in hand-written HIP
kernels.  Parameters are one flat FP32 vector laid out W0, b0, W1, b1, W2, b2,
W3, b3 with W_l of shape (in, out), y = x W + b.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import DdmError, load_library

IN, HIDDEN, OUT = 1536, 256, 68
LAYERS = ((IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, OUT))
NPARAMS = sum(i * o + o for i, o in LAYERS)


def unpack(p: np.ndarray):
    """Flat vector -> [(W, b), ...] views."""
    out, o = [], 0
    for i, n in LAYERS:
        W = p[o:o + i * n].reshape(i, n)
        o += i * n
        b = p[o:o + n]
        o += n
        out.append((W, b))
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights for the GELU layers, LeCun-normal for the linear head, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, dtype=np.float32)
    for li, (W, b) in enumerate(unpack(p)):
        gain = 2.0 if li < len(LAYERS) - 1 else 1.0
        W[:] = rng.standard_normal(W.shape) * np.sqrt(gain / W.shape[0])
    return p


def _declare(L):
    if getattr(L, "_mlp_declared", False):
        return L
    vp = C.c_void_p
    L.ddm_mlp_create.argtypes = [C.c_int32, C.POINTER(vp)]
    L.ddm_mlp_destroy.argtypes = [vp]
    L.ddm_mlp_destroy.restype = None
    L.ddm_mlp_last_error.argtypes = [vp]
    L.ddm_mlp_last_error.restype = C.c_char_p
    L.ddm_mlp_set_params.argtypes = [vp, vp]
    L.ddm_mlp_get_params.argtypes = [vp, vp]
    L.ddm_mlp_forward.argtypes = [vp, vp, C.c_int32, vp]
    L.ddm_mlp_loss_grad.argtypes = [vp, vp, vp, vp, C.c_int32, vp, vp]
    L.ddm_mlp_train_step.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_float, vp]
    for n in ("ddm_mlp_create", "ddm_mlp_set_params", "ddm_mlp_get_params", "ddm_mlp_forward",
              "ddm_mlp_loss_grad", "ddm_mlp_train_step"):
        getattr(L, n).restype = C.c_int
    L._mlp_declared = True
    return L


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise DdmError(f"expected shape {shape}, got {a.shape}")
    return a


class Mlp:
    """One network on the device: parameters, Adam state, activation buffers for max_batch rows."""

    def __init__(self, max_batch: int = 1024, params: np.ndarray | None = None):
        self._lib = _declare(load_library())
        h = C.c_void_p()
        if self._lib.ddm_mlp_create(int(max_batch), C.byref(h)) != 0:
            raise DdmError("ddm_mlp_create failed")
        self._h = h
        self.max_batch = int(max_batch)
        self.set_params(init_params() if params is None else params)

    def close(self):
        if self._h:
            self._lib.ddm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise DdmError(self._lib.ddm_mlp_last_error(self._h).decode())

    def set_params(self, p: np.ndarray) -> None:
        """Load parameters (also resets the Adam moments and step count)."""
        p = _f32(p, (NPARAMS,))
        self._check(self._lib.ddm_mlp_set_params(self._h, p.ctypes.data))

    def get_params(self) -> np.ndarray:
        p = np.zeros(NPARAMS, dtype=np.float32)
        self._check(self._lib.ddm_mlp_get_params(self._h, p.ctypes.data))
        return p

    def forward(self, X: np.ndarray) -> np.ndarray:
        """Inference on (M, 1536) features, M any size (chunked by max_batch)."""
        X = _f32(X)
        if X.ndim != 2 or X.shape[1] != IN:
            raise DdmError(f"features must be (M, {IN})")
        M = X.shape[0]
        Y = np.zeros((M, OUT), dtype=np.float32)
        for s in range(0, M, self.max_batch):
            e = min(M, s + self.max_batch)
            xs = np.ascontiguousarray(X[s:e])
            ys = np.zeros((e - s, OUT), dtype=np.float32)
            self._check(self._lib.ddm_mlp_forward(self._h, xs.ctypes.data, e - s, ys.ctypes.data))
            Y[s:e] = ys
        return Y

    @staticmethod
    def _batch(X, Tg, wt):
        X = _f32(X)
        if X.ndim != 2 or X.shape[1] != IN:
            raise DdmError(f"features must be (M, {IN})")
        return X, _f32(Tg, (X.shape[0], OUT)), _f32(wt, (OUT,))

    def loss_grad(self, X, Tg, wt):
        """(loss, flat gradient) of 0.5/M sum wt_j (y_ij - t_ij)^2, no update."""
        X, Tg, wt = self._batch(X, Tg, wt)
        loss = C.c_float()
        g = np.zeros(NPARAMS, dtype=np.float32)
        self._check(self._lib.ddm_mlp_loss_grad(self._h, X.ctypes.data, Tg.ctypes.data, wt.ctypes.data,
                                                X.shape[0], C.byref(loss), g.ctypes.data))
        return float(loss.value), g

    def train_step(self, X, Tg, wt, lr: float) -> float:
        """One Adam step on a batch (M <= max_batch); returns the loss before the update."""
        X, Tg, wt = self._batch(X, Tg, wt)
        loss = C.c_float()
        self._check(self._lib.ddm_mlp_train_step(self._h, X.ctypes.data, Tg.ctypes.data, wt.ctypes.data,
                                                 X.shape[0], float(lr), C.byref(loss)))
        return float(loss.value)
