"""Python view of the SGS model training library (lib/libhonee_sgs_train.so): the FP32 6-20-6 MLP trained with Adam
on the GPU, and the conversion of nodal training fields into model-space samples and of trained weights into the
closure's parameter set (the layout pyhonee.sgs.load_dd_model returns)."""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from .driver import LIB_DIR, load_libs

_fp, _dp = ct.POINTER(ct.c_float), ct.POINTER(ct.c_double)
_lib = None


class HoneeMiMlpTrainOptions(ct.Structure):
    _fields_ = [("lr", ct.c_float), ("beta1", ct.c_float), ("beta2", ct.c_float), ("eps", ct.c_float), ("batch", ct.c_int), ("epochs", ct.c_int),
                ("seed", ct.c_ulonglong)]


def lib():
    global _lib
    if _lib is None:
        load_libs()  # libceed first: the training library shares the HIP runtime
        path = os.path.join(os.environ.get("HONEE_MI355X_LIB_DIR", LIB_DIR), "libhonee_sgs_train.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ct.CDLL(path)
        L.HoneeMiMlpLastError.restype = ct.c_char_p
        L.HoneeMiMlpCreate.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_float, ct.c_ulonglong, ct.c_void_p]
        for name in ("HoneeMiMlpSetParams", "HoneeMiMlpGetParams"):
            getattr(L, name).argtypes = [ct.c_void_p, _fp]
        L.HoneeMiMlpNumParams.argtypes = [ct.c_void_p]
        L.HoneeMiMlpLossGrad.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, _fp, _fp]
        L.HoneeMiMlpAdamStep.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, _fp]
        L.HoneeMiMlpTrain.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, _fp]
        L.HoneeMiMlpPredict.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int, ct.c_void_p]
        L.HoneeMiMlpDestroy.argtypes = [ct.c_void_p]
        L.HoneeMiSgsModelSamples.argtypes = [ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                             ct.c_void_p]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _chk(ierr, what):
    if ierr:
        raise RuntimeError(f"{what} failed: {lib().HoneeMiMlpLastError().decode()}")


def options(lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, batch=4096, epochs=1, seed=0) -> HoneeMiMlpTrainOptions:
    return HoneeMiMlpTrainOptions(lr, beta1, beta2, eps, batch, epochs, seed)


class Mlp:
    """The FP32 leaky-ReLU network on the GPU: dims (6, 20, 6), the reference closure's shape, or (6, 20, 20, 6) with a
    second hidden layer."""

    def __init__(self, alpha=0.01, seed=0, dims=(6, 20, 6)):
        self.alpha, self.dims = float(alpha), tuple(dims)
        self.h = ct.c_void_p()
        if len(self.dims) == 3:
            ni, nh, no = self.dims
            nh2 = 0
        elif len(self.dims) == 4:
            ni, nh, nh2, no = self.dims
        else:
            raise ValueError(f"dims {dims}: one or two hidden layers")
        _chk(lib().HoneeMiMlpCreate2(ni, nh, nh2, no, ct.c_float(self.alpha), ct.c_ulonglong(seed), ct.byref(self.h)), "HoneeMiMlpCreate2")
        self.nparams = lib().HoneeMiMlpNumParams(self.h)

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, dtype=np.float32)
        _chk(lib().HoneeMiMlpGetParams(self.h, out.ctypes.data_as(_fp)), "GetParams")
        return out

    @params.setter
    def params(self, theta):
        t = _f32(theta)
        _chk(lib().HoneeMiMlpSetParams(self.h, t.ctypes.data_as(_fp)), "SetParams")

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        loss, g = ct.c_float(0), np.zeros(self.nparams, dtype=np.float32)
        _chk(lib().HoneeMiMlpLossGrad(self.h, X.shape[0], X.ctypes.data, T.ctypes.data, 0, ct.byref(loss), g.ctypes.data_as(_fp)), "LossGrad")
        return loss.value, g

    def adam_step(self, X, T, opts: HoneeMiMlpTrainOptions):
        X, T = _f32(X), _f32(T)
        loss = ct.c_float(0)
        _chk(lib().HoneeMiMlpAdamStep(self.h, X.shape[0], X.ctypes.data, T.ctypes.data, 0, ct.byref(opts), ct.byref(loss)), "AdamStep")
        return loss.value

    def train(self, X, T, opts: HoneeMiMlpTrainOptions) -> np.ndarray:
        X, T = _f32(X), _f32(T)
        hist = np.zeros(opts.epochs, dtype=np.float32)
        _chk(lib().HoneeMiMlpTrain(self.h, X.shape[0], X.ctypes.data, T.ctypes.data, 0, ct.byref(opts), hist.ctypes.data_as(_fp)), "Train")
        return hist

    def predict(self, X) -> np.ndarray:
        X = _f32(X)
        Y = np.zeros((X.shape[0], self.dims[-1]), dtype=np.float32)
        _chk(lib().HoneeMiMlpPredict(self.h, X.shape[0], X.ctypes.data, 0, Y.ctypes.data), "Predict")
        return Y

    def closure_model(self, bounds) -> dict:
        """Parameters for the fused closure (pyhonee.sgs.load_dd_model layout) with the output bounds the training
        targets were normalized with."""
        th = self.params.astype(np.float64)
        layers, o = [], 0
        for n_in, n_out in zip(self.dims[:-1], self.dims[1:]):
            layers.append((th[o:o + n_out * n_in].reshape(n_out, n_in), th[o + n_out * n_in:o + n_out * n_in + n_out]))
            o += n_out * n_in + n_out
        no = self.dims[-1]
        out = {"W1": layers[0][0], "b1": layers[0][1], "W2": layers[-1][0], "b2": layers[-1][1],
               "bounds": np.asarray(bounds, dtype=np.float64).reshape(no, 2), "alpha": self.alpha}
        if len(layers) == 3:  # second hidden layer
            out["Wh"], out["bh"] = layers[1]
        return out

    def close(self):
        if self.h:
            lib().HoneeMiMlpDestroy(ct.byref(self.h))
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def model_samples(train, filtered, grad_velo, aniso, gas):
    """Model inputs X [n][6] (FP32) and eigenframe stress S [n][6] (FP64) of nodal training fields, on the GPU."""
    train, filtered, grad_velo, aniso = (np.ascontiguousarray(a, dtype=np.float64) for a in (train, filtered, grad_velo, aniso))
    n = train.shape[0]
    X, S = np.zeros((n, 6), dtype=np.float32), np.zeros((n, 6))
    _chk(lib().HoneeMiSgsModelSamples(n, 0, train.ctypes.data, filtered.ctypes.data, grad_velo.ctypes.data, aniso.ctypes.data, ct.byref(gas),
                                      X.ctypes.data, S.ctypes.data), "SgsModelSamples")
    return X, S


def normalize_targets(S, bounds=None):
    """Per-output min-max bounds [6][2] of S and the targets mapped to [0, 1] (the inverse of the closure's
    de-normalization (hi - lo) (t - 1) + hi)."""
    S = np.asarray(S, dtype=np.float64)
    if bounds is None:
        bounds = np.stack([S.min(0), S.max(0)], axis=1)
    lo, hi = bounds[:, 0], bounds[:, 1]
    return bounds, (S - hi) / (hi - lo) + 1.0
