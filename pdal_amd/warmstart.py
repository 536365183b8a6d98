"""ctypes bindings of libpdal_ws.so (include/pdal_warmstart.h): the MLP that
predicts POD coefficients of LP solutions and the POD expansion into a
primal-dual start point."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class WarmStartLib:
    def __init__(self, path: os.PathLike | str):
        path = Path(path)
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built")
        self.path = path
        L = self.lib = C.CDLL(str(path), mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L.pdal_mlp_create.restype = C.c_void_p
        L.pdal_mlp_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_ulonglong]
        L.pdal_mlp_free.restype = None
        L.pdal_mlp_free.argtypes = [C.c_void_p]
        L.pdal_mlp_param_count.restype = C.c_longlong
        L.pdal_mlp_param_count.argtypes = [C.c_void_p]
        for name in ("pdal_mlp_get_params", "pdal_mlp_set_params"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [C.c_void_p, _fp]
        L.pdal_mlp_forward.restype = C.c_int
        L.pdal_mlp_forward.argtypes = [C.c_void_p, _fp, C.c_int, _fp]
        L.pdal_mlp_loss_grad.restype = C.c_int
        L.pdal_mlp_loss_grad.argtypes = [C.c_void_p, _fp, _fp, C.c_int, _fp, _fp]
        L.pdal_mlp_train.restype = C.c_int
        L.pdal_mlp_train.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _fp]
        L.pdal_pod_expand.restype = C.c_int
        L.pdal_pod_expand.argtypes = [_fp, _fp, C.c_int, C.c_int, _fp, C.c_int, _dp]

    def create(self, widths, seed: int = 0) -> "Mlp":
        return Mlp(self, widths, seed)

    def pod_expand(self, basis, mean, coef) -> np.ndarray:
        basis, mean, coef = _f32(basis), _f32(mean), _f32(coef)
        k, dim = basis.shape
        batch = coef.shape[0]
        out = np.empty((batch, dim), dtype=np.float64)
        rc = self.lib.pdal_pod_expand(basis.ctypes.data_as(_fp), mean.ctypes.data_as(_fp), dim, k,
                                      coef.ctypes.data_as(_fp), batch, out.ctypes.data_as(_dp))
        if rc != 0:
            raise RuntimeError(f"pdal_pod_expand failed ({rc})")
        return out


class Mlp:
    def __init__(self, lib: WarmStartLib, widths, seed: int):
        self.lib = lib
        self.widths = [int(w) for w in widths]
        arr = (C.c_int * len(self.widths))(*self.widths)
        self.h = lib.lib.pdal_mlp_create(len(self.widths) - 1, arr, C.c_ulonglong(seed))
        if not self.h:
            raise RuntimeError("pdal_mlp_create failed")
        self.nparam = int(lib.lib.pdal_mlp_param_count(self.h))

    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        if self.lib.lib.pdal_mlp_get_params(self.h, out.ctypes.data_as(_fp)) != 0:
            raise RuntimeError("pdal_mlp_get_params failed")
        return out

    def set_params(self, p) -> None:
        p = _f32(p)
        if p.size != self.nparam or self.lib.lib.pdal_mlp_set_params(self.h, p.ctypes.data_as(_fp)) != 0:
            raise RuntimeError("pdal_mlp_set_params failed")

    def forward(self, X) -> np.ndarray:
        X = _f32(X)
        out = np.empty((X.shape[0], self.widths[-1]), dtype=np.float32)
        if self.lib.lib.pdal_mlp_forward(self.h, X.ctypes.data_as(_fp), X.shape[0], out.ctypes.data_as(_fp)) != 0:
            raise RuntimeError("pdal_mlp_forward failed")
        return out

    def loss_grad(self, X, Y):
        X, Y = _f32(X), _f32(Y)
        loss = C.c_float()
        g = np.empty(self.nparam, dtype=np.float32)
        rc = self.lib.lib.pdal_mlp_loss_grad(self.h, X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp), X.shape[0],
                                             C.byref(loss), g.ctypes.data_as(_fp))
        if rc != 0:
            raise RuntimeError("pdal_mlp_loss_grad failed")
        return float(loss.value), g.astype(np.float64)

    def train(self, X, Y, batch: int = 256, epochs: int = 100, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        X, Y = _f32(X), _f32(Y)
        hist = np.zeros(epochs, dtype=np.float32)
        rc = self.lib.lib.pdal_mlp_train(self.h, X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp), X.shape[0], batch,
                                         epochs, C.c_float(lr), C.c_ulonglong(seed), hist.ctypes.data_as(_fp))
        if rc != 0:
            raise RuntimeError("pdal_mlp_train failed")
        return hist

    def close(self) -> None:
        if self.h:
            self.lib.lib.pdal_mlp_free(self.h)
            self.h = None


def load_ws() -> WarmStartLib:
    """The MI355X warm-start library; raises FileNotFoundError if it is not built."""
    return WarmStartLib(LIB_DIR / "libpdal_ws.so")
