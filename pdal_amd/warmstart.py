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
        L.pdal_pod_create.restype = C.c_void_p
        L.pdal_pod_create.argtypes = [_fp, _fp, C.c_int, C.c_int, _dp, _dp]
        L.pdal_pod_free.restype = None
        L.pdal_pod_free.argtypes = [C.c_void_p]
        L.pdal_pod_expand_device.restype = C.c_int
        L.pdal_pod_expand_device.argtypes = [C.c_void_p, _fp, C.c_int, C.c_void_p, C.c_longlong]
        L.pdal_pod_expand_device_noalloc.restype = C.c_int
        L.pdal_pod_expand_device_noalloc.argtypes = [C.c_void_p, _fp, C.c_int, C.c_void_p, C.c_longlong]

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


class DevicePod:
    """A POD basis kept on the device (pdal_pod_create); expand() writes
    clip(mean + coef . basis) into a device address."""

    def __init__(self, lib: WarmStartLib, basis, mean, lo=None, hi=None):
        self.lib = lib
        basis, mean = _f32(basis), _f32(mean)
        self.k, self.dim = basis.shape
        lo = None if lo is None else np.ascontiguousarray(lo, dtype=np.float64)
        hi = None if hi is None else np.ascontiguousarray(hi, dtype=np.float64)
        self.h = lib.lib.pdal_pod_create(basis.ctypes.data_as(_fp), mean.ctypes.data_as(_fp), self.dim, self.k,
                                         None if lo is None else lo.ctypes.data_as(_dp),
                                         None if hi is None else hi.ctypes.data_as(_dp))
        if not self.h:
            raise RuntimeError("pdal_pod_create failed")

    def expand(self, coef, out_device: int, out_capacity: int) -> None:
        """out_capacity: doubles available at out_device (the call writes
        coef.shape[0] * dim of them and refuses more)."""
        coef = _f32(coef)
        if coef.ndim != 2 or coef.shape[1] != self.k:
            raise ValueError("coefficient count does not match the basis")
        if coef.shape[0] * self.dim > out_capacity:
            raise ValueError(f"expansion of {coef.shape[0]} x {self.dim} exceeds the {out_capacity}-double buffer")
        if self.lib.lib.pdal_pod_expand_device(self.h, coef.ctypes.data_as(_fp), coef.shape[0],
                                               C.c_void_p(out_device), int(out_capacity)) != 0:
            raise RuntimeError("pdal_pod_expand_device failed")

    def expand_noalloc(self, coef, out_device: int, out_capacity: int) -> int:
        """expand() through pdal_pod_expand_device_noalloc (no device
        allocation; batch x k <= 131072).  Returns the entry point's code."""
        coef = _f32(coef)
        if coef.ndim != 2 or coef.shape[1] != self.k:
            raise ValueError("coefficient count does not match the basis")
        return int(self.lib.lib.pdal_pod_expand_device_noalloc(self.h, coef.ctypes.data_as(_fp), coef.shape[0],
                                                               C.c_void_p(out_device), int(out_capacity)))

    def close(self) -> None:
        if self.h:
            self.lib.lib.pdal_pod_free(self.h)
            self.h = None


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

    def _check(self, X, Y=None) -> None:
        """The library reads rows of widths[0] inputs and widths[-1] targets:
        other shapes would be read past their ends."""
        if X.ndim != 2 or X.shape[1] != self.widths[0]:
            raise ValueError(f"inputs of shape {X.shape}: the network takes rows of {self.widths[0]}")
        if Y is not None and (Y.ndim != 2 or Y.shape != (X.shape[0], self.widths[-1])):
            raise ValueError(f"targets of shape {Y.shape}: ({X.shape[0]}, {self.widths[-1]}) expected")

    def forward(self, X) -> np.ndarray:
        X = _f32(X)
        self._check(X)
        out = np.empty((X.shape[0], self.widths[-1]), dtype=np.float32)
        if self.lib.lib.pdal_mlp_forward(self.h, X.ctypes.data_as(_fp), X.shape[0], out.ctypes.data_as(_fp)) != 0:
            raise RuntimeError("pdal_mlp_forward failed")
        return out

    def loss_grad(self, X, Y):
        X, Y = _f32(X), _f32(Y)
        self._check(X, Y)
        loss = C.c_float()
        g = np.empty(self.nparam, dtype=np.float32)
        rc = self.lib.lib.pdal_mlp_loss_grad(self.h, X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp), X.shape[0],
                                             C.byref(loss), g.ctypes.data_as(_fp))
        if rc != 0:
            raise RuntimeError("pdal_mlp_loss_grad failed")
        return float(loss.value), g.astype(np.float64)

    def train(self, X, Y, batch: int = 256, epochs: int = 100, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        X, Y = _f32(X), _f32(Y)
        self._check(X, Y)
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


class LearnedWarmStart:
    """Warm start for a fixed-A LP family: theta -> POD coefficients (+ log
    primal weight) -> (x_hat, y_hat, pw_hat).

    fit(): POD of the training solutions (mean + the leading kx modes of x and
    ky modes of y, from the eigendecomposition of the snapshot Gram matrix in
    FP64 on the host).  The coefficients are predicted by an affine head
    (least squares in FP64 on [theta, 1]) plus the MLP (widths[0] = len(theta),
    widths[-1] outputs), trained on the device on what the head leaves.  While
    a family's optimal basis does not change, its solution is an affine
    function of theta (x_B = A_B^-1 b(theta), y = A_B^-T c_B(theta) for
    b, c affine in theta), which the head fits exactly; the network carries
    the part that changes of active set make nonlinear.  The network's
    targets are the head's residuals divided by one scale per block (x modes,
    y modes, primal weight): the POD modes are orthonormal, so its loss is the
    solution-space error of each block, and modes that only hold the training
    solves' own KKT noise weigh as little as they are worth; a block the head
    fits to the training solves' own accuracy (residual RMS below 1e-6 of the
    block's spread) gets no network output at all.  `head=False`
    gives the network alone, on per-mode standardized coefficients.  When the
    training solves' final primal weights are given, the last output predicts
    their logarithm: a PDHG start near the solution pays off only together with
    a primal weight near the one the solve ends with.  predict(): head + MLP
    forward + POD expansion on the device; x_hat is clipped to the variable
    bounds given to fit()."""

    def __init__(self, lib: WarmStartLib, widths=(64, 512, 512, 512, 256), kx: int = 128, seed: int = 0,
                 head: bool = True):
        if kx >= widths[-1] - 1:
            raise ValueError("kx must leave room for the dual modes")
        self.lib = lib
        self.widths = list(widths)
        self.kx = kx
        self.seed = seed
        self.head = head
        self.mlp = None
        self._dev = None  # device-resident bases of predict_into

    @staticmethod
    def _pod(S: np.ndarray, k: int):
        mean = S.mean(axis=0)
        C0 = S - mean
        gram = C0 @ C0.T
        w, V = np.linalg.eigh(gram)
        order = np.argsort(w)[::-1][:k]
        w, V = np.maximum(w[order], 1e-300), V[:, order]
        modes = (C0.T @ V) / np.sqrt(w)  # dim x k, orthonormal columns
        energy = float(w.sum() / max(np.trace(gram), 1e-300))
        return mean, modes.T.copy(), energy  # basis k x dim

    @staticmethod
    def _design(thetas: np.ndarray) -> np.ndarray:
        return np.hstack([thetas, np.ones((thetas.shape[0], 1))])

    def fit(self, thetas, X, Y, pw=None, lb=None, ub=None, epochs: int = 300, batch: int = 256,
            lr: float = 1e-3) -> dict:
        thetas = np.asarray(thetas, dtype=np.float64)
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        out = self.widths[-1]
        self.has_pw = pw is not None
        ky_room = out - self.kx - (1 if self.has_pw else 0)
        kx, ky = min(self.kx, X.shape[0] - 1), min(ky_room, Y.shape[0] - 1)
        self.mx, self.Bx, ex = self._pod(X, kx)
        self.my, self.By, ey = self._pod(Y, ky)
        coef = np.zeros((X.shape[0], out))
        coef[:, :kx] = (X - self.mx) @ self.Bx.T
        coef[:, self.kx: self.kx + ky] = (Y - self.my) @ self.By.T  # fewer snapshots than modes: the rest learn zeros
        if self.has_pw:
            coef[:, -1] = np.log(np.asarray(pw, dtype=np.float64))
        self.kx_eff, self.ky_eff = kx, ky
        if self.head:
            D = self._design(thetas)
            self.Wlin = np.linalg.lstsq(D, coef, rcond=None)[0]
            resid = coef - D @ self.Wlin
            self.cmu = np.zeros(out)
            scale = np.ones(out)
            blocks = [slice(0, self.kx), slice(self.kx, out - 1 if self.has_pw else out)]
            if self.has_pw:
                blocks.append(slice(out - 1, out))
            self.exact_blocks = []
            for name, blk in zip(("x", "y", "pw"), blocks):
                scale[blk] = np.sqrt(np.mean(resid[:, blk] ** 2)) + 1e-300
                spread = np.sqrt(np.mean((coef[:, blk] - coef[:, blk].mean(axis=0)) ** 2))
                if scale[blk][0] <= 1e-6 * spread:
                    # the head fits this block to the training solves' own accuracy: what is left
                    # is their KKT noise, which the network would only learn to reproduce
                    scale[blk] = 0.0
                    self.exact_blocks.append(name)
            self.csd = scale
        else:
            self.Wlin = None
            resid = coef
            self.cmu = coef.mean(axis=0)
            self.csd = coef.std(axis=0) + 1e-12
        target = np.divide(resid - self.cmu, self.csd, out=np.zeros_like(resid), where=self.csd > 0)
        self.lb = None if lb is None else np.asarray(lb, dtype=np.float64)
        self.ub = None if ub is None else np.asarray(ub, dtype=np.float64)
        if self.mlp is not None:
            self.mlp.close()
        if self._dev is not None:
            for d in self._dev:
                d.close()
            self._dev = None
        self.mlp = self.lib.create(self.widths, seed=self.seed)
        hist = self.mlp.train(thetas, target, batch=min(batch, thetas.shape[0]), epochs=epochs, lr=lr,
                              seed=self.seed)
        # reconstruction error of the POD itself on the training set (the floor of any prediction)
        recon_x = self.mx + ((X - self.mx) @ self.Bx.T) @ self.Bx
        info = {"pod_energy_x": ex, "pod_energy_y": ey, "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                "pod_rel_err_x": float(np.linalg.norm(recon_x - X) / np.linalg.norm(X)), "epochs": epochs}
        if self.head:
            info["head_rel_resid"] = float(np.linalg.norm(resid) / max(np.linalg.norm(coef), 1e-300))
            info["head_exact_blocks"] = list(self.exact_blocks)
        return info

    def predict_coef(self, thetas, mlp: bool = True) -> np.ndarray:
        """Predicted coefficients (FP64); mlp=False: the affine head alone."""
        thetas = np.asarray(thetas, dtype=np.float64)
        coef = np.zeros((thetas.shape[0], self.widths[-1]))
        if self.head:
            coef += self._design(thetas) @ self.Wlin
        elif not mlp:
            raise ValueError("no affine head to predict with")
        if mlp:
            coef += self.mlp.forward(thetas.astype(np.float32)).astype(np.float64) * self.csd + self.cmu
        return coef

    def predict(self, thetas, mlp: bool = True):
        """(x_hat, y_hat, pw_hat): pw_hat is None when fit() had no primal weights."""
        coef = self.predict_coef(thetas, mlp)
        cx = coef[:, : self.kx_eff].astype(np.float32)
        cy = coef[:, self.kx: self.kx + self.ky_eff].astype(np.float32)
        xh = self.lib.pod_expand(self.Bx.astype(np.float32), self.mx.astype(np.float32), cx)
        yh = self.lib.pod_expand(self.By.astype(np.float32), self.my.astype(np.float32), cy)
        if self.lb is not None:
            xh = np.maximum(xh, self.lb)
        if self.ub is not None:
            xh = np.minimum(xh, self.ub)
        pw = np.exp(coef[:, -1]) if self.has_pw else None
        return xh, yh, pw

    def predict_into(self, batch, thetas, mlp: bool = True):
        """predict() with the expansion written straight into the start
        buffers of `batch` (an LpBatch of len(thetas) members, in order) and
        handed to it on the device, with the predicted primal weights.  The
        bases stay on the device between calls.  Returns pw_hat."""
        thetas = np.asarray(thetas, dtype=np.float64)
        count = len(batch.sizes)
        if thetas.ndim != 2 or thetas.shape[0] != count:
            raise ValueError(f"{thetas.shape[0] if thetas.ndim else 0} parameter rows for a batch of {count}")
        dim_x, dim_y = self.Bx.shape[1], self.By.shape[1]
        if any(mi != dim_y or ni != dim_x for mi, ni in batch.sizes):
            raise ValueError(f"every member must have m = {dim_y}, n = {dim_x} (the POD dimensions)")
        coef = self.predict_coef(thetas, mlp)
        if self._dev is None:
            self._dev = (DevicePod(self.lib, self.Bx, self.mx, self.lb, self.ub), DevicePod(self.lib, self.By, self.my))
        px, py = batch.start_buffers()
        cap_x = sum(ni for _, ni in batch.sizes)
        cap_y = sum(mi for mi, _ in batch.sizes)
        self._dev[0].expand(coef[:, : self.kx_eff], px, cap_x)
        self._dev[1].expand(coef[:, self.kx: self.kx + self.ky_eff], py, cap_y)
        batch.set_starts_from_buffers()
        pw = np.exp(coef[:, -1]) if self.has_pw else None
        if pw is not None:
            batch.set_primal_weights(pw)
        return pw

    def close(self) -> None:
        if self.mlp is not None:
            self.mlp.close()
            self.mlp = None
        if self._dev is not None:
            for d in self._dev:
                d.close()
            self._dev = None


def solve_warm(lp_lib, instances: list, params, x_hat, y_hat, pw_hat=None):
    """Batched solve of `instances` from predicted starts (rows of x_hat /
    y_hat, one per instance) and, when given, predicted primal weights.  The
    FP64 KKT test of the solver judges every warm answer; instances the warm
    pass leaves unsolved are solved again from zero (the cold fallback), and
    their reported iterations charge both passes.  Returns (results, info)."""
    from .lp import LpBatch
    with LpBatch(lp_lib, instances, params) as b:
        b.set_starts(np.asarray(x_hat, dtype=np.float64).ravel(), np.asarray(y_hat, dtype=np.float64).ravel())
        if pw_hat is not None:
            b.set_primal_weights(pw_hat)
        b.solve()
        res = [b.result(k) for k in range(len(instances))]
        warm_s = b.stats()["solve_s"]
    failed = [k for k, r in enumerate(res) if r.status != "OPTIMAL"]
    cold_s = 0.0
    if failed:
        with LpBatch(lp_lib, [instances[k] for k in failed], params) as b:
            b.solve()
            cold_s = b.stats()["solve_s"]
            for j, k in enumerate(failed):
                r = b.result(j)
                r.iterations += res[k].iterations
                res[k] = r
    return res, {"fallback": len(failed), "warm_solve_s": warm_s, "fallback_solve_s": cold_s}
