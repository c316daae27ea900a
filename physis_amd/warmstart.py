"""ctypes binding of libphysis_warmstart_hip (include/physis/physis_warmstart.h):
the U-Net warm start of the Poisson solve, and helpers that train it on
solutions of the Jacobi solve and use it as the initial guess.

No CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import runtime as rt

WARMSTART_SO = os.path.join(rt.LIB_DIR, "libphysis_warmstart_hip.so")
_lib = None
U16P = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        rt.runtime()
        if not os.path.exists(WARMSTART_SO):
            raise RuntimeError(f"{WARMSTART_SO} is missing: run "
                               "`python -m physis_amd.build`")
        L = C.CDLL(WARMSTART_SO)
        G = rt.PSGridPtr
        L.PSUNetNew.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                C.c_uint64]
        L.PSUNetNew.restype = C.c_void_p
        L.PSUNetFree.argtypes = [C.c_void_p]
        L.PSUNetNumParams.argtypes = [C.c_void_p]
        L.PSUNetNumParams.restype = C.c_size_t
        L.PSUNetGetParams.argtypes = [C.c_void_p, F32P]
        L.PSUNetSetParams.argtypes = [C.c_void_p, F32P]
        L.PSUNetPredict.argtypes = [C.c_void_p, G, G, C.c_int, C.c_int, G]
        L.PSUNetLossGrad.argtypes = [C.c_void_p, G, G, G, C.c_int, C.c_int,
                                     C.c_void_p]
        L.PSUNetLossGrad.restype = C.c_double
        L.PSUNetTrainStep.argtypes = [C.c_void_p, G, G, G, C.c_int, C.c_int,
                                      C.c_float]
        L.PSUNetTrainStep.restype = C.c_double
        L.PSUNetSetLossWeights.argtypes = [C.c_void_p, C.c_float, C.c_float]
        L.PSUNetReflect.argtypes = [G, C.c_int, C.c_int, G, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_int]
        L.PSConv3dForwardHost.argtypes = [U16P, U16P, F32P, C.c_int] + \
            [C.c_int] * 6 + [U16P]
        L.PSConv3dBackwardDataHost.argtypes = [U16P, U16P] + [C.c_int] * 6 + \
            [U16P]
        L.PSConv3dBackwardFilterHost.argtypes = [U16P, U16P] + \
            [C.c_int] * 6 + [F32P, F32P]
        _lib = L
    return _lib


class UNet:
    """The warm-start network for problems of nx x ny x nzs."""

    def __init__(self, nx, ny, nzs, max_batch, out_scale, seed=0):
        L = library()
        self.handle = L.PSUNetNew(nx, ny, nzs, max_batch, out_scale, seed)
        if not self.handle:
            raise RuntimeError("PSUNetNew failed")
        self.L = L
        self.max_batch = max_batch
        self.out_scale = out_scale
        self.shape = (nx, ny, nzs)

    def free(self):
        if self.handle:
            self.L.PSUNetFree(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(self.L.PSUNetNumParams(self.handle))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        self.L.PSUNetGetParams(self.handle, out)
        return out

    def set_params(self, p: np.ndarray) -> None:
        self.L.PSUNetSetParams(self.handle,
                               np.ascontiguousarray(p, np.float32))

    def predict(self, rhs, bnd, first, count, p) -> None:
        self.L.PSUNetPredict(self.handle, rhs.handle, bnd.handle, first,
                             count, p.handle)

    def set_loss_weights(self, w4: float, w16: float) -> None:
        """Weights of the smooth-error terms (the error's means over 4^3 and
        16^3 cells) in the training loss; 0, 0 is the mean square error."""
        self.L.PSUNetSetLossWeights(self.handle, w4, w16)

    def loss_grad(self, rhs, bnd, target, first, count, want_grad=True):
        g = np.empty(self.num_params, np.float32) if want_grad else None
        loss = self.L.PSUNetLossGrad(self.handle, rhs.handle, bnd.handle,
                                     target.handle, first, count,
                                     g.ctypes.data if g is not None else None)
        return loss, g

    def train_step(self, rhs, bnd, target, first, count, lr) -> float:
        return self.L.PSUNetTrainStep(self.handle, rhs.handle, bnd.handle,
                                      target.handle, first, count, lr)

    def train(self, rhs, bnd, target, nproblems, steps, batch, lr=1e-3,
              seed=0, lr_final=None, log=None):
        """Adam over random contiguous windows of `batch` training problems;
        lr decays linearly to lr_final.  Returns the loss history."""
        rng = np.random.default_rng(seed)
        lr_final = lr * 0.1 if lr_final is None else lr_final
        hist = []
        t0 = time.perf_counter()
        for s in range(steps):
            first = int(rng.integers(0, nproblems - batch + 1))
            a = lr + (lr_final - lr) * s / max(steps - 1, 1)
            hist.append(self.train_step(rhs, bnd, target, first, batch, a))
            if log and (s % log == 0 or s == steps - 1):
                print(f"  train step {s}: loss {hist[-1]:.4e} "
                      f"({time.perf_counter() - t0:.1f} s)", flush=True)
        return hist


def reflect(src, first, count, dst, dst_first, shape, flips) -> None:
    """Problems [first, first + count) of src (problems of shape = (nx, ny,
    nzs) stacked along z) into problems [dst_first, ...) of dst with the
    axes in `flips` (bit 0 x, 1 y, 2 z) reversed."""
    library().PSUNetReflect(src.handle, first, count, dst.handle, dst_first,
                            *shape, flips)


def conv_forward_host(x, w, b, relu, n, d, h, wd, cin, cout):
    y = np.empty(n * d * h * wd * cout, np.uint16)
    library().PSConv3dForwardHost(np.ascontiguousarray(x, np.uint16),
                                  np.ascontiguousarray(w, np.uint16),
                                  np.ascontiguousarray(b, np.float32),
                                  int(relu), n, d, h, wd, cin, cout, y)
    return y


def conv_backward_data_host(dy, w, n, d, h, wd, cin, cout):
    dx = np.empty(n * d * h * wd * cin, np.uint16)
    library().PSConv3dBackwardDataHost(np.ascontiguousarray(dy, np.uint16),
                                       np.ascontiguousarray(w, np.uint16),
                                       n, d, h, wd, cin, cout, dx)
    return dx


def conv_backward_filter_host(x, dy, n, d, h, wd, cin, cout):
    dw = np.empty(cout * 27 * cin, np.float32)
    db = np.empty(cout, np.float32)
    library().PSConv3dBackwardFilterHost(np.ascontiguousarray(x, np.uint16),
                                         np.ascontiguousarray(dy, np.uint16),
                                         n, d, h, wd, cin, cout, dw, db)
    return dw, db
