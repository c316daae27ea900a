"""ctypes binding of libphysis_poisson_hip (include/physis/physis_poisson.h):
the precision tier (mixed-precision iterative refinement) and the plain FP32
Jacobi solve of the 19-point Himeno operator, over runtime grids.

No CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import runtime as rt

POISSON_SO = os.path.join(rt.LIB_DIR, "libphysis_poisson_hip.so")

PS_STORAGE_FP32, PS_STORAGE_BF16, PS_STORAGE_FP16 = 0, 1, 2
STORAGE = {"fp32": PS_STORAGE_FP32, "bf16": PS_STORAGE_BF16,
           "fp16": PS_STORAGE_FP16}
PS_POISSON_E, PS_POISSON_RHS, PS_POISSON_COEF0 = 0, 1, 2


class PSPoissonStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("stagnated", C.c_int),
                ("outer", C.c_int), ("sweeps", C.c_long),
                ("rel_residual", C.c_double), ("seconds", C.c_double),
                ("history", C.POINTER(C.c_double)), ("history_cap", C.c_int),
                ("history_len", C.c_int), ("progress_seconds", C.c_double)]

    _SCALARS = ("converged", "stagnated", "outer", "sweeps", "rel_residual",
                "seconds")

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k in self._SCALARS}


def _stats(history: int, progress: float = 0.0):
    st = PSPoissonStats()
    st.progress_seconds = progress
    buf = None
    if history:
        buf = np.zeros(3 * history, np.float64)
        st.history = buf.ctypes.data_as(C.POINTER(C.c_double))
        st.history_cap = history
    return st, buf


def _result(st, buf) -> dict:
    d = st.as_dict()
    if buf is not None:
        d["history"] = buf[:3 * st.history_len].reshape(-1, 3).tolist()
    return d


def time_to(history, level: float):
    """(sweeps, seconds) at which a trace first reached `level`, linearly
    interpolated in log10(residual) between checks; None if never."""
    prev = None
    for sw, rel, t in history:
        if rel <= level:
            if prev is None or prev[1] <= level:
                return sw, t
            f = (np.log10(prev[1]) - np.log10(level)) / \
                (np.log10(prev[1]) - np.log10(rel))
            return prev[0] + f * (sw - prev[0]), prev[2] + f * (t - prev[2])
        prev = (sw, rel, t)
    return None


_lib = None


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        rt.runtime()
        if not os.path.exists(POISSON_SO):
            raise RuntimeError(f"{POISSON_SO} is missing: run "
                               "`python -m physis_amd.build`")
        L = C.CDLL(POISSON_SO)
        G = rt.PSGridPtr
        L.PSPoissonNew.argtypes = [G * 12, C.c_int, C.c_int]
        L.PSPoissonNew.restype = C.c_void_p
        L.PSPoissonFree.argtypes = [C.c_void_p]
        L.PSPoissonSolveMixed.argtypes = [C.c_void_p, G, C.c_float,
                                          C.c_double, C.c_int, C.c_long,
                                          C.POINTER(PSPoissonStats)]
        L.PSPoissonSolvePlain.argtypes = [C.c_void_p, G, C.c_float,
                                          C.c_double, C.c_int, C.c_long,
                                          C.POINTER(PSPoissonStats)]
        L.PSPoissonInnerSweeps.argtypes = [C.c_void_p, C.c_int, C.c_float]
        L.PSPoissonPlainSweeps.argtypes = [C.c_void_p, G, C.c_int, C.c_float]
        L.PSPoissonResidual.argtypes = [C.c_void_p, G, G, C.c_double,
                                        C.c_double, C.c_void_p]
        L.PSPoissonSetField.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.PSPoissonGetField.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        _lib = L
    return _lib


class PoissonSolver:
    """A solver over 12 FP32 coefficient grids (a0 .. c2, bnd, wrk1) holding
    nbatch problems stacked along z."""

    def __init__(self, coef_grids, nbatch: int = 1, storage: str = "bf16"):
        L = library()
        self.storage = STORAGE[storage]
        arr = (rt.PSGridPtr * 12)(*[g.handle for g in coef_grids])
        self.handle = L.PSPoissonNew(arr, nbatch, self.storage)
        if not self.handle:
            raise RuntimeError("PSPoissonNew failed")
        self.coefs = list(coef_grids)  # keep the masters alive
        self.shape = coef_grids[0].shape
        self.nbatch = nbatch
        self.L = L

    @property
    def field_dtype(self):
        return np.float32 if self.storage == PS_STORAGE_FP32 else np.uint16

    def free(self) -> None:
        if self.handle:
            self.L.PSPoissonFree(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def solve_mixed(self, p, omega=0.8, tol=1e-6, inner=100,
                    max_sweeps=10 ** 7, history=0, progress=0.0) -> dict:
        """Stats dict; with history=N also "history": up to N
        (sweeps, rel_residual, seconds) rows, one per residual check;
        progress > 0 prints a line to stderr about that often (seconds)."""
        st, buf = _stats(history, progress)
        self.L.PSPoissonSolveMixed(self.handle, p.handle, omega, tol, inner,
                                   max_sweeps, C.byref(st))
        return _result(st, buf)

    def solve_plain(self, p, omega=0.8, tol=1e-6, check=100,
                    max_sweeps=10 ** 7, history=0, progress=0.0) -> dict:
        st, buf = _stats(history, progress)
        self.L.PSPoissonSolvePlain(self.handle, p.handle, omega, tol, check,
                                   max_sweeps, C.byref(st))
        return _result(st, buf)

    def inner_sweeps(self, n: int, omega: float = 0.8) -> None:
        self.L.PSPoissonInnerSweeps(self.handle, n, omega)

    def plain_sweeps(self, p, n: int, omega: float = 0.8) -> None:
        self.L.PSPoissonPlainSweeps(self.handle, p.handle, n, omega)

    def residual(self, p, p_out=None, inv_sigma=0.0, sigma=0.0) -> np.ndarray:
        """(nbatch, 3): sum r^2, max |r|, sum (bnd wrk1)^2 per problem."""
        sums = np.zeros(3 * self.nbatch, np.float64)
        self.L.PSPoissonResidual(self.handle, p.handle,
                                 p_out.handle if p_out is not None else None,
                                 inv_sigma, sigma, sums.ctypes.data)
        return sums.reshape(self.nbatch, 3)

    def set_field(self, field: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.field_dtype).reshape(-1)
        if a.size != int(np.prod(self.shape)):
            raise ValueError("field size mismatch")
        if self.L.PSPoissonSetField(self.handle, field, a.ctypes.data) != 0:
            raise ValueError(f"field {field} cannot be set")

    def get_field(self, field: int) -> np.ndarray:
        out = np.empty(int(np.prod(self.shape)), self.field_dtype)
        if self.L.PSPoissonGetField(self.handle, field, out.ctypes.data) != 0:
            raise ValueError(f"bad field {field}")
        return out
