"""ctypes access to the oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path with it, never the
other way round.  Built by `make -C oracle` (also by __graft_entry__.build).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "oracle", "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_outputs.json")

_lib = None

F = C.c_float
FP = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
IP = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Domain(C.Structure):
    _fields_ = [("min", C.c_int * 3), ("max", C.c_int * 3)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")],
                           check=True)
        L = C.CDLL(SO)
        L.oracle_reduce_float.argtypes = [C.c_int, FP, C.c_int64]
        L.oracle_reduce_float.restype = C.c_float
        L.oracle_reduce_int.argtypes = [C.c_int, IP, C.c_int64]
        L.oracle_reduce_int.restype = C.c_int
        L.oracle_reduce_long.argtypes = [C.c_int, C.c_void_p, C.c_int64]
        L.oracle_reduce_long.restype = C.c_long
        L.oracle_reduce_double.argtypes = [C.c_int, C.c_void_p, C.c_int64]
        L.oracle_reduce_double.restype = C.c_double
        L.oracle_reduce_float_f64.argtypes = [C.c_int, FP, C.c_int64]
        L.oracle_reduce_float_f64.restype = C.c_double
        L.oracle_abs_sum_f64.argtypes = [FP, C.c_int64]
        L.oracle_abs_sum_f64.restype = C.c_double
        L.oracle_diffusion7_sweep.argtypes = [FP, FP, C.c_int, C.c_int,
                                              C.c_int, C.POINTER(Domain),
                                              C.c_int] + [F] * 7
        L.oracle_diffusion7_run.argtypes = [FP, FP, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_int] + [F] * 7
        L.oracle_diffusion7_run_omp.argtypes = (
            [FP, FP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int] + [F] * 7 +
            [C.c_int])
        L.oracle_diffusion7_run_omp.restype = C.c_int
        L.oracle_himeno_sweep.argtypes = [FP, FP, C.c_void_p, C.c_int,
                                          C.c_int, C.c_int, F,
                                          C.POINTER(C.c_double)]
        L.oracle_himeno_run.argtypes = [FP, FP, C.c_void_p, C.c_int, C.c_int,
                                        C.c_int, C.c_int, F, C.c_void_p]
        L.oracle_himeno_run_omp.argtypes = [FP, FP, C.c_void_p, C.c_int,
                                            C.c_int, C.c_int, C.c_int, F,
                                            C.c_int]
        L.oracle_himeno_run_omp.restype = C.c_int
        for n in ("oracle_test_7pt", "oracle_test_7pt_periodic",
                  "oracle_test_7pt_neumann", "oracle_test_asymmetric",
                  "oracle_test_asymmetric_periodic"):
            getattr(L, n).argtypes = [FP, FP, C.c_int]
        L.oracle_test_27pt.argtypes = [FP, FP, C.c_int, C.c_int]
        L.oracle_test_27pt_int.argtypes = [IP, IP, C.c_int]
        L.oracle_diffusion_coefs.argtypes = [C.c_int, C.c_int, C.c_int, FP]
        L.oracle_diffusion_init.argtypes = [FP, C.c_int, C.c_int, C.c_int,
                                            C.c_int]
        _lib = L
    return _lib


def coef_array(coefs):
    """The 12 Himeno coefficient grids as a const float *const * argument."""
    arr = (C.c_void_p * 12)(*[c.ctypes.data for c in coefs])
    return arr


def dump_f(a: np.ndarray) -> bytes:
    """The reference tests' dump: printf("%f\\n", v) per element."""
    return ("\n".join(np.char.mod("%f", a.astype(np.float64))) + "\n").encode()


def sha256(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def himeno_inputs(nx, ny, nz, seed):
    """The arrays himeno_ref draws: a0..a3 b0..b2 c0..c2 bnd wrk1 p, each
    x fastest, consecutive in the seeded stream."""
    from physis_amd.seeded import uniform01
    n = nx * ny * nz
    u = uniform01(13 * n, seed=seed)
    arrs = [np.ascontiguousarray(u[i * n:(i + 1) * n]) for i in range(13)]
    return arrs[:12], arrs[12]
