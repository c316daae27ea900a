"""Host-side mirror of the Physis DSL over libphysis_rt_hip's C ABI.

Python programs (the tests, the benchmark) drive the HIP target the way a
Physis program drives its runtime: create grids, copy data in, run the
translated program's functions, reduce, copy out.  Names and argument order
follow the DSL (reference include/physis/physis_user.h:46-108 and
physis_common.h:78-103); each call goes straight to the C entry point that
generated code would call (include/physis/physis_hip.h).

The shared libraries are the ones built in-tree by physis_amd/build.py.
There is no CPU fallback: a missing library raises at import of the
function that needs it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
RUNTIME_SO = os.path.join(LIB_DIR, "libphysis_rt_hip.so")

PS_MAX, PS_MIN, PS_SUM, PS_PROD = 0, 1, 2, 3
PS_INT, PS_LONG, PS_FLOAT, PS_DOUBLE = 0, 1, 2, 3
_DTYPES = {PS_INT: np.int32, PS_LONG: np.int64, PS_FLOAT: np.float32,
           PS_DOUBLE: np.float64}
_TYPE_OF = {np.dtype(v): k for k, v in _DTYPES.items()}


class PSDomain(C.Structure):
    """__PSDomain (physis_common.h:88-96)."""
    _fields_ = [("min", C.c_int * 3), ("max", C.c_int * 3),
                ("local_min", C.c_int * 3), ("local_max", C.c_int * 3)]


class PSGridDev(C.Structure):
    _fields_ = [("dim", C.c_int * 3), ("p", C.c_void_p)]


class PSGridStruct(C.Structure):
    """__PSGrid of the HIP target (include/physis/physis_hip.h)."""
    _fields_ = [("p", C.c_void_p), ("dim", C.c_int * 3),
                ("elm_size", C.c_int), ("num_dims", C.c_int),
                ("num_elms", C.c_int64), ("dev", C.POINTER(PSGridDev)),
                ("type", C.c_int)]


class PSGridTypeInfo(C.Structure):
    _fields_ = [("type", C.c_int), ("size", C.c_int),
                ("num_members", C.c_int), ("members", C.c_void_p)]


PSGridPtr = C.POINTER(PSGridStruct)

_rt: Optional[C.CDLL] = None


def runtime() -> C.CDLL:
    """The loaded libphysis_rt_hip (raises if it was not built)."""
    global _rt
    if _rt is None:
        if not os.path.exists(RUNTIME_SO):
            raise RuntimeError(
                f"{RUNTIME_SO} is missing: build it with "
                "`python -m physis_amd.build` (no CPU fallback exists)")
        lib = C.CDLL(RUNTIME_SO, mode=C.RTLD_GLOBAL)
        lib.PSInit.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        lib.PSFinalize.argtypes = []
        lib.__PSGridNew.argtypes = [C.POINTER(PSGridTypeInfo), C.c_int,
                                    C.c_int * 3, C.c_void_p]
        lib.__PSGridNew.restype = PSGridPtr
        lib.__PSGridFree.argtypes = [PSGridPtr, C.c_void_p]
        lib.__PSGridCopyin.argtypes = [PSGridPtr, C.c_void_p, C.c_void_p]
        lib.__PSGridCopyout.argtypes = [PSGridPtr, C.c_void_p, C.c_void_p]
        lib.__PSGridCopyoutSubgrid.argtypes = [PSGridPtr, C.c_void_p,
                                               C.c_int * 3, C.c_int * 3]
        lib.__PSGridCopyinSubgrid.argtypes = [PSGridPtr, C.c_void_p,
                                              C.c_int * 3, C.c_int * 3]
        for name in ("__PSReduceGridFloat", "__PSReduceGridDouble",
                     "__PSReduceGridInt", "__PSReduceGridLong"):
            getattr(lib, name).argtypes = [C.c_void_p, C.c_int, PSGridPtr]
        lib.PSDomain3DNew.argtypes = [C.c_int] * 6
        lib.PSDomain3DNew.restype = PSDomain
        lib.PSDomain2DNew.argtypes = [C.c_int] * 4
        lib.PSDomain2DNew.restype = PSDomain
        lib.PSDomain1DNew.argtypes = [C.c_int] * 2
        lib.PSDomain1DNew.restype = PSDomain
        lib.__PSHipStream.restype = C.c_void_p
        lib.__PSHipNumCUs.restype = C.c_int
        _rt = lib
    return _rt


def PSInit() -> None:
    """PSInit(&argc, &argv, ...) without program arguments."""
    runtime().PSInit(None, None, 3)


def PSFinalize() -> None:
    runtime().PSFinalize()


class Grid:
    """A device grid handle (__PSGrid*) with its element dtype."""

    def __init__(self, handle: PSGridPtr, dtype: np.dtype, shape):
        self.handle = handle
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape)  # (nx, ny, nz), x fastest

    @property
    def num_elms(self) -> int:
        return int(np.prod(self.shape))

    def __repr__(self) -> str:
        return f"Grid({self.shape}, {self.dtype})"


def PSGridNew(dims, dtype=np.float32) -> Grid:
    """PSGrid{1,2,3}D<T>New(nx[, ny[, nz]]); raises if allocation fails
    (the C call returns INVALID_GRID)."""
    dims = list(dims)
    dt = np.dtype(dtype)
    ti = PSGridTypeInfo(_TYPE_OF[dt], dt.itemsize, 0, None)
    d = (C.c_int * 3)(*(dims + [1] * (3 - len(dims))))
    h = runtime().__PSGridNew(C.byref(ti), len(dims), d, None)
    if not h:
        raise MemoryError(f"__PSGridNew({dims}, {dt}) returned INVALID_GRID")
    return Grid(h, dt, dims + [1] * (3 - len(dims)))


def PSGrid3DFloatNew(nx: int, ny: int, nz: int) -> Grid:
    return PSGridNew([nx, ny, nz], np.float32)


def PSGridFree(g: Grid) -> None:
    if g.handle:
        runtime().__PSGridFree(g.handle, None)
        g.handle = None


def PSGridCopyin(g: Grid, src: np.ndarray) -> None:
    a = np.ascontiguousarray(src, dtype=g.dtype).reshape(-1)
    if a.size != g.num_elms:
        raise ValueError(f"copyin of {a.size} elements into {g}")
    runtime().__PSGridCopyin(g.handle, a.ctypes.data, None)


def PSGridCopyout(g: Grid, dst: Optional[np.ndarray] = None) -> np.ndarray:
    """Returns the grid as a flat array (x fastest)."""
    if dst is None:
        dst = np.empty(g.num_elms, dtype=g.dtype)
    if dst.dtype != g.dtype or dst.size != g.num_elms or \
            not dst.flags.c_contiguous:
        raise ValueError("copyout destination does not match the grid")
    runtime().__PSGridCopyout(g.handle, dst.ctypes.data, None)
    return dst


def PSGridCopyoutSubgrid(g: Grid, offset, size) -> np.ndarray:
    out = np.empty(int(np.prod(size)), dtype=g.dtype)
    runtime().__PSGridCopyoutSubgrid(g.handle, out.ctypes.data,
                                     (C.c_int * 3)(*offset),
                                     (C.c_int * 3)(*size))
    return out


def PSGridCopyinSubgrid(g: Grid, src: np.ndarray, offset, size) -> None:
    a = np.ascontiguousarray(src, dtype=g.dtype).reshape(-1)
    if a.size != int(np.prod(size)):
        raise ValueError("sub-grid source size mismatch")
    runtime().__PSGridCopyinSubgrid(g.handle, a.ctypes.data,
                                    (C.c_int * 3)(*offset),
                                    (C.c_int * 3)(*size))


def PSDomain3DNew(minx, maxx, miny, maxy, minz, maxz) -> PSDomain:
    return runtime().PSDomain3DNew(minx, maxx, miny, maxy, minz, maxz)


def PSReduce(op: int, g: Grid):
    """PSReduce(&v, op, g): the whole-grid reduction in the element type."""
    buf = np.zeros(1, dtype=g.dtype)
    fn = {np.dtype(np.float32): "__PSReduceGridFloat",
          np.dtype(np.float64): "__PSReduceGridDouble",
          np.dtype(np.int32): "__PSReduceGridInt",
          np.dtype(np.int64): "__PSReduceGridLong"}[g.dtype]
    getattr(runtime(), fn)(buf.ctypes.data, op, g.handle)
    return buf[0]


def hip_stream() -> int:
    return runtime().__PSHipStream() or 0


def num_cus() -> int:
    return runtime().__PSHipNumCUs()


def load_program(name: str) -> C.CDLL:
    """A translated Physis program built as libphysis_prog_<name>.so."""
    runtime()
    path = os.path.join(LIB_DIR, f"libphysis_prog_{name}.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `python -m physis_amd.build`")
    return C.CDLL(path)
