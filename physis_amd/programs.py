"""Typed ctypes bindings of the translated benchmark programs.

Each program is a Physis DSL source under physis_amd/programs/ whose
translation (<name>.hip.cpp) is built into libphysis_prog_<name>.so; the
functions bound here are the program's own C functions (the ones a Physis
user would call), taking grid handles from physis_amd.runtime.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import runtime as rt

_F = C.c_float


class Diffusion7:
    """physis_amd/programs/diffusion7.c: 7-point diffusion, clamped or
    periodic boundaries."""

    def __init__(self):
        L = rt.load_program("diffusion7")
        G = rt.PSGridPtr
        L.diffusion7_run.argtypes = [G, G, C.c_int] + [_F] * 7
        L.diffusion7_periodic_run.argtypes = [G, G, C.c_int] + [_F] * 7
        L.diffusion7_sweep_domain.argtypes = ([G, G] + [C.c_int] * 7 +
                                              [_F] * 7)
        self.lib = L

    def run(self, g1, g2, count: int, coef, periodic: bool = False) -> None:
        """count sweeps g1 -> g2 -> g1 over the whole grid (count even).
        coef = (ce, cw, cn, cs, ct, cb, cc)."""
        fn = self.lib.diffusion7_periodic_run if periodic else \
            self.lib.diffusion7_run
        fn(g1.handle, g2.handle, count, *[float(c) for c in coef])

    def sweep_domain(self, g1, g2, dom, coef, periodic: bool = False) -> None:
        """One sweep g1 -> g2 over dom = (x0, x1, y0, y1, z0, z1)."""
        self.lib.diffusion7_sweep_domain(g1.handle, g2.handle, *dom,
                                         int(periodic),
                                         *[float(c) for c in coef])


class RedBlack7:
    """physis_amd/programs/redblack7.c: red-black Gauss-Seidel relaxation of
    the 7-point diffusion operator, in place over the grid's interior."""

    def __init__(self):
        L = rt.load_program("redblack7")
        L.redblack7_run.argtypes = [rt.PSGridPtr, C.c_int] + [_F] * 7
        self.lib = L

    def run(self, g, count: int, coef) -> None:
        """count iterations (red then black sweep) of grid g.
        coef = (ce, cw, cn, cs, ct, cb, cc)."""
        self.lib.redblack7_run(g.handle, count, *[float(c) for c in coef])


class Udt7:
    """physis_amd/programs/udt7.c: a 7-point stencil over a grid of points
    with members (p, q): q <- weighted 7-point sum of p over the interior."""

    def __init__(self):
        L = rt.load_program("udt7")
        L.udt7_new.argtypes = [C.c_int] * 3
        L.udt7_new.restype = C.c_void_p
        L.udt7_copyin.argtypes = [C.c_void_p, C.c_void_p]
        L.udt7_copyout.argtypes = [C.c_void_p, C.c_void_p]
        L.udt7_free.argtypes = [C.c_void_p]
        L.udt7_run.argtypes = [C.c_void_p, C.c_int] + [_F] * 7
        L.udt7_whole_run.argtypes = [C.c_void_p, C.c_void_p] + [_F] * 7
        self.lib = L

    def new(self, shape, pq: np.ndarray):
        """A grid of shape (nx, ny, nz) holding the points pq [n][2]."""
        g = self.lib.udt7_new(*shape)
        if not g:
            raise RuntimeError("udt7_new failed")
        self.lib.udt7_copyin(g, np.ascontiguousarray(pq, np.float32).ctypes.data)
        return g

    def copyout(self, g, n: int) -> np.ndarray:
        out = np.empty((n, 2), np.float32)
        self.lib.udt7_copyout(g, out.ctypes.data)
        return out

    def free(self, g) -> None:
        self.lib.udt7_free(g)

    def run(self, g, count: int, coef) -> None:
        self.lib.udt7_run(g, count, *[float(c) for c in coef])

    def whole_run(self, g, g2, coef) -> None:
        """One sweep g2 <- (p, 7-point(p)) over the interior, through
        struct-valued gets and emit."""
        self.lib.udt7_whole_run(g, g2, *[float(c) for c in coef])


class Stencil2D:
    """physis_amd/programs/stencil2d.c: 2-D 5- and 9-point maps over the
    interior of 2-D float grids."""

    def __init__(self):
        L = rt.load_program("stencil2d")
        G = rt.PSGridPtr
        L.stencil2d_run.argtypes = [G, G, C.c_int, C.c_int, _F]
        L.stencil2d_sweep.argtypes = [G, G, C.c_int, _F]
        self.lib = L

    def run(self, a, b, points: int, count: int, w: float) -> None:
        """count pairs of sweeps a -> b -> a (points = 5 or 9)."""
        self.lib.stencil2d_run(a.handle, b.handle, points, count, float(w))

    def sweep(self, a, b, points: int, w: float) -> None:
        """One sweep a -> b."""
        self.lib.stencil2d_sweep(a.handle, b.handle, points, float(w))


class Mixed7:
    """physis_amd/programs/mixed7.c: a 3-D 7-point map weighted by 1-D
    coefficient grids along each axis."""

    def __init__(self):
        L = rt.load_program("mixed7")
        G = rt.PSGridPtr
        L.mixed7_run.argtypes = [G] * 5
        self.lib = L

    def run(self, g, g2, cx, cy, cz) -> None:
        """One sweep g2 <- kernel(g) over the interior."""
        self.lib.mixed7_run(g.handle, g2.handle, cx.handle, cy.handle,
                            cz.handle)


class Himeno:
    """physis_amd/programs/himeno.c: 19-point Himeno Jacobi sweeps."""

    COEFS = ("a0", "a1", "a2", "a3", "b0", "b1", "b2", "c0", "c1", "c2",
             "bnd", "wrk1")

    def __init__(self):
        L = rt.load_program("himeno")
        G = rt.PSGridPtr
        L.himeno_run.argtypes = [C.c_int] + [G] * 14 + [_F]
        L.himeno_run_gosa.argtypes = ([C.c_int] + [G] * 14 + [_F] +
                                      [C.c_void_p, C.c_void_p])
        self.lib = L

    def run(self, nn: int, p0, p1, coefs, omega: float = 0.8) -> None:
        self.lib.himeno_run(nn, p0.handle, p1.handle,
                            *[c.handle for c in coefs], float(omega))

    def run_gosa(self, nn: int, p0, p1, coefs,
                 omega: float = 0.8) -> np.ndarray:
        """nn sweeps; returns the float residual sum of each sweep."""
        even = np.zeros(nn // 2, np.float32)
        odd = np.zeros(nn // 2, np.float32)
        self.lib.himeno_run_gosa(nn, p0.handle, p1.handle,
                                 *[c.handle for c in coefs], float(omega),
                                 even.ctypes.data, odd.ctypes.data)
        out = np.empty(nn // 2 * 2, np.float32)
        out[0::2] = even
        out[1::2] = odd
        return out
