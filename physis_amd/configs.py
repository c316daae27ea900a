"""Synthetic inputs of the benchmark configurations (shared by the tests
and bench.py).  All data come from physis_amd.seeded; the same host arrays
are copied to the HIP grids and handed to the oracle.

Readings of the configuration text, stated once here:
  * "100 PSStencilRun steps" = 100 sweeps = PSStencilRun(m12, m21, 50), as
    the reference benchmark drives it (diffusion3d_physis.c:68-72);
  * Himeno coefficients: a0..a2 = 1, a3 = 1/6, c0..c2 = 1, each multiplied
    by (1 + 0.01*(2u-1)); b0..b2 = 0 + 0.01*(2u-1) additively, because a
    relative perturbation of 0 is 0; wrk1 = 0; bnd = 1 inside, 0 on the
    faces; p = (k/(nz-1))^2 + 1e-3*u with k the z index.
"""
from __future__ import annotations

import numpy as np

from .seeded import uniform01

# 7-point diffusion coefficients (ce, cw, cn, cs, ct, cb, cc)
DIFFUSION_COEF = (0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4)


def diffusion_input(nx: int, ny: int, nz: int, seed: int = 0) -> np.ndarray:
    return uniform01(nx * ny * nz, seed=seed)


def himeno_inputs(nx: int, ny: int, nz: int, seed: int = 0):
    """(coefs[12] in the order a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd wrk1, p)."""
    n = nx * ny * nz
    one, pct = np.float32(1.0), np.float32(0.01)

    def noise(k):
        u = uniform01(n, seed=seed, start=k * n)
        return (np.float32(2.0) * u - one).astype(np.float32)

    coefs = []
    for k in range(3):  # a0 a1 a2
        coefs.append((one * (one + pct * noise(k))).astype(np.float32))
    coefs.append((np.float32(1.0 / 6.0) * (one + pct * noise(3)))
                 .astype(np.float32))
    for k in range(3):  # b0 b1 b2
        coefs.append((pct * noise(4 + k)).astype(np.float32))
    for k in range(3):  # c0 c1 c2
        coefs.append((one * (one + pct * noise(7 + k))).astype(np.float32))
    bnd = np.zeros((nz, ny, nx), np.float32)
    bnd[1:-1, 1:-1, 1:-1] = 1.0
    coefs.append(bnd.reshape(-1))
    coefs.append(np.zeros(n, np.float32))
    k = np.arange(nz, dtype=np.float32)
    prof = (k / np.float32(nz - 1)) ** 2
    p = np.broadcast_to(prof[:, None, None], (nz, ny, nx)).reshape(-1)
    p = (p + np.float32(1e-3) * uniform01(n, seed=seed, start=10 * n)) \
        .astype(np.float32)
    return coefs, p
