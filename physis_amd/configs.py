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

from concurrent.futures import ThreadPoolExecutor

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


def gaussian_rhs(nx: int, ny: int, nz: int, count: int, sig_lo: float,
                 sig_hi: float, seed: int = 0, stream: int = 0) -> np.ndarray:
    """float32[nz*ny*nx]: a sum of `count` 3-D Gaussians
    a * exp(-|r - c|^2 / (2 s^2)) with centre c uniform over the grid,
    s uniform in [sig_lo, sig_hi] cells and amplitude a = +1 or -1 (the
    reading of "amplitudes +-1"), all drawn from the seeded stream (five
    draws per Gaussian, starting at draw 5*count*stream).  Evaluated in FP64
    as separable products, rounded once to float32."""
    u = uniform01(5 * count, seed=seed, start=5 * count * stream) \
        .astype(np.float64)
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    z = np.arange(nz, dtype=np.float64)
    terms = []
    for g in range(count):
        cx, cy, cz = u[5 * g] * nx, u[5 * g + 1] * ny, u[5 * g + 2] * nz
        s = sig_lo + (sig_hi - sig_lo) * u[5 * g + 3]
        a = 1.0 if u[5 * g + 4] >= 0.5 else -1.0
        k = -0.5 / (s * s)
        gx = np.exp(k * (x - cx) ** 2)
        gy = np.exp(k * (y - cy) ** 2)
        gz = np.exp(k * (z - cz) ** 2)
        terms.append((a, gz, gy[:, None] * gx[None, :]))
    out = np.empty((nz, ny, nx), np.float32)

    # Planes are independent: a thread pool sums the Gaussians plane by
    # plane, in the same order for every plane, so the result does not
    # depend on the split.
    def planes(zr):
        acc = np.zeros((len(zr), ny, nx), np.float64)
        for a, gz, gxy in terms:
            acc += a * gz[zr][:, None, None] * gxy[None]
        out[zr] = acc.astype(np.float32)

    step = max(1, (1 << 21) // max(1, nx * ny))
    ranges = [np.arange(z0, min(nz, z0 + step)) for z0 in range(0, nz, step)]
    if len(ranges) == 1:
        planes(ranges[0])
    else:
        from .seeded import _threads
        with ThreadPoolExecutor(_threads()) as ex:
            list(ex.map(planes, ranges))
    return out.reshape(-1)


def poisson_problem(nx: int, ny: int, nzs: int, nbatch: int = 1,
                    noise: float = 0.01, gaussians=(16, 16),
                    sigma=(4.0, 32.0), seed: int = 0,
                    consistent: bool = False):
    """Coefficient grids (a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd wrk1) of nbatch
    Poisson problems of nx x ny x nzs stacked along z.

    The operator is the Himeno one of himeno_inputs with relative noise
    `noise` (b additive); bnd = 0 on the faces of every problem; wrk1 of
    problem q is a sum of gaussians[0] .. gaussians[1] Gaussians (the count
    drawn per problem) with widths in `sigma`, masked by bnd.

    consistent=True reads "coefficients with +-noise" as a conductance
    field: the six face coefficients a0..a2, c0..c2 carry the noise, b = 0
    (a relative perturbation of 0), and a3 = 1 / (a0+a1+a2+c0+c1+c2), so
    every row balances and the operator is a diagonally dominant M-matrix.
    Perturbing a3 independently acts as a random potential of size ~noise
    that makes the operator indefinite once its smallest eigenvalue
    (~3 pi^2 / n^2) falls below it: with 5% noise Jacobi stalls at 128^3.
    """
    nz = nzs * nbatch
    n = nx * ny * nz
    one, pct = np.float32(1.0), np.float32(noise)

    def nz_(k):
        u = uniform01(n, seed=seed, start=k * n)
        return (np.float32(2.0) * u - one).astype(np.float32)

    coefs = []
    for k in range(3):
        coefs.append((one * (one + pct * nz_(k))).astype(np.float32))
    coefs.append((np.float32(1.0 / 6.0) * (one + pct * nz_(3)))
                 .astype(np.float32))
    for k in range(3):
        coefs.append((pct * nz_(4 + k)).astype(np.float32))
    for k in range(3):
        coefs.append((one * (one + pct * nz_(7 + k))).astype(np.float32))
    if consistent:
        tot = (coefs[0] + coefs[1] + coefs[2] + coefs[7] + coefs[8] +
               coefs[9]).astype(np.float32)
        coefs[3] = (one / tot).astype(np.float32)
        for k in range(4, 7):
            coefs[k] = np.zeros(n, np.float32)
    bnd = np.zeros((nbatch, nzs, ny, nx), np.float32)
    bnd[:, 1:-1, 1:-1, 1:-1] = 1.0
    bnd = bnd.reshape(-1)
    coefs.append(bnd)
    lo, hi = gaussians
    counts = uniform01(nbatch, seed=seed + 1)
    rhs = np.empty(n, np.float32)
    m = nx * ny * nzs
    for q in range(nbatch):
        c = lo + min(int(counts[q] * (hi - lo + 1)), hi - lo)
        rhs[q * m:(q + 1) * m] = gaussian_rhs(nx, ny, nzs, c, *sigma,
                                              seed=seed + 2, stream=q)
    coefs.append((rhs * bnd).astype(np.float32))
    return coefs
