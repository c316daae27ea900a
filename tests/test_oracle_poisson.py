"""Pins the Poisson-extension oracle (oracle/poisson.c) on the CPU.

What each check fixes:
  * 16-bit conversions: torch's bfloat16 and numpy's float16 casts (both
    round to nearest even) on random, tie, subnormal and overflow values;
  * the FP32-storage sweep: bit equality with oracle_himeno_sweep, itself
    pinned to the reference's original Himeno benchmark (test_oracle.py);
  * the FP64 residual: omega * a3 * r equals the pinned Himeno update p' - p
    up to FP32 rounding (a wrong sign, a p*a3 for p/a3 or a dropped bnd fails);
  * the FP64 Jacobi solve: scipy's direct sparse solve of the operator
    assembled as a matrix from the stencil's offset table;
  * batches: problems stacked along z solve exactly as when solved alone;
  * the refinement schedule: it reaches the tolerance in BF16, FP16 and FP32
    storage and its solution agrees with the FP64 Jacobi solution within the
    bound the residual tolerance implies.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
from physis_amd.configs import himeno_inputs, poisson_problem


def _specials32():
    rng = np.random.default_rng(1)
    v = (rng.standard_normal(20000) *
         10.0 ** rng.uniform(-45, 38, 20000)).astype(np.float32)
    bits = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32)
    ties = (bits & np.uint32(0xFFFF0000)) | np.uint32(0x8000)
    extra = np.array([0.0, -0.0, np.inf, -np.inf, 65504.0, 65519.99, 65520.0,
                      6.1035156e-05, 5.9604645e-08, 2.9802322e-08,
                      2.9802326e-08, 1e-8, -3e-6, 3.4028235e38],
                     np.float32)
    return np.concatenate([v, bits.view(np.float32), ties.view(np.float32),
                           extra])


def test_bf16_conversion_matches_torch():
    torch = pytest.importorskip("torch")
    L = O.lib()
    v = _specials32()
    v = v[~np.isnan(v)]
    ref = torch.from_numpy(v).to(torch.bfloat16).view(torch.int16).numpy() \
        .view(np.uint16)
    got = np.array([L.oracle_f32_to_bf16(float(x)) for x in v], np.uint16)
    assert np.array_equal(got, ref)
    allbits = np.arange(65536, dtype=np.uint16)
    back = torch.from_numpy(allbits.view(np.int16)).view(torch.bfloat16) \
        .float().numpy()
    ok = ~np.isnan(back)
    mine = np.array([L.oracle_bf16_to_f32(int(h)) for h in allbits],
                    np.float32)
    assert np.array_equal(mine[ok].view(np.uint32), back[ok].view(np.uint32))


def test_f16_conversion_matches_numpy():
    L = O.lib()
    v = _specials32()
    v = v[~np.isnan(v)]
    with np.errstate(over="ignore"):
        ref = v.astype(np.float16).view(np.uint16)
    got = np.array([L.oracle_f32_to_f16(float(x)) for x in v], np.uint16)
    assert np.array_equal(got, ref)
    allbits = np.arange(65536, dtype=np.uint16)
    back = allbits.view(np.float16).astype(np.float32)
    ok = ~np.isnan(back)
    mine = np.array([L.oracle_f16_to_f32(int(h)) for h in allbits], np.float32)
    assert np.array_equal(mine[ok].view(np.uint32), back[ok].view(np.uint32))


def test_fp32_storage_sweep_is_the_himeno_sweep():
    shape = (18, 14, 12)
    nx, ny, nz = shape
    coefs, p = himeno_inputs(nx, ny, nz, seed=5)
    coefs[11] = (np.random.default_rng(2).uniform(-0.1, 0.1, p.size)
                 .astype(np.float32) * coefs[10])
    out = O.poisson_sweep(0, p, coefs, shape, 1)
    ref = p.copy()
    O.lib().oracle_himeno_sweep(p, ref, O.coef_array(coefs), nx, ny, nz,
                                np.float32(0.8), None)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("storage", [1, 2])
def test_storage_sweep_is_fp32_arithmetic_on_rounded_inputs(storage):
    """The 16-bit sweep equals the FP32 sweep of the widened inputs, rounded
    once to the storage type."""
    shape = (16, 12, 10)
    coefs, p = himeno_inputs(*shape, seed=3)
    fields = [O.to_storage(storage, c) for c in coefs]
    ps = O.to_storage(storage, p)
    out = O.poisson_sweep(storage, ps, fields, shape, 1)
    wide = O.poisson_sweep(0, O.from_storage(storage, ps),
                           [O.from_storage(storage, f) for f in fields],
                           shape, 1)
    assert np.array_equal(out, O.to_storage(storage, wide))


def test_residual_is_the_scaled_himeno_update():
    shape = (20, 16, 14)
    nx, ny, nz = shape
    coefs, p = himeno_inputs(nx, ny, nz, seed=9)
    coefs[11] = (np.random.default_rng(4).uniform(-0.5, 0.5, p.size)
                 .astype(np.float32) * coefs[10])
    _, _, sums = O.poisson_residual(0, p, coefs, shape, 1)
    # r itself: through the RHS write at sigma = 1 in FP32 storage
    _, rhs, _ = O.poisson_residual(0, p, coefs, shape, 1, sigma=1.0)
    newp = p.copy()
    O.lib().oracle_himeno_sweep(p, newp, O.coef_array(coefs), nx, ny, nz,
                                np.float32(0.8), None)
    upd = newp.astype(np.float64) - p
    pred = 0.8 * coefs[3].astype(np.float64) * rhs
    assert np.max(np.abs(pred - upd)) < 2e-6
    assert np.isclose(sums[0, 0], np.sum(rhs.astype(np.float64) ** 2),
                      rtol=1e-6)
    assert np.isclose(sums[0, 2], np.sum((coefs[11].astype(np.float64) *
                                          coefs[10]) ** 2), rtol=1e-12)


def _assemble(coefs, shape):
    """Sparse matrix of A over every point (bnd = 0 rows are identity)."""
    sp = pytest.importorskip("scipy.sparse")
    nx, ny, nz = shape
    n = nx * ny * nz
    c = [a.astype(np.float64) for a in coefs]
    # (dx, dy, dz, coefficient index, sign) of every off-centre term of s0
    table = [(0, 0, 1, 0, 1), (0, 1, 0, 1, 1), (1, 0, 0, 2, 1),
             (0, 1, 1, 4, 1), (0, -1, 1, 4, -1), (0, 1, -1, 4, -1),
             (0, -1, -1, 4, 1),
             (1, 1, 0, 5, 1), (1, -1, 0, 5, -1), (-1, 1, 0, 5, -1),
             (-1, -1, 0, 5, 1),
             (1, 0, 1, 6, 1), (1, 0, -1, 6, -1), (-1, 0, 1, 6, -1),
             (-1, 0, -1, 6, 1),
             (0, 0, -1, 7, 1), (0, -1, 0, 8, 1), (-1, 0, 0, 9, 1)]
    rows, cols, vals = [], [], []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                o = x + nx * (y + ny * z)
                if c[10][o] == 0.0:
                    rows.append(o), cols.append(o), vals.append(1.0)
                    continue
                rows.append(o), cols.append(o), vals.append(1.0 / c[3][o])
                for dx, dy, dz, k, sgn in table:
                    q = (x + dx) + nx * ((y + dy) + ny * (z + dz))
                    rows.append(o), cols.append(q), vals.append(-sgn * c[k][o])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_fp64_jacobi_converges_to_the_direct_solution():
    spla = pytest.importorskip("scipy.sparse.linalg")
    shape = (9, 8, 7)
    coefs = poisson_problem(*shape, nbatch=1, gaussians=(2, 2),
                            sigma=(1.5, 3.0), seed=3)
    A = _assemble(coefs, shape)
    b = coefs[11].astype(np.float64)  # 0 where bnd = 0 (Dirichlet 0)
    x = spla.spsolve(A.tocsc(), b)
    p, rel, sweeps = O.poisson_jacobi_fp64(coefs, np.zeros(b.size), shape, 1,
                                           tol=1e-13, check=20)
    assert rel <= 1e-13
    assert np.max(np.abs(p - x)) <= 1e-11 * np.max(np.abs(x))
    _, _, sums = O.poisson_residual(0, x, coefs, shape, 1)
    assert np.sqrt(sums[0, 0] / sums[0, 2]) < 1e-12


def test_batched_problems_are_independent():
    nx, ny, nzs, nb = 10, 9, 8, 3
    coefs = poisson_problem(nx, ny, nzs, nbatch=nb, gaussians=(2, 5),
                            sigma=(1.5, 3.0), seed=4)
    shape = (nx, ny, nzs * nb)
    p, rel, _ = O.poisson_jacobi_fp64(coefs, np.zeros(nx * ny * nzs * nb),
                                      shape, nb, tol=1e-12, check=20)
    m = nx * ny * nzs
    for q in range(nb):
        sub = [np.ascontiguousarray(c[q * m:(q + 1) * m]) for c in coefs]
        pq, _, _ = O.poisson_jacobi_fp64(sub, np.zeros(m), (nx, ny, nzs), 1,
                                         tol=1e-12, check=20)
        assert np.max(np.abs(p[q * m:(q + 1) * m] - pq)) <= \
            1e-9 * np.max(np.abs(pq))


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_mixed_refinement_reaches_tolerance(storage):
    shape = (24, 20, 16)
    coefs = poisson_problem(*shape, nbatch=1, gaussians=(4, 8),
                            sigma=(2.0, 5.0), seed=11)
    p0 = np.zeros(np.prod(shape))
    p, st = O.poisson_solve_mixed(storage, coefs, p0, shape, 1, 0.8, 1e-6,
                                  40, 100000)
    assert st["rel_residual"] <= 1e-6 and not st["stagnated"]
    ref, rel, sweeps = O.poisson_jacobi_fp64(coefs, p0, shape, 1, tol=1e-12,
                                             check=40)
    # plain Jacobi needs about as many sweeps: refinement with Jacobi inner
    # sweeps is Jacobi in exact arithmetic
    _, _, s6 = O.poisson_jacobi_fp64(coefs, p0, shape, 1, tol=1e-6, check=40)
    assert abs(st["sweeps"] - s6) <= 0.25 * s6
    err = np.linalg.norm(p - ref) / np.linalg.norm(ref)
    assert err < 1e-4, err
