"""Pins the oracle (oracle/liboracle.so) before anything is compared with it.

Every oracle function is checked against the reference's own outputs
(tests/golden/ref_outputs.json, written by tests/golden/make_ref_outputs.py
from the reference's hand-written test oracles, its C++ diffusion baseline
and its original Himeno benchmark) or against what the mathematics fixes.
CPU only.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle_lib import (Domain, GOLDEN, ROOT, coef_array, dump_f,
                        himeno_inputs, lib, sha256)

GOLD = json.load(open(GOLDEN))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def iota(n, dtype=np.float32):
    return np.arange(n, dtype=np.int64).astype(dtype)


# --------------------------------------------------------- system tests
# Host logic of each reference test program restated around the oracle's
# kernel (tests/system_tests/test_cases/<name>.c main()).
def run_system_test(name):
    L = lib()
    if name == "test_7-pt-neumann-cond":
        n = 8
        g1, g2 = iota(n ** 3), np.zeros(n ** 3, np.float32)
        L.oracle_test_7pt_neumann(g1, g2, n)
        return dump_f(g2)
    if name == "test_27-pt-reduction":
        n = 16
        g1, g2 = iota(n ** 3, np.int32), iota(n ** 3, np.int32)
        L.oracle_test_27pt_int(g1, g2, n)
        v1 = L.oracle_reduce_int(2, g2, g2.size)
        L.oracle_test_27pt_int(g2, g1, n)
        v2 = L.oracle_reduce_int(2, g1, g1.size)
        return f"{v1}\n{v2}\n".encode()
    if name == "test_reduction-3d-sum":
        g = iota(16 ** 3)
        return ("%f\n" % L.oracle_reduce_float(2, g, g.size)).encode()
    n = 32
    g1, g2 = iota(n ** 3), iota(n ** 3)
    if name == "test_redblack":
        # the unweighted in-place sum: integer values below 2^24, so the
        # sum is exact in any operand order
        L.oracle_redblack7_run(g1, n, n, n, 1, np.ones(7, np.float32))
        return dump_f(g1)
    if name == "test_7-pt-multi-iterations":
        for _ in range(5):
            L.oracle_test_7pt(g1, g2, n)
            L.oracle_test_7pt(g2, g1, n)
        return dump_f(g1)
    fn = {
        "test_7-pt": lambda: L.oracle_test_7pt(g1, g2, n),
        "test_7-pt-periodic": lambda: L.oracle_test_7pt_periodic(g1, g2, n),
        "test_asymmetric": lambda: L.oracle_test_asymmetric(g1, g2, n),
        "test_asymmetric-periodic":
            lambda: L.oracle_test_asymmetric_periodic(g1, g2, n),
        "test_27-pt": lambda: L.oracle_test_27pt(g1, g2, n, 0),
        "test_27-pt-periodic": lambda: L.oracle_test_27pt(g1, g2, n, 1),
    }[name]
    fn()
    return dump_f(g2)


# The system tests the oracle restates; the fixture holds every reference
# oracle's output, which the GPU runs of the unchanged programs are compared
# with directly.
ORACLE_SYSTEM_TESTS = [
    "test_7-pt", "test_7-pt-periodic", "test_7-pt-neumann-cond",
    "test_7-pt-multi-iterations", "test_asymmetric", "test_asymmetric-periodic",
    "test_27-pt", "test_27-pt-periodic", "test_27-pt-reduction",
    "test_reduction-3d-sum", "test_redblack"]


@pytest.mark.parametrize("name", ORACLE_SYSTEM_TESTS)
def test_system_test_golden(name):
    out = run_system_test(name)
    rec = GOLD["system_tests"][name]
    assert out.count(b"\n") == rec["lines"]
    assert sha256(out) == rec["sha256"]


def test_survey_hash_prefixes_match_fixture():
    # The same outputs were hashed independently when the reference was
    # surveyed; the fixture must agree with those prefixes.
    prefixes = {"test_7-pt": "da79101f27c3ea7e",
                "test_7-pt-periodic": "c8d38cff03c27f9a",
                "test_7-pt-neumann-cond": "e2d33bbe536e90c4",
                "test_27-pt-reduction": "2d33eff5c36e21aa",
                "test_reduction-3d-sum": "10f2d295e0bf62e7",
                "test_7-pt-double-type": "da79101f27c3ea7e",
                "test_7-pt-int-type": "64f0e0f649602d4f",
                "test_7-pt-type-mix": "33bdddf7ac356fd0",
                "test_asymmetric": "7b16fb365f996516",
                "test_asymmetric-periodic": "9e2737a3082d07f6",
                "test_27-pt": "5e0a49abe433f1e3",
                "test_27-pt-periodic": "0297608095678316",
                "test_9-pt-reduction": "2867d20e81cabdc8",
                "test_9-pt-periodic-reduction": "5d71ccfd64e69d43",
                "test_mixed-dim": "edc2a7b52a06c799",
                "test_mixed-dim2": "e099c965c83f6738",
                "test_mixed-dim3": "d4115d0f100b26d1",
                "test_stencil-hole": "39d9002f7ace1d12",
                "test_3-pt-periodic": "099f372c4f1a1571",
                "test_5-pt-periodic": "0c07bafbbe7a6802",
                "test_redblack": "9d38b80f37027e4f"}
    for k, v in prefixes.items():
        assert GOLD["system_tests"][k]["sha256"].startswith(v)


# ---------------------------------------------------- clamped diffusion
def diffusion_state(n, count):
    L = lib()
    coef = np.zeros(8, np.float32)
    L.oracle_diffusion_coefs(n, n, n, coef)
    f = np.zeros(n ** 3, np.float32)
    L.oracle_diffusion_init(f, n, n, n, count)
    return coef, f


@pytest.mark.parametrize("key", sorted(GOLD["diffusion_baseline"]))
def test_diffusion_matches_reference_baseline(key):
    """Bit-level (%f dump) equality with the reference's Baseline::RunKernel
    (examples/diffusion-benchmark/baseline.cc:22-49) and its RMS accuracy
    against the closed-form solution (baseline.cc:51-60)."""
    n, count = map(int, key.split("x"))
    L = lib()
    coef, f1 = diffusion_state(n, 0)
    f2 = np.zeros_like(f1)
    ce, cw, cn, cs, ct, cb, cc = (float(c) for c in coef[:7])
    L.oracle_diffusion7_run(f1, f2, n, n, n, count, 0, ce, cw, cn, cs, ct, cb,
                            cc)
    assert sha256(dump_f(f1)) == GOLD["diffusion_baseline"][key]["sha256"]
    _, ref = diffusion_state(n, count)
    err = np.float32(0.0)
    for d in (ref - f1):
        err = np.float32(err + np.float32(d * d))
    acc = np.sqrt(np.float32(err / np.float32(n ** 3)))
    assert "%e" % acc == GOLD["diffusion_baseline"][key]["accuracy"]


def test_diffusion_periodic_matches_system_test_kernel_on_sum_stencil():
    """With cc = ce = ... = 1 and the operand order of test_7-pt-periodic the
    periodic diffusion body is that test's kernel up to operand order; for
    integer-valued data below 2^24 every order is exact, so both agree."""
    L = lib()
    n = 32
    g = iota(n ** 3)
    a, b = np.zeros_like(g), np.zeros_like(g)
    d = Domain((0, 0, 0), (n, n, n))
    L.oracle_diffusion7_sweep(g, a, n, n, n, C.byref(d), 1, *([1.0] * 7))
    L.oracle_test_7pt_periodic(g, b, n)
    assert np.array_equal(a, b)


def test_diffusion_conserves_and_bounds():
    """Coefficients summing to one with clamped or periodic boundaries keep
    the total (up to rounding) and obey the maximum principle."""
    L = lib()
    from physis_amd.seeded import uniform01
    n = 20
    for periodic in (0, 1):
        g1 = uniform01(n ** 3, seed=3)
        s0, mx, mn = g1.astype(np.float64).sum(), g1.max(), g1.min()
        g2 = np.zeros_like(g1)
        L.oracle_diffusion7_run(g1, g2, n, n, n, 20, periodic, *([0.1] * 6),
                                0.4)
        assert abs(g1.astype(np.float64).sum() - s0) < 1e-5 * s0
        assert g1.max() <= mx and g1.min() >= mn


def test_omp_shape_is_bit_identical():
    L = lib()
    from physis_amd.seeded import uniform01
    n = 24
    for periodic in (0, 1):
        a1 = uniform01(n * 20 * 17, seed=5)
        b1 = a1.copy()
        a2, b2 = np.zeros_like(a1), np.zeros_like(a1)
        L.oracle_diffusion7_run(a1, a2, n, 20, 17, 6, periodic, 0.1, 0.2,
                                0.05, 0.1, 0.15, 0.1, 0.3)
        L.oracle_diffusion7_run_omp(b1, b2, n, 20, 17, 6, periodic, 0.1, 0.2,
                                    0.05, 0.1, 0.15, 0.1, 0.3, 5)
        assert np.array_equal(a1, b1)


# -------------------------------------------------------------- Himeno
@pytest.mark.parametrize("key", sorted(GOLD["himeno_original"]))
def test_himeno_matches_original_benchmark(key):
    """p after nn sweeps equals, bit for bit, the reference's original Himeno
    jacobi() (examples/himeno/himenobmtxpa_original.c:298-346) run on the
    same inputs; its float gosa agrees with the oracle's FP64 residual."""
    dims, rest = key.split("s")
    nx, ny, nz, nn = map(int, dims.split("x"))
    seed = int(rest)
    coefs, p = himeno_inputs(nx, ny, nz, seed)
    p0, p1 = p.copy(), p.copy()
    gosa = np.zeros(nn)
    lib().oracle_himeno_run(p0, p1, coef_array(coefs), nx, ny, nz, nn,
                            np.float32(0.8), gosa.ctypes.data)
    body = "".join("%08x\n" % u for u in p0.view(np.uint32))
    assert sha256(body.encode()) == GOLD["himeno_original"][key]["p_bits_sha256"]
    ref = float(GOLD["himeno_original"][key]["gosa"])
    assert abs(ref - gosa[-1]) <= 1e-5 * abs(gosa[-1])


def himeno_one(p, coefs, nx, ny, nz, omega=1.0):
    out = p.copy()
    g = C.c_double(0)
    lib().oracle_himeno_sweep(p, out, coef_array(coefs), nx, ny, nz,
                              np.float32(omega), C.byref(g))
    return out, g.value


def himeno_coefs(n3, a=(0, 0, 0, 1), b=(0, 0, 0), c=(0, 0, 0)):
    vals = list(a) + list(b) + list(c) + [1.0, 0.0]
    return [np.full(n3, v, np.float32) for v in vals]


def test_himeno_constant_is_fixed_point():
    nx, ny, nz = 9, 8, 7
    coefs = himeno_coefs(nx * ny * nz, a=(1, 1, 1, 1.0 / 6), c=(1, 1, 1))
    p = np.full(nx * ny * nz, 0.75, np.float32)
    out, g = himeno_one(p, coefs, nx, ny, nz)
    assert np.array_equal(out, p) and g == 0.0


@pytest.mark.parametrize("term", range(12))
def test_himeno_each_term_reads_its_neighbour(term):
    """Each coefficient multiplies one fixed neighbour pattern: with p a
    polynomial that separates the patterns and only that coefficient set,
    s0 (recovered from p' with a3 = 1, bnd = 1, omega = 1) must equal the
    pattern's value.  Catches a dropped term, a sign or an axis mix-up."""
    nx, ny, nz = 7, 6, 5
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                          indexing="ij")
    x, y, z = (v.transpose(2, 1, 0).reshape(-1).astype(np.float32)
               for v in (x, y, z))
    # (coefficient index in a0..c2 + wrk1, p, expected s0 as a function)
    cases = [
        (0, z, z + 1), (1, y, y + 1), (2, x, x + 1),
        (4, y * z, np.float32(4)), (5, x * y, np.float32(4)),
        (6, x * z, np.float32(4)),
        (7, z, z - 1), (8, y, y - 1), (9, x, x - 1),
        (10, None, None), (11, None, None), (3, None, None)]
    idx, p, want = cases[term]
    n3 = nx * ny * nz
    coefs = [np.zeros(n3, np.float32) for _ in range(12)]
    coefs[3][:] = 1.0   # a3
    coefs[10][:] = 1.0  # bnd
    if idx == 10:   # bnd = 0 freezes p
        coefs[10][:] = 0.0
        coefs[0][:] = 1.0
        p = z.copy()
        out, _ = himeno_one(p, coefs, nx, ny, nz)
        assert np.array_equal(out, p)
        return
    if idx == 11:   # wrk1 adds into s0
        coefs[11][:] = 2.5
        p = np.zeros(n3, np.float32)
        want = np.float32(2.5)
    elif idx == 3:  # a3 scales s0: s0 = p(z+1) with a3 = 0.5
        coefs[0][:] = 1.0
        coefs[3][:] = 0.5
        p = z.copy()
        out, _ = himeno_one(p, coefs, nx, ny, nz)
        inner = (x > 0) & (x < nx - 1) & (y > 0) & (y < ny - 1) & \
                (z > 0) & (z < nz - 1)
        assert np.array_equal(out[inner], (0.5 * (z + 1))[inner])
        return
    else:
        coefs[idx][:] = 1.0
    out, _ = himeno_one(p, coefs, nx, ny, nz)
    # omega = 1, a3 = 1, bnd = 1: p' = p + (s0 - p) = s0
    inner = (x > 0) & (x < nx - 1) & (y > 0) & (y < ny - 1) & \
            (z > 0) & (z < nz - 1)
    want = np.broadcast_to(np.asarray(want, np.float32), p.shape)
    assert np.array_equal(out[inner], want[inner])
    assert np.array_equal(out[~inner], p[~inner])


def test_himeno_omp_shape_is_bit_identical():
    nx, ny, nz = 17, 11, 13
    coefs, p = himeno_inputs(nx, ny, nz, 11)
    a0, a1 = p.copy(), p.copy()
    b0, b1 = p.copy(), p.copy()
    ga, gb = np.zeros(4), np.zeros(4)
    lib().oracle_himeno_run(a0, a1, coef_array(coefs), nx, ny, nz, 4,
                            np.float32(0.8), ga.ctypes.data)
    lib().oracle_himeno_run_omp(b0, b1, coef_array(coefs), nx, ny, nz, 4,
                                np.float32(0.8), 4, gb.ctypes.data)
    assert np.array_equal(a0, b0)
    # the same FP64 terms summed in another grouping
    np.testing.assert_allclose(gb, ga, rtol=1e-12)
    assert np.all(ga > 0)


# ----------------------------------------------------------- reductions
def test_reduction_semantics():
    L = lib()
    f = np.array([-3.0, -1.0, -2.0], np.float32)
    assert L.oracle_reduce_float(0, f, 3) == -1.0   # MAX of all-negative data
    assert L.oracle_reduce_float(1, f, 3) == -3.0
    assert L.oracle_reduce_float(3, f, 3) == -6.0
    i = np.array([2 ** 31 - 1, 1], np.int32)
    assert L.oracle_reduce_int(2, i, 2) == -2 ** 31  # two's-complement wrap
    g = iota(16 ** 3)
    assert L.oracle_reduce_float(2, g, g.size) == 8386560.0
    assert L.oracle_reduce_float_f64(2, g, g.size) == 8386560.0


def test_redblack_oracle_is_order_free():
    """Every neighbour of a point is of the other colour, so a colour sweep
    in any order gives the REFERENCE loop order's bits: a vectorised sweep
    (each colour computed from the array as the previous colour left it)
    equals the oracle's in-place loops with the diffusion weights."""
    nx, ny, nz = 11, 8, 9
    rng = np.random.default_rng(3)
    g = rng.random(nx * ny * nz).astype(np.float32)
    coef = np.array([0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4], np.float32)
    o = g.copy()
    lib().oracle_redblack7_run(o, nx, ny, nz, 2, coef)
    a = g.reshape(nz, ny, nx).copy()
    ce, cw, cn, cs, ct, cb, cc = coef
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx),
                          indexing="ij")
    inner = ((x > 0) & (x < nx - 1) & (y > 0) & (y < ny - 1) & (z > 0) &
             (z < nz - 1))
    for _ in range(2):
        for c in range(2):
            m = inner & ((x + y + z + c) % 2 == 0)
            v = np.zeros_like(a)
            s = (slice(1, -1),) * 3
            v[s] = (cc * a[s] + cw * a[1:-1, 1:-1, :-2] +
                    ce * a[1:-1, 1:-1, 2:] + cs * a[1:-1, 2:, 1:-1] +
                    cn * a[1:-1, :-2, 1:-1] + cb * a[:-2, 1:-1, 1:-1] +
                    ct * a[2:, 1:-1, 1:-1])
            a[m] = v[m]
    assert np.array_equal(a.reshape(-1), o)
    assert not np.array_equal(o, g)


def test_udt7_oracle_matches_reference_output():
    """oracle_udt7_sweep with unit weights is the reference's
    test_user-defined-type-7-pt (p = i, q = 0, one map over [1, N-1)^3;
    integer sums below 2^24 are exact in any order): its "%f %f" dump of
    (p, q) hashes to the reference oracle's output."""
    n = 32
    p = np.arange(n ** 3, dtype=np.float32)
    q = np.zeros(n ** 3, np.float32)
    lib().oracle_udt7_sweep(p, q, n, n, n, np.ones(7, np.float32))
    rows = np.stack([p, q], 1).astype(np.float64)
    dump = "".join("%f %f\n" % (a, b) for a, b in rows.tolist()).encode()
    assert sha256(dump) == \
        GOLD["system_tests"]["test_user-defined-type-7-pt"]["sha256"]


@pytest.mark.parametrize("points", [5, 9])
def test_stencil2d_oracle_matches_reference_output(points):
    """oracle_stencil2d_run with w = 1 and one sweep is the reference's
    test_5-pt-2d / test_9-pt-2d (input i, output grid initialised to the
    input, one map over [1, N-1)^2; integer sums below 2^24 are exact): its
    "%f" dump hashes to the reference oracle's output."""
    n = 32
    a = np.arange(n * n, dtype=np.float32)
    b = a.copy()
    lib().oracle_stencil2d_run(a, b, n, n, points, 1.0, 1)
    dump = "".join("%f\n" % v for v in b.astype(np.float64).tolist()).encode()
    assert sha256(dump) == \
        GOLD["system_tests"][f"test_{points}-pt-2d"]["sha256"]


@pytest.mark.parametrize("points", [5, 9])
def test_stencil2d_oracle_linear_field(points):
    """A linear field x + 1000 y on a non-square grid: every interior point
    sums to points * f exactly (the neighbour offsets cancel), points outside
    keep their value, and a second sweep maps back into a; a transposed
    index or a wrong neighbour breaks it."""
    nx, ny = 37, 23
    f = (np.arange(nx)[None, :] + 1000.0 * np.arange(ny)[:, None]).astype(
        np.float32).reshape(-1)
    a, b = f.copy(), np.full_like(f, -1.0)
    lib().oracle_stencil2d_run(a, b, nx, ny, points, 1.0, 1)
    want = np.full((ny, nx), -1.0, np.float32)
    want[1:-1, 1:-1] = points * f.reshape(ny, nx)[1:-1, 1:-1]
    assert np.array_equal(b.reshape(ny, nx), want)
    lib().oracle_stencil2d_run(a, b, nx, ny, points, 0.5, 2)
    assert np.array_equal(b.reshape(ny, nx)[1:-1, 1:-1],
                          (points * f.reshape(ny, nx)[1:-1, 1:-1] *
                           np.float32(0.5)).astype(np.float32))


def test_mixed7_oracle_matches_reference_output():
    """oracle_mixed7_sweep is the reference's test_mixed-dim2 (g = i, output
    grid initialised to g, coefficients 1 + (i % 2) on every axis, one map
    over [1, N-1)^3; integer sums below 2^24 are exact): its "%f" dump
    hashes to the reference oracle's output."""
    n = 32
    g = np.arange(n ** 3, dtype=np.float32)
    o = g.copy()
    c = (1 + np.arange(n) % 2).astype(np.float32)
    lib().oracle_mixed7_sweep(g, o, n, n, n, c, c, c)
    dump = "".join("%f\n" % v for v in o.astype(np.float64).tolist()).encode()
    assert sha256(dump) == GOLD["system_tests"]["test_mixed-dim2"]["sha256"]


def test_mixed7_oracle_axis_weights():
    """Distinct weights per axis on a non-cubic grid: with g = 1 the interior
    equals 1 + the sum of the six neighbouring coefficients, so a swapped
    axis or coefficient index shows."""
    nx, ny, nz = 9, 7, 5
    g = np.ones(nx * ny * nz, np.float32)
    o = np.full_like(g, -1.0)
    cx = np.arange(nx, dtype=np.float32)
    cy = 100 * np.arange(ny, dtype=np.float32)
    cz = 10000 * np.arange(nz, dtype=np.float32)
    lib().oracle_mixed7_sweep(g, o, nx, ny, nz, cx, cy, cz)
    o = o.reshape(nz, ny, nx)
    z, y, x = 2, 3, 4
    assert o[z, y, x] == 1 + cx[x - 1] + cx[x + 1] + cy[y - 1] + cy[y + 1] + \
        cz[z - 1] + cz[z + 1]
    assert o[0, 0, 0] == -1.0 and o[z, y, 0] == -1.0
