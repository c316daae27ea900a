"""HIP run kernels against the oracle, through the programs' C ABI.

Bit-exact for every sweep (the kernels keep the stencil's operand order and
never contract a*b+c); reductions within the stated bounds.  Sizes: the
64^3 benchmark case and odd shapes in full, the 512^3 and 512x256x256
benchmark shapes for the whole run where the oracle (OpenMP restatement of
the MPI-OpenMP target, bit-identical to the REF order) finishes in seconds.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_lib import Domain, coef_array, lib

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def d7(ps):
    from physis_amd.programs import Diffusion7
    return Diffusion7()


@pytest.fixture(scope="module")
def him(ps):
    from physis_amd.programs import Himeno
    return Himeno()


def grids(ps, shape, data, n=2):
    gs = [ps.PSGridNew(shape) for _ in range(n)]
    for g in gs:
        ps.PSGridCopyin(g, data)
    return gs


def oracle_run7(a, shape, count, periodic, coef):
    g1 = a.copy()
    g2 = a.copy()
    lib().oracle_diffusion7_run_omp(g1, g2, *shape, count, int(periodic),
                                    *coef, THREADS)
    return g1


def assert_bits(got, want, what):
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (f"{what}: {bad.size} mismatches, first at "
                           f"{bad[0]}: {got[bad[0]]!r} vs {want[bad[0]]!r}")


def test_config1_64cubed_100_sweeps(ps, d7):
    """64^3 clamped diffusion, cc=0.4, 100 sweeps, then PSReduce SUM."""
    from physis_amd.configs import DIFFUSION_COEF, diffusion_input
    shape = (64, 64, 64)
    a = diffusion_input(*shape, seed=0)
    g1, g2 = grids(ps, shape, a)
    d7.run(g1, g2, 100, DIFFUSION_COEF)
    got = ps.PSGridCopyout(g1)
    want = oracle_run7(a, shape, 100, False, DIFFUSION_COEF)
    assert_bits(got, want, "64^3 x100")
    s = ps.PSReduce(ps.PS_SUM, g1)
    s64 = lib().oracle_reduce_float_f64(2, want, want.size)
    assert s == np.float32(s64)   # FP64 sum rounded once to float
    ps.PSGridFree(g1)
    ps.PSGridFree(g2)


@pytest.mark.parametrize("shape", [(37, 23, 19), (64, 48, 40), (8, 8, 8),
                                   (300, 5, 7), (4, 130, 9), (260, 33, 3)])
@pytest.mark.parametrize("periodic", [False, True])
def test_shapes(ps, d7, shape, periodic):
    """Grid extents that are / are not multiples of the vector width, tile
    width and z chunk."""
    from physis_amd.seeded import uniform01
    coef = (0.11, 0.09, 0.12, 0.08, 0.105, 0.095, 0.39)
    a = uniform01(int(np.prod(shape)), seed=17)
    g1, g2 = grids(ps, shape, a)
    d7.run(g1, g2, 6, coef, periodic=periodic)
    want = oracle_run7(a, shape, 6, periodic, coef)
    assert_bits(ps.PSGridCopyout(g1), want, f"{shape} periodic={periodic}")
    ps.PSGridFree(g1)
    ps.PSGridFree(g2)


@pytest.mark.parametrize("shape", [(36, 26, 22), (260, 30, 20), (8, 8, 8),
                                   (64, 13, 5)])
@pytest.mark.parametrize("pairs", [1, 2, 3, 5])
@pytest.mark.parametrize("periodic", [False, True])
def test_two_sweep_passes_leave_both_grids(ps, d7, shape, pairs, periodic):
    """PSStencilRun(m12, m21, pairs) of the 7-point kernels runs as two-sweep
    passes when x is a multiple of 4 (an odd pair count starts with one plain
    pair): g1 must hold sweep 2*pairs and g2 sweep 2*pairs - 1, bit-exact,
    across tile edges in x (260 = 256 + 4), y (not a multiple of the 12-row
    tile) and z chunks."""
    from physis_amd.seeded import uniform01
    coef = (0.11, 0.09, 0.12, 0.08, 0.105, 0.095, 0.39)
    print(f"two-sweep {shape} pairs={pairs} periodic={periodic}", flush=True)
    a = uniform01(int(np.prod(shape)), seed=23)
    g1, g2 = grids(ps, shape, a)
    d7.run(g1, g2, 2 * pairs, coef, periodic=periodic)
    o1, o2 = a.copy(), a.copy()
    lib().oracle_diffusion7_run_omp(o1, o2, *shape, 2 * pairs, int(periodic),
                                    *coef, THREADS)
    assert_bits(ps.PSGridCopyout(g1), o1, f"g1 {shape} x{2 * pairs}")
    assert_bits(ps.PSGridCopyout(g2), o2, f"g2 {shape} x{2 * pairs - 1}")
    ps.PSGridFree(g1)
    ps.PSGridFree(g2)


@pytest.mark.parametrize("dom", [(1, 31, 1, 31, 1, 31), (5, 17, 0, 32, 3, 4),
                                 (0, 32, 7, 8, 0, 32), (13, 14, 2, 30, 9, 25)])
@pytest.mark.parametrize("periodic", [False, True])
def test_subdomains(ps, d7, dom, periodic):
    """Points outside the domain keep the emitted grid's old contents."""
    from physis_amd.seeded import uniform01
    shape = (32, 32, 32)
    coef = (0.1, 0.2, 0.05, 0.15, 0.1, 0.1, 0.3)
    a = uniform01(32 ** 3, seed=2)
    b = uniform01(32 ** 3, seed=3)
    g1 = ps.PSGridNew(shape)
    g2 = ps.PSGridNew(shape)
    ps.PSGridCopyin(g1, a)
    ps.PSGridCopyin(g2, b)
    d7.sweep_domain(g1, g2, dom, coef, periodic=periodic)
    want = b.copy()
    d = Domain((dom[0], dom[2], dom[4]), (dom[1], dom[3], dom[5]))
    lib().oracle_diffusion7_sweep(a, want, *shape, C.byref(d), int(periodic),
                                  *coef)
    assert_bits(ps.PSGridCopyout(g2), want, f"domain {dom}")
    ps.PSGridFree(g1)
    ps.PSGridFree(g2)


@pytest.mark.parametrize("periodic", [False, True])
def test_config2_512cubed_200_sweeps(ps, d7, periodic):
    """512^3, 200 sweeps with PSReduce SUM and MAX after every 50 sweeps:
    every checkpointed field bit-exact, MAX exact, SUM = FP64 sum rounded
    to float; the maximum principle and conservation hold."""
    from physis_amd.configs import DIFFUSION_COEF, diffusion_input
    shape = (512, 512, 512)
    a = diffusion_input(*shape, seed=0)
    total0 = lib().oracle_reduce_float_f64(2, a, a.size)
    g1, g2 = grids(ps, shape, a)
    o1, o2 = a.copy(), a.copy()
    for step in range(4):
        d7.run(g1, g2, 50, DIFFUSION_COEF, periodic=periodic)
        lib().oracle_diffusion7_run_omp(o1, o2, *shape, 50, int(periodic),
                                        *DIFFUSION_COEF, THREADS)
        s = ps.PSReduce(ps.PS_SUM, g1)
        m = ps.PSReduce(ps.PS_MAX, g1)
        assert m == lib().oracle_reduce_float(0, o1, o1.size)
        s64 = lib().oracle_reduce_float_f64(2, o1, o1.size)
        assert s == np.float32(s64)
        assert abs(s64 - total0) <= 1e-5 * total0
        assert m <= a.max()
    assert_bits(ps.PSGridCopyout(g1), o1, "512^3 x200")
    assert_bits(ps.PSGridCopyout(g2), o2, "512^3 x199")
    ps.PSGridFree(g1)
    ps.PSGridFree(g2)


# --------------------------------------------------------------- Himeno
def himeno_grids(ps, shape, coefs, p):
    cg = []
    for c in coefs:
        g = ps.PSGridNew(shape)
        ps.PSGridCopyin(g, c)
        cg.append(g)
    p0, p1 = grids(ps, shape, p)
    return cg, p0, p1


@pytest.mark.parametrize("shape", [(33, 17, 20), (64, 32, 32), (19, 9, 11),
                                   (128, 64, 24)])
def test_himeno_shapes_with_gosa(ps, him, shape):
    from physis_amd.configs import himeno_inputs
    coefs, p = himeno_inputs(*shape, seed=4)
    cg, p0, p1 = himeno_grids(ps, shape, coefs, p)
    nn = 6
    gosa = him.run_gosa(nn, p0, p1, cg)
    o0, o1 = p.copy(), p.copy()
    og = np.zeros(nn)
    lib().oracle_himeno_run(o0, o1, coef_array(coefs), *shape, nn,
                            np.float32(0.8), og.ctypes.data)
    assert_bits(ps.PSGridCopyout(p0), o0, f"himeno {shape}")
    # float per-thread/per-block partial sums, folded in FP64
    np.testing.assert_allclose(gosa, og, rtol=2e-5)
    for g in cg + [p0, p1]:
        ps.PSGridFree(g)


def test_himeno_matches_reference_original_inputs(ps, him):
    """The seeded inputs of the reference-pinned oracle case (the oracle
    equals the original Himeno bit for bit on these)."""
    from oracle_lib import himeno_inputs
    shape = (33, 17, 20)
    coefs, p = himeno_inputs(*shape, 7)
    cg, p0, p1 = himeno_grids(ps, shape, coefs, p)
    him.run(6, p0, p1, cg)
    o0, o1 = p.copy(), p.copy()
    lib().oracle_himeno_run(o0, o1, coef_array(coefs), *shape, 6,
                            np.float32(0.8), None)
    assert_bits(ps.PSGridCopyout(p0), o0, "himeno original inputs")
    for g in cg + [p0, p1]:
        ps.PSGridFree(g)


def test_config3_512x256x256(ps, him):
    """The config-3 shape for 20 sweeps with per-sweep gosa (the full 400
    sweeps are run by bench.py)."""
    from physis_amd.configs import himeno_inputs
    shape = (512, 256, 256)
    coefs, p = himeno_inputs(*shape, seed=0)
    cg, p0, p1 = himeno_grids(ps, shape, coefs, p)
    nn = 20
    gosa = him.run_gosa(nn, p0, p1, cg)
    o0, o1 = p.copy(), p.copy()
    og = np.zeros(nn)
    lib().oracle_himeno_run_omp(o0, o1, coef_array(coefs), *shape, nn,
                                np.float32(0.8), THREADS, og.ctypes.data)
    assert_bits(ps.PSGridCopyout(p0), o0, "himeno 512x256x256 x20")
    # float per-thread/per-block partial sums, folded in FP64
    np.testing.assert_allclose(gosa, og, rtol=2e-5)
    for g in cg + [p0, p1]:
        ps.PSGridFree(g)
