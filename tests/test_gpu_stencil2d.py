"""2-D 5- and 9-point maps (physis_amd/programs/stencil2d.c), which run on
the 2.5D kernel over the grids' rows view (nx x 1 x ny, y marched as z), on
the GPU against oracle_stencil2d_run, bit for bit on both grids: ragged
shapes, x extents that are not multiples of 4 (one point per lane), single
sweeps and pairs, and 16384^2."""
import numpy as np
import pytest

from oracle_lib import lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def s2d(ps):
    from physis_amd.programs import Stencil2D
    return Stencil2D()


def _grids(ps, a, b):
    ny, nx = a.shape
    ga, gb = ps.PSGridNew((nx, ny)), ps.PSGridNew((nx, ny))
    ps.PSGridCopyin(ga, a)
    ps.PSGridCopyin(gb, b)
    return ga, gb


@pytest.mark.parametrize("points", [5, 9])
@pytest.mark.parametrize("shape,pairs", [((3, 3), 1), ((7, 300), 2),
                                         ((300, 260), 3), ((301, 130), 2),
                                         ((1029, 517), 2),
                                         ((16384, 16384), 1)])
def test_stencil2d_matches_oracle(ps, s2d, points, shape, pairs):
    nx, ny = shape
    print(f"2-D {points}-point {shape} x {pairs} pairs", flush=True)
    rng = np.random.default_rng(nx * 7 + ny + points)
    a = rng.standard_normal((ny, nx)).astype(np.float32)
    b = rng.standard_normal((ny, nx)).astype(np.float32)
    w = 1.0 / points
    ga, gb = _grids(ps, a, b)
    try:
        s2d.run(ga, gb, points, pairs, w)
        got_a, got_b = ps.PSGridCopyout(ga), ps.PSGridCopyout(gb)
    finally:
        ps.PSGridFree(ga)
        ps.PSGridFree(gb)
    wa, wb = a.reshape(-1).copy(), b.reshape(-1).copy()
    lib().oracle_stencil2d_run(wa, wb, nx, ny, points, w, 2 * pairs)
    assert np.array_equal(got_a.reshape(-1).view(np.uint32), wa.view(np.uint32))
    assert np.array_equal(got_b.reshape(-1).view(np.uint32), wb.view(np.uint32))


@pytest.mark.parametrize("points", [5, 9])
def test_stencil2d_single_sweep(ps, s2d, points):
    nx, ny = 515, 77
    rng = np.random.default_rng(points)
    a = rng.standard_normal((ny, nx)).astype(np.float32)
    b = np.zeros((ny, nx), np.float32)
    ga, gb = _grids(ps, a, b)
    try:
        s2d.sweep(ga, gb, points, 0.25)
        got_a, got_b = ps.PSGridCopyout(ga), ps.PSGridCopyout(gb)
    finally:
        ps.PSGridFree(ga)
        ps.PSGridFree(gb)
    wa, wb = a.reshape(-1).copy(), b.reshape(-1).copy()
    lib().oracle_stencil2d_run(wa, wb, nx, ny, points, 0.25, 1)
    assert np.array_equal(got_a.reshape(-1).view(np.uint32), wa.view(np.uint32))
    assert np.array_equal(got_b.reshape(-1).view(np.uint32), wb.view(np.uint32))
