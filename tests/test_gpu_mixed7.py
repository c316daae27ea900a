"""A 3-D 7-point map weighted by 1-D coefficient grids
(physis_amd/programs/mixed7.c; the 2.5D kernel with the 1-D grids read
through bounds-checked gets) on the GPU against oracle_mixed7_sweep, bit for
bit, at ragged shapes, an x extent that is not a multiple of 4, and 256^3."""
import numpy as np
import pytest

from oracle_lib import lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed(ps):
    from physis_amd.programs import Mixed7
    return Mixed7()


@pytest.mark.parametrize("shape", [(37, 13, 9), (64, 48, 40), (256, 256, 256)])
def test_mixed7_matches_oracle(ps, mixed, shape):
    print(f"mixed-dimension 7-point {shape}", flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(nx + ny + nz)
    a = rng.standard_normal(nx * ny * nz).astype(np.float32)
    b = rng.standard_normal(nx * ny * nz).astype(np.float32)
    cs = [rng.uniform(0.5, 1.5, m).astype(np.float32) for m in shape]
    g, g2 = ps.PSGridNew(shape), ps.PSGridNew(shape)
    cg = [ps.PSGridNew((m,)) for m in shape]
    try:
        ps.PSGridCopyin(g, a)
        ps.PSGridCopyin(g2, b)
        for h, c in zip(cg, cs):
            ps.PSGridCopyin(h, c)
        mixed.run(g, g2, *cg)
        got = ps.PSGridCopyout(g2)
        got_g = ps.PSGridCopyout(g)
    finally:
        for h in [g, g2] + cg:
            ps.PSGridFree(h)
    want = b.copy()
    lib().oracle_mixed7_sweep(a, want, nx, ny, nz, *cs)
    assert np.array_equal(got_g.view(np.uint32), a.view(np.uint32))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
