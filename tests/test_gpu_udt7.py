"""A 7-point stencil over a grid of user-defined points
(physis_amd/programs/udt7.c) on the GPU against oracle_udt7_sweep, bit for
bit, at ragged shapes and at 512^3 (the member p is read, q written; p must
come back unchanged)."""
import numpy as np
import pytest

from oracle_lib import lib
from physis_amd.configs import DIFFUSION_COEF, diffusion_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def udt(ps):
    from physis_amd.programs import Udt7
    return Udt7()


@pytest.mark.parametrize("shape", [(37, 13, 9), (64, 64, 64), (512, 512, 512)])
def test_udt7_matches_oracle(ps, udt, shape):
    print(f"user-type 7-point {shape}", flush=True)
    n = int(np.prod(shape))
    pq = np.empty((n, 2), np.float32)
    pq[:, 0] = diffusion_input(*shape, seed=4)
    pq[:, 1] = diffusion_input(*shape, seed=5)
    g = udt.new(shape, pq)
    try:
        udt.run(g, 2, DIFFUSION_COEF)
        got = udt.copyout(g, n)
    finally:
        udt.free(g)
    want = pq[:, 1].copy()
    lib().oracle_udt7_sweep(np.ascontiguousarray(pq[:, 0]), want, *shape,
                            np.asarray(DIFFUSION_COEF, np.float32))
    assert np.array_equal(got[:, 0].view(np.uint32), pq[:, 0].view(np.uint32))
    assert np.array_equal(got[:, 1].view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("shape", [(37, 13, 9), (512, 512, 512)])
def test_udt7_whole_points_match_oracle(ps, udt, shape):
    """Struct-valued gets and emit (read and written member by member on
    the 2.5D kernel): g2's interior gets (p, 7-point(p)), its other points
    keep their values, g is unchanged."""
    print(f"user-type 7-point, whole points {shape}", flush=True)
    n = int(np.prod(shape))
    pq = np.empty((n, 2), np.float32)
    pq[:, 0] = diffusion_input(*shape, seed=6)
    pq[:, 1] = diffusion_input(*shape, seed=7)
    pq2 = np.empty((n, 2), np.float32)
    pq2[:, 0] = diffusion_input(*shape, seed=8)
    pq2[:, 1] = diffusion_input(*shape, seed=9)
    g, g2 = udt.new(shape, pq), udt.new(shape, pq2)
    try:
        udt.whole_run(g, g2, DIFFUSION_COEF)
        got, got2 = udt.copyout(g, n), udt.copyout(g2, n)
    finally:
        udt.free(g)
        udt.free(g2)
    want_q = pq2[:, 1].copy()
    lib().oracle_udt7_sweep(np.ascontiguousarray(pq[:, 0]), want_q, *shape,
                            np.asarray(DIFFUSION_COEF, np.float32))
    nx, ny, nz = shape
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    want_p = np.where(inner, pq[:, 0], pq2[:, 0])
    assert np.array_equal(got.view(np.uint32), pq.view(np.uint32))
    assert np.array_equal(got2[:, 0].view(np.uint32), want_p.view(np.uint32))
    assert np.array_equal(got2[:, 1].view(np.uint32), want_q.view(np.uint32))
