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
