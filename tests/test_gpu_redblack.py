"""Red-black relaxation (physis_amd/programs/redblack7.c) on the GPU against
the oracle's REFERENCE-order in-place sweeps (oracle_redblack7_run), bit for
bit: ragged shapes whose tiles, z chunks and colours fall unevenly, and the
full 512^3 grid of the benchmark."""
import numpy as np
import pytest

from oracle_lib import lib
from physis_amd.configs import DIFFUSION_COEF, diffusion_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rb(ps):
    from physis_amd.programs import RedBlack7
    return RedBlack7()


# x extents that are multiples of 4 run one pass per iteration (odd counts
# end with a copy back), the others two colour sweeps
@pytest.mark.parametrize("shape,count", [((130, 66, 40), 3), ((36, 9, 7), 2),
                                         ((64, 64, 64), 1), ((33, 17, 20), 2),
                                         ((128, 96, 80), 3),
                                         ((512, 512, 512), 2),
                                         ((512, 512, 512), 3)])
def test_redblack_matches_oracle(ps, rb, shape, count):
    print(f"red-black {shape} x{count}", flush=True)
    a = diffusion_input(*shape, seed=9)
    g = ps.PSGridNew(shape)
    try:
        ps.PSGridCopyin(g, a)
        rb.run(g, count, DIFFUSION_COEF)
        got = ps.PSGridCopyout(g)
    finally:
        ps.PSGridFree(g)
    want = a.copy()
    lib().oracle_redblack7_run(want, *shape, count,
                               np.asarray(DIFFUSION_COEF, np.float32))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the faces keep their values
    w3 = want.reshape(shape[::-1])
    a3 = a.reshape(shape[::-1])
    assert np.array_equal(w3[0], a3[0]) and np.array_equal(w3[:, :, -1],
                                                           a3[:, :, -1])
