"""libphysis_rt_hip grid management through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_new_is_zero_filled_and_round_trips(ps):
    g = ps.PSGridNew((17, 5, 3))
    assert np.all(ps.PSGridCopyout(g) == 0)
    a = np.arange(17 * 5 * 3, dtype=np.float32)
    ps.PSGridCopyin(g, a)
    assert np.array_equal(ps.PSGridCopyout(g), a)
    ps.PSGridFree(g)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
def test_dtypes_and_dims(ps, dtype):
    for dims in ([10], [6, 7], [5, 4, 3]):
        g = ps.PSGridNew(dims, dtype)
        a = np.arange(int(np.prod(dims))).astype(dtype)
        ps.PSGridCopyin(g, a)
        assert np.array_equal(ps.PSGridCopyout(g), a)
        ps.PSGridFree(g)


@pytest.mark.parametrize("off,size", [((0, 0, 0), (16, 12, 10)),
                                      ((3, 2, 1), (5, 4, 3)),
                                      ((0, 0, 4), (16, 12, 2)),
                                      ((0, 5, 0), (16, 1, 10)),
                                      ((15, 11, 9), (1, 1, 1))])
def test_subgrid_copy(ps, off, size):
    shape = (16, 12, 10)
    a = np.arange(np.prod(shape), dtype=np.float32)
    g = ps.PSGridNew(shape)
    ps.PSGridCopyin(g, a)
    ref = a.reshape(shape[::-1])[off[2]:off[2] + size[2],
                                 off[1]:off[1] + size[1],
                                 off[0]:off[0] + size[0]]
    got = ps.PSGridCopyoutSubgrid(g, off, size)
    assert np.array_equal(got, ref.reshape(-1))
    ps.PSGridCopyinSubgrid(g, -got, off, size)
    full = ps.PSGridCopyout(g).reshape(shape[::-1])
    exp = a.reshape(shape[::-1]).copy()
    exp[off[2]:off[2] + size[2], off[1]:off[1] + size[1],
        off[0]:off[0] + size[0]] *= -1
    assert np.array_equal(full, exp)
    ps.PSGridFree(g)


def test_oversized_allocation_returns_invalid_grid(ps):
    # 2^40 floats cannot be allocated: __PSGridNew returns INVALID_GRID
    with pytest.raises(MemoryError):
        ps.PSGridNew((1 << 14, 1 << 13, 1 << 13))
    g = ps.PSGridNew((8, 8, 8))   # the runtime is still usable
    ps.PSGridFree(g)
