"""__PSReduceGridNoAlloc (the reduction whose partials live in a fixed
device array of the library) against the allocating __PSReduceGrid* entry
points, through the C ABI: bit-identical results for every element type and
operator, on two calls in a row, the second over a smaller grid."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# PSType codes (include/physis/types.h) and the allocating entry point
TYPES = {np.dtype(np.float32): (2, "__PSReduceGridFloat"),
         np.dtype(np.float64): (3, "__PSReduceGridDouble"),
         np.dtype(np.int32): (0, "__PSReduceGridInt"),
         np.dtype(np.int64): (1, "__PSReduceGridLong")}


def _no_alloc(ps):
    fn = ps.runtime().__PSReduceGridNoAlloc
    fn.argtypes = [C.c_void_p, C.c_int, ps.PSGridPtr, C.c_int]
    fn.restype = None
    return fn


def _data(n, dtype, seed):
    rng = np.random.default_rng(seed)
    if np.dtype(dtype).kind == "f":
        return (rng.standard_normal(n) * 3).astype(dtype)
    return rng.integers(-1000, 1000, n).astype(dtype)


def _both(ps, data, shape, op):
    """(allocating result, no-allocation result) over one grid."""
    ptype, name = TYPES[data.dtype]
    g = ps.PSGridNew(shape, data.dtype)
    ps.PSGridCopyin(g, data)
    want = np.zeros(1, data.dtype)
    getattr(ps.runtime(), name)(want.ctypes.data, op, g.handle)
    got = np.full(1, 7, data.dtype)
    _no_alloc(ps)(got.ctypes.data, op, g.handle, ptype)
    ps.PSGridFree(g)
    return want, got


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32,
                                   np.int64])
@pytest.mark.parametrize("op", ["PS_SUM", "PS_MAX", "PS_MIN", "PS_PROD"])
def test_no_alloc_reduce_matches_allocating_entry_point(ps, dtype, op):
    # a large grid (many stage-1 blocks) and then a smaller one, so that the
    # second call runs over partials the first call left in the array
    for shape, seed in (((160, 96, 90), 1), ((37, 5, 3), 2)):
        n = int(np.prod(shape))
        data = _data(n, dtype, seed)
        if op == "PS_PROD":
            # products of values near 1 stay finite (and exact for ints)
            data = (np.ones(n, dtype) if np.dtype(dtype).kind == "i"
                    else (1 + 1e-4 * data).astype(dtype))
            data[::97] = 2 if np.dtype(dtype).kind == "i" else data[::97]
        want, got = _both(ps, data, shape, getattr(ps, op))
        assert want.view(np.uint8).tobytes() == got.view(np.uint8).tobytes(), \
            (shape, op, want, got)


def test_no_alloc_reduce_sum_against_numpy(ps):
    # the float sum is the FP64 sum rounded once to float, whatever array
    # state an earlier call left behind
    for shape in ((128, 128, 64), (9, 1, 1)):
        data = np.arange(int(np.prod(shape)), dtype=np.float32) % 1000
        _, got = _both(ps, data, shape, ps.PS_SUM)
        assert got[0] == np.float32(data.astype(np.float64).sum())
