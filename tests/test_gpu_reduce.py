"""reduce_hip (PSReduce) against the reference fixtures and the oracle."""
import numpy as np
import pytest

from oracle_lib import lib

pytestmark = pytest.mark.gpu


def reduce_of(ps, data, op, shape=None):
    shape = shape or (data.size,)
    g = ps.PSGridNew(shape, data.dtype)
    ps.PSGridCopyin(g, data)
    v = ps.PSReduce(op, g)
    ps.PSGridFree(g)
    return v


def test_reference_sum_fixture(ps):
    # tests/system_tests/test_cases/test_reduction-3d-sum.c: N=16, value i
    v = reduce_of(ps, np.arange(16 ** 3, dtype=np.float32), ps.PS_SUM,
                  (16, 16, 16))
    assert "%f" % v == "8386560.000000"


@pytest.mark.parametrize("op,ref", [("PS_MAX", 63.0), ("PS_MIN", 0.0)])
def test_reference_minmax_fixture(ps, op, ref):
    # test_reduction-3d-{max,min}.c: N=4, value i, checked against a loop
    v = reduce_of(ps, np.arange(64, dtype=np.float32), getattr(ps, op),
                  (4, 4, 4))
    assert v == ref


def test_reference_prod_fixture(ps):
    # test_reduction-3d-prod.c multiplies 1..N^3 style values; with small
    # exact integers the product is exact in any order
    data = np.array([1, 2, 3, 1, 2, 1, 1, 5] * 8, np.float32)
    assert reduce_of(ps, data, ps.PS_PROD, (4, 4, 4)) == np.float32(
        2.0 ** 16 * 3.0 ** 8 * 5.0 ** 8)


def test_int_wraps_like_reference(ps):
    # test_27-pt-reduction.c's second value overflows int32 in the REF loop
    data = np.full(1000, 2 ** 30, np.int32)
    want = lib().oracle_reduce_int(2, data, data.size)
    assert reduce_of(ps, data, ps.PS_SUM) == want
    assert want == np.int32((1000 * 2 ** 30) % 2 ** 32 - 2 ** 32
                            if (1000 * 2 ** 30) % 2 ** 32 >= 2 ** 31
                            else (1000 * 2 ** 30) % 2 ** 32)


@pytest.mark.parametrize("n", [1, 3, 4, 5, 1023, 4097, 1 << 20, 3_000_001])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32,
                                   np.int64])
def test_all_ops_types_sizes(ps, n, dtype):
    rng = np.random.default_rng(n)
    if np.dtype(dtype).kind == "f":
        data = (rng.standard_normal(n) * 3).astype(dtype)
    else:
        data = rng.integers(-1000, 1000, n).astype(dtype)
    assert reduce_of(ps, data, ps.PS_MAX) == data.max()
    assert reduce_of(ps, data, ps.PS_MIN) == data.min()
    s = reduce_of(ps, data, ps.PS_SUM)
    if dtype == np.float32:
        assert s == np.float32(data.astype(np.float64).sum(dtype=np.float64)) \
            or abs(s - data.astype(np.float64).sum()) <= \
            1e-12 * np.abs(data.astype(np.float64)).sum() + \
            np.spacing(np.float32(abs(s)))
    elif dtype == np.float64:
        # double folds in a fixed tree order: |err| <= (k+log2 n) u sum|x|
        bound = 64 * np.finfo(np.float64).eps * np.abs(data).sum()
        assert abs(s - data.sum()) <= bound
    else:
        assert s == data.sum(dtype=dtype)


def test_all_negative_max_uses_lowest_identity(ps):
    # the reference CUDA runtime seeds MAX with FLT_MIN (runtime/reduce.h:
    # 61-63); the REF target, and this runtime, return the true maximum
    data = -np.arange(1, 1000, dtype=np.float32)
    assert reduce_of(ps, data, ps.PS_MAX) == -1.0
