"""The reference's system tests, translated by physisc-hip and run on the
GPU: each dump must hash to the reference oracle's output
(tests/golden/ref_outputs.json, the same check tests/system_tests/
run_system_tests.sh.cmake:833-883 performs with diff)."""
import ctypes as C
import json

import numpy as np
import pytest

from oracle_lib import GOLDEN, dump_f, sha256

pytestmark = pytest.mark.gpu
GOLD = json.load(open(GOLDEN))["system_tests"]

FLOAT_TESTS = {
    "test_7-pt": ("systest_7pt", 32 ** 3),
    "test_7-pt-periodic": ("systest_7pt_periodic", 32 ** 3),
    "test_7-pt-neumann-cond": ("systest_7pt_neumann", 8 ** 3),
    "test_7-pt-multi-iterations": ("systest_7pt_multi_iterations", 32 ** 3),
    "test_asymmetric": ("systest_asymmetric", 32 ** 3),
    "test_asymmetric-periodic": ("systest_asymmetric_periodic", 32 ** 3),
    "test_27-pt": ("systest_27pt", 32 ** 3),
    "test_27-pt-periodic": ("systest_27pt_periodic", 32 ** 3),
    "test_redblack": ("systest_redblack", 32 ** 3),
    # red then black as two maps of one run: the same sweep as test_redblack
    "test_redblack#separated": ("systest_redblack_separated", 32 ** 3),
    "test_redblack-periodic": ("systest_redblack_periodic", 32 ** 3),
    "test_3-pt-1d": ("systest_3pt_1d", 1024),
    "test_5-pt-2d": ("systest_5pt_2d", 32 ** 2),
    "test_mixed-dim3": ("systest_mixed_dim3", 32 * 34 * 32),
    "test_stencil-hole": ("systest_stencil_hole", 32 ** 3),
}


@pytest.fixture(scope="module")
def prog(ps):
    return ps.load_program("systests")


@pytest.mark.parametrize("name", sorted(FLOAT_TESTS))
def test_float_dump_matches_reference(prog, name):
    fn, n = FLOAT_TESTS[name]
    out = np.zeros(n, np.float32)
    getattr(prog, fn)(out.ctypes.data_as(C.c_void_p))
    assert sha256(dump_f(out)) == GOLD[name.split("#")[0]]["sha256"], name


def test_27pt_reduction_matches_reference(prog):
    v1, v2 = C.c_int(), C.c_int()
    prog.systest_27pt_reduction(C.byref(v1), C.byref(v2))
    assert [str(v1.value), str(v2.value)] == GOLD["test_27-pt-reduction"]["values"]


def test_reduction_3d_sum_matches_reference(prog):
    v = C.c_float()
    prog.systest_reduction_3d_sum(C.byref(v))
    assert sha256(("%f\n" % v.value).encode()) == \
        GOLD["test_reduction-3d-sum"]["sha256"]


def _dump_rows(a):
    """The reference dump of struct points: the members of each point as
    "%f" separated by spaces, one point per line."""
    return "".join(" ".join("%f" % v for v in row) + "\n"
                   for row in a.tolist()).encode()


def test_user_type_three_members(prog):
    # test_user-defined-type1: dump "%f %f %f" of x, y, z
    out = np.zeros((32 ** 3, 3), np.float32)
    prog.systest_udt_vec3(out.ctypes.data_as(C.c_void_p))
    assert sha256(_dump_rows(out)) == GOLD["test_user-defined-type1"]["sha256"]


def test_user_type_array_member(prog):
    # test_user-defined-type-multi-dim-member: dump of p[1][0]
    out = np.zeros((32 ** 3, 3, 2), np.float32)
    prog.systest_udt_tensor(out.ctypes.data_as(C.c_void_p))
    assert sha256(dump_f(np.ascontiguousarray(out[:, 1, 0]))) == \
        GOLD["test_user-defined-type-multi-dim-member"]["sha256"]
    # the other components keep their inputs
    i = np.arange(32 ** 3, dtype=np.float64)
    for j, k in ((0, 0), (0, 1), (1, 1), (2, 0), (2, 1)):
        assert np.array_equal(out[:, j, k], (i + j + k).astype(np.float32))


def test_user_type_whole_point_periodic(prog):
    # test_user-defined-type-7-pt-periodic-complex: dump "%f %f" of r, i
    out = np.zeros((32 ** 3, 2), np.float32)
    prog.systest_udt_complex(out.ctypes.data_as(C.c_void_p))
    assert sha256(_dump_rows(out)) == \
        GOLD["test_user-defined-type-7-pt-periodic-complex"]["sha256"]


def test_host_set_and_get(prog):
    # PSGridSet of i at (x, y, z) in x-fastest order, a map doubling it,
    # PSGridGet of single elements in host code
    prog.systest_host_get.restype = C.c_float
    prog.systest_host_get.argtypes = [C.c_int] * 3
    for x, y, z in ((0, 0, 0), (3, 2, 1), (1, 2, 0)):
        assert prog.systest_host_get(x, y, z) == 2.0 * (x + 4 * y + 12 * z)


def test_shadowed_index_reads_the_shifted_point(prog):
    # kshadow reads g at (x + 1, y, z) through a block-scope x = x + 2; with
    # the REF linear rule that is element i + 1, and 0 past the last one
    nx, ny, nz = 40, 6, 5
    a = np.arange(1, nx * ny * nz + 1, dtype=np.float32)
    out = np.zeros_like(a)
    prog.systest_shadowed_index(a.ctypes.data_as(C.c_void_p),
                                out.ctypes.data_as(C.c_void_p), nx, ny, nz)
    expect = np.zeros_like(a)
    expect[:-1] = a[1:]
    assert np.array_equal(out, expect)


def test_seven_reducing_maps_keep_their_own_partials(prog):
    sums = np.zeros(7, np.float32)
    prog.systest_seven_reductions(sums.ctypes.data_as(C.c_void_p))
    g = (np.arange(16 ** 3) % 7).astype(np.float64)
    full, half = g.sum(), g[:16 * 16 * 8].sum()
    expect = [(i + 1) * (full if i % 2 == 0 else half) for i in range(7)]
    assert sums.tolist() == expect
