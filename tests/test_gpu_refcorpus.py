"""The reference's own system-test programs, unchanged, translated by
physisc-hip and run on the GPU (the executables `make -C oracle ref-hip`
builds from tests/system_tests/test_cases/test_*.c where the reference lies).

This is the reference's system-test procedure (tests/system_tests/
run_system_tests.sh.cmake:833-883) applied to the HIP target: a program must
exit with status 0 (the self-checking tests exit 1 on a wrong result), and
where the reference has a hand-written REFERENCE-target oracle
(test_cases/<name>.manual.ref.c) its stdout must equal that oracle's, byte
for byte (compared by sha256 against tests/golden/ref_outputs.json).  The
module test is checked against the oracle of its base name, as that script
does (`${1%%.*}`, :785-790).
"""
import hashlib
import json
import os
import subprocess

import pytest

from oracle_lib import GOLDEN

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE_DIR = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "hip")
GOLD = json.load(open(GOLDEN))["system_tests"]

# Every test program of the reference's system-test corpus (test_cplusplus is
# the one C++ source, test_cplusplus.cc).
CORPUS = [
    "test_01", "test_02", "test_03", "test_08", "test_09", "test_10",
    "test_15", "test_16", "test_27-pt", "test_27-pt-periodic",
    "test_27-pt-reduction", "test_3-pt-1d", "test_3-pt-periodic",
    "test_5-pt-2d", "test_5-pt-periodic", "test_7-pt", "test_7-pt.module",
    "test_7-pt-double-type", "test_7-pt-int-type",
    "test_7-pt-multi-iterations", "test_7-pt-neumann-cond",
    "test_7-pt-periodic", "test_7-pt-type-mix", "test_9-pt-2d",
    "test_9-pt-periodic-reduction", "test_9-pt-reduction",
    "test_asymmetric", "test_asymmetric-periodic", "test_mixed-dim",
    "test_mixed-dim2", "test_mixed-dim3", "test_multi-kernels",
    "test_param_name", "test_redblack", "test_redblack-periodic",
    "test_redblack-separated", "test_reduction-2d", "test_reduction-3d-int",
    "test_reduction-3d-long", "test_reduction-3d-max",
    "test_reduction-3d-min", "test_reduction-3d-prod",
    "test_reduction-3d-sum", "test_stencil-hole",
    "test_user-defined-type-7-pt", "test_user-defined-type-7-pt-periodic",
    "test_user-defined-type-7-pt-periodic-complex",
    "test_user-defined-type-array-member-copy",
    "test_user-defined-type-copyin-copyout",
    "test_user-defined-type-copyin-copyout-two-members",
    "test_user-defined-type-kernel-copy",
    "test_user-defined-type-multi-dim-member",
    "test_user-defined-type-multi-members",
    "test_user-defined-type-transpose", "test_user-defined-type1",
    "test_user-defined-type2", "test_user-defined-type3",
    "test_user-defined-type5", "test_cplusplus",
]


def _require(exe):
    """The corpus executables are part of the build (__graft_entry__.build()
    makes them wherever the reference sources are) and travel with the
    tree; a missing one is a failed build, not a skipped test."""
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: build with __graft_entry__.build() "
                    "(make -C oracle ref-hip) before the GPU tests")


# Programs physisc-hip does not translate (none at present): it stops with an
# error naming the feature.
UNSUPPORTED = set()


@pytest.mark.parametrize("name", CORPUS)
def test_reference_program_on_hip(name):
    if name in UNSUPPORTED:
        pytest.skip(f"{name}: not supported by physisc-hip yet")
    exe = os.path.join(EXE_DIR, name + ".exe")
    _require(exe)
    print(f"running {name}", flush=True)
    proc = subprocess.run([exe], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, timeout=90)
    assert proc.returncode == 0, (name, proc.returncode,
                                  proc.stderr.decode()[-2000:])
    gold = GOLD.get(name.split(".")[0])
    if gold is not None:
        lines = proc.stdout.count(b"\n")
        assert lines == gold["lines"], (name, lines, gold["lines"])
        assert hashlib.sha256(proc.stdout).hexdigest() == gold["sha256"], name


# The reference's example programs that test features (examples/test_*.c),
# built as ex_<name> by the same recipe; test_periodic_boundary.c is left
# out: it uses PS_GRID_CIRCULAR, which the reference's headers never define.
EXAMPLES = ["test_double_buffering", "test_global_variable", "test_set",
            "test_staggered_grid"]


def _staggered_expected():
    # examples/test_staggered_grid.c: g1 (2^3) = sum of the 8 corners of each
    # cell of g2 (3^3, values i), in the program's operand order
    import numpy as np
    n = 2
    g2 = np.arange(27, dtype=np.float32).reshape(3, 3, 3)   # [z][y][x]
    out = []
    for z in range(n):
        for y in range(n):
            for x in range(n):
                v = np.float32(0)
                terms = [g2[z, y, x], g2[z, y, x + 1], g2[z, y + 1, x],
                         g2[z + 1, y, x], g2[z, y + 1, x + 1],
                         g2[z + 1, y, x + 1], g2[z + 1, y + 1, x],
                         g2[z + 1, y + 1, x + 1]]
                v = terms[0]
                for t in terms[1:]:
                    v = np.float32(v + t)
                out.append("%f\n" % v)
    return "".join(out).encode()


@pytest.mark.parametrize("name", EXAMPLES)
def test_reference_example_on_hip(name):
    exe = os.path.join(EXE_DIR, "ex_" + name + ".exe")
    _require(exe)
    print(f"running example {name}", flush=True)
    proc = subprocess.run([exe], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, timeout=90)
    err = proc.stderr.decode()
    # these programs report mismatches on stderr ("Error: mismatch ...")
    assert proc.returncode == 0 and "Error" not in err, (name, err[-2000:])
    if name == "test_staggered_grid":
        assert proc.stdout == _staggered_expected()


# The reference's Physis diffusion benchmark (examples/diffusion-benchmark:
# diffusion3d_physis.c translated by physisc-hip, linked with the benchmark's
# own driver main.cc / baseline.cc / diffusion3d.cc): its --dump output and
# its accuracy against the closed-form solution must equal those of the
# reference's C++ baseline at the same size and count (the fixture's
# diffusion_baseline records, made from the baseline built in place).
@pytest.mark.parametrize("size,count", [(32, 4), (24, 10)])
def test_reference_diffusion_benchmark_on_hip(size, count, tmp_path):
    exe = os.path.join(EXE_DIR, "diffusion3d_physis.exe")
    _require(exe)
    print(f"running the diffusion benchmark {size}^3 x {count}", flush=True)
    proc = subprocess.run([exe, "--count", str(count), "--size", str(size),
                           "--dump"], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, timeout=90, cwd=tmp_path)
    assert proc.returncode == 0, proc.stderr.decode()[-2000:]
    gold = json.load(open(GOLDEN))["diffusion_baseline"][f"{size}x{count}"]
    dump = (tmp_path / "diffusion3d_result.physis.out").read_bytes()
    assert hashlib.sha256(dump).hexdigest() == gold["sha256"]
    acc = [l for l in proc.stdout.decode().splitlines()
           if l.startswith("Accuracy")]
    assert acc and acc[0].split(":")[1].strip() == gold["accuracy"]
