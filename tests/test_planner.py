"""Run-kernel launch planning on the host (no GPU): z chunks are capped so
that every raw buffer window stays within the 2 GiB reach of a 32-bit
buffer range, and grids whose planes cannot fit one window are rejected.

The shapes are those that a single-chunk plan would have windowed at 2-4 GiB
(4096x4096x64 and 8192x8192x16 FP32, 2048x2048x255 FP64); the reference's
32-bit PSIndex (include/physis/physis_common.h:33-55) admits all of them.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
REACH = 2 ** 31 - 1


@pytest.fixture(scope="module")
def planner(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("planner") / "planner_check")
    subprocess.run([HIPCC, "-std=c++17", "-O1", f"-I{ROOT}/include",
                    os.path.join(ROOT, "tests", "planner_check.cc"), "-o",
                    out], check=True, capture_output=True)
    return out


def _rows(text):
    return [tuple(int(v) for v in line.split())
            for line in text.strip().splitlines()]


def test_windows_stay_within_reach(planner):
    rows = _rows(subprocess.run([planner], check=True, capture_output=True,
                                text=True).stdout)
    assert len(rows) == 6
    for nx, ny, nz, elm, margin, tz, chunk, blocks, window in rows:
        assert window <= REACH, (nx, ny, nz, elm, window)
        assert tz * chunk >= nz and (tz - 1) * chunk < nz
        assert chunk >= 1
        # the window formula the run kernels use: chunk + margin per side
        assert window == (chunk + 2 * margin) * nx * ny * elm
    big = {(r[0], r[1], r[2], r[3], r[4]): r for r in rows}
    # these shapes need several chunks only because of the reach
    assert big[(8192, 8192, 16, 4, 2)][6] <= 4
    assert big[(4096, 4096, 64, 4, 2)][6] <= 28
    assert big[(2048, 2048, 255, 8, 2)][6] <= 60
    # ordinary shapes are planned for occupancy alone
    assert big[(512, 512, 512, 4, 2)][5] >= 1


def test_plane_beyond_reach_is_rejected(planner):
    # a 1 GiB plane with one halo plane per side needs a 3 GiB window
    r = subprocess.run([planner, "16384", "16384", "4", "4", "1", "256", "14",
                        "256"], capture_output=True, text=True)
    assert r.returncode != 0
    assert "2 GiB reach of a buffer window" in r.stderr
