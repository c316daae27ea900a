"""Streaming-bandwidth ceilings of this GPU (libphysis_ceiling_hip): a
512 MiB copy (one read and one write stream, 1 GiB moved) and the sum of 13
arrays of 512x256x256 floats into a 14th (the 19-point Himeno sweep's
streams, 1.88 GB moved), each over a range of launch sizes, timed with HIP
events on the runtime stream.  Prints one JSON line with the best rate of
each and writes nothing else.

Usage: python -m physis_amd.ceilings [--reps R]
Under rocprofv3 --kernel-trace --stats its kernels are CopyKernel and
Sum13Kernel.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys

from . import runtime as ps

CEILING_SO = os.path.join(ps.LIB_DIR, "libphysis_ceiling_hip.so")


def _hip():
    L = C.CDLL("libamdhip64.so")
    for n in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize",
              "hipEventElapsedTime"):
        getattr(L, n).restype = C.c_int
    return L


def _time(hip, stream, fn, reps):
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0
    assert hip.hipEventCreate(C.byref(e1)) == 0
    fn()
    assert hip.hipEventRecord(e0, stream) == 0
    for _ in range(reps):
        fn()
    assert hip.hipEventRecord(e1, stream) == 0
    assert hip.hipEventSynchronize(e1) == 0
    t = C.c_float()
    assert hip.hipEventElapsedTime(C.byref(t), e0, e1) == 0
    return t.value / reps


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    if not os.path.exists(CEILING_SO):
        raise RuntimeError(f"{CEILING_SO} is missing: run "
                           "`python -m physis_amd.build`")
    ps.PSInit()
    L = C.CDLL(CEILING_SO)
    L.PSHipStreamCopy.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_int,
                                  C.c_void_p]
    L.PSHipStreamSum13.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_int,
                                   C.c_void_p]
    hip = _hip()
    stream = C.c_void_p(ps.hip_stream())
    cus = ps.num_cus()
    out = {"cus": cus}
    # copy: 512^3 floats
    n = 512 ** 3
    a, b = ps.PSGridNew((512, 512, 512)), ps.PSGridNew((512, 512, 512))
    pa, pb = a.handle.contents.p, b.handle.contents.p
    best = None
    for per_cu in (2, 4, 8, 16, 32):
        ms = _time(hip, stream, lambda: L.PSHipStreamCopy(
            pa, pb, n, cus * per_cu, stream), args.reps)
        gbs = 2 * 4 * n / (ms * 1e-3) / 1e9
        print(f"[ceilings] copy {cus * per_cu} blocks: {ms * 1e3:.1f} us "
              f"{gbs:.0f} GB/s", file=sys.stderr, flush=True)
        if best is None or gbs > best["GBps"]:
            best = {"GBps": round(gbs, 1), "us": round(ms * 1e3, 1),
                    "blocks": cus * per_cu, "bytes": 2 * 4 * n}
    out["copy_512MiB"] = best
    ps.PSGridFree(a)
    ps.PSGridFree(b)
    # 13 read streams + 1 write stream of 512x256x256 floats
    shape = (512, 256, 256)
    n = 512 * 256 * 256
    grids = [ps.PSGridNew(shape) for _ in range(14)]
    srcs = (C.c_void_p * 13)(*[g.handle.contents.p for g in grids[:13]])
    dst = grids[13].handle.contents.p
    best = None
    for per_cu in (2, 4, 8, 16, 32):
        ms = _time(hip, stream, lambda: L.PSHipStreamSum13(
            srcs, dst, n, cus * per_cu, stream), args.reps)
        gbs = 14 * 4 * n / (ms * 1e-3) / 1e9
        print(f"[ceilings] sum13 {cus * per_cu} blocks: {ms * 1e3:.1f} us "
              f"{gbs:.0f} GB/s", file=sys.stderr, flush=True)
        if best is None or gbs > best["GBps"]:
            best = {"GBps": round(gbs, 1), "us": round(ms * 1e3, 1),
                    "blocks": cus * per_cu, "bytes": 14 * 4 * n}
    out["sum13_512x256x256"] = best
    for g in grids:
        ps.PSGridFree(g)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
