"""Workload for rocprofv3 measurements of the hot-path kernels.

Run under the profiler, e.g.
    rocprofv3 --pmc FETCH_SIZE -d OUT -o fetch --output-format csv -- \\
        python3 -m physis_amd.profile_kernels
It issues, in order: a device-to-device copy of one 512 MiB grid (the
runtime's plain copy kernel, whose FETCH_SIZE / WRITE_SIZE calibrate the
counters' units), at 512^3 24 two-sweep launches of the 7-point clamped
kernel, five single sweeps of it, 10 periodic sweeps (five two-sweep
launches), 3 PSReduce SUMs at 512^3, 10 Himeno sweeps at 512x256x256 (five
two-sweep launches) and three pairs of single Himeno sweeps, and the
precision tier at 512^3: 10 BF16 inner sweeps (one pair of single sweeps,
four two-sweep launches) and 3 fused FP64 update + residual passes.
Then `python3 -m physis_amd.profile_summary OUT_fetch OUT_write` writes the
per-kernel HBM traffic summary.

`python3 -m physis_amd.profile_kernels diffusion` runs only the copy and the
7-point launches (the workload of the instruction-counter profiles).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import runtime as ps
from .configs import DIFFUSION_COEF, diffusion_input, himeno_inputs
from .programs import Diffusion7, Himeno

COPY_BYTES = 512 * 2 ** 20


def main(which: str = "all") -> None:
    ps.PSInit()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyDtoD.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    hip.hipDeviceSynchronize.restype = C.c_int
    shape = (512, 512, 512)
    a = diffusion_input(*shape, seed=0)
    g1, g2 = ps.PSGridNew(shape), ps.PSGridNew(shape)
    ps.PSGridCopyin(g1, a)
    for _ in range(3):
        assert hip.hipMemcpyDtoD(g2.handle.contents.p, g1.handle.contents.p,
                                 COPY_BYTES) == 0
    assert hip.hipDeviceSynchronize() == 0
    d7 = Diffusion7()
    # 24 two-sweep launches, the last of which also writes u' (as one in 25
    # of the benchmark's)
    d7.run(g1, g2, 48, DIFFUSION_COEF)
    for _ in range(5):                       # the one-sweep kernel
        d7.sweep_domain(g1, g2, (0, 512, 0, 512, 0, 512), DIFFUSION_COEF)
    d7.run(g1, g2, 10, DIFFUSION_COEF, periodic=True)
    for _ in range(3):
        ps.PSReduce(ps.PS_SUM, g1)
    ps.PSGridFree(g1)
    ps.PSGridFree(g2)
    if which == "diffusion":
        ps.PSFinalize()
        print("profile workload done (diffusion)")
        return
    hs = (512, 256, 256)
    coefs, p = himeno_inputs(*hs, seed=0)
    grids = []
    for c in coefs + [p, p]:
        g = ps.PSGridNew(hs)
        ps.PSGridCopyin(g, c)
        grids.append(g)
    him = Himeno()
    him.run(10, grids[12], grids[13], grids[:12])   # five two-sweep passes
    for _ in range(3):                              # single-sweep pairs
        him.run(2, grids[12], grids[13], grids[:12])
    for g in grids:
        ps.PSGridFree(g)
    # precision tier at 512^3: 10 BF16 inner sweeps, 3 fused FP64
    # update + residual passes
    from .configs import poisson_problem
    from .poisson import PoissonSolver
    coefs = poisson_problem(*shape, nbatch=1, gaussians=(16, 16),
                            sigma=(4.0, 32.0), seed=0)
    grids = []
    for c in coefs:
        g = ps.PSGridNew(shape)
        ps.PSGridCopyin(g, c)
        grids.append(g)
    del coefs
    solver = PoissonSolver(grids, 1, "bf16")
    solver.inner_sweeps(10)
    p = ps.PSGridNew(shape, np.float64)
    q = ps.PSGridNew(shape, np.float64)
    for _ in range(3):
        solver.residual(p, q, 0.5, 1.0)
    solver.free()
    for g in grids + [p, q]:
        ps.PSGridFree(g)
    ps.PSFinalize()
    print("profile workload done", float(np.float32(0)))


if __name__ == "__main__":
    import sys
    main(sys.argv[1] if len(sys.argv) > 1 else "all")
