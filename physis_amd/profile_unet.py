"""Workload for rocprofv3 measurements of the U-Net warm-start kernels.

Run under the profiler, e.g.
    rocprofv3 --kernel-trace --stats -d OUT -o unet --output-format csv -- \\
        python3 -m physis_amd.profile_unet
It builds the network for 128^3 problems with batch 8 (the warm-start
benchmark's shape) and issues 4 predictions of 8 problems (the first one
warms up), then 3 training steps of batch 8.  Then
    python3 -m physis_amd.profile_summary --unet OUT profiles/OUT.md CMD
writes each convolution kernel's time against its roofline.
"""
from __future__ import annotations

import numpy as np

from . import runtime as ps
from .configs import poisson_problem
from .warmstart import UNet

N = 128
BATCH = 8


def main() -> None:
    ps.PSInit()
    gen = dict(noise=0.05, gaussians=(4, 24), sigma=(2.0, 16.0),
               consistent=True)
    coefs = poisson_problem(N, N, N, nbatch=BATCH, seed=7, **gen)
    shape = (N, N, N * BATCH)
    rhs, bnd = ps.PSGridNew(shape), ps.PSGridNew(shape)
    ps.PSGridCopyin(rhs, coefs[11])
    ps.PSGridCopyin(bnd, coefs[10])
    target = ps.PSGridNew(shape)
    ps.PSGridCopyin(target, coefs[11] * 0.5)
    del coefs
    p = ps.PSGridNew(shape, np.float64)
    net = UNet(N, N, N, BATCH, 1.0, seed=1)
    for _ in range(4):
        net.predict(rhs, bnd, 0, BATCH, p)
    for _ in range(3):
        net.train_step(rhs, bnd, target, 0, BATCH, 1e-3)
    ps.PSGridCopyout(p)  # synchronises
    net.free()
    for g in (rhs, bnd, target, p):
        ps.PSGridFree(g)
    print("profile_unet: done", flush=True)


if __name__ == "__main__":
    main()
