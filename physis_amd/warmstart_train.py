"""Trains the warm-start U-Net for batches of 128^3 Poisson problems (face
coefficients with 5% seeded noise and a3 balancing them, right-hand sides of
4-24 Gaussians with sigma 2-16 cells) and writes its parameters.

    python -m physis_amd.warmstart_train [--ntrain 512] [--steps 6000]
        [--smooth W4,W16] [--out physis_amd/weights/unet128.npz]

Training data: ntrain problems (generator seeds 1000, 1064, ...; 64 per
batched solve) solved by the BF16 refinement to relative residual 1e-5.
Each solved problem also yields its seven reflections (PSUNetReflect: the
reflected right-hand side and solution solve the reflected problem, whose
coefficients are drawn from the same distribution), so there are
8 x ntrain pairs without more solves.  Adam over random windows of `batch`
problems of one random reflection, lr decaying linearly from 1e-3 to 1e-4,
on the mean square error plus the smooth-error terms weighted W4, W16
(PSUNetLossGrad).

Writes OUT (numpy arrays: params float32, out_scale, n) and OUT with the
suffix .json (sizes, loss and the cost of the solves and of the training).
The benchmark problems (seed 7) are not among the training problems.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import runtime as ps
from .configs import poisson_problem

N = 128
GEN = dict(noise=0.05, gaussians=(4, 24), sigma=(2.0, 16.0), consistent=True)
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                           "weights", "unet128.npz")


def _grid(a, shape, dtype=np.float32):
    g = ps.PSGridNew(shape, dtype)
    ps.PSGridCopyin(g, a)
    return g


def solve_training_set(ntrain: int, tol: float = 1e-5, chunk: int = 64):
    """(rhs, bnd, solution) host arrays of ntrain problems and the seconds
    the solves took."""
    from .poisson import PoissonSolver
    m = N ** 3
    rhs = np.empty(ntrain * m, np.float32)
    bnd = np.empty(ntrain * m, np.float32)
    sol = np.empty(ntrain * m, np.float32)
    secs = 0.0
    for c0 in range(0, ntrain, chunk):
        nb = min(chunk, ntrain - c0)
        coefs = poisson_problem(N, N, N, nbatch=nb, seed=1000 + c0, **GEN)
        shape = (N, N, N * nb)
        grids = [_grid(c, shape) for c in coefs]
        solver = PoissonSolver(grids, nb, "bf16")
        p = _grid(np.zeros(nb * m), shape, np.float64)
        st = solver.solve_mixed(p, tol=tol, inner=200)
        secs += st["seconds"]
        print(f"[train] solutions {c0}..{c0 + nb}: {st['sweeps']} sweeps, "
              f"{st['seconds']:.1f} s, rel {st['rel_residual']:.2e}",
              flush=True)
        rhs[c0 * m:(c0 + nb) * m] = coefs[11]
        bnd[c0 * m:(c0 + nb) * m] = coefs[10]
        sol[c0 * m:(c0 + nb) * m] = ps.PSGridCopyout(p)
        solver.free()
        for g in grids + [p]:
            ps.PSGridFree(g)
    return rhs, bnd, sol, secs


def training_grids(rhs, bnd, sol, ntrain: int):
    """Device grids of the training set: rg[r], tg[r] the right-hand sides
    and solutions under reflection r (0 = as solved), and bnd, which every
    reflection leaves unchanged."""
    from .warmstart import reflect
    shape = (N, N, N * ntrain)
    rg = [_grid(rhs, shape)]
    tg = [_grid(sol, shape)]
    bg = _grid(bnd, shape)
    for r in range(1, 8):
        a, b = ps.PSGridNew(shape), ps.PSGridNew(shape)
        reflect(rg[0], 0, ntrain, a, 0, (N, N, N), r)
        reflect(tg[0], 0, ntrain, b, 0, (N, N, N), r)
        rg.append(a)
        tg.append(b)
    return rg, tg, bg


def fit(rg, tg, bg, ntrain: int, scale: float, steps: int, batch: int,
        seed: int = 1, lr: float = 1e-3, lr_final: float = 1e-4,
        w4: float = 0.0, w16: float = 0.0):
    """Adam over random windows of `batch` problems of one random
    reflection; returns (params, loss history, seconds)."""
    from .warmstart import UNet
    net = UNet(N, N, N, batch, scale, seed=seed)
    net.set_loss_weights(w4, w16)
    rng = np.random.default_rng(seed)
    hist = []
    t1 = time.perf_counter()
    for s in range(steps):
        r = int(rng.integers(0, 8))
        first = int(rng.integers(0, ntrain - batch + 1))
        a = lr + (lr_final - lr) * s / max(steps - 1, 1)
        hist.append(net.train_step(rg[r], bg, tg[r], first, batch, a))
        if s % max(steps // 20, 1) == 0 or s == steps - 1:
            print(f"[train] step {s}: loss {hist[-1]:.4e} "
                  f"({time.perf_counter() - t1:.1f} s)", flush=True)
    secs = time.perf_counter() - t1
    params = net.get_params()
    net.free()
    return params, hist, secs


def train(ntrain: int, steps: int, batch: int, out: str, seed: int = 1,
          lr: float = 1e-3, lr_final: float = 1e-4, w4: float = 0.0,
          w16: float = 0.0) -> dict:
    t0 = time.perf_counter()
    rhs, bnd, sol, solve_s = solve_training_set(ntrain)
    data_s = time.perf_counter() - t0
    scale = float(np.max(np.abs(sol)))
    rg, tg, bg = training_grids(rhs, bnd, sol, ntrain)
    del rhs, sol, bnd
    params, hist, train_s = fit(rg, tg, bg, ntrain, scale, steps, batch,
                                seed, lr, lr_final, w4, w16)
    for g in rg + tg + [bg]:
        ps.PSGridFree(g)
    meta = {"problem": "128^3, face coefficients +-5%, a3 = 1/sum, "
                       "4-24 Gaussians sigma 2-16, BF16 refinement to 1e-5",
            "ntrain_solved": ntrain, "pairs": 8 * ntrain,
            "augmentation": "the 8 axis reflections of each solved problem",
            "steps": steps, "batch": batch, "lr": [lr, lr_final],
            "loss": {"w4": w4, "w16": w16},
            "params": int(params.size), "out_scale": scale,
            "solve_seconds": round(solve_s, 2),
            "data_seconds": round(data_s, 2),
            "train_seconds": round(train_s, 2),
            "loss_first": float(np.mean(hist[:20])),
            "loss_last": float(np.mean(hist[-50:]))}
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    np.savez(out, params=params, out_scale=np.float32(scale),
             n=np.int32(N))
    with open(os.path.splitext(out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta), flush=True)
    return meta


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ntrain", type=int, default=512)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--smooth", default="0,0",
                    help="W4,W16: weights of the smooth-error loss terms")
    a = ap.parse_args(argv)
    w4, w16 = (float(v) for v in a.smooth.split(","))
    ps.PSInit()
    train(a.ntrain, a.steps, a.batch, a.out, w4=w4, w16=w16)
    ps.PSFinalize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
