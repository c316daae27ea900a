"""Time-to-residual benchmarks of the Poisson extensions on one GPU.

    python -m physis_amd.solve_bench precision [--n 512] [--sweeps 20000]
    python -m physis_amd.solve_bench precision-full [--n 256] [--runs bf16,fp16]
    python -m physis_amd.solve_bench warmstart [--ntrain 256] [--steps 1500]

precision       the 19-point operator at n^3 (coefficients with 1% noise,
                right-hand side = 16 Gaussians, sigma 4..32 cells scaled by
                n/512): a fixed number of sweeps of plain FP32 Jacobi and of
                the BF16 / FP16 refinement, reporting the per-sweep time and
                the measured convergence rate, from which the time to
                relative residual 1e-6 is extrapolated (a full solve at
                512^3 takes ~10^6 sweeps);
precision-full  the same problem solved to 1e-6 at a size where that is
                affordable (FP64-verified residual, sweeps, seconds);
warmstart       64 problems of 128^3 (face coefficients with 5% noise, a
                consistent diagonal, 4..24 Gaussians, sigma 2..16): train
                the U-Net on `ntrain` other
                problems solved by the Jacobi solve, then solve the 64 to
                1e-6 from zero (plain FP32 and BF16 refinement) and from the
                network's prediction.

Each prints one JSON line per result.  Inputs are synthetic and seeded.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import runtime as ps
from .configs import poisson_problem
from .poisson import PoissonSolver


def _emit(d: dict) -> None:
    print(json.dumps(d), flush=True)


def _grids(arrays, shape, dtype=np.float32):
    out = []
    for a in arrays:
        g = ps.PSGridNew(shape, dtype)
        ps.PSGridCopyin(g, a)
        out.append(g)
    return out


def _free(grids):
    for g in grids:
        ps.PSGridFree(g)


def _rate(r0: float, r1: float, sweeps: int) -> float:
    """Residual reduction per sweep (log10) between two residuals."""
    return (math.log10(r1) - math.log10(r0)) / sweeps


def precision(n: int, sweeps: int, tol: float = 1e-6, inner: int = 200) -> None:
    s = n / 512.0
    coefs = poisson_problem(n, n, n, nbatch=1, noise=0.01, gaussians=(16, 16),
                            sigma=(4.0 * s, 32.0 * s), seed=0)
    shape = (n, n, n)
    grids = _grids(coefs, shape)
    del coefs
    # plain FP32: residual at start, after sweeps/2 and after sweeps
    solver = PoissonSolver(grids, 1, "fp32")
    p = _grids([np.zeros(n ** 3, np.float32)], shape)[0]
    st = solver.solve_plain(p, tol=tol, check=sweeps // 2, max_sweeps=sweeps // 2)
    r_half, t_half = st["rel_residual"], st["seconds"]
    st2 = solver.solve_plain(p, tol=tol, check=sweeps // 2,
                             max_sweeps=sweeps // 2)
    per = st2["seconds"] / st2["sweeps"]
    rate = _rate(r_half, st2["rel_residual"], st2["sweeps"])
    need = (math.log10(tol) - math.log10(st2["rel_residual"])) / rate \
        if rate < 0 else float("inf")
    _emit({"bench": "precision", "n": n, "storage": "plain_fp32",
           "sweeps": st["sweeps"] + st2["sweeps"],
           "rel_residual": st2["rel_residual"], "ms_per_sweep": per * 1e3,
           "log10_rate_per_sweep": rate,
           "extrapolated_sweeps_to_tol": st["sweeps"] + st2["sweeps"] + need,
           "extrapolated_seconds_to_tol":
               t_half + st2["seconds"] + need * per})
    solver.free()
    _free([p])
    for storage in ("bf16", "fp16"):
        solver = PoissonSolver(grids, 1, storage)
        p = _grids([np.zeros(n ** 3)], shape, np.float64)[0]
        st = solver.solve_mixed(p, tol=tol, inner=inner, max_sweeps=sweeps // 2)
        st2 = solver.solve_mixed(p, tol=tol, inner=inner,
                                 max_sweeps=sweeps // 2)
        per = st2["seconds"] / st2["sweeps"]
        rate = _rate(st["rel_residual"], st2["rel_residual"], st2["sweeps"])
        need = (math.log10(tol) - math.log10(st2["rel_residual"])) / rate \
            if rate < 0 else float("inf")
        _emit({"bench": "precision", "n": n, "storage": f"mixed_{storage}",
               "sweeps": st["sweeps"] + st2["sweeps"],
               "rel_residual": st2["rel_residual"], "ms_per_sweep": per * 1e3,
               "log10_rate_per_sweep": rate,
               "extrapolated_sweeps_to_tol":
                   st["sweeps"] + st2["sweeps"] + need,
               "extrapolated_seconds_to_tol":
                   st["seconds"] + st2["seconds"] + need * per})
        solver.free()
        _free([p])
    _free(grids)


def precision_full(n: int, tol: float = 1e-6,
                   runs=("plain_fp32", "bf16", "fp16"),
                   max_sweeps: int = 10 ** 7, inner: int = 200,
                   consistent: bool = False) -> None:
    s = n / 512.0
    coefs = poisson_problem(n, n, n, nbatch=1, noise=0.01, gaussians=(16, 16),
                            sigma=(4.0 * s, 32.0 * s), seed=0,
                            consistent=consistent)
    shape = (n, n, n)
    grids = _grids(coefs, shape)
    del coefs
    base = None
    if "plain_fp32" in runs:
        solver = PoissonSolver(grids, 1, "fp32")
        p = _grids([np.zeros(n ** 3, np.float32)], shape)[0]
        st = solver.solve_plain(p, tol=tol, check=500, progress=30,
                                max_sweeps=max_sweeps)
        _emit({"bench": "precision-full", "n": n, "storage": "plain_fp32",
               "consistent": consistent, **st})
        base = st["seconds"] if st["converged"] else None
        solver.free()
        _free([p])
    for storage in ("bf16", "fp16"):
        if storage not in runs:
            continue
        solver = PoissonSolver(grids, 1, storage)
        p = _grids([np.zeros(n ** 3)], shape, np.float64)[0]
        st = solver.solve_mixed(p, tol=tol, inner=inner, progress=30,
                                max_sweeps=max_sweeps)
        st["inner_sweeps"] = inner
        if base:
            st["speedup_vs_plain_fp32"] = base / st["seconds"]
        _emit({"bench": "precision-full", "n": n,
               "storage": f"mixed_{storage}", "consistent": consistent, **st})
        solver.free()
        _free([p])
    _free(grids)


def warmstart(ntrain: int, steps: int, batch: int, lr: float,
              tol: float = 1e-6, n: int = 128, ntest: int = 64) -> None:
    from .warmstart import UNet
    m = n ** 3
    gen = dict(noise=0.05, gaussians=(4, 24), sigma=(2.0, 16.0),
               consistent=True)
    # training data: problems solved by the Jacobi solve (BF16 refinement),
    # 64 at a time
    chunk = 64
    t0 = time.perf_counter()
    rhs_all = np.empty(ntrain * m, np.float32)
    bnd_all = np.empty(ntrain * m, np.float32)
    sol_all = np.empty(ntrain * m, np.float32)
    solve_s = 0.0
    for c0 in range(0, ntrain, chunk):
        nb = min(chunk, ntrain - c0)
        coefs = poisson_problem(n, n, n, nbatch=nb, seed=1000 + c0, **gen)
        shape = (n, n, n * nb)
        grids = _grids(coefs, shape)
        solver = PoissonSolver(grids, nb, "bf16")
        p = _grids([np.zeros(nb * m)], shape, np.float64)[0]
        st = solver.solve_mixed(p, tol=1e-5, inner=200, progress=30)
        solve_s += st["seconds"]
        print(f"  training solutions {c0}..{c0 + nb}: {st}", flush=True)
        rhs_all[c0 * m:(c0 + nb) * m] = coefs[11]
        bnd_all[c0 * m:(c0 + nb) * m] = coefs[10]
        sol_all[c0 * m:(c0 + nb) * m] = ps.PSGridCopyout(p)
        solver.free()
        _free(grids + [p])
    gen_s = time.perf_counter() - t0
    scale = float(np.max(np.abs(sol_all)))
    tshape = (n, n, n * ntrain)
    rg, bg, tg = _grids([rhs_all, bnd_all, sol_all], tshape)
    del rhs_all, sol_all, bnd_all
    net = UNet(n, n, n, batch, scale, seed=1)
    t0 = time.perf_counter()
    hist = net.train(rg, bg, tg, ntrain, steps=steps, batch=batch, lr=lr,
                     log=max(steps // 10, 1))
    ps_sync = time.perf_counter() - t0
    _emit({"bench": "warmstart-train", "ntrain": ntrain, "steps": steps,
           "batch": batch, "params": net.num_params,
           "data_seconds": gen_s, "data_solve_seconds": solve_s,
           "train_seconds": ps_sync, "loss_first": hist[0],
           "loss_last": float(np.mean(hist[-20:])), "out_scale": scale})
    _free([rg, bg, tg])
    # the 64 benchmark problems
    coefs = poisson_problem(n, n, n, nbatch=ntest, seed=7, **gen)
    shape = (n, n, n * ntest)
    grids = _grids(coefs, shape)
    res = {}
    H = 8192
    solver = PoissonSolver(grids, ntest, "fp32")
    p = _grids([np.zeros(ntest * m, np.float32)], shape)[0]
    res["plain_fp32_zero"] = solver.solve_plain(p, tol=tol, check=500,
                                                history=H, progress=30)
    solver.free()
    _free([p])
    solver = PoissonSolver(grids, ntest, "bf16")
    p = _grids([np.zeros(ntest * m)], shape, np.float64)[0]
    res["mixed_bf16_zero"] = solver.solve_mixed(p, tol=tol, inner=200,
                                                history=H, progress=30)
    # warm start: prediction (timed) then the same solve
    t0 = time.perf_counter()
    for f in range(0, ntest, batch):
        net.predict(grids[11], grids[10], f, min(batch, ntest - f), p)
    t_pred = time.perf_counter() - t0
    st = solver.solve_mixed(p, tol=tol, inner=200, history=H, progress=30)
    st["predict_seconds"] = t_pred
    res["mixed_bf16_warm"] = st
    solver.free()
    # equal-accuracy comparisons: at 1e-6, and at the residual plain FP32
    # from zero reached (it may stop above 1e-6)
    from .poisson import time_to
    base = res["plain_fp32_zero"]
    levels = sorted({tol, max(base["rel_residual"], tol)})
    for k, v in res.items():
        hist = v.pop("history")
        out = {"bench": "warmstart", "problems": ntest, "n": n, "run": k,
               **v, "initial_rel_residual": hist[0][1] if hist else None}
        for lv in levels:
            t = time_to(hist, lv)
            if t is not None:
                pre = v.get("predict_seconds", 0.0)
                out[f"to_{lv:.3g}"] = {"sweeps": t[0], "seconds": t[1] + pre}
        _emit(out)
    net.free()
    _free(grids + [p])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["precision", "precision-full",
                                     "warmstart"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--sweeps", type=int, default=20000)
    ap.add_argument("--ntrain", type=int, default=256)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-sweeps", type=int, default=10 ** 7)
    ap.add_argument("--inner", type=int, default=200,
                    help="inner sweeps per refinement step")
    ap.add_argument("--consistent", action="store_true",
                    help="precision-full: a3 = 1 / sum of the face "
                         "coefficients (b = 0) instead of independent noise")
    ap.add_argument("--runs", default="plain_fp32,bf16,fp16",
                    help="precision-full: runs to make (comma-separated)")
    a = ap.parse_args(argv)
    ps.PSInit()
    if a.what == "precision":
        precision(a.n or 512, a.sweeps, inner=a.inner)
    elif a.what == "precision-full":
        precision_full(a.n or 256, runs=tuple(a.runs.split(",")),
                       max_sweeps=a.max_sweeps, inner=a.inner,
                       consistent=a.consistent)
    else:
        warmstart(a.ntrain, a.steps, a.batch, a.lr)
    ps.PSFinalize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
