"""The learned warm start at its own workload: a batch of 64 Poisson
problems of 128^3 (4-24 seeded Gaussians with sigma 2-16 cells on the
right-hand side, +-5% noise on the face coefficients, a3 balancing them),
the 8/16/32-channel U-Net, training batches of 8 problems.

Full-size checks:
  * the prediction over all 64 problems is finite and exactly 0 where
    bnd = 0 (Dirichlet faces), and the prediction for one whole 128^3
    problem equals the oracle's (oracle/unet_ref.py, PyTorch CPU FP32 with
    BF16 rounding at the same points) within 1e-2 relative L2;
  * training steps on the GPU lower the loss against solutions of the
    problems themselves (the BF16 refinement run to 1e-3);
  * from the trained network's prediction, the refinement solves unseen
    problems to relative residual 1e-6, verified by the oracle's FP64
    residual of the returned solutions.
  * the committed network (trained on 512 other problems and their
    reflections) saves at least 15% of the sweeps to 1e-6 on 8 of them;
    a network trained for a few seconds here is not expected to save any.
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from physis_amd.configs import poisson_problem

torch = pytest.importorskip("torch")
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu

N = 128
NB = 64
NTEST = 8
SHAPE = (N, N, N * NB)


@pytest.fixture(scope="module")
def batch(ps):
    print("generating 64 problems of 128^3", flush=True)
    coefs = poisson_problem(N, N, N, nbatch=NB, noise=0.05,
                            gaussians=(4, 24), sigma=(2.0, 16.0), seed=5,
                            consistent=True)
    grids = []
    for c in coefs:
        g = ps.PSGridNew(SHAPE, np.float32)
        ps.PSGridCopyin(g, c)
        grids.append(g)
    yield coefs, grids
    for g in grids:
        ps.PSGridFree(g)


@pytest.fixture(scope="module")
def trained(ps, batch):
    """A network trained for 60 steps on the first 56 problems against
    their BF16 refinement solutions to 1e-3; (net, loss history)."""
    from physis_amd.poisson import PoissonSolver
    from physis_amd.warmstart import UNet
    coefs, grids = batch
    n = int(np.prod(SHAPE))
    p64 = ps.PSGridNew(SHAPE, np.float64)
    ps.PSGridCopyin(p64, np.zeros(n))
    solver = PoissonSolver(grids, NB, "bf16")
    tgt = ps.PSGridNew(SHAPE)
    try:
        st = solver.solve_mixed(p64, tol=1e-3, inner=100)
        print("training targets", st, flush=True)
        assert st["converged"]
        sol = ps.PSGridCopyout(p64).astype(np.float32)
        ps.PSGridCopyin(tgt, sol)
        scale = float(np.max(np.abs(sol)))
        net = UNet(N, N, N, 8, scale, seed=3)
        hist = net.train(grids[11], grids[10], tgt, NB - NTEST, steps=60,
                         batch=8, lr=1e-3, log=20)
    finally:
        solver.free()
        ps.PSGridFree(p64)
        ps.PSGridFree(tgt)
    yield net, hist
    net.free()


def test_prediction_over_the_batch(ps, batch):
    from physis_amd.warmstart import UNet
    coefs, grids = batch
    print("prediction over 64 x 128^3", flush=True)
    net = UNet(N, N, N, 8, 2.0, seed=11)
    pg = ps.PSGridNew(SHAPE)
    try:
        for f0 in range(0, NB, 8):  # the network runs batches of 8
            net.predict(grids[11], grids[10], f0, 8, pg)
        got = ps.PSGridCopyout(pg)
        assert np.all(np.isfinite(got))
        assert np.all(got[coefs[10] == 0] == 0)
        assert np.count_nonzero(got) > 0.5 * np.count_nonzero(coefs[10])
        # one whole problem (the 6th) against the oracle
        m = N ** 3
        q = 5
        sl = slice(q * m, (q + 1) * m)
        ref = U.predict(net.get_params(), coefs[11][sl], coefs[10][sl],
                        (N, N, N), 1, 2.0)
        err = np.linalg.norm(got[sl] - ref) / np.linalg.norm(ref)
        print("prediction rel err vs oracle", err, flush=True)
        assert err <= 1e-2
    finally:
        net.free()
        ps.PSGridFree(pg)


def test_training_lowers_the_loss(ps, trained):
    _, hist = trained
    first, last = np.mean(hist[:5]), np.mean(hist[-5:])
    print("loss first/last", first, last, flush=True)
    assert np.all(np.isfinite(hist))
    assert last < 0.7 * first


def test_solve_from_prediction_reaches_1e6(ps, batch, trained):
    from physis_amd.poisson import PoissonSolver
    coefs, _ = batch
    net, _ = trained
    m = N ** 3
    lo = (NB - NTEST) * m
    tcoefs = [np.ascontiguousarray(c[lo:]) for c in coefs]
    tshape = (N, N, N * NTEST)
    print("solving 8 unseen 128^3 problems from the prediction", flush=True)
    tgrids = []
    for c in tcoefs:
        g = ps.PSGridNew(tshape)
        ps.PSGridCopyin(g, c)
        tgrids.append(g)
    pw = ps.PSGridNew(tshape, np.float64)
    ts = PoissonSolver(tgrids, NTEST, "bf16")
    try:
        net.predict(tgrids[11], tgrids[10], 0, NTEST, pw)
        guess = ps.PSGridCopyout(pw)
        assert np.all(guess[tcoefs[10] == 0] == 0)
        _, _, s0 = O.poisson_residual(0, guess, tcoefs, tshape, NTEST)
        rel0 = np.sqrt(s0[:, 0] / s0[:, 2])
        print("relative residual of the prediction", rel0, flush=True)
        assert np.all(np.isfinite(rel0))
        st = ts.solve_mixed(pw, tol=1e-6, inner=100)
        print("warm start", st, flush=True)
        assert st["converged"]
        p = ps.PSGridCopyout(pw)
        _, _, sums = O.poisson_residual(0, p, tcoefs, tshape, NTEST)
        rel = np.sqrt(sums[:, 0] / sums[:, 2])
        print("FP64 relative residuals", rel, flush=True)
        assert np.all(rel <= 1e-6 * (1 + 1e-9))
        assert math.isclose(float(np.max(rel)), st["rel_residual"],
                            rel_tol=1e-9)
    finally:
        ts.free()
        for g in tgrids + [pw]:
            ps.PSGridFree(g)


def test_committed_weights_save_sweeps(ps, batch):
    """The committed network (physis_amd/weights/unet128.npz, trained on
    other problems by physis_amd/warmstart_train.py) on 8 problems of this
    batch: the BF16 refinement from its prediction reaches 1e-6 in at least
    15% fewer sweeps than from zero (measured 24% over 64 problems)."""
    from physis_amd.poisson import PoissonSolver, time_to
    from physis_amd.warmstart import UNet
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    w = np.load(os.path.join(root, "physis_amd", "weights", "unet128.npz"))
    coefs, _ = batch
    m = N ** 3
    lo = (NB - NTEST) * m
    tshape = (N, N, N * NTEST)
    tgrids = []
    for c in coefs:
        g = ps.PSGridNew(tshape)
        ps.PSGridCopyin(g, np.ascontiguousarray(c[lo:]))
        tgrids.append(g)
    ts = PoissonSolver(tgrids, NTEST, "bf16")
    net = UNet(N, N, N, 8, float(w["out_scale"]), seed=0)
    p = ps.PSGridNew(tshape, np.float64)
    try:
        net.set_params(w["params"])
        st0 = ts.solve_mixed(p, tol=1e-6, inner=200, history=8192)
        net.predict(tgrids[11], tgrids[10], 0, NTEST, p)
        st1 = ts.solve_mixed(p, tol=1e-6, inner=200, history=8192)
        assert st0["converged"] and st1["converged"]
        s0 = time_to(st0.pop("history"), 1e-6)[0]
        s1 = time_to(st1.pop("history"), 1e-6)[0]
        print(f"sweeps to 1e-6: zero {s0:.0f}, prediction {s1:.0f}, saving "
              f"{1 - s1 / s0:.3f}", flush=True)
        assert s1 <= 0.85 * s0
    finally:
        net.free()
        ts.free()
        for g in tgrids + [p]:
            ps.PSGridFree(g)
