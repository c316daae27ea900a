"""GPU checks of the learned warm start (libphysis_warmstart_hip) against the
oracle (oracle/unet_ref.py, PyTorch CPU FP32 with BF16 rounding at the same
storage points).

Tolerances (BF16 activations, FP32 accumulation in a different order):
  * convolution outputs (BF16): |gpu - ref| <= 2^-7 |ref| + 1e-5 max|ref|
    (one BF16 rounding of each side);
  * filter gradients (FP32 sums over all positions): 1e-4 of the largest;
  * whole-network gradient: relative L2 error <= 5e-2 and cosine >= 0.998
    (BF16 rounding points of gradients differ between the two);
  * loss: 1e-2 relative.
The end-to-end test trains the network on solutions of the Jacobi solve and
checks that starting from its prediction needs fewer sweeps than starting
from zero.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402
from physis_amd.configs import poisson_problem  # noqa: E402

pytestmark = pytest.mark.gpu


def _close_bf16(got_bits, ref_bits):
    g, r = U.bf16_float(got_bits), U.bf16_float(ref_bits)
    tol = 2.0 ** -7 * np.abs(r) + 1e-5 * np.max(np.abs(r))
    bad = np.abs(g - r) > tol
    assert not bad.any(), (int(bad.sum()), np.max(np.abs(g - r)))


# (n, d, h, w, cin, cout).  The convolution stages x tiles of 64, 32 or 16
# positions: widths below, at and above each tile, ragged last tiles and
# rows not a multiple of 4.
CONV_DIMS = [(2, 4, 5, 20, 8, 8), (1, 3, 4, 8, 8, 16), (1, 3, 4, 9, 16, 32),
             (1, 3, 5, 7, 48, 16), (1, 4, 4, 16, 24, 8), (1, 2, 3, 33, 32, 32),
             (1, 3, 3, 6, 16, 16), (1, 2, 4, 128, 8, 8),
             (1, 2, 5, 96, 24, 8), (1, 2, 3, 40, 16, 16), (2, 2, 3, 64, 8, 24)]


@pytest.mark.parametrize("dims", CONV_DIMS)
@pytest.mark.parametrize("relu", [False, True])
def test_conv_forward(ps, dims, relu):
    from physis_amd.warmstart import conv_forward_host
    print(f"conv forward {dims} relu={relu}", flush=True)
    n, d, h, wd, cin, cout = dims
    rng = np.random.default_rng(3)
    x = U.bf16_bits(rng.standard_normal(n * d * h * wd * cin))
    w = U.bf16_bits(rng.standard_normal(cout * 27 * cin) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    got = conv_forward_host(x, w, b, relu, *dims)
    _close_bf16(got, U.conv_forward(x, w, b, relu, *dims))


@pytest.mark.parametrize("dims", CONV_DIMS)
def test_conv_backward(ps, dims):
    from physis_amd.warmstart import (conv_backward_data_host,
                                      conv_backward_filter_host)
    print(f"conv backward {dims}", flush=True)
    n, d, h, wd, cin, cout = dims
    rng = np.random.default_rng(4)
    x = U.bf16_bits(rng.standard_normal(n * d * h * wd * cin))
    dy = U.bf16_bits(rng.standard_normal(n * d * h * wd * cout))
    w = U.bf16_bits(rng.standard_normal(cout * 27 * cin) * 0.2)
    rdx, rdw, rdb = U.conv_backward(x, dy, w, *dims)
    if cout % 8 == 0:
        _close_bf16(conv_backward_data_host(dy, w, *dims), rdx)
    dw, db = conv_backward_filter_host(x, dy, *dims)
    assert np.max(np.abs(dw - rdw)) <= 1e-4 * np.max(np.abs(rdw))
    assert np.max(np.abs(db - rdb)) <= 1e-4 * np.max(np.abs(rdb))


# Rows of a multiple of 32 positions take the LDS-staged filter-gradient
# kernel (the widths above take the gathering one): both channel layouts of
# every level, a batch of two, and several blocks.
WGRAD_LDS_DIMS = [(1, 3, 4, 32, 8, 8), (2, 3, 4, 32, 8, 16),
                  (1, 2, 3, 64, 24, 8), (1, 3, 3, 32, 48, 16),
                  (1, 2, 3, 32, 16, 48), (1, 2, 2, 32, 32, 32),
                  (1, 8, 9, 64, 8, 8)]


@pytest.mark.parametrize("dims", WGRAD_LDS_DIMS)
def test_wgrad_lds_rows(ps, dims):
    from physis_amd.warmstart import conv_backward_filter_host
    print(f"filter gradient {dims}", flush=True)
    n, d, h, wd, cin, cout = dims
    rng = np.random.default_rng(6)
    x = U.bf16_bits(rng.standard_normal(n * d * h * wd * cin))
    dy = U.bf16_bits(rng.standard_normal(n * d * h * wd * cout))
    w = U.bf16_bits(rng.standard_normal(cout * 27 * cin) * 0.2)
    _, rdw, rdb = U.conv_backward(x, dy, w, *dims)
    dw, db = conv_backward_filter_host(x, dy, *dims)
    assert np.max(np.abs(dw - rdw)) <= 1e-4 * np.max(np.abs(rdw))
    assert np.max(np.abs(db - rdb)) <= 1e-4 * np.max(np.abs(rdb))


def test_wgrad_spans_several_blocks(ps):
    """More positions than one filter-gradient block (256 positions at this
    size): partials of several blocks, the last one ragged, through the
    gathering kernel (rows of 40)."""
    from physis_amd.warmstart import conv_backward_filter_host
    dims = (1, 9, 16, 40, 8, 8)  # 5760 positions
    n, d, h, wd, cin, cout = dims
    rng = np.random.default_rng(5)
    x = U.bf16_bits(rng.standard_normal(n * d * h * wd * cin))
    dy = U.bf16_bits(rng.standard_normal(n * d * h * wd * cout))
    w = U.bf16_bits(rng.standard_normal(cout * 27 * cin) * 0.2)
    _, rdw, rdb = U.conv_backward(x, dy, w, *dims)
    dw, db = conv_backward_filter_host(x, dy, *dims)
    assert np.max(np.abs(dw - rdw)) <= 1e-4 * np.max(np.abs(rdw))
    assert np.max(np.abs(db - rdb)) <= 1e-4 * np.max(np.abs(rdb))


def _problem_grids(ps, coefs, shape):
    rhs = ps.PSGridNew(shape)
    ps.PSGridCopyin(rhs, coefs[11])
    bnd = ps.PSGridNew(shape)
    ps.PSGridCopyin(bnd, coefs[10])
    return rhs, bnd


@pytest.mark.parametrize("w4,w16", [(0.0, 0.0), (4.0, 16.0)])
def test_network_loss_and_gradient(ps, w4, w16):
    """Whole-network loss and gradient against the oracle, with the mean
    square error and with the smooth-error terms."""
    from physis_amd.warmstart import UNet
    nx = ny = nzs = 32
    nb, scale = 2, 4.0
    shape = (nx, ny, nzs * nb)
    coefs = poisson_problem(nx, ny, nzs, nbatch=nb, gaussians=(3, 6),
                            sigma=(2.0, 6.0), seed=41)
    rng = np.random.default_rng(6)
    target = (rng.standard_normal(int(np.prod(shape))).astype(np.float32) *
              coefs[10] * 3.0)
    rhs, bnd = _problem_grids(ps, coefs, shape)
    tg = ps.PSGridNew(shape)
    ps.PSGridCopyin(tg, target)
    net = UNet(nx, ny, nzs, nb, scale, seed=7)
    try:
        params = net.get_params()
        assert params.size == U.num_params()
        net.set_loss_weights(w4, w16)
        loss, grad = net.loss_grad(rhs, bnd, tg, 0, nb)
        rloss, rgrad = U.loss_grad(params, coefs[11], coefs[10], target,
                                   (nx, ny, nzs), nb, scale, w4=w4, w16=w16)
        print("loss", loss, rloss, "rel", abs(loss - rloss) / rloss,
              flush=True)
        # measured: loss 2.4e-5 relative, gradient 1.7e-3 relative L2 (BF16
        # activations in a different summation order); bounds about 5x
        assert abs(loss - rloss) <= 1.5e-4 * rloss
        err = np.linalg.norm(grad - rgrad) / np.linalg.norm(rgrad)
        cos = float(grad @ rgrad / np.linalg.norm(grad) / np.linalg.norm(rgrad))
        print("grad rel err", err, "cos", cos, flush=True)
        assert err <= 1e-2 and cos >= 0.9999
        # every layer's slice agrees too (a wrong layer cannot hide)
        off = 0
        for ci, co in U.LAYERS:
            n = co * 27 * ci + co
            g, r = grad[off:off + n], rgrad[off:off + n]
            lerr = np.linalg.norm(g - r) / np.linalg.norm(r)
            print(f"layer {ci}->{co}: grad rel err {lerr:.3e}", flush=True)
            assert lerr <= 1.5e-2, (ci, co)  # measured <= 2.5e-3
            off += n
        # prediction = out_scale * out * bnd
        pg = ps.PSGridNew(shape)
        net.predict(rhs, bnd, 0, nb, pg)
        got = ps.PSGridCopyout(pg)
        ref = U.predict(params, coefs[11], coefs[10], (nx, ny, nzs), nb, scale)
        assert np.linalg.norm(got - ref) <= 1e-2 * np.linalg.norm(ref)
        assert np.all(got[coefs[10] == 0] == 0)
        ps.PSGridFree(pg)
    finally:
        net.free()
        for g in (rhs, bnd, tg):
            ps.PSGridFree(g)


def test_warm_start_saves_sweeps(ps):
    """Train on Jacobi solutions of 64 problems and their 8 axis
    reflections, then solve 8 unseen ones from the prediction and from zero:
    the warm start needs at least 10% fewer sweeps (measured 12.7%; the
    training is deterministic).  The committed 128^3 network's saving of
    at least 15% is tested in test_gpu_warmstart128.py."""
    from physis_amd.poisson import PoissonSolver
    from physis_amd.warmstart import UNet, reflect
    nx = ny = nzs = 32
    ntrain, ntest = 64, 8
    nb = ntrain + ntest
    shape = (nx, ny, nzs * nb)
    coefs = poisson_problem(nx, ny, nzs, nbatch=nb, noise=0.05,
                            gaussians=(4, 24), sigma=(1.0, 4.0), seed=51)
    grids = []
    for c in coefs:
        g = ps.PSGridNew(shape)
        ps.PSGridCopyin(g, c)
        grids.append(g)
    n = int(np.prod(shape))
    p64 = ps.PSGridNew(shape, np.float64)
    ps.PSGridCopyin(p64, np.zeros(n))
    solver = PoissonSolver(grids, nb, "bf16")
    try:
        st = solver.solve_mixed(p64, tol=1e-5, inner=50)
        print("training solutions", st, flush=True)
        assert st["converged"]
        sol = ps.PSGridCopyout(p64).astype(np.float32)
        tgt = ps.PSGridNew(shape)
        ps.PSGridCopyin(tgt, sol)
        scale = float(np.max(np.abs(sol)))
        # the training problems under the 8 reflections, stacked
        ashape = (nx, ny, nzs * ntrain * 8)
        aug = [ps.PSGridNew(ashape) for _ in range(3)]
        for r in range(8):
            for src, dst in zip((grids[11], grids[10], tgt), aug):
                reflect(src, 0, ntrain, dst, r * ntrain, (nx, ny, nzs), r)
        net = UNet(nx, ny, nzs, 8, scale, seed=3)
        hist = net.train(aug[0], aug[1], aug[2], ntrain * 8, steps=3000,
                         batch=8, lr=1e-3, lr_final=1e-4, log=500)
        for g in aug:
            ps.PSGridFree(g)
        first, last = np.mean(hist[:10]), np.mean(hist[-50:])
        print("loss first/last", first, last, flush=True)
        assert last < 0.5 * first
        # unseen problems: a solver over just the test problems
        m = nx * ny * nzs
        tcoefs = [np.ascontiguousarray(c[ntrain * m:]) for c in coefs]
        tshape = (nx, ny, nzs * ntest)
        tgrids = []
        for c in tcoefs:
            g = ps.PSGridNew(tshape)
            ps.PSGridCopyin(g, c)
            tgrids.append(g)
        ts = PoissonSolver(tgrids, ntest, "bf16")
        pz = ps.PSGridNew(tshape, np.float64)
        ps.PSGridCopyin(pz, np.zeros(ntest * m))
        zero = ts.solve_mixed(pz, tol=1e-6, inner=50)
        pw = ps.PSGridNew(tshape, np.float64)
        net.predict(tgrids[11], tgrids[10], 0, ntest, pw)
        warm = ts.solve_mixed(pw, tol=1e-6, inner=50)
        print("zero start", zero, flush=True)
        print("warm start", warm, flush=True)
        assert zero["converged"] and warm["converged"]
        saving = 1.0 - warm["sweeps"] / zero["sweeps"]
        print("sweep saving", saving, flush=True)
        assert saving >= 0.10
        ts.free()
        net.free()
        for g in tgrids + [pz, pw, tgt]:
            ps.PSGridFree(g)
    finally:
        solver.free()
        for g in grids + [p64]:
            ps.PSGridFree(g)


def test_adam_steps_match_the_oracle(ps):
    """Three training steps: each Adam update equals the one computed on the
    host from the GPU's own gradient (FP32 Adam, 1e-6 relative), and the
    gradient at the updated parameters still matches the oracle's."""
    from physis_amd.warmstart import UNet
    nx = ny = nzs = 32
    nb, scale = 4, 2.0
    shape = (nx, ny, nzs * nb)
    coefs = poisson_problem(nx, ny, nzs, nbatch=nb, gaussians=(3, 6),
                            sigma=(2.0, 6.0), seed=43)
    rng = np.random.default_rng(8)
    target = (rng.standard_normal(int(np.prod(shape))).astype(np.float32) *
              coefs[10])
    rhs, bnd = _problem_grids(ps, coefs, shape)
    tg = ps.PSGridNew(shape)
    ps.PSGridCopyin(tg, target)
    net = UNet(nx, ny, nzs, 2, scale, seed=9)
    m = v = None
    m3 = nx * ny * nzs
    try:
        for step, first in enumerate((0, 2, 1), start=1):
            p0 = net.get_params()
            _, g = net.loss_grad(rhs, bnd, tg, first, 2)
            lr = 1e-3
            net.train_step(rhs, bnd, tg, first, 2, lr)
            p1 = net.get_params()
            g = g.astype(np.float32)
            m = (np.float32(0.1) * g if m is None else
                 np.float32(0.9) * m + np.float32(0.1) * g)
            v = (np.float32(0.001) * g * g if v is None else
                 np.float32(0.999) * v + np.float32(0.001) * g * g)
            c1 = np.float32(1 - 0.9 ** step)
            c2 = np.float32(1 - 0.999 ** step)
            want = p0 - np.float32(lr) * (m / c1) / (np.sqrt(v / c2) +
                                                     np.float32(1e-8))
            d = np.max(np.abs(p1 - want))
            print(f"step {step}: max |param - host Adam| {d:.3e}", flush=True)
            assert d <= 1e-6 * max(1.0, float(np.max(np.abs(p0))))
            _, g1 = net.loss_grad(rhs, bnd, tg, first, 2)
            sl = slice(first * m3, (first + 2) * m3)
            _, r1 = U.loss_grad(p1, coefs[11][sl], coefs[10][sl], target[sl],
                                (nx, ny, nzs), 2, scale)
            cos = float(g1 @ r1 / np.linalg.norm(g1) / np.linalg.norm(r1))
            print(f"step {step}: gradient cosine vs oracle {cos:.6f}",
                  flush=True)
            assert cos >= 0.998
    finally:
        net.free()
        for gg in (rhs, bnd, tg):
            ps.PSGridFree(gg)
