"""Pins the U-Net oracle (oracle/unet_ref.py) on the CPU.

  * its convolution equals a direct loop over the 27 taps written from the
    weight layout in include/physis/physis_warmstart.h (W[co][tap][ci],
    tap = 9 dz + 3 dy + dx, offsets d - 1, zero padding) — a transposed
    weight, a flipped tap order or a shifted offset fails;
  * its autograd gradient, on an unrounded FP64 copy of the network, equals
    central finite differences of the loss.
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


def _direct_conv(x, w, b, n, d, h, wd, cin, cout):
    """x [n][d][h][w][cin], w [cout][27][cin] -> y [n][d][h][w][cout]."""
    x = x.reshape(n, d, h, wd, cin).astype(np.float64)
    w = w.reshape(cout, 27, cin).astype(np.float64)
    xp = np.zeros((n, d + 2, h + 2, wd + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    y = np.zeros((n, d, h, wd, cout)) + b
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        win = xp[:, dz:dz + d, dy:dy + h, dx:dx + wd]  # offset (dz-1, ...)
        y += np.einsum("ndhwc,oc->ndhwo", win, w[:, tap, :])
    return y.reshape(-1)


@pytest.mark.parametrize("dims", [(1, 4, 5, 6, 8, 8), (2, 3, 4, 5, 16, 24)])
def test_conv_matches_direct_taps(dims):
    n, d, h, wd, cin, cout = dims
    rng = np.random.default_rng(0)
    x = U.bf16_bits(rng.standard_normal(n * d * h * wd * cin))
    w = U.bf16_bits(rng.standard_normal(cout * 27 * cin) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    got = U.bf16_float(U.conv_forward(x, w, b, False, *dims))
    ref = _direct_conv(U.bf16_float(x), U.bf16_float(w), b, *dims)
    # one bfloat16 rounding of an FP32 sum
    assert np.all(np.abs(got - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-6)


def test_backward_matches_direct_loops():
    """dx, dw and db against direct loops over the taps."""
    dims = (1, 4, 4, 4, 8, 16)
    n, d, h, wd, cin, cout = dims
    rng = np.random.default_rng(1)
    x = U.bf16_bits(rng.standard_normal(n * d * h * wd * cin))
    dy = U.bf16_bits(rng.standard_normal(n * d * h * wd * cout))
    w = U.bf16_bits(rng.standard_normal(cout * 27 * cin) * 0.2)
    dx, dw, db = U.conv_backward(x, dy, w, *dims)
    # dx[p][ci] = sum_tap sum_co dy[p - off(tap)][co] w[co][tap][ci]
    wf = U.bf16_float(w).reshape(cout, 27, cin).astype(np.float64)
    gp = np.zeros((n, d + 2, h + 2, wd + 2, cout))
    gp[:, 1:-1, 1:-1, 1:-1] = U.bf16_float(dy).reshape(n, d, h, wd, cout)
    dxr = np.zeros((n, d, h, wd, cin))
    for tap in range(27):
        dz, dyy, dx_ = tap // 9, (tap // 3) % 3, tap % 3
        win = gp[:, 2 - dz:2 - dz + d, 2 - dyy:2 - dyy + h, 2 - dx_:2 - dx_ + wd]
        dxr += np.einsum("ndhwo,oc->ndhwc", win, wf[:, tap, :])
    got = U.bf16_float(dx).reshape(-1)
    dxr = dxr.reshape(-1)
    assert np.all(np.abs(got - dxr) <= 2.0 ** -8 * np.abs(dxr) + 1e-6)
    xf, dyf = U.bf16_float(x).astype(np.float64), U.bf16_float(dy)
    # dw[co][tap][ci] = sum_p dy[p][co] x[p + off(tap)][ci]
    xp = np.zeros((n, d + 2, h + 2, wd + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = xf.reshape(n, d, h, wd, cin)
    g = dyf.reshape(n, d, h, wd, cout).astype(np.float64)
    ref = np.zeros((cout, 27, cin))
    for tap in range(27):
        dz, dyy, dx = tap // 9, (tap // 3) % 3, tap % 3
        win = xp[:, dz:dz + d, dyy:dyy + h, dx:dx + wd]
        ref[:, tap, :] = np.einsum("ndhwo,ndhwc->oc", g, win)
    np.testing.assert_allclose(dw, ref.reshape(-1), rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(db, g.sum(axis=(0, 1, 2, 3)), rtol=1e-5)


def test_network_gradient_matches_finite_differences():
    shape, nb, scale = (16, 16, 16), 1, 3.0
    rng = np.random.default_rng(2)
    n = int(np.prod(shape))
    params = rng.standard_normal(U.num_params()) * 0.15
    rhs = rng.standard_normal(n)
    bnd = np.ones(n)
    target = rng.standard_normal(n)
    loss, grad = U.loss_grad(params, rhs, bnd, target, shape, nb, scale,
                             rounding=False, dtype=torch.float64)
    # one weight of every layer, a bias, and the head
    offs, off = [], 0
    for ci, co in U.LAYERS:
        offs += [off + 5 % (co * 27 * ci), off + co * 27 * ci]
        off += co * 27 * ci + co
    offs += [off, off + 8]
    eps = 1e-5
    for i in offs:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = U.loss_grad(pp, rhs, bnd, target, shape, nb, scale, False,
                            torch.float64)
        lm, _ = U.loss_grad(pm, rhs, bnd, target, shape, nb, scale, False,
                            torch.float64)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])


@pytest.mark.parametrize("w4,w16", [(4.0, 16.0), (0.5, 0.0), (0.0, 2.0)])
def test_smooth_loss_terms_closed_forms(w4, w16):
    """The smooth-error terms of the loss on errors whose cell means are
    known: a constant error c has every cell mean c (loss (1 + w4 + w16)
    c^2); an error c (-1)^(x+y+z) has every 4^3 and 16^3 cell mean 0
    (loss c^2); an error c (-1)^(floor(x/4)) has 4^3 cell means +-c and
    16^3 cell means 0 (loss (1 + w4) c^2).  A wrong cell size, a sum in
    place of a mean or a term dropped fails one of them."""
    shape, nb, scale = (32, 16, 16), 2, 2.0
    nx, ny, nzs = shape
    rng = np.random.default_rng(5)
    n = nx * ny * nzs * nb
    params = rng.standard_normal(U.num_params()) * 0.1
    rhs = rng.standard_normal(n)
    bnd = np.ones(n)
    out, _ = U.forward(params, rhs, bnd, shape, nb, scale,
                       rounding=False, dtype=torch.float64)
    out = out.detach().numpy().reshape(-1)
    z, y, x = np.meshgrid(np.arange(nzs * nb), np.arange(ny), np.arange(nx),
                          indexing="ij")
    c = 0.25
    for err, want in ((np.full(n, c), (1 + w4 + w16) * c * c),
                      (c * (-1.0) ** (x + y + z).reshape(-1), c * c),
                      (c * (-1.0) ** (x // 4).reshape(-1), (1 + w4) * c * c)):
        target = (out - err) * scale  # d = out - target / scale = err
        loss, _ = U.loss_grad(params, rhs, bnd, target, shape, nb, scale,
                              rounding=False, dtype=torch.float64, w4=w4,
                              w16=w16)
        assert abs(loss - want) <= 1e-12 * want, (loss, want)


def test_smooth_loss_gradient_matches_finite_differences():
    shape, nb, scale = (16, 16, 16), 1, 3.0
    rng = np.random.default_rng(3)
    n = int(np.prod(shape))
    params = rng.standard_normal(U.num_params()) * 0.15
    rhs = rng.standard_normal(n)
    bnd = np.ones(n)
    target = rng.standard_normal(n) + 2.0
    kw = dict(rounding=False, dtype=torch.float64, w4=3.0, w16=5.0)
    _, grad = U.loss_grad(params, rhs, bnd, target, shape, nb, scale, **kw)
    last = U.num_params() - 1
    eps = 1e-5
    for i in (3, 1000, last - 4, last):
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = U.loss_grad(pp, rhs, bnd, target, shape, nb, scale, **kw)
        lm, _ = U.loss_grad(pm, rhs, bnd, target, shape, nb, scale, **kw)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])
