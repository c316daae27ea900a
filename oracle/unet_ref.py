"""Oracle of the U-Net warm start (include/physis/physis_warmstart.h), test
infrastructure only: used by tests/ to check libphysis_warmstart_hip, never
by the product.  The network has no counterpart in the reference; this is a
CPU FP32 restatement of the header's definition with PyTorch's CPU
operators, rounding to bfloat16 wherever the GPU stores BF16 (weights fed to
the convolutions, every activation, the network input), so that forward
values agree up to FP32 summation order.  Gradients come from autograd of the
same function (a cast to bfloat16 and back also rounds the gradient flowing
through it, as the GPU stores BF16 gradients).

Pinned by tests/test_oracle_unet.py: the convolution against a direct loop
over taps written from the header's weight layout, and the gradient of an
unrounded FP64 copy against central finite differences.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [(8, 8), (8, 8), (8, 16), (16, 16), (16, 32), (32, 32), (48, 16),
          (16, 16), (24, 8), (8, 8)]
POOL = 4


def num_params() -> int:
    return sum(co * 27 * ci + co for ci, co in LAYERS) + 9


def unpack(p):
    """Flat parameter tensor -> [(W [co, ci, 3, 3, 3], b [co])] and the head.

    W[co][tap][ci] with tap = 9 dz + 3 dy + dx is conv3d's
    weight[co, ci, dz, dy, dx] (cross-correlation, padding 1)."""
    out, off = [], 0
    for ci, co in LAYERS:
        n = co * 27 * ci
        w = p[off:off + n].reshape(co, 3, 3, 3, ci).permute(0, 4, 1, 2, 3)
        off += n
        out.append((w, p[off:off + co]))
        off += co
    return out, p[off:off + 9]


def _bf(t, rounding):
    return t.to(torch.bfloat16).to(t.dtype) if rounding else t


def forward(params, rhs, bnd, shape, nbatch, out_scale, rounding=True,
            dtype=torch.float32, requires_grad=False):
    """Network output `out` [nbatch, nzs, ny, nx] (before out_scale) and the
    parameter tensor (a leaf when requires_grad).  rhs, bnd: flat arrays of
    nbatch problems of shape (nx, ny, nzs) stacked along z."""
    nx, ny, nzs = shape
    p = torch.tensor(np.asarray(params), dtype=dtype,
                     requires_grad=requires_grad)
    layers, head = unpack(p)
    r = torch.as_tensor(np.asarray(rhs), dtype=dtype).reshape(nbatch, 1, nzs,
                                                              ny, nx)
    m = torch.as_tensor(np.asarray(bnd), dtype=dtype).reshape(nbatch, 1, nzs,
                                                              ny, nx)
    x = torch.cat([_bf(r, rounding), _bf(m, rounding),
                   torch.zeros((nbatch, 6, nzs, ny, nx), dtype=dtype)], 1)

    def conv(l, t):
        w, b = layers[l]
        return _bf(F.relu(F.conv3d(t, _bf(w, rounding), b, padding=1)),
                   rounding)

    def pool(t):
        return _bf(F.avg_pool3d(t, POOL), rounding)

    def up(t):
        for d in (2, 3, 4):
            t = t.repeat_interleave(POOL, dim=d)
        return t

    e1 = conv(1, conv(0, x))
    e2 = conv(3, conv(2, pool(e1)))
    e3 = conv(5, conv(4, pool(e2)))
    d2 = conv(7, conv(6, torch.cat([up(e3), e2], 1)))
    d1 = conv(9, conv(8, torch.cat([up(d2), e1], 1)))
    out = (d1 * head[:8].reshape(1, 8, 1, 1, 1)).sum(1) + head[8]
    return out, p


def loss_grad(params, rhs, bnd, target, shape, nbatch, out_scale,
              rounding=True, dtype=torch.float32, w4=0.0, w16=0.0):
    """(loss, grad[numpy]) of mean(d^2) + w4 mean(pool4(d)^2) +
    w16 mean(pool16(d)^2), d = out - target / out_scale, poolK the mean
    over K x K x K cells (the header's PSUNetLossGrad)."""
    out, p = forward(params, rhs, bnd, shape, nbatch, out_scale, rounding,
                     dtype, requires_grad=True)
    nx, ny, nzs = shape
    t = torch.as_tensor(np.asarray(target), dtype=dtype).reshape(
        nbatch, nzs, ny, nx) / out_scale
    d = out - t
    loss = (d ** 2).mean()
    if w4 or w16:
        d5 = d.reshape(nbatch, 1, nzs, ny, nx)
        loss = loss + w4 * (F.avg_pool3d(d5, POOL) ** 2).mean() + \
            w16 * (F.avg_pool3d(d5, POOL * POOL) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.numpy().copy()


def predict(params, rhs, bnd, shape, nbatch, out_scale):
    out, _ = forward(params, rhs, bnd, shape, nbatch, out_scale)
    m = torch.as_tensor(np.asarray(bnd), dtype=torch.float32)
    return ((out.detach().reshape(-1) * out_scale) * m).numpy()


# --------------------------------------------------------- single layers
def bf16_bits(a: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns (round to nearest even)."""
    t = torch.as_tensor(np.ascontiguousarray(a, np.float32))
    return t.to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)


def bf16_float(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(
        np.float32)


def _cl_to_cf(a, n, d, h, w, c):
    return torch.as_tensor(a.reshape(n, d, h, w, c)).permute(0, 4, 1, 2, 3)


def _cf_to_cl(t):
    return t.permute(0, 2, 3, 4, 1).contiguous().numpy().reshape(-1)


def _weight(wbits, cin, cout):
    w = bf16_float(wbits).reshape(cout, 3, 3, 3, cin)
    return torch.as_tensor(w).permute(0, 4, 1, 2, 3).contiguous()


def conv_forward(xbits, wbits, bias, relu, n, d, h, w, cin, cout):
    """BF16 bits of the layer output (channels-last)."""
    x = _cl_to_cf(bf16_float(xbits), n, d, h, w, cin)
    y = F.conv3d(x, _weight(wbits, cin, cout), torch.as_tensor(bias),
                 padding=1)
    if relu:
        y = F.relu(y)
    return bf16_bits(_cf_to_cl(y))


def conv_backward(xbits, dybits, wbits, n, d, h, w, cin, cout):
    """(dx bf16 bits, dw FP32 [co][27][ci] flat, db FP32)."""
    x = _cl_to_cf(bf16_float(xbits), n, d, h, w, cin).clone() \
        .requires_grad_(True)
    wt = _weight(wbits, cin, cout).clone().requires_grad_(True)
    b = torch.zeros(cout, requires_grad=True)
    y = F.conv3d(x, wt, b, padding=1)
    dy = _cl_to_cf(bf16_float(dybits), n, d, h, w, cout)
    y.backward(dy)
    dw = wt.grad.permute(0, 2, 3, 4, 1).contiguous().numpy().reshape(-1)
    return bf16_bits(_cf_to_cl(x.grad)), dw, b.grad.numpy().copy()
