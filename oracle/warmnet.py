"""CPU restatement of the warm-start network (msg123_amd/csrc/warmnet.hip).
TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference: this module is the
checker for the HIP kernels.  FP64 numpy; with `emulate=True` it rounds to
bf16 at exactly the points where the device stores bf16 (weights, input
features, activations, output gradient, propagated gradients), so the two
differ only by accumulation order and fp32-vs-fp64 accumulation.

Pinned by: central finite differences of the FP64 loss (tests/test_warmnet.py)
and a brute-force loop convolution on a tiny grid.
"""
from __future__ import annotations

import numpy as np

CIN = (4, 16, 16, 16)
COUT = (16, 16, 16, 1)
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def layout():
    """(weight offset, bias offset) of each layer in the flat parameter vector."""
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        w = 27 * co * ci
        out.append((off, off + w))
        off += w + co
    return out, off


LAYOUT, NPARAM = layout()


def bf16(a):
    """Round to bf16 (nearest even) via float32; returns float64 values."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(
        np.float64)


def unpack(p, l):
    wo, bo = LAYOUT[l]
    w = np.asarray(p[wo:bo], np.float64).reshape(27, COUT[l], CIN[l])
    return w, np.asarray(p[bo:bo + COUT[l]], np.float64)


def shift(x, dz, dy, dx):
    """y[z,y,x] = x[z+dz, y+dy, x+dx] with zeros outside (x: (nz, ny, nx, C))."""
    out = np.zeros_like(x)
    nz, ny, nx = x.shape[:3]
    zs, ys, xs = (slice(max(0, -d), n - max(0, d)) for d, n in ((dz, nz), (dy, ny), (dx, nx)))
    zd, yd, xd = (slice(max(0, d), n - max(0, -d)) for d, n in ((dz, nz), (dy, ny), (dx, nx)))
    out[zs, ys, xs] = x[zd, yd, xd]
    return out


def conv(x, w, b=None):
    """'same' 3x3x3 convolution: y[p, co] = sum_t sum_ci w[t, co, ci] x[p + off(t), ci] + b."""
    y = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += shift(x, dz, dy, dx) @ w[t].T
    return y if b is None else y + b


def forward(p, x, emulate=True):
    """x: (nz, ny, nx, 4) features.  Returns output (nz, ny, nx) and the activations."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    acts = [r(x)]
    h = acts[0]
    for l in range(4):
        w, b = unpack(p, l)
        z = conv(h, r(w), np.asarray(b, np.float32).astype(np.float64) if emulate else b)
        if l < 3:
            h = r(np.maximum(z, 0.0))
            acts.append(h)
        else:
            y = z[..., 0].astype(np.float32).astype(np.float64) if emulate else z[..., 0]
    return y, acts


def loss_grad(p, x, t, emulate=True):
    """Mean squared error and its gradient with respect to the flat parameters."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    y, acts = forward(p, x, emulate)
    n = y.size
    e = (y - t).astype(np.float32).astype(np.float64) if emulate else y - t
    loss = float(np.sum(e * e) / n)
    g = np.zeros(NPARAM)
    dy = r(2.0 * e / n)[..., None]                     # (nz, ny, nx, 1)
    for l in (3, 2, 1, 0):
        w, _ = unpack(p, l)
        xin = acts[l]
        wo, bo = LAYOUT[l]
        dw = np.zeros((27, COUT[l], CIN[l]))
        for k, (dz, dyy, dx) in enumerate(TAPS):
            xs = shift(xin, dz, dyy, dx)
            dw[k] = np.einsum("zyxo,zyxi->oi", dy, xs)
        g[wo:bo] = dw.ravel()
        g[bo:bo + COUT[l]] = dy.reshape(-1, COUT[l]).sum(axis=0)
        if l == 0:
            break
        # input gradient: mirrored taps, transposed weights, times relu'(a_l)
        wr = r(w)
        dx_ = np.zeros(xin.shape)
        for k, (dz, dyy, dx) in enumerate(TAPS):
            dx_ += shift(dy, -dz, -dyy, -dx) @ wr[k]
        dy = r(dx_ * (xin > 0))
    return loss, g


def features(hn, hm, top, lnks, r, L, rech_depth, nxy):
    """Features (bf16 values), h_lin and the increment scale, as k_features."""
    inc = r * (hn - hm)
    s = np.sqrt(np.sum((inc * L) ** 2) / hn.size) + 1.0e-12
    f = np.zeros((hn.size, 4))
    f[:, 0] = inc * L / s
    f[:, 1] = (hn - top) * L
    f[:nxy, 2] = rech_depth * 1.0e3
    f[:, 3] = (np.asarray(lnks, np.float32) + np.float32(11.512925)) / np.float32(1.5)
    return bf16(f), hn + inc, s


def init_params(seed):
    """The device's initial parameters (He-uniform from splitmix64, zero biases)."""
    mask = (1 << 64) - 1
    st = seed & mask
    p = np.zeros(NPARAM)
    for l in range(4):
        lim = np.sqrt(6.0 / (27 * CIN[l])) * (0.1 if l == 3 else 1.0)
        wo, bo = LAYOUT[l]
        for i in range(bo - wo):
            st = (st + 0x9E3779B97F4A7C15) & mask
            z = st
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
            z ^= z >> 31
            u = (z >> 11) * (1.0 / 9007199254740992.0)
            p[wo + i] = np.float32((2.0 * u - 1.0) * lim)
    return p
