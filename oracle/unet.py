"""FP64 oracle of the learned preconditioner's network (test infrastructure
only).

The reference has no learned preconditioner (its projection is BiCGSTAB with
the separable inverse, projection_3d.py:756-791), so nothing here restates
reference code: this is the NumPy statement of the network the GPU kernels
implement, used to check their forward values and gradients.  It is pinned
by central finite differences of the loss (every parameter block) and by
closed forms (a delta kernel is a copy, a shifted delta kernel is a shift,
pooling then upsampling a piecewise-constant field is the identity).

Network (two levels, grid dims even; y periodic, zero padding in x and z),
on the residual r and a density feature q = RHO_REF / rho - 1 (the
correction the separable inverse needs depends on the density field):
    h0  = relu(conv3([r, q], w0, b0))        2  -> C0
    d1  = avgpool2(h0)
    h1  = relu(conv3(d1, w1, b1))            C0 -> C1   (coarse grid)
    h1b = relu(conv3(h1, w2, b2))            C1 -> C1
    u0  = upsample2(h1b)
    h2  = relu(conv3([h0, u0], w3, b3))      C0+C1 -> C0
    c   = conv3(h2, w4, b4)                  C0 -> 1
conv3: y[co, k, j, i] = b[co] + sum_ci sum_t w[co, ci, t] x[ci, k+dk, j+dj, i+di]
with t = (dk+1) 9 + (dj+1) 3 + (di+1).
Loss: L = sum dz_k (c - t)^2 / sum dz_k t^2.
"""
from __future__ import annotations

import numpy as np

C0, C1 = 8, 16
LAYERS = (("w0", "b0", 2, C0), ("w1", "b1", C0, C1), ("w2", "b2", C1, C1),
          ("w3", "b3", C0 + C1, C0), ("w4", "b4", C0, 1))
TAPS = [(dk, dj, di) for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)]


def n_params():
    return sum(co * ci * 27 + co for _, _, ci, co in LAYERS)


def unpack(theta):
    """Flat parameter vector (per layer: w (co, ci, 27) then b (co)) -> dict."""
    p, o = {}, 0
    for wn, bn, ci, co in LAYERS:
        p[wn] = theta[o:o + co * ci * 27].reshape(co, ci, 27)
        o += co * ci * 27
        p[bn] = theta[o:o + co]
        o += co
    return p


def pack(p):
    return np.concatenate([np.concatenate([p[wn].ravel(), p[bn].ravel()])
                           for wn, bn, _, _ in LAYERS])


def init(seed=0):
    """He-style random weights, zero biases (the GPU training starts here)."""
    rng = np.random.default_rng(seed)
    p = {}
    for wn, bn, ci, co in LAYERS:
        p[wn] = rng.standard_normal((co, ci, 27)) * np.sqrt(2.0 / (27 * ci))
        p[bn] = np.zeros(co)
    p["w4"] *= 0.1
    return pack(p)


def _shift(x, dk, dj, di):
    """s[..., k, j, i] = x[..., k+dk, (j+dj) mod Ny, i+di], zero outside in k, i."""
    s = np.roll(x, -dj, axis=-2)
    out = np.zeros_like(s)
    Nz, Nx = x.shape[-3], x.shape[-1]
    ks, ke = max(0, -dk), min(Nz, Nz - dk)
    is_, ie = max(0, -di), min(Nx, Nx - di)
    out[..., ks:ke, :, is_:ie] = s[..., ks + dk:ke + dk, :, is_ + di:ie + di]
    return out


def conv3(x, w, b):
    co = w.shape[0]
    y = np.empty((co,) + x.shape[1:])
    y[:] = b[:, None, None, None]
    for t, (dk, dj, di) in enumerate(TAPS):
        y += np.tensordot(w[:, :, t], _shift(x, dk, dj, di), axes=(1, 0))
    return y


def conv3_T(g, w):
    """Adjoint of conv3 in x: dx[ci, p] = sum_t sum_co w[co, ci, t] g[co, p - off_t]."""
    dx = np.zeros((w.shape[1],) + g.shape[1:])
    for t, (dk, dj, di) in enumerate(TAPS):
        dx += np.tensordot(w[:, :, t].T, _shift(g, -dk, -dj, -di), axes=(1, 0))
    return dx


def conv3_dw(g, x):
    """dL/dw[co, ci, t] = sum_q g[co, q] x[ci, q + off_t]."""
    co, ci = g.shape[0], x.shape[0]
    dw = np.empty((co, ci, 27))
    for t, (dk, dj, di) in enumerate(TAPS):
        dw[:, :, t] = np.tensordot(g, _shift(x, dk, dj, di), axes=((1, 2, 3), (1, 2, 3)))
    return dw


def pool2(x):
    c, Nz, Ny, Nx = x.shape
    return x.reshape(c, Nz // 2, 2, Ny // 2, 2, Nx // 2, 2).mean(axis=(2, 4, 6))


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def forward(theta, r, q):
    """r, q: (Nz, Ny, Nx).  Returns (c, activations)."""
    p = unpack(theta)
    a = {}
    a["r"] = np.stack([r, q])
    a["h0"] = np.maximum(conv3(a["r"], p["w0"], p["b0"]), 0.0)
    a["d1"] = pool2(a["h0"])
    a["h1"] = np.maximum(conv3(a["d1"], p["w1"], p["b1"]), 0.0)
    a["h1b"] = np.maximum(conv3(a["h1"], p["w2"], p["b2"]), 0.0)
    a["cat"] = np.concatenate([a["h0"], up2(a["h1b"])])
    a["h2"] = np.maximum(conv3(a["cat"], p["w3"], p["b3"]), 0.0)
    c = conv3(a["h2"], p["w4"], p["b4"])[0]
    return c, a


def loss(theta, r, q, t, dz):
    c, _ = forward(theta, r, q)
    wz = dz[:, None, None]
    return float(np.sum(wz * (c - t) ** 2) / np.sum(wz * t ** 2))


def backward(theta, r, q, t, dz):
    """(L, dL/dtheta) for one sample."""
    p = unpack(theta)
    c, a = forward(theta, r, q)
    wz = dz[:, None, None]
    den = np.sum(wz * t ** 2)
    L = float(np.sum(wz * (c - t) ** 2) / den)
    g = {}
    gc = (2.0 * wz * (c - t) / den)[None]
    g["w4"], g["b4"] = conv3_dw(gc, a["h2"]), gc.sum(axis=(1, 2, 3))
    dh2 = conv3_T(gc, p["w4"]) * (a["h2"] > 0)
    g["w3"], g["b3"] = conv3_dw(dh2, a["cat"]), dh2.sum(axis=(1, 2, 3))
    dcat = conv3_T(dh2, p["w3"])
    du0 = dcat[C0:]
    c_, Nz, Ny, Nx = du0.shape
    dh1b = du0.reshape(c_, Nz // 2, 2, Ny // 2, 2, Nx // 2, 2).sum(axis=(2, 4, 6))
    dh1b = dh1b * (a["h1b"] > 0)
    g["w2"], g["b2"] = conv3_dw(dh1b, a["h1"]), dh1b.sum(axis=(1, 2, 3))
    dh1 = conv3_T(dh1b, p["w2"]) * (a["h1"] > 0)
    g["w1"], g["b1"] = conv3_dw(dh1, a["d1"]), dh1.sum(axis=(1, 2, 3))
    dd1 = conv3_T(dh1, p["w1"])
    dh0 = (dcat[:C0] + up2(dd1) / 8.0) * (a["h0"] > 0)
    g["w0"], g["b0"] = conv3_dw(dh0, a["r"]), dh0.sum(axis=(1, 2, 3))
    return L, pack(g)
