"""FP64 oracle of the learned preconditioner's network (test infrastructure
only).

The reference has no learned preconditioner (its projection is BiCGSTAB with
the separable inverse, projection_3d.py:756-791), so nothing here restates
reference code: this is the NumPy statement of the network the GPU kernels
implement, used to check their forward values and gradients.  It is pinned
by central finite differences of the loss (every layer), the adjoint
identity of the convolution, and closed forms (a delta kernel is a copy, a
shifted delta kernel is a shift, pooling an upsampled field is the
identity).

Network: a 4-level 3-D U-Net with 16/32/64/64 channels and 3x3x3 kernels on
levels L0 (the grid), L1 = L0/2, L2 = L0/4, L3 = L0/8 (grid dims divisible by
8; y periodic, zero padding in x and z).  Inputs are the residual r and the
density feature q = RHO_REF / rho - 1.  Activations are channel-last
(Nz, Ny, Nx, C).
    e0 = relu(conv([r, q]))                  L0   2 -> 16
    e1 = relu(conv(pool(e0)))                L1  16 -> 32
    e2 = relu(conv(pool(e1)))                L2  32 -> 64
    e3 = relu(conv(pool(e2)))                L3  64 -> 64
    d2 = relu(conv([e2, up(e3)]))            L2 128 -> 64
    d1 = relu(conv([e1, up(d2)]))            L1  96 -> 32
    d0 = relu(conv([e0, up(d1)]))            L0  48 -> 16
    c  = conv(d0)                            L0  16 -> 1
conv: y[p, co] = b[co] + sum_t sum_ci w[t, ci, co] x[p + off_t, ci], tap
t = (dk+1) 9 + (dj+1) 3 + (di+1); pool: mean of the 8 children; up: nearest.
Parameters: one flat vector, per layer w (27, ci, co) then b (co).
Loss: L = sum dz_k (c - t)^2 / sum dz_k t^2.

``bf16=True`` applies the GPU kernels' rounding model: weights, the inputs r
and q, every stored activation (after bias and relu, and after pooling) and
every stored gradient of an activation (after the relu mask) are rounded to
bfloat16 (round to nearest even); sums are exact here, float32 on the GPU.
The output c and the weight gradients are not rounded.
"""
from __future__ import annotations

import numpy as np

# (name, level, c1 (same-level input), c2 (coarse input, upsampled), co, relu)
LAYERS = (("e0", 0, 2, 0, 16, True), ("e1", 1, 16, 0, 32, True),
          ("e2", 2, 32, 0, 64, True), ("e3", 3, 64, 0, 64, True),
          ("d2", 2, 64, 64, 64, True), ("d1", 1, 32, 64, 32, True),
          ("d0", 0, 16, 32, 16, True), ("c", 0, 16, 0, 1, False))
TAPS = [(dk, dj, di) for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)]


def n_params():
    return sum(27 * (c1 + c2) * co + co for _, _, c1, c2, co, _ in LAYERS)


def unpack(theta):
    """Flat parameter vector -> {name: (w (27, ci, co), b (co))}."""
    p, o = {}, 0
    for name, _, c1, c2, co, _ in LAYERS:
        nw = 27 * (c1 + c2) * co
        p[name] = (theta[o:o + nw].reshape(27, c1 + c2, co), theta[o + nw:o + nw + co])
        o += nw + co
    return p


def pack(p):
    return np.concatenate([np.concatenate([p[n][0].ravel(), p[n][1].ravel()])
                           for n, *_ in LAYERS])


def init(seed=0):
    """He-style random weights, zero biases, a small output layer."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, _, c1, c2, co, _ in LAYERS:
        ci = c1 + c2
        w = rng.standard_normal((27, ci, co)) * np.sqrt(2.0 / (27 * ci))
        if name == "c":
            w *= 0.1
        p[name] = (w, np.zeros(co))
    return pack(p)


def bf16(a):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    u = np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shift(x, dk, dj, di):
    """s[k, j, i] = x[k+dk, (j+dj) mod Ny, i+di], zero outside in k and i."""
    s = np.roll(x, -dj, axis=1)
    out = np.zeros_like(s)
    Nz, Nx = x.shape[0], x.shape[2]
    ks, ke = max(0, -dk), min(Nz, Nz - dk)
    is_, ie = max(0, -di), min(Nx, Nx - di)
    out[ks:ke, :, is_:ie] = s[ks + dk:ke + dk, :, is_ + di:ie + di]
    return out


def conv(x, w, b=None):
    y = np.zeros(x.shape[:3] + (w.shape[2],))
    for t, (dk, dj, di) in enumerate(TAPS):
        y += shift(x, dk, dj, di) @ w[t]
    return y if b is None else y + b


def conv_T(g, w):
    """Adjoint of conv in x: dx[q, ci] = sum_t sum_co w[t, ci, co] g[q - off_t, co]."""
    dx = np.zeros(g.shape[:3] + (w.shape[1],))
    for t, (dk, dj, di) in enumerate(TAPS):
        dx += shift(g, -dk, -dj, -di) @ w[t].T
    return dx


def conv_dw(g, x):
    """dL/dw[t, ci, co] = sum_p x[p + off_t, ci] g[p, co]."""
    dw = np.empty((27, x.shape[3], g.shape[3]))
    for t, (dk, dj, di) in enumerate(TAPS):
        dw[t] = np.tensordot(shift(x, dk, dj, di), g, axes=((0, 1, 2), (0, 1, 2)))
    return dw


def pool(x):
    Nz, Ny, Nx, c = x.shape
    return x.reshape(Nz // 2, 2, Ny // 2, 2, Nx // 2, 2, c).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def sum8(x):
    """Adjoint of up: the sum over the 8 children."""
    Nz, Ny, Nx, c = x.shape
    return x.reshape(Nz // 2, 2, Ny // 2, 2, Nx // 2, 2, c).sum(axis=(1, 3, 5))


def forward(theta, r, q, bf=False):
    """r, q: (Nz, Ny, Nx).  Returns (c, activations, layer inputs)."""
    R = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    p = unpack(theta)
    W = {n: (R(w), b) for n, (w, b) in p.items()}
    a, xin = {}, {}
    x = R(np.stack([r, q], axis=-1))
    for name, lev, c1, c2, co, relu in LAYERS:
        if name == "e0":
            xi = x
        elif name in ("e1", "e2", "e3"):
            prev = {"e1": "e0", "e2": "e1", "e3": "e2"}[name]
            a["p" + prev] = R(pool(a[prev]))
            xi = a["p" + prev]
        elif name == "c":
            xi = a["d0"]
        else:
            skip, low = {"d2": ("e2", "e3"), "d1": ("e1", "d2"), "d0": ("e0", "d1")}[name]
            xi = np.concatenate([a[skip], up(a[low])], axis=-1)
        xin[name] = xi
        y = conv(xi, *W[name])
        if name == "c":
            return y[..., 0], a, xin
        a[name] = R(np.maximum(y, 0.0))


def loss(theta, r, q, t, dz, bf=False):
    c, _, _ = forward(theta, r, q, bf)
    wz = np.asarray(dz)[:, None, None]
    return float(np.sum(wz * (c - t) ** 2) / np.sum(wz * t ** 2))


def backward(theta, r, q, t, dz, bf=False):
    """(L, dL/dtheta) for one sample, with the rounding model of ``bf``."""
    R = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    p = unpack(theta)
    W = {n: R(w) for n, (w, _) in p.items()}
    c, a, xin = forward(theta, r, q, bf)
    wz = np.asarray(dz)[:, None, None]
    den = np.sum(wz * t ** 2)
    L = float(np.sum(wz * (c - t) ** 2) / den)
    g = {}
    gc = (2.0 * wz * (c - t) / den)[..., None]
    g["c"] = (conv_dw(gc, xin["c"]), gc.sum(axis=(0, 1, 2)))
    m = lambda name: a[name] > 0.0
    gd = {"d0": R(conv_T(gc, W["c"]) * m("d0"))}
    dx = {}
    for name in ("d0", "d1", "d2"):
        gg = gd[name]
        g[name] = (conv_dw(gg, xin[name]), gg.sum(axis=(0, 1, 2)))
        dx[name] = conv_T(gg, W[name])
        c1 = dict((l[0], l[2]) for l in LAYERS)[name]
        low = {"d0": "d1", "d1": "d2", "d2": "e3"}[name]
        gd[low] = R(sum8(dx[name][..., c1:]) * m(low))
    for name, skip_of in (("e3", "d2"), ("e2", "d1"), ("e1", "d0")):
        gg = gd[name]
        g[name] = (conv_dw(gg, xin[name]), gg.sum(axis=(0, 1, 2)))
        dpool = conv_T(gg, W[name])
        prev = {"e3": "e2", "e2": "e1", "e1": "e0"}[name]
        c1 = dict((l[0], l[2]) for l in LAYERS)[skip_of]
        gd[prev] = R((dx[skip_of][..., :c1] + up(dpool) / 8.0) * m(prev))
    gg = gd["e0"]
    g["e0"] = (conv_dw(gg, xin["e0"]), gg.sum(axis=(0, 1, 2)))
    return L, pack(g)
