"""The network oracle of the learned preconditioner, pinned by mathematics.

There is no reference counterpart (the reference preconditions with the
separable inverse only), so the oracle is checked by what fixes it: closed
forms of the building blocks, the adjoint identity of the convolution, and
central finite differences of the loss for every parameter block.
"""
import numpy as np
import pytest

from oracle import unet as U


def _case(seed=3, shape=(4, 6, 8)):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal(shape)
    q = rng.random(shape) * 3.0
    t = rng.standard_normal(shape)
    dz = 0.05 + 0.1 * rng.random(shape[0])
    theta = U.init(seed) + 0.05 * rng.standard_normal(U.n_params())
    return theta, r, q, t, dz


def test_conv_closed_forms():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 6, 8))
    w = np.zeros((2, 2, 27))
    w[0, 0, 13] = w[1, 1, 13] = 1.0           # centre tap: copy
    assert np.array_equal(U.conv3(x, w, np.zeros(2)), x)
    w = np.zeros((1, 2, 27))
    w[0, 1, 9 + 3 * 2 + 1] = 1.0              # (dk, dj, di) = (0, +1, 0): periodic y
    assert np.array_equal(U.conv3(x, w, np.zeros(1))[0], np.roll(x[1], -1, axis=1))
    w = np.zeros((1, 1, 27))
    w[0, 0, 9 + 3 + 2] = 1.0                  # (0, 0, +1): zero past the last x
    y = U.conv3(x[:1], w, np.zeros(1))[0]
    assert np.array_equal(y[:, :, :-1], x[0, :, :, 1:]) and not y[:, :, -1].any()
    z = rng.standard_normal((3, 2, 3, 4))
    assert np.array_equal(U.pool2(U.up2(z)), z)


def test_conv_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 4, 6, 8))
    g = rng.standard_normal((5, 4, 6, 8))
    w = rng.standard_normal((5, 3, 27))
    lhs = np.sum(U.conv3(x, w, np.zeros(5)) * g)
    rhs = np.sum(x * U.conv3_T(g, w))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gradient_finite_differences():
    theta, r, q, t, dz = _case()
    L, g = U.backward(theta, r, q, t, dz)
    assert abs(L - U.loss(theta, r, q, t, dz)) <= 1e-14 * L
    rng = np.random.default_rng(5)
    o = 0
    for wn, bn, ci, co in U.LAYERS:
        nw = co * ci * 27
        for idx in list(rng.choice(nw, 4, replace=False) + o) + [o + nw]:
            h = 1e-6
            tp, tm = theta.copy(), theta.copy()
            tp[idx] += h
            tm[idx] -= h
            fd = (U.loss(tp, r, q, t, dz) - U.loss(tm, r, q, t, dz)) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1e-3, abs(g[idx])), (wn, idx, fd, g[idx])
        o += nw + co


def test_pack_roundtrip():
    theta = U.init(2)
    assert np.array_equal(U.pack(U.unpack(theta)), theta)
    assert theta.size == U.n_params() == 16249
