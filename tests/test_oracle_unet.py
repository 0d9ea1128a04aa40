"""The network oracle of the learned preconditioner, pinned by mathematics.

There is no reference counterpart (the reference preconditions with the
separable inverse only), so the oracle is checked by what fixes it: closed
forms of the building blocks, the adjoint identities of the convolution and
of pooling/upsampling, and central finite differences of the loss for every
layer's weights and biases.  The bfloat16 rounding helper is checked against
known bit patterns.
"""
import numpy as np

from oracle import unet as U

SHAPE = (8, 8, 16)   # L3 = 1 x 1 x 2


def _case(seed=3, shape=SHAPE):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal(shape)
    q = rng.random(shape) * 3.0
    t = rng.standard_normal(shape)
    dz = 0.05 + 0.1 * rng.random(shape[0])
    theta = U.init(seed) + 0.02 * rng.standard_normal(U.n_params())
    return theta, r, q, t, dz


def test_conv_closed_forms():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((4, 6, 8, 2))
    w = np.zeros((27, 2, 2))
    w[13, 0, 0] = w[13, 1, 1] = 1.0                # centre tap: copy
    assert np.array_equal(U.conv(x, w), x)
    w = np.zeros((27, 2, 1))
    w[9 + 3 * 2 + 1, 1, 0] = 1.0                   # (dk, dj, di) = (0, +1, 0): periodic y
    assert np.array_equal(U.conv(x, w)[..., 0], np.roll(x[..., 1], -1, axis=1))
    w = np.zeros((27, 1, 1))
    w[9 + 3 + 2, 0, 0] = 1.0                       # (0, 0, +1): zero past the last x
    y = U.conv(x[..., :1], w)[..., 0]
    assert np.array_equal(y[:, :, :-1], x[:, :, 1:, 0]) and not y[:, :, -1].any()
    w = np.zeros((27, 1, 1))
    w[2 * 9 + 4, 0, 0] = 1.0                       # (+1, 0, 0): zero above the top
    y = U.conv(x[..., :1], w)[..., 0]
    assert np.array_equal(y[:-1], x[1:, :, :, 0]) and not y[-1].any()
    z = rng.standard_normal((2, 3, 4, 5))
    np.testing.assert_allclose(U.pool(U.up(z)), z, rtol=1e-15, atol=0)
    np.testing.assert_allclose(U.sum8(U.up(z)), 8.0 * z, rtol=1e-15, atol=0)


def test_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 6, 8, 3))
    g = rng.standard_normal((4, 6, 8, 5))
    w = rng.standard_normal((27, 3, 5))
    lhs = np.sum(U.conv(x, w) * g)
    assert abs(lhs - np.sum(x * U.conv_T(g, w))) <= 1e-12 * abs(lhs)
    # conv_dw is the gradient of <conv(x, w), g> in w
    assert abs(lhs - np.sum(w * U.conv_dw(g, x))) <= 1e-12 * abs(lhs)
    z = rng.standard_normal((2, 3, 4, 5))
    y = rng.standard_normal((4, 6, 8, 5))
    assert abs(np.sum(U.up(z) * y) - np.sum(z * U.sum8(y))) <= 1e-12 * abs(np.sum(z * U.sum8(y)))


def test_gradient_finite_differences():
    theta, r, q, t, dz = _case()
    L, g = U.backward(theta, r, q, t, dz)
    assert abs(L - U.loss(theta, r, q, t, dz)) <= 1e-14 * L
    rng = np.random.default_rng(5)
    o = 0
    for name, _, c1, c2, co, _ in U.LAYERS:
        nw = 27 * (c1 + c2) * co
        for idx in list(rng.choice(nw, 3, replace=False) + o) + [o + nw]:
            h = 1e-6
            tp, tm = theta.copy(), theta.copy()
            tp[idx] += h
            tm[idx] -= h
            fd = (U.loss(tp, r, q, t, dz) - U.loss(tm, r, q, t, dz)) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-5 * max(1e-4, abs(g[idx])), (name, idx, fd, g[idx])
        o += nw + co


def test_bf16_rounding_and_model():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins;
    # 1 + 3 * 2^-8 rounds up to 1 + 2^-6
    assert U.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert U.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert U.bf16(np.array([-3.0]))[0] == -3.0
    theta, r, q, t, dz = _case(4)
    c64, _, _ = U.forward(theta, r, q)
    c16, a, _ = U.forward(theta, r, q, bf=True)
    # every stored activation is a bfloat16 value
    for v in a.values():
        assert np.array_equal(U.bf16(v), v)
    assert np.linalg.norm(c16 - c64) <= 0.05 * np.linalg.norm(c64)
    L16, g16 = U.backward(theta, r, q, t, dz, bf=True)
    L64, g64 = U.backward(theta, r, q, t, dz)
    assert np.linalg.norm(g16 - g64) <= 0.1 * np.linalg.norm(g64)


def test_pack_roundtrip():
    theta = U.init(2)
    assert np.array_equal(U.pack(U.unpack(theta)), theta)
    assert theta.size == U.n_params() == 506161
