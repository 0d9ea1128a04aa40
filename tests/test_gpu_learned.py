"""GPU: the learned preconditioner's network against its oracle, its
training, and its use in flexible PCG.

Forward values and parameter gradients of fw_net_* (float32) are held to the
float64 oracle (oracle/unet.py, pinned by finite differences) at float32
accuracy.  Training on a learnable target must reduce the loss.  In flexible
PCG the learned mode must reach the same 1e-8 float64 residual as plain PCG.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import projection as OP
from oracle import unet as U

pytestmark = pytest.mark.gpu

SHAPE = (6, 4, 10)   # (nz, ny, nx), all even; coarse grid 3 x 2 x 5


def _theta(seed=1):
    rng = np.random.default_rng(seed)
    return (U.init(seed) + 0.05 * rng.standard_normal(U.n_params())).astype(np.float32)


def _tensors(seed=2):
    import torch
    rng = np.random.default_rng(seed)
    r = rng.standard_normal(SHAPE).astype(np.float32)
    q = (3.0 * rng.random(SHAPE)).astype(np.float32)
    t = rng.standard_normal(SHAPE).astype(np.float32)
    dz = 0.05 + 0.1 * rng.random(SHAPE[0])
    T = lambda a: torch.as_tensor(a, device="cuda")
    return r, q, t, dz, T(r), T(q), T(t), T(dz)


def test_net_forward_matches_oracle():
    from firewebapp_amd.learned import ConvNet
    th = _theta()
    r, q, _, _, R, Q, _, _ = _tensors()
    net = ConvNet(SHAPE[2], SHAPE[1], SHAPE[0], theta=th)
    c = net.forward(R, Q).double().cpu().numpy()
    want, _ = U.forward(th.astype(np.float64), r.astype(np.float64), q.astype(np.float64))
    err = np.max(np.abs(c - want)) / np.max(np.abs(want))
    print("forward max rel err", err, flush=True)
    assert err <= 2e-5


def test_net_gradient_matches_oracle():
    from firewebapp_amd.learned import ConvNet
    th = _theta(3)
    r, q, t, dz, R, Q, Tt, DZ = _tensors(4)
    net = ConvNet(SHAPE[2], SHAPE[1], SHAPE[0], theta=th)
    L = net.train_step(R, Q, Tt, DZ, 0.0, 1)
    g = net.grad().astype(np.float64)
    L_ref, g_ref = U.backward(th.astype(np.float64), r.astype(np.float64),
                              q.astype(np.float64), t.astype(np.float64), dz)
    assert abs(L - L_ref) <= 1e-5 * L_ref
    o = 0
    for wn, bn, ci, co in U.LAYERS:
        n = co * ci * 27 + co
        a, b = g[o:o + n], g_ref[o:o + n]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(wn, "grad rel L2", rel, flush=True)
        assert rel <= 1e-4, wn
        o += n
    assert np.array_equal(net.params, th)   # lr = 0 leaves the parameters


def test_training_reduces_loss():
    """A learnable target (a smoothing of r modulated by q): Adam must fit it."""
    import torch
    from firewebapp_amd.learned import ConvNet, train
    rng = np.random.default_rng(7)
    shape = (8, 8, 16)
    samples = []
    for _ in range(4):
        r = rng.standard_normal(shape)
        q = rng.random(shape)
        t = 0.5 * (r + 0.25 * (np.roll(r, 1, 2) + np.roll(r, -1, 2) + np.roll(r, 1, 1) +
                               np.roll(r, -1, 1))) * (1.0 + q)
        T = lambda a: torch.as_tensor(a, dtype=torch.float32, device="cuda")
        samples.append((T(r), T(q), T(t)))
    dz = torch.full((shape[0],), 0.1, dtype=torch.float64, device="cuda")
    net = ConvNet(shape[2], shape[1], shape[0], seed=0)
    hist = train(net, samples, dz, 400, lr=3e-3)
    first, last = np.mean(hist[:4]), np.mean(hist[-4:])
    print("loss", first, "->", last, flush=True)
    assert last < 0.3 * first


def test_learned_fpcg_reaches_tolerance():
    """FW_PRECOND_LEARNED (separable + network correction, flexible PCG)
    solves to the same float64 1e-8 residual as plain PCG."""
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, ptr, stream_ptr
    from firewebapp_amd.learned import ConvNet, attach
    from firewebapp_amd.physics_3d.fft_poisson_3d import create_handle
    rng = np.random.default_rng(11)
    Nz, Ny, Nx = 16, 8, 32
    nb = 3
    dz = np.concatenate([np.full(nb, 0.042), np.full(Nz - nb, 0.1)])
    zc = np.cumsum(dz) - 0.5 * dz
    da, db = np.empty(Nz), np.empty(Nz)
    da[:-1] = zc[1:] - zc[:-1]
    da[-1] = dz[-1]
    db[1:] = zc[1:] - zc[:-1]
    db[0] = dz[0]
    x = np.arange(Nx)[None, None, :]
    z = np.arange(Nz)[:, None, None]
    hot = np.exp(-((x - Nx / 2) ** 2) / 18.0 - z ** 2 / 30.0)
    rho = 101325.0 / (8.314 / 0.029 * (300.0 + 1200.0 * hot * np.ones((Nz, Ny, Nx))))
    b = rng.standard_normal((Nz, Ny, Nx))
    g = DeviceGrid(Nx, Ny, Nz, 0.1, 0.1, dz, da, db, dtype=torch.float32)
    R = torch.as_tensor(rho.astype(np.float32), device="cuda")
    B = torch.as_tensor(b, device="cuda")
    coef = OP.coefficients(rho.astype(np.float32).astype(np.float64), 0.1, 0.1, dz, da, db)
    h = create_handle(g, 1.0e-6, 1.0e-8, maxiter=2000)
    # a network whose output layer is zero: the learned mode is then the
    # separable preconditioner inside flexible PCG, which must converge as
    # plain PCG does (to the same residual, within a couple of iterations)
    th = U.unpack(U.init(5))
    th["w4"][:] = 0.0
    net = ConvNet(Nx, Ny, Nz, theta=U.pack(th).astype(np.float32))
    its = {}
    try:
        for mode in ("plain", "learned"):
            attach(h, net if mode == "learned" else None)
            X = torch.zeros_like(B)
            it, rel = C.c_int(), C.c_double()
            _lib.call("fw_proj_solve", h, ptr(R), ptr(B), ptr(X), 1e-8, C.byref(it),
                      C.byref(rel), stream_ptr())
            true_rel = np.linalg.norm(b - OP.matvec(coef, X.cpu().numpy())) / np.linalg.norm(b)
            assert true_rel <= 1.05e-8, (mode, true_rel, it.value)
            its[mode] = it.value
        print("iterations to 1e-8: plain", its["plain"], "learned (zero correction)",
              its["learned"], flush=True)
        assert abs(its["learned"] - its["plain"]) <= 2
    finally:
        _lib.lib().fw_proj_destroy(h)
        net.close()


def test_trained_network_gain():
    """What the trained network (firewebapp_amd/data/precond_net.npz, trained
    on 128 x 32 x 32 runs) gains over the method without it, on held-out
    projection systems of a run with another wind, moisture and seed:
    * on its objective, the error left by the separable inverse
      (sum dz (N - t)^2 / sum dz t^2, 1 for the method without it) must drop;
    * in flexible PCG, every system still reaches 1e-8 (a stalled round hands
      the solve to plain PCG); the iteration counts of both are printed."""
    import os
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import ptr, stream_ptr
    from firewebapp_amd.learned import (ConvNet, attach, c1_deck, collect_systems,
                                        make_samples)
    from firewebapp_amd.physics_3d.fft_poisson_3d import create_handle
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    path = os.path.join(os.path.dirname(__file__), "..", "firewebapp_amd", "data",
                        "precond_net.npz")
    if not os.path.exists(path):
        pytest.skip("no trained parameters committed")
    theta = np.load(path)["theta"]
    ri, w = c1_deck(5.0, 0.06, 9)
    sv = Spread3DSolver(Spread3DConfig.resolve(ri, wind_speed_m_s=w))
    systems = collect_systems(sv, 24, 6)
    nz, ny, nx = sv.shape
    net = ConvNet(nx, ny, nz, theta=theta)
    samples = make_samples(sv.dg, systems)
    losses = [net.train_step(r, q, t, sv.dg.dz, 0.0, 1) for r, q, t in samples]
    print("held-out loss of the trained network", np.round(losses, 3), flush=True)
    assert max(losses) < 1.0 and np.mean(losses) < 0.95
    h = create_handle(sv.dg, 1.0e-6, 1.0e-8, maxiter=3000)
    its = {}
    try:
        for mode in ("plain", "learned"):
            attach(h, net if mode == "learned" else None)
            its[mode] = []
            for rho, b in systems:
                X = torch.zeros_like(b)
                it, rel = C.c_int(), C.c_double()
                _lib.call("fw_proj_solve", h, ptr(rho), ptr(b), ptr(X), 1e-8, C.byref(it),
                          C.byref(rel), stream_ptr())
                assert rel.value <= 1e-8
                its[mode].append(it.value)
        fb = C.c_int()
        _lib.call("fw_proj_learned_stats", h, C.byref(fb))
        print("iterations to 1e-8: plain", its["plain"], "learned", its["learned"],
              "fallbacks", fb.value, flush=True)
    finally:
        _lib.lib().fw_proj_destroy(h)
        net.close()
        sv.close()
