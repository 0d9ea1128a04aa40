"""GPU: the learned preconditioner's U-Net against its oracle, its training,
and its use in flexible PCG.

Forward values and parameter gradients of fw_net_* (bfloat16 weights and
activations, float32 accumulation on the matrix cores) are held to the
float64 oracle (oracle/unet.py, pinned by finite differences) run with the
same bfloat16 rounding points: the remaining differences are float32 against
float64 sums and the rare bfloat16 rounding flips they cause.  Training on a
learnable target must reduce the loss.  In flexible PCG the learned mode must
reach the same 1e-8 float64 residual as plain PCG, and the trained network
must take fewer iterations than plain PCG on held-out projection systems.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import projection as OP
from oracle import unet as U

pytestmark = pytest.mark.gpu

SHAPE = (16, 8, 32)   # (nz, ny, nx), multiples of 8; L3 = 2 x 1 x 4


def _theta(seed=1):
    rng = np.random.default_rng(seed)
    return (U.init(seed) + 0.02 * rng.standard_normal(U.n_params())).astype(np.float32)


def _tensors(seed=2, shape=SHAPE):
    import torch
    rng = np.random.default_rng(seed)
    r = rng.standard_normal(shape).astype(np.float32)
    q = (3.0 * rng.random(shape)).astype(np.float32)
    t = rng.standard_normal(shape).astype(np.float32)
    dz = 0.05 + 0.1 * rng.random(shape[0])
    T = lambda a: torch.as_tensor(a, device="cuda")
    return r, q, t, dz, T(r), T(q), T(t), T(dz)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", [SHAPE, (16, 32, 128)], ids=["split-k", "m-blocks"])
def test_net_forward_matches_oracle(shape):
    """At 16 x 8 x 32 every convolution takes the split-K kernel; at
    16 x 32 x 128 the finest levels take the multi-block kernel."""
    from firewebapp_amd.learned import ConvNet
    th = _theta()
    r, q, _, _, R, Q, _, _ = _tensors(shape=shape)
    net = ConvNet(shape[2], shape[1], shape[0], theta=th)
    c = net.forward(R, Q).double().cpu().numpy()
    want, _, _ = U.forward(th.astype(np.float64), r.astype(np.float64), q.astype(np.float64),
                           bf=True)
    exact, _, _ = U.forward(th.astype(np.float64), r.astype(np.float64),
                            q.astype(np.float64))
    err, err64 = _rel(c, want), _rel(c, exact)
    print(f"forward rel L2: vs bf16 oracle {err:.2e}, vs float64 oracle {err64:.2e}", flush=True)
    assert err <= 2e-3
    assert err64 <= 5e-2          # the bfloat16 network is the float64 one, rounded
    net.close()


def test_net_gradient_matches_oracle():
    from firewebapp_amd.learned import ConvNet
    th = _theta(3)
    r, q, t, dz, R, Q, Tt, DZ = _tensors(4)
    net = ConvNet(SHAPE[2], SHAPE[1], SHAPE[0], theta=th)
    L = net.train_step(R, Q, Tt, DZ, 0.0, 1)
    g = net.grad().astype(np.float64)
    L_ref, g_ref = U.backward(th.astype(np.float64), r.astype(np.float64),
                              q.astype(np.float64), t.astype(np.float64), dz, bf=True)
    print("loss", L, L_ref, flush=True)
    assert abs(L - L_ref) <= 2e-3 * L_ref
    o = 0
    for name, _, c1, c2, co, _ in U.LAYERS:
        n = 27 * (c1 + c2) * co + co
        rel = _rel(g[o:o + n], g_ref[o:o + n])
        print(name, "grad rel L2", f"{rel:.2e}", flush=True)
        assert rel <= 1e-2, name
        o += n
    assert np.array_equal(net.params, th)   # lr = 0 leaves the parameters
    net.close()


def test_training_reduces_loss():
    """A learnable target (a local smoothing of r modulated by q): Adam must
    fit it."""
    import torch
    from firewebapp_amd.learned import ConvNet, train
    rng = np.random.default_rng(7)
    shape = (16, 16, 32)
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
    hist = train(net, samples, dz, 300, lr=2e-3)
    first, last = np.mean(hist[:4]), np.mean(hist[-4:])
    print("loss", first, "->", last, flush=True)
    assert last < 0.3 * first
    net.close()


def _slab_case():
    import torch
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
    return (Nz, Ny, Nx), dz, da, db, rho, b


def test_learned_fpcg_reaches_tolerance():
    """FW_PRECOND_LEARNED (separable + network correction, flexible PCG)
    with a zero output layer is the separable
    preconditioner inside flexible PCG: it must reach the same float64 1e-8
    residual as plain PCG, within a couple of iterations."""
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, ptr, stream_ptr
    from firewebapp_amd.learned import ConvNet, attach
    from firewebapp_amd.physics_3d.fft_poisson_3d import create_handle
    (Nz, Ny, Nx), dz, da, db, rho, b = _slab_case()
    g = DeviceGrid(Nx, Ny, Nz, 0.1, 0.1, dz, da, db, dtype=torch.float32)
    R = torch.as_tensor(rho.astype(np.float32), device="cuda")
    B = torch.as_tensor(b, device="cuda")
    coef = OP.coefficients(rho.astype(np.float32).astype(np.float64), 0.1, 0.1, dz, da, db)
    h = create_handle(g, 1.0e-6, 1.0e-8, maxiter=2000)
    p = U.unpack(U.init(5))
    p["c"] = (0.0 * p["c"][0], 0.0 * p["c"][1])
    net = ConvNet(Nx, Ny, Nz, theta=U.pack(p).astype(np.float32))
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
      (sum dz (N - t)^2 / sum dz t^2, 1 for the method without it) drops;
    * in flexible PCG (the network on every 4th direction, the handle's
      default) every system reaches 1e-8 and the mean iteration count is
      below plain PCG's."""
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
    data = np.load(path)
    theta = data["theta"]
    assert theta.size == U.n_params()
    ri, w = c1_deck(4.5, 0.065, 97)
    sv = Spread3DSolver(Spread3DConfig.resolve(ri, wind_speed_m_s=w))
    systems = collect_systems(sv, 60, 12)
    nz, ny, nx = sv.shape
    net = ConvNet(nx, ny, nz, theta=theta)
    samples = make_samples(sv.dg, systems[:2])
    losses = [net.train_step(r, q, t, sv.dg.dz, 0.0, 1) for r, q, t in samples]
    print("held-out loss of the trained network", np.round(losses, 3), flush=True)
    assert np.mean(losses) < 0.9
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
        assert np.mean(its["learned"]) < np.mean(its["plain"])
    finally:
        _lib.lib().fw_proj_destroy(h)
        net.close()
        sv.close()
