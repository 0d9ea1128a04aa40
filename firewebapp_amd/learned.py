"""Learned correction of the separable pressure preconditioner.

The reference preconditions its Krylov solve with the separable
constant-density inverse only (projection_3d.py:756-791).  This module adds
a learned correction and the pipeline that trains it on the GPU with the
library's own kernels (csrc/unet.hip):

* ``ConvNet``: the 4-level 3-D U-Net (16/32/64/64 channels, 3x3x3 kernels,
  bfloat16 weights and activations, float32 accumulation on the matrix
  cores) behind fw_net_*;
* ``collect_systems``: runs the step on seeded decks and keeps the
  projection systems (rho, b) it hands to its Krylov solver (fw_proj_rhs);
* ``make_samples``: for each system, the exact error e = x* - x_k of the
  PCG iterate at a sequence of tolerances (x* solved to 1e-10) and its
  residual r = A e; the training pair is (r / s, (e - M_sep r) / s),
  s = RMS(r): the network learns what the separable inverse misses on the
  residuals a Krylov solve actually meets, early and late, from r and the
  density feature q = RHO_REF / rho - 1;
* ``train``: Adam on those pairs (fw_net_train_step);
* ``attach``: FW_PRECOND_LEARNED on a projection handle, z = M_sep r +
  s N(r / s, q) inside flexible PCG.

Run ``python -m firewebapp_amd.learned --out <file.npz>`` on a GPU to train
and save the parameters.
"""
from __future__ import annotations

import argparse
import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from .device import ptr, stream_ptr

PRECOND_SEPARABLE, PRECOND_JACOBI, PRECOND_LEARNED = 1, 2, 3
# (level, same-level input channels, coarse input channels, output channels)
# of the network's 8 convolutions: e0 e1 e2 e3 d2 d1 d0 c
LAYERS = ((0, 2, 0, 16), (1, 16, 0, 32), (2, 32, 0, 64), (3, 64, 0, 64),
          (2, 64, 64, 64), (1, 32, 64, 32), (0, 16, 32, 16), (0, 16, 0, 1))
# PCG tolerances whose iterates give the training residuals of a system
SAMPLE_RTOLS = (1.0, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6)


def init_params(seed=0):
    """He-style random weights w (27, ci, co), zero biases, a small output
    layer; one flat float32 vector in layer order."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (_, c1, c2, co) in enumerate(LAYERS):
        ci = c1 + c2
        w = rng.standard_normal((27, ci, co)) * np.sqrt(2.0 / (27 * ci))
        if li == len(LAYERS) - 1:
            w *= 0.1
        out += [w.ravel(), np.zeros(co)]
    theta = np.concatenate(out).astype(np.float32)
    assert theta.size == _lib.lib().fw_net_nparams()
    return theta


class ConvNet:
    """fw_net_* handle on an (nz, ny, nx) grid (all dims multiples of 8)."""

    def __init__(self, nx, ny, nz, theta=None, seed=0):
        self.shape = (nz, ny, nx)
        th = np.ascontiguousarray(init_params(seed) if theta is None else theta, np.float32)
        h = C.c_void_p()
        _lib.call("fw_net_create", int(nx), int(ny), int(nz),
                  th.ctypes.data_as(C.c_void_p), C.byref(h))
        self.h = h

    @property
    def params(self):
        th = np.empty(_lib.lib().fw_net_nparams(), np.float32)
        _lib.call("fw_net_get_params", self.h, th.ctypes.data_as(C.c_void_p))
        return th

    @params.setter
    def params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        _lib.call("fw_net_set_params", self.h, th.ctypes.data_as(C.c_void_p))

    def forward(self, r, q):
        c = torch.empty_like(r)
        _lib.call("fw_net_forward", self.h, ptr(r), ptr(q), ptr(c), stream_ptr())
        return c

    def train_step(self, r, q, t, dz, lr, it):
        loss = C.c_double()
        _lib.call("fw_net_train_step", self.h, ptr(r), ptr(q), ptr(t), ptr(dz), float(lr),
                  int(it), C.byref(loss), stream_ptr())
        return loss.value

    def grad(self):
        g = np.empty(_lib.lib().fw_net_nparams(), np.float32)
        _lib.call("fw_net_grad_copy", self.h, g.ctypes.data_as(C.c_void_p))
        return g

    def close(self):
        if self.h:
            _lib.lib().fw_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def attach(proj_handle, net):
    """Use ``net`` as the learned correction of ``proj_handle`` (same grid)."""
    _lib.call("fw_proj_set_learned", proj_handle, net.h if net is not None else None)
    _lib.call("fw_proj_set_preconditioner", proj_handle,
              PRECOND_LEARNED if net is not None else PRECOND_SEPARABLE)


def c1_deck(wind, moisture, seed):
    """The 128 x 32 x 32 S4 production deck (n_z_bed 2 of 0.105 m, 30 air
    cells of 0.1 m), float32, with the given wind and moisture."""
    from .deck import RomInputs
    ovr = dict(
        lx=12.8, ly=3.2, lz=3.2, dx=0.1, dy=0.1, n_z_bed=2, wall_bl_n=1,
        wall_bl_first_dz=0.105, wall_bl_growth=1.0, atm_uniform_dz=0.1,
        dom_quadrature=4, initial_moisture_frac=moisture, fuel_depth_m=0.21,
        bulk_density_kg_m3=1.4762, sav_ratio_1_m=9770.0, bed_x_start=1.0,
        turbulence_model="k_epsilon", combustion_closure="edc", radiation_solver="dom",
        dom_subcycle_every=5, projection_method="fft_pcg", lagrangian_bed_enable=True,
        lagrangian_bed_n_per_cell=4, lagrangian_bed_drying_mode="combined",
        level_set_passive=True, solid_phase_ignition_enable=True,
        particle_jitter=0.10, particle_jitter_seed=seed, min_dt_s=1.0e-4,
        max_wall_time_s=1.0e9, dtype="float32")
    return RomInputs(outdoor_overrides=ovr), wind


def collect_systems(solver, n_steps, every=1):
    """Advance ``solver`` n_steps, keeping (rho, b) of the last projection of
    every ``every``-th step (float32 rho, float64 b, device)."""
    out = []
    for s in range(n_steps):
        if solver.step() is not None:
            break
        if (s + 1) % every == 0:
            b = torch.empty(solver.shape, dtype=torch.float64, device="cuda")
            _lib.call("fw_proj_rhs", solver.proj, ptr(b), stream_ptr())
            out.append((solver.f["rho"].clone(), b))
    return out


def density_feature(rho):
    """The network's density input q = RHO_REF / max(rho, 0.01) - 1 (float32)."""
    q = torch.empty_like(rho)
    _lib.call("fw_net_density_feature", ptr(rho), ptr(q), int(rho.numel()), stream_ptr())
    return q


def _solve(h, rho, b, rtol):
    x = torch.zeros_like(b)
    it, rel = C.c_int(), C.c_double()
    rc = _lib.lib().fw_proj_solve(h, ptr(rho), ptr(b), ptr(x), float(rtol), C.byref(it),
                                  C.byref(rel), stream_ptr())
    _lib.check(rc, "fw_proj_solve")
    return x, it.value


def make_samples(dgrid, systems, rtols=SAMPLE_RTOLS, keep=None):
    """(r / s, q, (e - M_sep r) / s) float32 triples: for each system and
    each tolerance, e = x* - x_tol (x* to 1e-10) and r = A e."""
    from .physics_3d.fft_poisson_3d import create_handle
    h = create_handle(dgrid, 1.0e-6, 1.0e-10, maxiter=4000)
    out = []
    try:
        for rho, b in systems:
            if not float(torch.linalg.vector_norm(b)) > 0.0:
                continue
            xs, _ = _solve(h, rho, b, 1.0e-10)
            q = density_feature(rho)
            for rt in rtols:
                xk = torch.zeros_like(b) if rt >= 1.0 else _solve(h, rho, b, rt)[0]
                e32 = (xs - xk).float().contiguous()
                r32 = torch.empty_like(e32)
                _lib.call("fw_proj_matvec", h, ptr(rho), ptr(e32), ptr(r32), stream_ptr())
                s = float(torch.sqrt(torch.mean(r32.double() ** 2)))
                if not s > 0.0:
                    continue
                zs = torch.empty_like(r32)
                _lib.call("fw_proj_precondition", h, ptr(r32), ptr(zs), stream_ptr())
                out.append(((r32 / s).contiguous(), q, ((e32 - zs) / s).contiguous()))
            if keep is not None and len(out) >= keep:
                break
    finally:
        _lib.lib().fw_proj_destroy(h)
    return out


def train(net, samples, dz, steps, lr=1.0e-3, log_every=0, lr_final=None, seed=0):
    """Adam over the samples in shuffled epochs; returns the loss history.
    With lr_final the rate decays geometrically from lr to lr_final."""
    hist = []
    rng = np.random.default_rng(seed)
    order = []
    for it in range(1, steps + 1):
        if not order:
            order = list(rng.permutation(len(samples)))
        r, q, t = samples[order.pop()]
        rate = lr if lr_final is None else lr * (lr_final / lr) ** ((it - 1) / max(steps - 1, 1))
        L = net.train_step(r, q, t, dz, rate, it)
        hist.append(L)
        if log_every and it % log_every == 0:
            print(f"# step {it}: loss {np.mean(hist[-log_every:]):.4f}", flush=True)
    return hist


# seeded 128 x 32 x 32 runs of the training set: (U10 m/s, moisture, seed),
# U10 1-7 m/s and moisture 2-12 %
TRAIN_RUNS = tuple((float(u), float(m), int(s)) for s, (u, m) in enumerate(
    [(1, 0.02), (2, 0.10), (3, 0.04), (4, 0.12), (5, 0.06), (6, 0.08), (7, 0.03),
     (1.5, 0.11), (2.5, 0.05), (3.5, 0.09), (4.5, 0.02), (5.5, 0.07), (6.5, 0.10),
     (7, 0.12), (3, 0.07), (5, 0.11)], start=1))


def scale_output(theta, tau):
    """Parameters of the network whose output is tau times theta's (the
    output layer is linear: its weights and bias scale)."""
    th = np.array(theta, np.float32, copy=True)
    o = 0
    for li, (_, c1, c2, co) in enumerate(LAYERS):
        n = 27 * (c1 + c2) * co + co
        if li == len(LAYERS) - 1:
            th[o:o + n] *= tau
        o += n
    return th


def target_scale(samples):
    """RMS of the targets over the samples: the network is trained on
    targets divided by it (unit scale), and its output layer is multiplied
    by it afterwards."""
    ss = [float(torch.mean(t.double() ** 2)) for _, _, t in samples]
    return float(np.sqrt(np.mean(ss)))


def train_c1(out_path, steps=20000, runs=TRAIN_RUNS, n_steps=180, every=1, lr=1.0e-3,
             lr_final=1.0e-4, seed=0, max_samples=20000):
    """Train on projection systems of seeded 128 x 32 x 32 runs; save the
    parameters and the training record to ``out_path`` (npz)."""
    from .spread_3d import Spread3DConfig, Spread3DSolver
    t0 = time.perf_counter()
    samples, dgrid = [], None
    for wind, moist, sd in runs:
        ri, w = c1_deck(wind, moist, sd)
        sv = Spread3DSolver(Spread3DConfig.resolve(ri, wind_speed_m_s=w))
        systems = collect_systems(sv, n_steps, every)
        dgrid = sv.dg
        samples += make_samples(sv.dg, systems)
        print(f"# U10 {wind} M {moist}: {len(systems)} systems, {len(samples)} samples "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
        sv.close()
        if len(samples) >= max_samples:
            break
    samples = samples[:max_samples]
    tau = target_scale(samples)
    samples = [(r, q, (t / tau).contiguous()) for r, q, t in samples]
    nz, ny, nx = samples[0][0].shape
    net = ConvNet(nx, ny, nz, seed=seed)
    t1 = time.perf_counter()
    hist = train(net, samples, dgrid.dz, steps, lr, log_every=max(steps // 20, 1),
                 lr_final=lr_final, seed=seed)
    th = scale_output(net.params, tau)
    np.savez(out_path, theta=th, loss=np.asarray(hist, np.float32), n_samples=len(samples),
             tau=tau, grid=np.asarray([nx, ny, nz]), data_seconds=t1 - t0,
             train_seconds=time.perf_counter() - t1)
    print(f"# trained on {len(samples)} pairs, loss {np.mean(hist[:50]):.4f} -> "
          f"{np.mean(hist[-200:]):.4f} ({time.perf_counter() - t1:.1f} s), saved {out_path}",
          flush=True)
    return th, hist


def forward_flops(nx, ny, nz):
    """Multiply-add flops (2 per MAC) of one network application."""
    return sum(2 * 27 * (c1 + c2) * co * (nx >> lv) * (ny >> lv) * (nz >> lv)
               for lv, c1, c2, co in LAYERS)


def time_forward(nx, ny, nz, theta=None, reps=50):
    """Mean wall time of one application on a (nz, ny, nx) grid (HIP events
    around reps back-to-back applications) and the achieved TFLOP/s."""
    net = ConvNet(nx, ny, nz, theta=theta)
    r = torch.randn((nz, ny, nx), device="cuda")
    q = torch.rand((nz, ny, nx), device="cuda")
    net.forward(r, q)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        net.forward(r, q)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    net.close()
    return ms, forward_flops(nx, ny, nz) / (ms * 1e-3) / 1e12


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned preconditioner")
    ap.add_argument("--out", help="train and save the parameters here (npz)")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--max-samples", type=int, default=20000)
    ap.add_argument("--runs", type=int, default=len(TRAIN_RUNS))
    ap.add_argument("--time-forward", metavar="NXxNYxNZ",
                    help="time the network's application on this grid instead")
    a = ap.parse_args(argv)
    if a.time_forward:
        nx, ny, nz = (int(v) for v in a.time_forward.split("x"))
        ms, tf = time_forward(nx, ny, nz)
        print(f"# network on {nx}x{ny}x{nz}: {ms:.3f} ms per application, "
              f"{forward_flops(nx, ny, nz) / 1e9:.2f} GFLOP, {tf:.1f} TFLOP/s "
              f"({100 * tf / 2500.0:.2f} % of the 2.5 PFLOP/s bf16 matrix peak)", flush=True)
        return 0
    if not a.out:
        ap.error("--out or --time-forward is required")
    train_c1(a.out, steps=a.steps, runs=TRAIN_RUNS[:a.runs], max_samples=a.max_samples)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
