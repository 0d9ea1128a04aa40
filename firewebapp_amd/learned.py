"""Learned correction of the separable pressure preconditioner.

The reference preconditions its Krylov solve with the separable
constant-density inverse only (projection_3d.py:756-791).  This module adds
a learned correction and the pipeline that trains it on the GPU with the
library's own kernels (csrc/learned.hip):

* ``ConvNet``: the two-level 3-D convolutional network (float32, y periodic,
  zero padding in x and z) behind fw_net_*;
* ``collect_systems``: runs the step on seeded decks and keeps the
  projection systems (rho, b) it hands to its Krylov solver (fw_proj_rhs);
* ``make_samples``: solves each system to 1e-8 with plain PCG and forms the
  training pair (r / s, (A^-1 r - M_sep r) / s), s = RMS(r): the network
  learns what the separable inverse misses, from r and the density feature
  q = RHO_REF / rho - 1;
* ``train``: Adam on those pairs (fw_net_train_step);
* ``attach``: FW_PRECOND_LEARNED on a projection handle, z = M_sep r +
  s N(r / s) inside flexible PCG.

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


def init_params(seed=0):
    """He-style random weights, zero biases, small output layer."""
    rng = np.random.default_rng(seed)
    layers = ((2, 8), (8, 16), (16, 16), (24, 8), (8, 1))
    out = []
    for li, (ci, co) in enumerate(layers):
        w = rng.standard_normal((co, ci, 27)) * np.sqrt(2.0 / (27 * ci))
        if li == len(layers) - 1:
            w *= 0.1
        out += [w.ravel(), np.zeros(co)]
    theta = np.concatenate(out).astype(np.float32)
    assert theta.size == _lib.lib().fw_net_nparams()
    return theta


class ConvNet:
    """fw_net_* handle on an (nz, ny, nx) grid (all dims even)."""

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


def make_samples(dgrid, systems, rtol=1.0e-8):
    """(r / s, q, (A^-1 r - M_sep r) / s) float32 triples for each system."""
    from .physics_3d.fft_poisson_3d import create_handle
    h = create_handle(dgrid, 1.0e-6, rtol, maxiter=2000)
    out = []
    try:
        for rho, b in systems:
            x = torch.zeros_like(b)
            it, rel = C.c_int(), C.c_double()
            _lib.call("fw_proj_solve", h, ptr(rho), ptr(b), ptr(x), rtol, C.byref(it),
                      C.byref(rel), stream_ptr())
            r32 = b.float()
            zs = torch.empty_like(r32)
            _lib.call("fw_proj_precondition", h, ptr(r32), ptr(zs), stream_ptr())
            s = float(torch.sqrt(torch.mean(b * b)))
            if not s > 0.0:
                continue
            out.append(((r32 / s).contiguous(), density_feature(rho),
                        ((x - zs.double()) / s).float().contiguous()))
    finally:
        _lib.lib().fw_proj_destroy(h)
    return out


def train(net, samples, dz, steps, lr=1.0e-3, log_every=0):
    """Adam over the samples in a fixed cyclic order; returns the loss history."""
    hist = []
    for it in range(1, steps + 1):
        r, q, t = samples[(it - 1) % len(samples)]
        L = net.train_step(r, q, t, dz, lr, it)
        hist.append(L)
        if log_every and it % log_every == 0:
            print(f"# step {it}: loss {np.mean(hist[-log_every:]):.4f}", flush=True)
    return hist


def train_c1(out_path, steps=2000, seeds=((4.0, 0.04, 1), (2.0, 0.08, 2), (6.0, 0.12, 3),
                                          (1.0, 0.02, 4)),
             n_steps=60, every=3, lr=1.0e-3, seed=0):
    """Train on projection systems of seeded 128 x 32 x 32 runs; save the
    parameters and the training record to ``out_path`` (npz)."""
    from .spread_3d import Spread3DConfig, Spread3DSolver
    t0 = time.perf_counter()
    samples, dgrid = [], None
    for wind, moist, sd in seeds:
        ri, w = c1_deck(wind, moist, sd)
        cfg = Spread3DConfig.resolve(ri, wind_speed_m_s=w)
        sv = Spread3DSolver(cfg)
        systems = collect_systems(sv, n_steps, every)
        samples += make_samples(sv.dg, systems)
        dgrid = sv.dg
        print(f"# U10 {wind} M {moist}: {len(systems)} systems", flush=True)
        sv.close()
    # mix the runs in the cyclic order
    samples = [samples[i] for i in np.random.default_rng(seed).permutation(len(samples))]
    nz, ny, nx = samples[0][0].shape
    net = ConvNet(nx, ny, nz, seed=seed)
    hist = train(net, samples, dgrid.dz, steps, lr, log_every=max(steps // 10, 1))
    th = net.params
    np.savez(out_path, theta=th, loss=np.asarray(hist), n_samples=len(samples),
             grid=np.asarray([nx, ny, nz]), seconds=time.perf_counter() - t0)
    print(f"# trained on {len(samples)} systems, loss {hist[0]:.4f} -> "
          f"{np.mean(hist[-50:]):.4f}, saved {out_path}", flush=True)
    return th, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned preconditioner")
    ap.add_argument("--out", required=True)
    ap.add_argument("--steps", type=int, default=2000)
    a = ap.parse_args(argv)
    train_c1(a.out, steps=a.steps)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
