"""Throughput of the Tier-3 spread outer step on one MI355X.

Workload (default): 256 x 64 x 48 cells at dx = dy = 0.1 m (25.6 x 6.4 x
4.51 m), float32 grid fields, grass bed 0.21 m deep in 5 layers from
x = 1 m (bulk density 1.4762 kg/m^3 = 0.31 kg/m^2, SAV 9770 1/m), moisture
drawn U(2 %, 10 %) from a seeded PCG64, 4 particles per bed cell with seeded
+-10 % load and depth jitter (314,880 particles), log-law inflow U10 = 4 m/s,
k-epsilon,
EDC, DOM S8 (80 directions) every 5th step, fft-PCG projection (rtol 1e-6,
outer divergence tolerance 1e-3), passive level set, solid-phase ignition.

One "step" is one outer step of the reference loop (spread_3d.py:1897-3683)
with its 10 chemistry/transport sub-steps.  ``value`` = cells x steps / wall
seconds over exactly K timed steps after W warm-up steps, inputs resident in
HBM.  The JSON line also carries:
* roofline: the dominant kernel's algorithmic bytes per launch / its mean
  launch time over the timed steps (every 4th launch), measured with HIP
  events that the launch itself records (hipExtLaunchKernel, on the stream
  the kernel is launched on: torch's current stream), against 8 TB/s;
  traffic is the HBM bytes per launch from the FETCH_SIZE and WRITE_SIZE
  counters;
* step_roofline: the whole step against the HBM roofline, with the
  phase-fused compulsory traffic per cell-step (each phase's distinct
  N-sized fields read once and written once, particles 72 B each);
* pcg_time_to_1e8: the second half of the metric.  The projection systems
  of the next few steps (b = (div - target)/dt and rho, as the step hands
  them to its Krylov solver) are solved again from x0 = 0 to
  ||b - A x||_2 / ||b||_2 <= 1e-8 (float64 residual, float32 inner PCG with
  iterative refinement): plain PCG (the separable preconditioner),
  Jacobi-PCG and the learned preconditioner (flexible PCG with the trained
  network's correction, handing a solve to plain PCG when a round stalls),
  iterations and wall time per system;
* cpu_baseline: the oracle's float64 NumPy step on one host core, timed on a
  bounded sample of the same workload (rank 0, N = 1 only).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "3D cell-steps/s at 256x64x48; learned vs plain PCG time-to-1e-8 residual"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # 256 x 64 x 48: n_z_bed 5 (dz 0.042) + 43 air cells of 0.1 m
    "spread3d_256x64x48_s8": dict(
        lx=25.6, ly=6.4, lz=4.51, dx=0.1, dy=0.1, n_z_bed=5, wall_bl_n=1,
        wall_bl_first_dz=0.042, wall_bl_growth=1.0, atm_uniform_dz=0.1,
        dom_quadrature=8, moisture_min=0.02, moisture_max=0.10, moisture_seed=1,
        particle_jitter=0.10, particle_jitter_seed=7),
    # 512 x 128 x 64: a lognormal fuel depth field (mean 0.21 m, CV 0.3,
    # correlation length 1 m) in 16 bed layers spanning 0.42 m (dz 0.02625;
    # deeper columns clipped) + 48 air cells of 0.1 m; U10 = 6 m/s with
    # +-15 % gusts at a 5 s period
    "spread3d_512x128x64_s8": dict(
        lx=51.2, ly=12.8, lz=5.22, dx=0.1, dy=0.1, n_z_bed=16, wall_bl_n=1,
        wall_bl_first_dz=0.02625, wall_bl_growth=1.0, atm_uniform_dz=0.1,
        dom_quadrature=8, moisture_min=0.02, moisture_max=0.10, moisture_seed=5,
        particle_jitter=0.10, particle_jitter_seed=11, fuel_depth_cv=0.3,
        fuel_depth_corr_length_m=1.0, fuel_depth_seed=3, fuel_depth_max_m=0.42,
        wind_gust_amplitude=0.15, wind_gust_period_s=5.0, wind_gust_seed=9),
    # 128 x 32 x 32: n_z_bed 2 (dz 0.105) + 30 air cells of 0.1 m
    "spread3d_128x32x32_s4": dict(
        lx=12.8, ly=3.2, lz=3.2, dx=0.1, dy=0.1, n_z_bed=2, wall_bl_n=1,
        wall_bl_first_dz=0.105, wall_bl_growth=1.0, atm_uniform_dz=0.1,
        dom_quadrature=4, initial_moisture_frac=0.04),
}
COMMON = dict(
    fuel_depth_m=0.21, bulk_density_kg_m3=1.4762, sav_ratio_1_m=9770.0,
    initial_moisture_frac=0.04, bed_x_start=1.0, turbulence_model="k_epsilon",
    combustion_closure="edc", radiation_solver="dom", dom_subcycle_every=5,
    projection_method="fft_pcg", lagrangian_bed_enable=True,
    lagrangian_bed_n_per_cell=4, lagrangian_bed_drying_mode="combined",
    lagrangian_bed_h_conv=250.0, level_set_passive=True,
    solid_phase_ignition_enable=True, solid_phase_ignition_t_s_k=1000.0,
    ignition_duration_s=3.0, ignition_q_mult=3.0, ignition_width_mult=3.0,
    min_dt_s=1.0e-4, max_wall_time_s=1.0e9)
WIND = {"spread3d_512x128x64_s8": 6.0}


def wind(name):
    """U10 of a workload (4 m/s unless it names another)."""
    return WIND.get(name, 4.0)


# the roofline kernel and its algorithmic fields per cell per launch:
# k_species3 reads Y_fuel, Y_O2, Y_H2O, rho, u, v, w, S_pyro part, S_dry part
# and writes the three species
ROOFLINE_KERNEL = ("k_species3", 12)
# HBM bytes per k_species3 launch at the default workload (float32), from the
# FETCH_SIZE + WRITE_SIZE counters scaled by a 256 MiB copy:
# profiles/step_256x64x48_s8_hbm_traffic.md
ROOFLINE_TRAFFIC = {("spread3d_256x64x48_s8", "float32"): 44060018}


# Phase-fused compulsory traffic of one outer step in field-passes (each
# phase's distinct N-sized fields read once plus each field written once),
# production configuration: 10 chemistry sub-steps of EDC 11, species x3 12,
# turbulent diffusion x3 7, gas update + solid conduction 11, gas
# turbulent diffusion 3; projection 8 + 26 per Krylov iteration + 12; DOM
# (9 + 5 M per source iteration + 6) once every dom_every steps.
STEP_PHASES = {"dt reduction": 4, "EoS": 2, "bed grid I/O": 8, "horizontal conduction": 3,
               "drag + momentum": 9, "k-epsilon": 11, "level set": 6,
               "gas-energy advection": 5, "diagnostics + EoS": 9}
SUBSTEP_PASSES = 11 + 12 + 7 + 11 + 3
PARTICLE_BYTES = 72
I_CG, I_DOM = 2, 2     # Krylov and DOM iterations the table is frozen at


def step_bytes_per_cell(esz, n_sub, M, dom_every, particles, cells):
    """Algorithmic HBM bytes per cell-step of the phase-fused definition."""
    passes = (sum(STEP_PHASES.values()) + (8 + 26 * I_CG + 12)
              + (9 + 5 * M * I_DOM + 6) / dom_every + SUBSTEP_PASSES * n_sub)
    return passes * esz + PARTICLE_BYTES * particles / cells


PCG_RTOL = 1.0e-8
# parameters of the learned preconditioner, trained on 128 x 32 x 32 runs by
# python -m firewebapp_amd.learned --out firewebapp_amd/data/precond_net.npz
NET_PARAMS = os.path.join(ROOT, "firewebapp_amd", "data", "precond_net.npz")


def pcg_time_to_rtol(sv, n_rhs, rtol):
    """Plain PCG vs Jacobi-PCG to rtol on the projection systems of the next
    n_rhs steps, solved from x0 = 0 with a handle of their own."""
    import ctypes as C
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.physics_3d.fft_poisson_3d import create_handle
    stream = torch.cuda.current_stream().cuda_stream
    seq = []
    for _ in range(n_rhs):
        sv.step()
        b = torch.empty(sv.shape, dtype=torch.float64, device="cuda")
        _lib.call("fw_proj_rhs", sv.proj, C.c_void_p(b.data_ptr()), C.c_void_p(stream))
        seq.append((sv.f["rho"].clone(), b))
    h = create_handle(sv.dg, 1.0e-6, rtol, maxiter=50000)
    out = {"rhs": n_rhs, "x0": "zero", "rtol": rtol}
    net = None
    modes = [("plain", 1), ("jacobi", 2)]
    if os.path.exists(NET_PARAMS):
        from firewebapp_amd.learned import ConvNet, attach
        nz, ny, nx = sv.shape
        net = ConvNet(nx, ny, nz, theta=np.load(NET_PARAMS)["theta"])
        modes.append(("learned", 3))
    else:
        out["learned"] = None
    try:
        for name, mode in modes:
            if mode == 3:
                attach(h, net)
            else:
                _lib.call("fw_proj_set_preconditioner", h, mode)
            its, ms, rel = [], [], []
            for rho, b in seq:
                x = torch.zeros_like(b)
                it, rr = C.c_int(), C.c_double()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _lib.call("fw_proj_solve", h, C.c_void_p(rho.data_ptr()),
                          C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), rtol,
                          C.byref(it), C.byref(rr), C.c_void_p(stream))
                torch.cuda.synchronize()
                ms.append(1e3 * (time.perf_counter() - t0))
                its.append(it.value)
                rel.append(rr.value)
            out[name] = {"iters_mean": float(np.mean(its)), "ms_mean": float(np.mean(ms)),
                         "iters": its, "max_rel_resid": float(max(rel))}
            if mode == 3:
                fb = C.c_int()
                _lib.call("fw_proj_learned_stats", h, C.byref(fb))
                out[name]["fallbacks_to_plain"] = fb.value
            print(f"# {name} PCG to {rtol:g}: {np.mean(its):.1f} iterations, "
                  f"{np.mean(ms):.2f} ms per system", file=sys.stderr, flush=True)
    finally:
        _lib.lib().fw_proj_destroy(h)
        if net is not None:
            net.close()
    return out


def deck(name, dtype):
    from firewebapp_amd.deck import RomInputs
    ovr = dict(COMMON)
    ovr.update(WORKLOADS[name])
    ovr["dtype"] = dtype
    return RomInputs(outdoor_overrides=ovr)


# every SPECIES_STRIDE-th k_species3 launch of the timed steps records its
# events (the event-recording launch costs extra host time per launch)
SPECIES_STRIDE = 4


def species_timer(on):
    """Start (reset) or stop the step driver's HIP-event timing of the k_species
    launches; the events are recorded on the stream the kernels run on."""
    from firewebapp_amd import _lib
    _lib.call("fw_spread_timer", SPECIES_STRIDE if on else 0)


def species_timer_read():
    import ctypes as C
    from firewebapp_amd import _lib
    ms, n = C.c_double(), C.c_int()
    _lib.call("fw_spread_timer_read", C.byref(ms), C.byref(n))
    return ms.value, n.value


def cpu_baseline(name, budget_s):
    """Oracle float64 step on one core over a bounded sample of the workload."""
    from threadpoolctl import threadpool_limits
    from firewebapp_amd.spread_3d import Spread3DConfig, bed_inputs, build_grid
    from oracle.step import OracleSpread3D
    cfg = Spread3DConfig.resolve(deck(name, "float64"), wind_speed_m_s=wind(name))
    g = build_grid(cfg)
    o = OracleSpread3D(cfg, g, bed_inputs(cfg, g))
    n = 0
    with threadpool_limits(1):
        t0 = time.perf_counter()
        while True:
            o.step()
            n += 1
            el = time.perf_counter() - t0
            if el >= budget_s or el * (n + 1) / n > 2.0 * budget_s:
                break
    cells = g.Nx * g.Ny * g.Nz
    return {"value": cells * n / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle float64 NumPy step, outer steps 1-{n} of {name} "
                      f"({el:.1f} s; no DOM solve in these steps, it runs every 5th)"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="spread3d_256x64x48_s8", choices=sorted(WORKLOADS))
    ap.add_argument("--dtype", default="float32", choices=("float32", "float64"))
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of oracle CPU work for cpu_baseline (0 = skip)")
    ap.add_argument("--pcg-rhs", type=int, default=4,
                    help="projection systems for pcg_time_to_1e8 (0 = skip)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (the "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the step is "
              f"single-GPU in this build", file=sys.stderr)
        return 2
    import torch
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    torch.cuda.init()
    if a.calibrate_copy_mib > 0:
        n = a.calibrate_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cfg = Spread3DConfig.resolve(deck(a.workload, a.dtype), wind_speed_m_s=wind(a.workload))
    sv = Spread3DSolver(cfg)
    g = sv.grid
    cells = g.Nx * g.Ny * g.Nz
    print(f"# workload {a.workload}: {g.Nx}x{g.Ny}x{g.Nz} cells, {sv.n_particles} "
          f"particles, moisture {cfg.moisture:.4f}", file=sys.stderr, flush=True)
    for i in range(a.warmup):
        sv.step()
    print(f"# warmup {a.warmup} steps done, t = {sv.t:.4f} s", file=sys.stderr, flush=True)
    species_timer(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        if sv.step() is not None:
            raise RuntimeError(f"run stopped early: {sv.stop_reason}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # device memory in use by this process (the library's own allocations,
    # the DOM intensities among them, are outside torch's allocator)
    free_b, total_b = torch.cuda.mem_get_info()
    hbm_used = total_b - free_b
    species_timer(False)
    print(f"# timed {a.steps} steps in {wall:.3f} s, t = {sv.t:.4f} s", file=sys.stderr,
          flush=True)
    kern, fields = ROOFLINE_KERNEL
    esz = 4 if a.dtype == "float32" else 8
    bytes_launch = fields * esz * cells
    tot_ms, launches = species_timer_read()
    ms = tot_ms / max(launches, 1)
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": cells * a.steps / wall, "unit": "cell-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if a.dtype == "float32" else "f64",
        "data": "synthetic deck (seeded moisture, particle load and depth jitter"
                + (", lognormal fuel depth field" if cfg.fuel_depth_cv > 0.0 else "")
                + (", inlet gusts" if cfg.gust_amplitude > 0.0 else "") + ")",
        "config": {"workload": a.workload, "cells": [g.Nx, g.Ny, g.Nz],
                   "particles": sv.n_particles, "dom": f"S{cfg.dom_quadrature}",
                   "dom_every": cfg.dom_subcycle_every, "n_sub": 10,
                   "projection": "fft-PCG rtol 1e-6", "parallelism": "single GPU",
                   "u10_m_s": cfg.wind_speed_m_s,
                   "hbm_used_gib": hbm_used / 2.0 ** 30},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": ROOFLINE_TRAFFIC.get((a.workload, a.dtype)), "bytes_per_launch": bytes_launch,
                     "mean_launch_ms": ms, "launches": launches},
        "cpu_baseline": None,
    }
    bpc = step_bytes_per_cell(esz, 10, sv.dom.M,
                              cfg.dom_subcycle_every, sv.n_particles, cells)
    ach = bpc * line["value"] / 1e9
    line["step_roofline"] = {
        "bound": "hbm", "bytes_per_cell_step": bpc, "achieved": ach, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
        "definition": f"phase-fused compulsory traffic at I_cg={I_CG}, I_dom={I_DOM}"}
    if a.pcg_rhs > 0:
        line["pcg_time_to_1e8"] = pcg_time_to_rtol(sv, a.pcg_rhs, PCG_RTOL)
    del sv
    torch.cuda.empty_cache()
    if a.cpu_budget > 0:
        print("# timing the oracle on the host", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_budget)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
