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
with its 10 chemistry/transport sub-steps.  The workload's configuration is
a 500-step run (1000 at 512 x 128 x 64, 200 at 128 x 32 x 32); the timed
window is its end, a developed fire: an untimed spin-up, W warm-up steps,
then exactly K timed steps ending at the configuration's last step
(--spinup overrides).  ``value`` = cells x K / wall seconds over those
steps, inputs resident in HBM.  The JSON line also carries:
* window: the step numbers and simulated times of the timed window;
* full_run: the rate over every step of the run (spin-up included, the autotuner's time not);
* early_window: the rate over steps W+1 .. W+K of a fresh run (the
  ignition transient, before the fire develops);
* roofline: k_dom_sweep_tile, the kernel with the largest share of the
  step's kernel time among the stencil passes: algorithmic bytes per launch
  (5 field passes per ordinate per source iteration over the sweep's
  launches) / its mean launch time over the timed steps, measured with HIP
  events that each launch itself records (hipExtLaunchKernelGGL, on the
  stream the kernel is launched on: torch's current stream), against
  8 TB/s; roofline_species3 the same for the sub-steps' species kernel
  (every 4th launch of 4 steps after the timed ones, run with the graph
  replay of the sub-step loop off);
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
  Jacobi-PCG and the learned preconditioner, iterations and wall time per
  system;
* sweep_c4: the 48-case batched wind x moisture sweep (300 steps each, 4
  cases in flight), whole-sweep cell-steps/s;
* autotune: the launch shapes the autotuner picked for the grid, with the
  timings behind the choice;
* cpu_baseline: the oracle's float64 NumPy step on one host core over
  steps 1 .. dom_every of the same workload (one DOM solve among them, the
  run's cadence), with the host's CPU model and CPU count (rank 0, N = 1).
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


# Outer steps of each workload's configuration: the timed window is placed
# at the end of that run (a developed fire), after an untimed spin-up.
CONFIG_STEPS = {"spread3d_256x64x48_s8": 500, "spread3d_512x128x64_s8": 1000,
                "spread3d_128x32x32_s4": 200}

# The roofline kernels.  The DOM sweep (k_dom_sweep_flow: one launch per
# source iteration; k_dom_sweep_tile: Ny/TS + Nz/TS - 1 launches) must, per
# source iteration, write every ordinate's intensities once and read its
# inflow rows, the one row in TS in y and in z that crosses a block edge
# (2/TS of a field); kappa and B are read once for all ordinates, since they
# are shared (further reads of them are the caches' work).  k_species3 reads
# Y_fuel, Y_O2, Y_H2O, rho, u, v, w, S_pyro part, S_dry part and writes the
# three species: 12 fields per cell per launch.
def dom_sweep_bytes(esz, M, cells, ts):
    """Compulsory bytes of one source iteration's sweep."""
    return esz * cells * (M * (1.0 + 2.0 / ts) + 2.0)


SPECIES3_FIELDS = 12
# HBM bytes per launch from the FETCH_SIZE + WRITE_SIZE counters scaled by a
# 256 MiB copy (float32), from the counter runs summarised in
# profiles/step_256x64x48_s8_hbm_traffic.md and
# profiles/step_512x128x64_s8_kernel_roofline.md; the DOM sweep's depend on
# its kernel and tile side (the last key)
ROOFLINE_TRAFFIC = {
    ("k_species3", "spread3d_256x64x48_s8", "float32", 0): 44010989,
    ("k_species3", "spread3d_512x128x64_s8", "float32", 0): 355120644,
    ("k_dom_sweep_flow", "spread3d_256x64x48_s8", "float32", 4): 485781011,
    ("k_dom_sweep_flow", "spread3d_512x128x64_s8", "float32", 4): 3676208589,
    ("k_dom_sweep_tile", "spread3d_256x64x48_s8", "float32", 4): 18176445,
    ("k_dom_sweep_tile", "spread3d_256x64x48_s8", "float32", 8): 36213937}


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


def _time_solves(h, systems, modes, net, rtol):
    """Iterations and wall time per system of fw_proj_solve from x0 = 0, per
    preconditioner mode (1 plain, 2 Jacobi, 3 learned)."""
    import ctypes as C
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import ptr, stream_ptr
    from firewebapp_amd.learned import attach
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for name, mode in modes:
        if mode == 3:
            attach(h, net)
        else:
            _lib.call("fw_proj_set_preconditioner", h, mode)
        its, ms, rel = [], [], []
        for rho, b in systems:
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
            attach(h, None)
    return out


def pcg_time_to_rtol(sv, n_rhs, rtol):
    """Plain PCG, Jacobi-PCG and the learned preconditioner to rtol from
    x0 = 0: on the projection systems of the bench workload's next n_rhs
    steps, and (the learned network's own comparison) on 5 systems of a
    held-out seeded 128 x 32 x 32 run (U10 4.5 m/s, moisture 6.5 %, a seed
    outside the training runs)."""
    import ctypes as C
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.learned import ConvNet, c1_deck, collect_systems
    from firewebapp_amd.physics_3d.fft_poisson_3d import create_handle
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    stream = torch.cuda.current_stream().cuda_stream
    seq = []
    for _ in range(n_rhs):
        sv.step()
        b = torch.empty(sv.shape, dtype=torch.float64, device="cuda")
        _lib.call("fw_proj_rhs", sv.proj, C.c_void_p(b.data_ptr()), C.c_void_p(stream))
        seq.append((sv.f["rho"].clone(), b))
    theta = np.load(NET_PARAMS)["theta"] if os.path.exists(NET_PARAMS) else None
    out = {"rhs": n_rhs, "x0": "zero", "rtol": rtol}
    modes = [("plain", 1), ("jacobi", 2)] + ([("learned", 3)] if theta is not None else [])
    h = create_handle(sv.dg, 1.0e-6, rtol, maxiter=50000)
    nz, ny, nx = sv.shape
    net = ConvNet(nx, ny, nz, theta=theta) if theta is not None else None
    try:
        out.update(_time_solves(h, seq, modes, net, rtol))
    finally:
        _lib.lib().fw_proj_destroy(h)
        if net is not None:
            net.close()
    for name in ("plain", "jacobi", "learned"):
        if name in out:
            print(f"# {name} PCG to {rtol:g}: {out[name]['iters_mean']:.1f} iterations, "
                  f"{out[name]['ms_mean']:.2f} ms per system", file=sys.stderr, flush=True)
    if theta is not None:
        ri, w = c1_deck(4.5, 0.065, 97)
        c1 = Spread3DSolver(Spread3DConfig.resolve(ri, wind_speed_m_s=w))
        held = collect_systems(c1, 60, 12)
        h = create_handle(c1.dg, 1.0e-6, rtol, maxiter=50000)
        net = ConvNet(*c1.shape[::-1], theta=theta)
        try:
            r = _time_solves(h, held, [("plain", 1), ("learned", 3)], net, rtol)
        finally:
            _lib.lib().fw_proj_destroy(h)
            net.close()
            c1.close()
        r["systems"] = "5 projection systems (steps 12-60) of a held-out 128x32x32 run"
        out["held_out_128x32x32"] = r
        print(f"# held-out 128x32x32 to {rtol:g}: plain {r['plain']['iters_mean']:.1f} "
              f"iterations {r['plain']['ms_mean']:.2f} ms, learned "
              f"{r['learned']['iters_mean']:.1f} iterations {r['learned']['ms_mean']:.2f} ms",
              file=sys.stderr, flush=True)
    return out


def deck(name, dtype):
    from firewebapp_amd.deck import RomInputs
    ovr = dict(COMMON)
    ovr.update(WORKLOADS[name])
    ovr["dtype"] = dtype
    return RomInputs(outdoor_overrides=ovr)


# Kernels timed with HIP events, in dom_every extra steps after the timed
# ones (so no event-recording launch slows the timed steps): every DOM sweep
# launch of the one DOM solve among them, with the solve issued in line
# (beside the flow work it shares the GPU with other kernels, which is not
# the kernel's own rate), and every SPECIES_STRIDE-th k_species3 launch; the
# step's graph replay is off meanwhile (a timed launch is never captured)
TIMER_SPECIES3, TIMER_DOM_SWEEP = 0, 1
SPECIES_STRIDE = 4


def sweep_c4():
    """The batched sweep: the 48 uneven cases of firewebapp_amd/data/sweep_c4.txt
    (64-256 x 32-64 x 32-48 cells, U10 1-7 m/s, moisture 2-12 %, two fuel
    loads, 300 steps each) run 4 at a time, each on its own stream and host
    thread; the figure is the sum of cells x steps over the wall time of the
    whole sweep."""
    import torch
    from firewebapp_amd.batch import run_batch
    deck = os.path.join(ROOT, "firewebapp_amd", "data", "sweep_c4.txt")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = run_batch(deck)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total = sum(r["cells"] * r["steps"] for r in recs)
    sizes = sorted({(r["Nx"], r["Ny"], r["Nz"]) for r in recs})
    print(f"# sweep: {len(recs)} cases in {wall:.2f} s, {total / wall:.3e} cell-steps/s",
          file=sys.stderr, flush=True)
    return {"cases": len(recs), "steps_each": 300, "concurrent": 4, "wall_s": wall,
            "cell_steps": total, "value": total / wall, "unit": "cell-steps/s",
            "grids": [list(s) for s in sizes]}


def tune_shapes(sv):
    """The autotuner's choice of launch shapes for the workload's grid,
    timed on the solver's current state (firewebapp_amd.autotune)."""
    from firewebapp_amd.autotune import tune
    c = tune(sv, retune=True)
    print(f"# autotune ({c['tuning_s']} s): projection {c['proj']}, DOM tile {c['dom_tile']}, "
          f"species CZ {c['species_cz']}", file=sys.stderr, flush=True)
    return c


def kernel_timer(which, stride):
    """Start (reset) with the given stride, or stop (0), the HIP-event
    timing of a kernel's launches; the events are recorded by the launches
    themselves, on the stream the kernels run on."""
    from firewebapp_amd import _lib
    _lib.call("fw_kernel_timer", which, stride)


def kernel_timer_read(which):
    import ctypes as C
    from firewebapp_amd import _lib
    ms, n = C.c_double(), C.c_int()
    _lib.call("fw_kernel_timer_read", which, C.byref(ms), C.byref(n))
    return ms.value, n.value


def host_cpu():
    """CPU model name and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_baseline(name, dom_every):
    """Oracle float64 step on one core over outer steps 1 .. dom_every of the
    workload: dom_every - 1 steps without and one with the DOM solve, the
    step cadence of the run."""
    from threadpoolctl import threadpool_limits
    from firewebapp_amd.spread_3d import Spread3DConfig, bed_inputs, build_grid
    from oracle.step import OracleSpread3D
    cfg = Spread3DConfig.resolve(deck(name, "float64"), wind_speed_m_s=wind(name))
    g = build_grid(cfg)
    o = OracleSpread3D(cfg, g, bed_inputs(cfg, g))
    times = []
    with threadpool_limits(1):
        for _ in range(dom_every):
            t0 = time.perf_counter()
            o.step()
            times.append(time.perf_counter() - t0)
    el = sum(times)
    cells = g.Nx * g.Ny * g.Nz
    model, ncpu, usable = host_cpu()
    return {"value": cells * dom_every / el, "unit": "cell-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle float64 NumPy step on 1 thread, outer steps 1-{dom_every} of "
                      f"{name} ({dom_every - 1} without and 1 with the DOM solve, the "
                      f"run's cadence; {el:.1f} s, the DOM step {times[-1]:.1f} s)",
            "cpu_model": model, "host_cpus": ncpu, "usable_cpus": usable,
            "step_s": times}


def _roofline(kernel, bytes_launch, ms, launches, workload, dtype, shape_key=0):
    """shape_key: the launch shape the kernel ran at, when its traffic
    depends on it (the DOM tile side; 0 otherwise); traffic is None when no
    counter figure exists for that shape."""
    achieved = bytes_launch / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = ROOFLINE_TRAFFIC.get((kernel, workload, dtype, shape_key))
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": bytes_launch, "mean_launch_ms": ms, "launches": launches}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--spinup", type=int, default=-1,
                    help="untimed steps before the warm-up (default: so that the timed "
                         "window ends at the configuration's last step, -1)")
    ap.add_argument("--workload", default="spread3d_256x64x48_s8", choices=sorted(WORKLOADS))
    ap.add_argument("--dtype", default="float32", choices=("float32", "float64"))
    ap.add_argument("--cpu-budget", type=float, default=1.0,
                    help="0 skips cpu_baseline (the oracle on one core over the first "
                         "dom_every steps)")
    ap.add_argument("--pcg-rhs", type=int, default=4,
                    help="projection systems for pcg_time_to_1e8 (0 = skip)")
    ap.add_argument("--early-window", type=int, default=1,
                    help="also time the first warmup + steps of a fresh run (0 = skip)")
    ap.add_argument("--sweep", type=int, default=1,
                    help="1: also run the 48-case wind x moisture sweep "
                         "(firewebapp_amd/data/sweep_c4.txt, 300 steps each, 4 cases in "
                         "flight) and report its whole-sweep cell-steps/s as sweep_c4")
    ap.add_argument("--tune", type=int, default=1,
                    help="1: run the autotuner on the spun-up state before the warm-up "
                         "(untimed) and report its choice; 0: the default launch shapes")
    ap.add_argument("--dom-tile", type=int, default=0, choices=(0, 2, 4, 8),
                    help="DOM sweep tile side to use after the autotuner (0: the tuner's "
                         "choice, or 4 with --tune 0); the counter runs fix it")
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
    n_config = CONFIG_STEPS[a.workload]
    spinup = a.spinup if a.spinup >= 0 else max(0, n_config - a.warmup - a.steps)
    sv = Spread3DSolver(cfg)
    g = sv.grid
    cells = g.Nx * g.Ny * g.Nz
    print(f"# workload {a.workload}: {g.Nx}x{g.Ny}x{g.Nz} cells, {sv.n_particles} "
          f"particles, moisture {cfg.moisture:.4f}; spin-up {spinup}, warm-up {a.warmup}, "
          f"timed {a.steps} steps", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_run0 = time.perf_counter()
    tuned = None
    tune_s = 0.0   # the autotuner's wall time, kept out of full_run

    def tune():
        nonlocal tuned, tune_s
        torch.cuda.synchronize()
        tt = time.perf_counter()
        tuned = tune_shapes(sv)
        torch.cuda.synchronize()
        tune_s += time.perf_counter() - tt

    # the spin-up and warm-up in native calls of up to 100 steps (progress
    # lines between them), the autotuner after the spin-up
    done, total = 0, spinup + a.warmup
    while done < total:
        chunk = min(100 - done % 100, total - done)
        if a.tune and done < spinup < done + chunk:
            chunk = spinup - done
        if sv.advance(chunk) is not None:
            raise RuntimeError(f"run stopped early at step {sv.step_no}: {sv.stop_reason}")
        done += chunk
        if done % 100 == 0:
            print(f"# step {done}, t = {sv.t:.4f} s", file=sys.stderr, flush=True)
        if done == spinup and a.tune:
            tune()
    if spinup == 0 and a.tune:
        tune()
    t_win0 = sv.t
    print(f"# spin-up and warm-up done: {spinup + a.warmup} steps, t = {sv.t:.4f} s",
          file=sys.stderr, flush=True)
    TS = int(tuned["dom_tile"]) if tuned else 4     # the sweep's tile side
    if a.dom_tile:
        import ctypes as C
        from firewebapp_amd import _lib
        act = C.c_int()
        _lib.call("fw_dom_set_tiling", sv.dom_h, a.dom_tile, C.byref(act))
        TS = act.value
        if tuned:
            tuned["dom_tile"] = TS
    import ctypes as C
    from firewebapp_amd import _lib
    nl = C.c_int()
    _lib.call("fw_dom_sweep_launches", sv.dom_h, C.byref(nl))
    sweep_launches_per_iter = nl.value
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if sv.advance(a.steps) is not None:
        raise RuntimeError(f"run stopped early: {sv.stop_reason}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    # device memory in use by this process (the library's own allocations,
    # the DOM intensities among them, are outside torch's allocator)
    free_b, total_b = torch.cuda.mem_get_info()
    hbm_used = total_b - free_b
    print(f"# timed {a.steps} steps in {wall:.3f} s, t = {t_win0:.4f} .. {sv.t:.4f} s",
          file=sys.stderr, flush=True)
    graph_stats = sv.graph_stats()
    graphs, sv.fields.graphs = sv.fields.graphs, None
    beside = os.environ.get("FIREWAVE_DOM_BESIDE")
    os.environ["FIREWAVE_DOM_BESIDE"] = "0"
    kernel_timer(TIMER_SPECIES3, SPECIES_STRIDE)
    kernel_timer(TIMER_DOM_SWEEP, 1)
    timing_steps = [sv.step_no + 1, sv.step_no + cfg.dom_subcycle_every]
    for i in range(cfg.dom_subcycle_every):
        sv.step()
    torch.cuda.synchronize()
    kernel_timer(TIMER_SPECIES3, 0)
    kernel_timer(TIMER_DOM_SWEEP, 0)
    if beside is None:
        del os.environ["FIREWAVE_DOM_BESIDE"]
    else:
        os.environ["FIREWAVE_DOM_BESIDE"] = beside
    sv.fields.graphs = graphs
    esz = 4 if a.dtype == "float32" else 8
    M = sv.dom.M
    dom_bytes = dom_sweep_bytes(esz, M, cells, TS) / sweep_launches_per_iter
    dom_ms, dom_n = kernel_timer_read(TIMER_DOM_SWEEP)
    sp_ms, sp_n = kernel_timer_read(TIMER_SPECIES3)
    total_steps = spinup + a.warmup + a.steps
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
        "window": {"spinup_steps": spinup, "warmup_steps": a.warmup,
                   "timed_steps": [spinup + a.warmup + 1, total_steps],
                   "config_steps": n_config, "t_start_s": t_win0, "t_end_s": sv.t},
        "full_run": {"steps": total_steps,
                     "value": cells * total_steps / (t1 - t_run0 - tune_s),
                     "ms_per_step": 1e3 * (t1 - t_run0 - tune_s) / total_steps},
        "roofline": _roofline("k_dom_sweep_flow" if sweep_launches_per_iter == 1
                              else "k_dom_sweep_tile", dom_bytes, dom_ms / max(dom_n, 1),
                              dom_n, a.workload, a.dtype, shape_key=TS),
        "roofline_species3": _roofline("k_species3", SPECIES3_FIELDS * esz * cells,
                                       sp_ms / max(sp_n, 1), sp_n, a.workload, a.dtype),
        "cpu_baseline": None,
        "autotune": tuned,
        "graphs": {"substep_loop": "one hipGraph launch per step" if graph_stats else "off",
                   "replays": graph_stats[0] if graph_stats else 0,
                   "captures": graph_stats[1] if graph_stats else 0},
    }
    line["roofline"]["definition"] = (
        f"per source iteration: {M} ordinates x (1 field written + 2/{TS} of a field of "
        f"inflow rows read) + kappa and B once, x {cells} cells x {esz} B, "
        f"/ {sweep_launches_per_iter} launches")
    line["roofline"]["sampled"] = (f"every sweep launch of the DOM solve in steps "
                                   f"{timing_steps[0]}-{timing_steps[1]} (after the timed "
                                   f"steps; the solve in line)")
    bpc = step_bytes_per_cell(esz, 10, M, cfg.dom_subcycle_every, sv.n_particles, cells)
    ach = bpc * line["value"] / 1e9
    line["step_roofline"] = {
        "bound": "hbm", "bytes_per_cell_step": bpc, "achieved": ach, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
        "definition": f"phase-fused compulsory traffic at I_cg={I_CG}, I_dom={I_DOM}"}
    if a.pcg_rhs > 0:
        line["pcg_time_to_1e8"] = pcg_time_to_rtol(sv, a.pcg_rhs, PCG_RTOL)
    sv.close()
    del sv
    torch.cuda.empty_cache()
    if a.early_window:
        # the former headline window: the first warmup + steps of a fresh run
        sv = Spread3DSolver(cfg)
        for i in range(a.warmup):
            sv.step()
        torch.cuda.synchronize()
        te0, tw0 = time.perf_counter(), sv.t
        sv.advance(a.steps)
        torch.cuda.synchronize()
        we = time.perf_counter() - te0
        line["early_window"] = {"timed_steps": [a.warmup + 1, a.warmup + a.steps],
                                "t_start_s": tw0, "t_end_s": sv.t,
                                "value": cells * a.steps / we,
                                "ms_per_step": 1e3 * we / a.steps}
        print(f"# early window: {1e3 * we / a.steps:.3f} ms per step", file=sys.stderr,
              flush=True)
        sv.close()
        del sv
        torch.cuda.empty_cache()
    if a.sweep:
        line["sweep_c4"] = sweep_c4()
    if a.cpu_budget > 0:
        print("# timing the oracle on the host", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(a.workload, cfg.dom_subcycle_every)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
