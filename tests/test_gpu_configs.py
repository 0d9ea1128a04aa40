"""GPU parity at the benchmark configurations, through the native step.

* 128x32x32 (S4, float64, 12 steps, two DOM solves) against the oracle's
  whole step: every field to relative L2 1e-8 with both projections solved
  to 1e-12, particle state, alive flags and particle cell indices exact.
* float32 production build against the float64 build (itself pinned to the
  oracle) at 128x32x32 S4 and 256x64x48 S8 (seeded moisture, load and depth
  jitter), 200 steps: relative L2 of the velocity vector, T_g, T_s, the three
  species, omega, q_rad, q_rad_gas and the DOM intensities <= 1e-4, the
  particle temperatures <= 1e-4, the same alive flags and the same T_s-front
  column (the front within one cell: here, none).
* 512x128x64 (S8, lognormal depth field, gusts): the same measures after 10
  steps (one DOM solve).
* three full-size cases of the 48-case sweep table (the smallest, the
  largest and the one with the most elongated plan), run concurrently,
  equal bit for bit to each case run alone over the table's 300 steps; the
  smallest in float64 against the oracle over 10 steps.

The oracle runs on the host in the same process, so each oracle comparison
costs some seconds of CPU time; every step of it is printed.
"""
import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

C1, C2, C5 = "spread3d_128x32x32_s4", "spread3d_256x64x48_s8", "spread3d_512x128x64_s8"


def _solver(name, dtype, **kw):
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    cfg = Spread3DConfig.resolve(bench.deck(name, dtype), wind_speed_m_s=bench.wind(name),
                                 **kw)
    return cfg, Spread3DSolver(cfg)


def _oracle_compare(cfg, sv, n_steps, tol=1e-8):
    """Run sv and the oracle n_steps from the same deck; assert the fields,
    the particle state, the alive flags and the particle cells agree."""
    from firewebapp_amd.spread_3d import bed_inputs, build_grid
    from oracle import bed as Bd
    from oracle.step import OracleSpread3D
    g = build_grid(cfg)
    o = OracleSpread3D(cfg, g, bed_inputs(cfg, g))
    for s in range(n_steps):
        sv.step()
        o.step()
        print(f"step {s + 1}: t gpu {sv.t:.6f} oracle {o.t:.6f}", flush=True)
    st = sv.state_numpy()
    assert sv.step_no == o.step_no == n_steps
    speed = float(np.sqrt(np.linalg.norm(o.u) ** 2 + np.linalg.norm(o.v) ** 2 +
                          np.linalg.norm(o.w) ** 2))
    errs = {}
    for name in ("u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "rho", "k", "eps",
                 "nu_t", "omega", "q_rad", "q_rad_gas", "S_pyro"):
        a, b = st[name], getattr(o, name)
        den = speed if name in ("u", "v", "w") else max(np.linalg.norm(b), 1e-300)
        errs[name] = float(np.linalg.norm(a - b) / den)
    print("rel L2 vs oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for name, e in errs.items():
        assert e < tol, (name, e)
    n = o.n_alloc
    assert sv.n_particles == n
    for k in ("x", "y", "z"):
        np.testing.assert_array_equal(st["p_" + k], o.buf[k][:n])
    np.testing.assert_array_equal(st["p_alive"], o.buf["alive"][:n])
    for k in ("T_s", "m_solid", "m_water", "m_char"):
        a, b = st["p_" + k], o.buf[k][:n]
        assert np.linalg.norm(a - b) <= tol * max(np.linalg.norm(b), 1e-300), k
    cells = sv.ix.cell.cpu().numpy()
    want = np.empty(n, np.int64)
    for p in range(n):
        i, j, k = Bd.locate_cell(o.buf["x"][p], o.buf["y"][p], o.buf["z"][p], g.dx, g.dy,
                                 g.z_face, g.Nz, g.Nx, g.Ny)
        want[p] = -1 if min(i, j, k) < 0 else (k * g.Ny + j) * g.Nx + i
    np.testing.assert_array_equal(cells, want)
    np.testing.assert_allclose(sv.diag["t"], o.diag["t"], rtol=1e-12)
    return o


def test_c1_fp64_vs_oracle():
    cfg, sv = _solver(C1, "float64", projection_cg_rtol=1e-12)
    try:
        g = sv.grid
        assert (g.Nx, g.Ny, g.Nz) == (128, 32, 32) and sv.dom.M == 24
        o = _oracle_compare(cfg, sv, 12)
        assert float(np.max(np.abs(o.q_rad))) > 0.0        # the DOM ran (steps 5, 10)
        assert float(np.max(o.T_s)) > cfg.T_amb + 100.0    # the ignition heated the bed
    finally:
        sv.close()


def fp32_errors(a, b):
    """Relative L2 of the float32 build a against the float64 build b, per
    field; the velocity as one vector field."""
    import torch
    out = {}
    ua = torch.stack([a.f[k].double() for k in ("u", "v", "w")])
    ub = torch.stack([b.f[k].double() for k in ("u", "v", "w")])
    out["velocity"] = float((ua - ub).norm() / ub.norm())
    for k in ("T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "omega", "q_rad", "q_rad_gas"):
        x, y = a.f[k].double(), b.f[k].double()
        out[k] = float((x - y).norm() / max(float(y.norm()), 1e-300))
    Ia, Ib = a.dom.I_set.double(), b.dom.I_set.double()
    out["intensity"] = float((Ia - Ib).norm() / max(float(Ib.norm()), 1e-300))
    pa, pb = a.P.t["T_s"], b.P.t["T_s"]
    out["particle_T_s"] = float((pa - pb).norm() / pb.norm())
    out["alive_differ"] = int((a.P.t["alive"] != b.P.t["alive"]).sum())
    out["front_cells_apart"] = abs(int(a.state.red[13]) - int(b.state.red[13]))
    return out


def _fp32_vs_fp64(name, n_steps, tol=1e-4):
    _, a = _solver(name, "float32")
    _, b = _solver(name, "float64")
    try:
        for s in range(n_steps):
            assert a.step() is None and b.step() is None
            if (s + 1) % 50 == 0:
                print(f"{name} step {s + 1}", flush=True)
        e = fp32_errors(a, b)
        print(f"{name} float32 vs float64 after {n_steps} steps:",
              {k: (f"{v:.2e}" if isinstance(v, float) else v) for k, v in e.items()},
              flush=True)
        for k, v in e.items():
            if isinstance(v, float):
                assert v <= tol, (k, v)
        assert e["alive_differ"] == 0
        assert e["front_cells_apart"] <= 1
        assert float(a.f["T_s"].max()) > 600.0            # the bed is burning
    finally:
        a.close()
        b.close()


def test_c1_fp32_vs_fp64_per_field():
    _fp32_vs_fp64(C1, 200)


def test_c2_fp32_vs_fp64_per_field():
    _fp32_vs_fp64(C2, 200)


def test_c5_fp32_vs_fp64_per_field():
    _fp32_vs_fp64(C5, 10)


def _sweep_cases():
    """The smallest, the largest and the most elongated (largest Nx / Ny)
    cases of the 48-case sweep table, as (case, Nx, Ny, Nz)."""
    import os
    from firewebapp_amd.batch import batch_table
    from firewebapp_amd.deck import load_text_input
    from firewebapp_amd.spread_3d import Spread3DConfig, build_grid
    deck = os.path.join(os.path.dirname(bench.__file__), "firewebapp_amd", "data",
                        "sweep_c4.txt")
    ri = load_text_input(deck)
    rows = []
    for c in batch_table(ri):
        g = build_grid(Spread3DConfig.resolve(c.inputs(ri), wind_speed_m_s=c.wind_speed_m_s))
        rows.append((c, g.Nx, g.Ny, g.Nz))
    assert len(rows) == 48
    small = min(rows, key=lambda r: r[1] * r[2] * r[3])
    large = max(rows, key=lambda r: r[1] * r[2] * r[3])
    long_ = max((r for r in rows if r not in (small, large)), key=lambda r: r[1] / r[2])
    return ri, [small, large, long_]


def test_c4_sweep_cases_batched_equal_alone():
    from firewebapp_amd.batch import run_batch
    from firewebapp_amd.deck import RomInputs
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    ri, picked = _sweep_cases()
    print("sweep cases:", [(c.index, nx, ny, nz) for c, nx, ny, nz in picked], flush=True)
    assert picked[0][1:] == (64, 32, 32) and picked[1][1:] == (256, 64, 48)
    # a three-case table with the picked cases' own parameters
    sub = RomInputs(outdoor_overrides=dict(ri.outdoor_overrides))
    keys = sorted(picked[0][0].overrides)
    sub.batch_overrides = {"n_cases": "3"}
    for k in keys:
        sub.batch_overrides[k] = ", ".join(repr(c.overrides[k]) for c, *_ in picked)
    sub.batch_overrides["wind_speed_m_s"] = ", ".join(repr(c.wind_speed_m_s)
                                                      for c, *_ in picked)
    recs = run_batch(sub, concurrent=3)
    for (case, nx, ny, nz), rec in zip(picked, recs):
        cfg = Spread3DConfig.resolve(case.inputs(ri), wind_speed_m_s=case.wind_speed_m_s)
        sv = Spread3DSolver(cfg)
        try:
            sv.run()
            alone = sv.result()
        finally:
            sv.close()
        got = rec["result"]
        print(f"case {case.index}: {nx}x{ny}x{nz} steps {rec['steps']} "
              f"front {rec['front_x']:.2f} m", flush=True)
        assert (rec["Nx"], rec["Ny"], rec["Nz"]) == (nx, ny, nz)
        assert rec["steps"] == alone.n_steps == 300
        assert got.diag_t == alone.diag_t and got.front_x == alone.front_x
        for k in ("T_g", "T_s", "u", "v", "w", "rho", "Y_fuel", "Y_O2", "Y_H2O"):
            assert np.array_equal(getattr(got.state_final, k),
                                  getattr(alone.state_final, k)), k


def test_c4_smallest_case_fp64_vs_oracle():
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    ri, picked = _sweep_cases()
    case = picked[0][0]
    inp = case.inputs(ri)
    inp.outdoor_overrides["dtype"] = "float64"
    cfg = Spread3DConfig.resolve(inp, wind_speed_m_s=case.wind_speed_m_s,
                                 projection_cg_rtol=1e-12)
    sv = Spread3DSolver(cfg)
    try:
        assert (sv.grid.Nx, sv.grid.Ny, sv.grid.Nz) == (64, 32, 32) and sv.dom.M == 80
        _oracle_compare(cfg, sv, 10)
    finally:
        sv.close()
