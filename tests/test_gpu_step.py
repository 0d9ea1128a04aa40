"""GPU whole-step parity through the C ABI.

* float64 build on the reference's integration deck (run to t_end = 0.6 s, 217
  steps) against the reference's own run (integration_reference.json): the
  step count, the front history and ROS exactly, every per-step maximum and
  the final field statistics to 1e-6 relative.
* float64 build on a small periodic-y deck (Ny = 4) against the oracle's
  whole step after 25 steps (fields to relative L2 1e-8, velocity components
  relative to the speed scale, particles 1e-8).
  Both projections solve to 1e-12 here, so the two Krylov methods (the
  oracle's BiCGSTAB, the GPU's PCG) reach the same pressure.
* float32 production build on the integration deck: fields within relative
  L2 1e-4 of the reference run's statistics, the front within one cell.
* the step glue kernels, the level-set update and the fused reduction against
  NumPy restatements / the golden flame_front vectors.
"""
import json
import os

import numpy as np
import pytest

from tests.golden_io import GOLDEN_DIR, arr, case_ids, cases, field

pytestmark = pytest.mark.gpu

DECK = os.path.join(GOLDEN_DIR, "integration_case.txt")


def _ref():
    with open(os.path.join(GOLDEN_DIR, "integration_reference.json")) as f:
        return json.load(f)


def _solver(deck, **kw):
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    cfg = Spread3DConfig.resolve(deck, **kw)
    return cfg, Spread3DSolver(cfg)


def _with(deck_path, **extra):
    from firewebapp_amd.deck import load_text_input
    ri = load_text_input(deck_path)
    ri.outdoor_overrides.update(extra)
    return ri


def _run_to_end(sv):
    while sv.t < sv.cfg.max_wall_time_s:
        if sv.step() is not None:
            break
    return sv


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


@pytest.fixture(scope="module")
def fp64_integration():
    _, sv = _solver(_with(DECK, dtype="float64"), wind_speed_m_s=4.0)
    return _run_to_end(sv)


def test_integration_fp64_history(fp64_integration):
    sv, ref = fp64_integration, _ref()
    print("steps", sv.step_no, "ros", sv.ros(), ref["ros_m_s"], flush=True)
    assert len(sv.diag["t"]) == ref["n_steps"]
    assert sv.front_t == ref["front_t"]
    np.testing.assert_allclose(sv.front_x, ref["front_x"], rtol=1e-12)
    assert _rel(sv.ros(), ref["ros_m_s"]) < 1e-12
    np.testing.assert_allclose(sv.diag["t"], ref["diag_t"], rtol=1e-12)
    for key, mine in (("diag_Tg_max", "Tg"), ("diag_Ts_max", "Ts"), ("diag_Y_max", "Y"),
                      ("diag_omega_max", "om")):
        a, b = np.asarray(sv.diag[mine]), np.asarray(ref[key])
        err = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))
        print(key, "max rel err", err, flush=True)
        assert err < 1e-6, key


def test_integration_fp64_fields(fp64_integration):
    st, ref = fp64_integration.state_numpy(), _ref()
    for name in ("T_g", "T_s", "Y_fuel", "Y_O2", "rho", "u"):
        for stat, fn in (("max", np.max), ("min", np.min), ("mean", np.mean)):
            r = ref[f"{name}_{stat}"]
            e = _rel(float(fn(st[name])), r)
            print(name, stat, e, flush=True)
            if name == "Y_fuel" and stat == "min":
                assert abs(float(fn(st[name])) - r) < 1e-12
            else:
                assert e < 1e-6, (name, stat)


def _small_deck():
    return _with(DECK, ly=0.4, dtype="float64", n_steps=25)


def test_small3d_fp64_vs_oracle():
    from firewebapp_amd.grid import Grid3D
    from oracle.step import OracleSpread3D
    cfg, sv = _solver(_small_deck(), wind_speed_m_s=4.0, projection_cg_rtol=1e-12)
    assert sv.grid.Ny == 4
    sv.run()
    g = Grid3D.build(Lx=cfg.Lx, Ly=cfg.Ly, Lz=cfg.Lz, dx=cfg.dx, dy=cfg.dy,
                     h_bed=cfg.h_bed, n_z_bed=cfg.n_z_bed, **cfg.mesh_kwargs)
    o = OracleSpread3D(cfg, g).run(cfg.n_steps)
    st = sv.state_numpy()
    assert sv.step_no == o.step_no
    # velocity components are measured against the speed scale |(u, v, w)|:
    # v and w are small here, so their own norms would magnify rounding
    speed = float(np.sqrt(np.linalg.norm(o.u) ** 2 + np.linalg.norm(o.v) ** 2 +
                          np.linalg.norm(o.w) ** 2))
    errs = {}
    for name in ("u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "rho", "k",
                 "eps", "nu_t", "omega", "q_rad"):
        a, b = st[name], getattr(o, name)
        den = speed if name in ("u", "v", "w") else max(np.linalg.norm(b), 1e-300)
        errs[name] = float(np.linalg.norm(a - b) / den)
    for k in ("T_s", "m_solid", "m_water", "m_char"):
        a, b = st["p_" + k], o.buf[k][:o.n_alloc]
        errs["p_" + k] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    print("rel L2 vs oracle:", errs, flush=True)
    for name, e in errs.items():
        assert e < 1e-8, (name, e)
    assert np.array_equal(st["p_alive"], o.buf["alive"][:o.n_alloc])
    assert np.array_equal(st["phi"], o.phi)
    np.testing.assert_allclose(sv.diag["t"], o.diag["t"], rtol=1e-12)


def test_integration_fp32_within_tolerance():
    _, sv = _solver(DECK, wind_speed_m_s=4.0)
    _run_to_end(sv)
    ref = _ref()
    assert len(sv.diag["t"]) == ref["n_steps"]
    assert sv.front_t == ref["front_t"]
    assert abs(sv.front_x[-1] - ref["front_x"][-1]) < sv.grid.dx
    st = sv.state_numpy()
    for name in ("T_g", "T_s", "Y_O2", "rho", "u"):
        for stat, fn in (("max", np.max), ("mean", np.mean)):
            e = _rel(float(fn(st[name])), ref[f"{name}_{stat}"])
            print("fp32", name, stat, e, flush=True)
            assert e < 1e-4, (name, stat)
    for key, mine in (("diag_Tg_max", "Tg"), ("diag_Ts_max", "Ts")):
        a, b = np.asarray(sv.diag[mine]), np.asarray(ref[key])
        err = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))
        print("fp32", key, err, flush=True)
        assert err < 1e-4, key


def test_advance_matches_step():
    """fw_spread_run (advance: many steps per native call) against one
    fw_spread_step per call: the same steps, diagnostics, front points and
    fields, bit for bit (float32 build, DOM solves at steps 5, 10, ...)."""
    _, a = _solver(DECK, wind_speed_m_s=4.0)
    _, b = _solver(DECK, wind_speed_m_s=4.0)
    for _ in range(33):
        assert a.step() is None
    assert b.advance(12) is None
    assert b.advance(21) is None
    assert b.step_no == a.step_no == 33
    assert a.diag == b.diag
    assert (a.front_t, a.front_x) == (b.front_t, b.front_x)
    sa, sb = a.state_numpy(), b.state_numpy()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


# ---- level set and glue kernels ---------------------------------------------

@pytest.mark.parametrize("c", cases("flame_front"), ids=case_ids("flame_front"))
def test_levelset_kernel_golden(c):
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, device, ptr, stream_ptr
    from oracle import levelset as LS
    dz = arr(c, "dz_arr")
    g = DeviceGrid(c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], dz, dz, dz,
                   dtype=torch.float64)
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device())
    phi = d(field(c, "phi_init"))
    vn = d(field(c, "v_n_in"))
    out = torch.empty_like(phi)
    _lib.call("fw_levelset_update", g.ref, ptr(phi), None, ptr(vn), ptr(out), c["dt"], 0,
              stream_ptr())
    assert np.array_equal(out.cpu().numpy(), field(c, "phi_evolved"))
    phi0 = out.clone()
    cur, nxt = out, torch.empty_like(out)
    dtau = LS.DT_REINIT_FRAC * min(c["dx"], c["dy"], float(dz.min()))
    for _ in range(LS.REINIT_SUBSTEPS):
        _lib.call("fw_levelset_update", g.ref, ptr(cur), ptr(phi0), None, ptr(nxt), dtau, 1,
                  stream_ptr())
        cur, nxt = nxt, cur
    assert np.array_equal(cur.cpu().numpy(), field(c, "phi_reinit"))
    assert LS.front_x(cur.cpu().numpy(), 1, c["ny"] // 2, c["dx"]) == c["front_x"]


def test_glue_kernels_fp64():
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, device, ptr, stream_ptr
    from oracle import levelset as LS
    rng = np.random.default_rng(11)
    nz, ny, nx, nzb = 7, 3, 10, 3
    dz = 0.02 + 0.05 * rng.random(nz)
    g = DeviceGrid(nx, ny, nz, 0.1, 0.12, dz, dz, dz, n_z_bed=nzb, dtype=torch.float64)
    keep = []   # device copies live until the end: no block is reused mid-call

    def d(a):
        t = torch.as_tensor(np.ascontiguousarray(a), device=device())
        keep.append(t)
        return t
    sh = (nz, ny, nx)
    T_amb = 300.0
    Tg = 250.0 + 1200.0 * rng.random(sh)
    rho = 0.3 + rng.random(sh)
    YO2 = 0.2 * rng.random(sh)
    t_Tg, t_rho, t_rp, t_YO = d(Tg), d(rho), d(np.zeros(sh)), d(YO2)
    _lib.call("fw_step_prologue", g.ref, ptr(t_Tg), ptr(t_rho), ptr(t_rp), ptr(t_YO), T_amb,
              stream_ptr())
    assert np.array_equal(t_rp.cpu().numpy(), rho)
    assert np.array_equal(t_rho.cpu().numpy(),
                          101325.0 / ((8.314 / 0.029) * np.maximum(Tg, T_amb)))
    YO2e = YO2.copy()
    YO2e[:, :, 0] = 0.232
    assert np.array_equal(t_YO.cpu().numpy(), YO2e)
    # Q_solid_ext on the bed layers
    q_rad = 4.0e3 * rng.normal(size=sh)
    Qx = d(np.zeros(sh))
    _lib.call("fw_bed_external_heat", g.ref, ptr(d(q_rad)), ptr(Qx), nzb, 2, 5, 2.4e5, 1.0e5,
              stream_ptr())
    e = np.zeros(sh)
    e[:nzb, :, 2:5] += 2.4e5
    for k in range(nzb):
        e[k] += np.clip(q_rad[k] / dz[k], -1.0e5, 1.0e5)
    assert np.array_equal(Qx.cpu().numpy(), e)
    # S_pyro and div_target
    Sp, Sd = 0.1 * rng.random(sh), 0.05 * rng.random(sh)
    rp = rho * (1.0 + 0.01 * rng.normal(size=sh))
    S_out, tgt = d(np.zeros(sh)), d(np.zeros(sh))
    _lib.call("fw_div_target", g.ref, ptr(d(Sp)), ptr(d(Sd)), ptr(d(rho)), ptr(d(rp)), 0.01,
              ptr(S_out), ptr(tgt), stream_ptr())
    S = Sp.copy()
    S += Sd
    assert np.array_equal(S_out.cpu().numpy(), S)
    assert np.array_equal(tgt.cpu().numpy(), (S - (rho - rp) / 0.01) / np.maximum(rho, 0.1))
    # sub-step gas energy update
    Ts = 300.0 + 900.0 * rng.random(sh)
    Yw = 0.3 * rng.random(sh)
    Qgc = 1.0e4 * rng.normal(size=sh)
    qrg = 1.0e3 * rng.normal(size=sh)
    Tg2 = d(Tg)
    Qc = d(np.zeros(sh))
    _lib.call("fw_gas_energy_update", g.ref, ptr(Tg2), ptr(d(Ts)), ptr(d(rho)), ptr(d(Yw)),
              ptr(d(Sp)), ptr(d(Sd)), ptr(d(Qgc)), ptr(d(qrg)), ptr(Qc), nzb, 2, 5, 6.0e4,
              1.0e-4, T_amb, stream_ptr())
    Qe = np.zeros(sh)
    Qe[:nzb, :, 2:5] += 6.0e4
    for k in range(nz):
        Qe[k] += qrg[k] / dz[k]
    cp = (1.0 - Yw) * 1100.0 + Yw * 2000.0
    deb = (Sp * np.maximum(Tg - Ts, 0.0) + Sd * np.maximum(Tg - 373.15, 0.0)) * 2000.0
    Te = Tg + (Qe - Qgc - deb) * (1.0e-4 / (np.maximum(rho, 1.0e-3) * cp))
    Te = np.maximum(Te, T_amb)
    assert np.array_equal(Qc.cpu().numpy(), Qe)
    assert np.array_equal(Tg2.cpu().numpy(), Te)
    # the fused reduction
    u, v, w = rng.normal(size=sh), rng.normal(size=sh), rng.normal(size=sh)
    nut, om, YF = rng.random(sh), 2e-3 * rng.random(sh), 2e-3 * rng.random(sh)
    alpha = np.zeros(sh)
    alpha[:nzb, :, 1:8] = 0.003
    Ts2 = Ts.copy()
    Ts2[1, 2, 6] = 700.0
    work = torch.empty(14 * 256, dtype=torch.float64, device=device())
    out = torch.empty(14, dtype=torch.float64, device=device())
    _lib.call("fw_step_reduce", g.ref, ptr(d(u)), ptr(d(v)), ptr(d(w)), ptr(d(nut)),
              ptr(d(Tg)), ptr(d(Ts2)), ptr(d(Sp)), ptr(d(Qe)), ptr(d(om)), ptr(d(YF)),
              ptr(d(alpha)), 600.0, 600.0, nzb, ptr(work), ptr(out), stream_ptr())
    r = out.cpu().numpy()
    burning = (Ts2[:nzb] >= 600.0) & (alpha[:nzb] > 0.0)
    cols = np.where(burning.any(axis=(0, 1)))[0]
    exp = [np.abs(u).max(), np.abs(v).max(), np.abs(w).max(), nut.max(), Tg.max(), 0.0,
           Ts2.max(), Sp.max(), Qe.max(), om.max(), YF.max(), float((Ts2 >= 600.0).sum()),
           float(LS.flame_active(om, Tg, YF).sum()), float(cols[-1])]
    assert list(r) == exp


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_fused_substep_kernels_match_single(dtype):
    """k_species3, k_turb_diff3 and k_gas_solid reuse the per-cell code of the
    single-field kernels: results are bit-identical in both precisions."""
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, device, ptr, stream_ptr
    import ctypes as C
    td = torch.float64 if dtype == "float64" else torch.float32
    rng = np.random.default_rng(21)
    nz, ny, nx, nzb = 9, 5, 23, 3
    dz = 0.02 + 0.05 * rng.random(nz)
    da, db = dz * 1.1, dz * 0.9
    g = DeviceGrid(nx, ny, nz, 0.1, 0.12, dz, da, db, n_z_bed=nzb, dtype=td)
    sh = (nz, ny, nx)
    keep = []

    def d(a):
        t = torch.as_tensor(np.ascontiguousarray(a), dtype=td, device=device())
        keep.append(t)
        return t

    Ys = [d(0.3 * rng.random(sh)), d(0.232 * rng.random(sh)), d(0.2 * rng.random(sh))]
    rho, u, v, w = d(0.4 + rng.random(sh)), d(rng.normal(size=sh)), \
        d(rng.normal(size=sh)), d(rng.normal(size=sh))
    Sp, Sd = d(0.05 * rng.random(sh)), d(0.02 * rng.random(sh))
    y_in, y_max = (0.0, 0.232, 0.0), (1.0, 0.232, 1.0)
    single = [torch.empty_like(Ys[0]) for _ in range(3)]
    for q in range(3):
        _lib.call("fw_species_transport_dilution", g.ref, ptr(Ys[q]), ptr(single[q]),
                  ptr(rho), ptr(u), ptr(v), ptr(w), ptr(Sp), ptr(Sd), q + 1, 1e-3, 1e-5,
                  y_in[q], y_max[q], stream_ptr())
    fused = [torch.empty_like(Ys[0]) for _ in range(3)]
    VP3 = C.c_void_p * 3
    D3 = C.c_double * 3
    _lib.call("fw_species_transport3", g.ref, VP3(*[ptr(t) for t in Ys]),
              VP3(*[ptr(t) for t in fused]), ptr(rho), ptr(u), ptr(v), ptr(w), ptr(Sp),
              ptr(Sd), 1e-3, 1e-5, D3(*y_in), D3(*y_max), stream_ptr())
    for a, b in zip(single, fused):
        assert torch.equal(a, b)
    nu = d(0.05 * rng.random(sh))
    lo, hi = (0.0, 0.0, 1.0), (1.0, 0.232, 0.0)
    single = [torch.empty_like(Ys[0]) for _ in range(3)]
    for q in range(3):
        _lib.call("fw_turbulent_diffusion_pass", g.ref, ptr(Ys[q]), ptr(single[q]), ptr(nu),
                  0.7, 2e-4, lo[q], hi[q], stream_ptr())
    fused = [torch.empty_like(Ys[0]) for _ in range(3)]
    _lib.call("fw_turbulent_diffusion_pass3", g.ref, VP3(*[ptr(t) for t in Ys]),
              VP3(*[ptr(t) for t in fused]), ptr(nu), 0.7, 2e-4, D3(*lo), D3(*hi),
              stream_ptr())
    for a, b in zip(single, fused):
        assert torch.equal(a, b)
    Tg = d(250.0 + 1200.0 * rng.random(sh))
    alpha = np.zeros(sh)
    alpha[:nzb] = 0.003 * rng.random((nzb, ny, nx))
    alpha = d(alpha)
    Ts = d(300.0 + 900.0 * rng.random(sh))
    Yw, Qgc, qrg = d(0.3 * rng.random(sh)), d(1e4 * rng.normal(size=sh)), \
        d(1e3 * rng.normal(size=sh))
    Tg1, Qc1, Ts1 = Tg.clone(), torch.empty_like(Tg), torch.empty_like(Tg)
    _lib.call("fw_gas_energy_update", g.ref, ptr(Tg1), ptr(Ts), ptr(rho), ptr(Yw), ptr(Sp),
              ptr(Sd), ptr(Qgc), ptr(qrg), ptr(Qc1), nzb, 2, 7, 6.0e4, 1e-4, 300.0,
              stream_ptr())
    _lib.call("fw_solid_conduction_vertical", g.ref, ptr(Ts), ptr(Ts1), ptr(alpha), 0.2,
              500.0, 1300.0, 1e-4, 300.0, 1.0e4, stream_ptr())
    Tg2, Qc2, Ts2 = Tg.clone(), torch.empty_like(Tg), torch.empty_like(Tg)
    _lib.call("fw_gas_update_solid_conduction", g.ref, ptr(Tg2), ptr(Ts), ptr(Ts2),
              ptr(alpha), ptr(rho), ptr(Yw), ptr(Sp), ptr(Sd), ptr(Qgc), ptr(qrg), ptr(Qc2),
              nzb, 2, 7, 6.0e4, 1e-4, 300.0, 0.2, 500.0, 1300.0, 300.0, 1.0e4, stream_ptr())
    assert torch.equal(Tg1, Tg2) and torch.equal(Qc1, Qc2) and torch.equal(Ts1, Ts2)


def test_fused_edc_matches_own_launch(monkeypatch):
    """The next sub-step's EDC fused into the temperature diffusion pass
    against its own launch (FIREWAVE_FUSE_EDC=0), at 256x64x48 (whole
    256-cell rows, the case the fusion takes): 12 steps, fields, particles
    and diagnostics bit for bit, float32 and float64."""
    import bench
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    name = "spread3d_256x64x48_s8"
    for dtype in ("float32", "float64"):
        states = []
        for fuse in ("1", "0"):
            monkeypatch.setenv("FIREWAVE_FUSE_EDC", fuse)
            cfg = Spread3DConfig.resolve(bench.deck(name, dtype), wind_speed_m_s=bench.wind(name))
            sv = Spread3DSolver(cfg)
            assert sv.advance(12) is None
            states.append((sv.state_numpy(), dict(sv.diag)))
            sv.close()
        (sa, da), (sb, db) = states
        assert da == db, dtype
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (dtype, k)
