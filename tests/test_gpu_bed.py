"""GPU parity of the Lagrangian bed through the C ABI vs the golden vectors.

Bit-exact: particle cell indices (both locate conventions), alive flags,
n_alive / n_burned, the conduction grid.  1e-12 relative: particle state,
the 8 per-cell sources, the aggregates, the diagnostics (exp/pow of two
math libraries differ by an ulp).  The FP32 grid variant (float32 T_g, Y_O2,
Q_ext and sources; float64 particles) at 1e-5.
"""
import numpy as np
import pytest

from oracle import bed as Bd
from tests.golden_io import arr, case_ids, cases, field, relerr

pytestmark = pytest.mark.gpu


def _lb():
    from firewebapp_amd.physics_3d import lagrangian_bed_3d
    return lagrangian_bed_3d


def _buf(c):
    lb = _lb()
    buf = lb.allocate_bed_particle_buffers(c["n_max"])
    for k in ("x", "y", "z", "m_solid", "m_water", "m_char", "T_s", "m_water_0",
              "sav", "m_char_max"):
        buf[k][:] = arr(c, "in_" + k)
    buf["alive"][:] = np.asarray(c["in_alive"], dtype=np.int8)
    return buf


def _step(c, buf, cast=lambda a: a):
    lb = _lb()
    sh = (c["nz"], c["ny"], c["nx"])
    src = [cast(np.zeros(sh)) for _ in range(8)]
    na, nb, diag = np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(16)
    lb.step_bed_particles(
        buf["x"], buf["y"], buf["z"], buf["alive"], buf["m_solid"], buf["m_water"],
        buf["m_char"], buf["T_s"], buf["m_water_0"], buf["sav"],
        cast(field(c, "T_g")), cast(field(c, "Y_O2")), cast(field(c, "Q_solid_ext")),
        c["n_per_cell"], *src, c["dx"], c["dy"], arr(c, "dz_arr"), arr(c, "z_face"),
        c["h_conv"], c["rho_solid_true"], c["cp_solid"], c["eps_solid"], c["T_amb"],
        c["view_factor"], c["view_factor_geometric"], c["h_bed"], c["kappa_bed_eff"],
        c["dt"], True, True, True, True, c["drying_mode"], c["char_ox_flux_cap"],
        c["char_ox_ash_exp"], buf["m_char_max"], na, nb, diag)
    return src, int(na[0]), int(nb[0]), diag


NAMES = ("S_pyro", "S_drying", "Q_pyro", "Q_drying", "Y_F_source", "Q_char",
         "Q_smold", "Q_g_conv")


def test_locate_bit_exact():
    import torch
    from firewebapp_amd.device import DeviceGrid
    lb = _lb()
    rng = np.random.default_rng(3)
    nz, ny, nx = 6, 5, 9
    dz = 0.03 + 0.05 * rng.random(nz)
    zf = np.concatenate([[0.0], np.cumsum(dz)])
    n = 4000
    x = rng.uniform(-0.5, nx * 0.1 + 0.5, n)
    y = rng.uniform(-0.3, ny * 0.1 + 0.3, n)
    z = rng.uniform(-0.05, zf[-1] + 0.05, n)
    x[:50] = np.arange(50) * 0.1          # exact face positions
    z[50:60] = zf[:10] if zf.size >= 10 else z[50:60]
    g = DeviceGrid(nx, ny, nz, 0.1, 0.1, dz, dz, dz, z_face=zf, dtype=torch.float64)
    P = lb.DeviceParticles({"x": x, "y": y, "z": z, "alive": np.ones(n, np.int8),
                            **{k: np.zeros(n) for k in ("m_solid", "m_water", "m_char",
                                                          "T_s", "m_water_0", "sav",
                                                          "m_char_max")}})
    cs, ca = lb._locate(g, P)
    i, j, k = Bd.locate_cell(x, y, z, 0.1, 0.1, zf, nz, nx, ny)
    want = np.where((i < 0) | (j < 0) | (k < 0), -1, (k * ny + j) * nx + i)
    assert np.array_equal(cs.cpu().numpy(), want)
    ia, ja, ka = Bd.agg_cell(x, y, z, 0.1, 0.1, zf, nz, nx, ny)
    wa = np.where((ia < 0) | (ka < 0), -1, (ka * ny + ja) * nx + ia)
    assert np.array_equal(ca.cpu().numpy(), wa)


@pytest.mark.parametrize("c", cases("lagrangian_bed"), ids=case_ids("lagrangian_bed"))
def test_bed_step_fp64(c):
    lb = _lb()
    sh = (c["nz"], c["ny"], c["nx"])
    buf = _buf(c)
    src, na, nb, diag = _step(c, buf)
    assert (na, nb) == (c["out_n_alive"], c["out_n_burned"])
    assert np.array_equal(buf["alive"], np.asarray(c["out_alive"], np.int8))
    for k in ("m_solid", "m_water", "m_char", "T_s", "m_char_max"):
        assert relerr(buf[k], arr(c, "out_" + k)) <= 1e-12, k
    for name, s in zip(NAMES, src):
        assert relerr(s, field(c, "out_" + name)) <= 1e-12, name
    assert relerr(diag, arr(c, "out_diag")) <= 1e-12
    Tg = np.full(sh, 300.0)
    lb.aggregate_particles_to_T_s_grid(buf["x"], buf["y"], buf["z"], buf["alive"],
                                       buf["m_solid"], buf["m_water"], buf["m_char"],
                                       buf["T_s"], c["dx"], c["dy"], arr(c, "z_face"),
                                       Tg, 300.0)
    assert relerr(Tg, field(c, "out_T_s_grid")) <= 1e-12
    M = np.zeros(sh)
    lb.aggregate_particles_to_M_local_grid(buf["x"], buf["y"], buf["z"], buf["alive"],
                                           buf["m_solid"], buf["m_water"], c["dx"],
                                           c["dy"], arr(c, "z_face"), M)
    assert relerr(M, field(c, "out_M_local")) <= 1e-12
    Tc = field(c, "cond_T_in")
    pT = buf["T_s"].copy()
    lb.step_horizontal_solid_conduction_scatter(
        buf["x"], buf["y"], buf["z"], buf["alive"], buf["m_solid"], buf["m_water"],
        buf["m_char"], pT, Tc, field(c, "alpha_s"), c["dx"], c["dy"], arr(c, "z_face"),
        c["k_solid"], c["rho_solid_true"], c["cp_solid"], c["n_z_bed"], c["dt"])
    assert np.array_equal(Tc, field(c, "cond_T_out"))
    assert relerr(pT, arr(c, "cond_part_T_out")) <= 1e-12


@pytest.mark.parametrize("c", cases("lagrangian_bed"), ids=case_ids("lagrangian_bed"))
def test_bed_step_fp32_grid(c):
    import torch
    buf = _buf(c)
    cast = lambda a: torch.as_tensor(a.astype(np.float32), device="cuda")
    src, na, nb, diag = _step(c, buf, cast)
    assert (na, nb) == (c["out_n_alive"], c["out_n_burned"])
    for k in ("m_solid", "m_water", "m_char", "T_s"):
        assert relerr(buf[k], arr(c, "out_" + k)) <= 1e-5, k
    for name, s in zip(NAMES, src):
        assert relerr(s.double().cpu().numpy(), field(c, "out_" + name)) <= 1e-5, name
