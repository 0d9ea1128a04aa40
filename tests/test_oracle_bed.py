"""Oracle Lagrangian bed vs the golden vectors + invariants.

Standards (docs/kerneltest.mjs:645-653 and SOLVER_PORT.md:89 of the
reference): initialisation and conduction bit-exact; the particle step ~1e-15
relative (exp() of two libms may differ by an ulp): 1e-12 here; alive flags
and counts exact.  Mass conservation of the step: solid+char+water lost by
the particles equals the volatile + water-vapour sources (times dt V).
"""
import numpy as np
import pytest

from oracle import bed as Bd
from tests.golden_io import arr, case_ids, cases, field, relerr


def _buf_from(c, prefix="in_"):
    n = c["n_max"]
    buf = Bd.allocate_bed_particle_buffers(n)
    for k in ("x", "y", "z", "m_solid", "m_water", "m_char", "T_s", "m_water_0",
              "sav", "m_char_max"):
        buf[k][:] = arr(c, prefix + k)
    buf["alive"][:] = np.asarray(c[prefix + "alive"], dtype=np.int8)
    return buf


def _step(c, buf):
    sh = (c["nz"], c["ny"], c["nx"])
    return Bd.step_bed_particles(
        buf, field(c, "T_g"), field(c, "Y_O2"), field(c, "Q_solid_ext"),
        c["n_per_cell"], c["dx"], c["dy"], arr(c, "dz_arr"), arr(c, "z_face"),
        c["h_conv"], c["rho_solid_true"], c["cp_solid"], c["eps_solid"],
        c["T_amb"], c["view_factor"], c["view_factor_geometric"], c["h_bed"],
        c["kappa_bed_eff"], c["dt"], True, True, True, True, c["drying_mode"],
        c["char_ox_flux_cap"], c["char_ox_ash_exp"])


@pytest.mark.parametrize("c", cases("lagrangian_bed"), ids=case_ids("lagrangian_bed"))
def test_init_bit_exact(c):
    buf = Bd.allocate_bed_particle_buffers(c["n_max"])
    n = Bd.initialize_bed_particles_from_alpha_s(
        buf, field(c, "alpha_s"), c["rho_b_dry"], c["moisture_frac"], c["T_amb"],
        c["dx"], c["dy"], arr(c, "dz_arr"), c["n_z_bed"], c["n_per_cell"], c["sav"])
    assert n == c["n_alloc"]
    for k in ("x", "y", "z", "m_solid", "m_water"):
        assert np.array_equal(buf[k], arr(c, "init_" + k)), k
    assert np.array_equal(buf["alive"], np.asarray(c["init_alive"], np.int8))


@pytest.mark.parametrize("c", cases("lagrangian_bed"), ids=case_ids("lagrangian_bed"))
def test_step_aggregate_conduction(c):
    sh = (c["nz"], c["ny"], c["nx"])
    buf = _buf_from(c)
    src, n_alive, n_burned, diag = _step(c, buf)
    assert n_alive == c["out_n_alive"] and n_burned == c["out_n_burned"]
    assert np.array_equal(buf["alive"], np.asarray(c["out_alive"], np.int8))
    for k in ("m_solid", "m_water", "m_char", "T_s", "m_char_max"):
        assert relerr(buf[k], arr(c, "out_" + k)) <= 1e-12, k
    for k in src:
        assert relerr(src[k], field(c, "out_" + k)) <= 1e-12, k
    assert relerr(diag, arr(c, "out_diag")) <= 1e-12
    Tg = np.full(sh, 300.0)
    Bd.aggregate_T_s_grid(buf, c["dx"], c["dy"], arr(c, "z_face"), Tg)
    assert relerr(Tg, field(c, "out_T_s_grid")) <= 1e-12
    M = np.zeros(sh)
    Bd.aggregate_M_local_grid(buf, c["dx"], c["dy"], arr(c, "z_face"), M)
    assert relerr(M, field(c, "out_M_local")) <= 1e-12
    # conduction on the golden input grid: bit-exact
    Tc = field(c, "cond_T_in")
    part = Bd.allocate_bed_particle_buffers(c["n_max"])
    for k in ("x", "y", "z"):
        part[k][:] = buf[k]
    part["alive"][:] = buf["alive"]
    part["T_s"][:] = buf["T_s"]
    Bd.horizontal_conduction(part, Tc, field(c, "alpha_s"), c["dx"], c["dy"],
                             arr(c, "z_face"), c["k_solid"], c["rho_solid_true"],
                             c["cp_solid"], c["n_z_bed"], c["dt"])
    assert np.array_equal(Tc, field(c, "cond_T_out"))
    assert relerr(part["T_s"], arr(c, "cond_part_T_out")) <= 1e-12


@pytest.mark.parametrize("c", cases("lagrangian_bed"), ids=case_ids("lagrangian_bed"))
def test_step_mass_conservation(c):
    buf = _buf_from(c)
    m0 = buf["m_solid"] + buf["m_water"] + buf["m_char"]
    alive0 = buf["alive"].copy()
    src, *_ = _step(c, buf)
    m1 = buf["m_solid"] + buf["m_water"] + buf["m_char"]
    dz = arr(c, "dz_arr")
    V = (c["dx"] * c["dy"] * dz).reshape(-1, 1, 1)
    i, j, k = Bd.locate_cell(arr(c, "in_x"), arr(c, "in_y"), arr(c, "in_z"), c["dx"],
                             c["dy"], arr(c, "z_face"), c["nz"], c["nx"], c["ny"])
    ok = (alive0 != 0) & (i >= 0) & (j >= 0) & (k >= 0)
    lost = float(np.sum((m0 - m1)[ok]))
    # gas-phase mass sources: pyrolysate (S_pyro / eta = dm_pyro incl. char) is
    # split: eta to gas, char stays; water to vapour; char/smoulder oxidised
    gas = float(np.sum((src["S_pyro"] + src["S_drying"]) * V)) * c["dt"]
    ox = float(np.sum((src["Q_char"] / Bd.HOC_CHAR + src["Q_smold"] / Bd.HOC_SMOLD) * V)) * c["dt"]
    assert abs(lost - (gas + ox)) <= 1e-9 * max(lost, 1e-30)
