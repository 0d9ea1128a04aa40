"""Oracle vs the reference's golden vectors: transport stencils.

Standards (docs/kerneltest.mjs of the reference): pure stencils bit-exact,
kernels with sqrt/tanh/pow 1e-9 relative.
"""
import numpy as np
import pytest

from oracle import transport as T
from tests.golden_io import arr, case_ids, cases, field, relerr, vectors


def test_minmod_and_face_helpers():
    for h in vectors()["muscl"]["helpers"]:
        if h["fn"] == "minmod":
            got = float(T.minmod(*h["args"]))
        else:
            got = float(T.muscl_face_value(*h["args"]))
        assert got == h["want"], h


@pytest.mark.parametrize("c", cases("muscl"), ids=case_ids("muscl"))
def test_muscl_advection(c):
    rhs = field(c, "rhs_in")
    T.advect_3d_scalar_muscl(field(c, "phi"), field(c, "u"), field(c, "v"),
                             field(c, "w"), c["dx"], c["dy"],
                             arr(c, "d_face_above"), arr(c, "d_face_below"),
                             rhs, c["phi_inlet"])
    assert np.array_equal(rhs, field(c, "rhs_out"))


@pytest.mark.parametrize("c", cases("species"), ids=case_ids("species"))
def test_species(c):
    Y = field(c, "Y_in")
    T.step_species_transport(Y, field(c, "rho"), field(c, "u"), field(c, "v"),
                             field(c, "w"), field(c, "S"), c["dt"], c["dx"],
                             c["dy"], arr(c, "dz_arr"), arr(c, "d_face_above"),
                             arr(c, "d_face_below"), c["D"], c["Y_inlet"])
    assert np.array_equal(Y, field(c, "Y_out"))
    assert int(((Y == 0.0) | (Y == 1.0)).sum()) == c["n_clipped"]


@pytest.mark.parametrize("c", cases("momentum"), ids=case_ids("momentum"))
def test_momentum(c):
    u, v, w = field(c, "u_in"), field(c, "v_in"), field(c, "w_in")
    sh2 = (c["nz"], c["ny"])
    T.step_tentative_velocity(u, v, w, field(c, "rho"), field(c, "T_g"),
                              field(c, "Fx"), field(c, "Fy"), field(c, "Fz"),
                              c["dt"], c["dx"], c["dy"], arr(c, "dz_arr"),
                              arr(c, "d_face_above"), arr(c, "d_face_below"),
                              c["T_amb"], field(c, "u_inlet", sh2),
                              field(c, "v_inlet", sh2), field(c, "w_inlet", sh2))
    assert np.array_equal(u, field(c, "u_out"))
    assert np.array_equal(v, field(c, "v_out"))
    assert np.array_equal(w, field(c, "w_out"))


@pytest.mark.parametrize("c", cases("drag"), ids=case_ids("drag"))
def test_drag(c):
    sh = (c["nz"], c["ny"], c["nx"])
    F = [np.full(sh, -7.0) for _ in range(3)]
    T.step_drag_force(field(c, "u"), field(c, "v"), field(c, "w"),
                      field(c, "rho"), field(c, "alpha_s"), c["sigma_sav"],
                      F[0], F[1], F[2], c["C_D"])
    for key, f in zip(("Fx", "Fy", "Fz"), F):
        assert relerr(f, field(c, key)) <= 1e-15, key


@pytest.mark.parametrize("c", cases("solid_conduction"),
                         ids=case_ids("solid_conduction"))
def test_solid_conduction(c):
    Ts = field(c, "T_s_in")
    T.step_solid_conduction_vertical(Ts, field(c, "alpha_s"), arr(c, "dz_arr"),
                                     arr(c, "d_face_above"),
                                     arr(c, "d_face_below"), c["k_solid"],
                                     c["rho_solid"], c["cp_solid"], c["dt"])
    assert np.array_equal(Ts, field(c, "T_s_out"))


@pytest.mark.parametrize("c", cases("turb_diff"), ids=case_ids("turb_diff"))
def test_turb_diff(c):
    f = field(c, "field_in")
    n = T.apply_turbulent_diffusion(f, field(c, "nu_t"), c["sc_t"], c["dt"],
                                    c["dx"], c["dy"], arr(c, "dz_arr"),
                                    arr(c, "d_face_above"),
                                    arr(c, "d_face_below"))
    assert n == c["n_sub"] and n > 1
    assert np.array_equal(f, field(c, "field_out"))


@pytest.mark.parametrize("c", cases("kepsilon"), ids=case_ids("kepsilon"))
def test_kepsilon(c):
    sh = (c["nz"], c["ny"], c["nx"])
    k, e = field(c, "k_in"), field(c, "eps_in")
    nu = np.zeros(sh)
    T.step_k_epsilon(k, e, nu, field(c, "u"), field(c, "v"), field(c, "w"),
                     field(c, "T_g"), field(c, "rho"), field(c, "alpha_s"),
                     c["sigma_sav"], c["dt"], c["dx"], c["dy"],
                     arr(c, "dz_arr"), arr(c, "d_face_above"),
                     arr(c, "d_face_below"), c["T_amb"],
                     field(c, "u_inlet", (c["nz"], c["ny"])),
                     field(c, "k_wall_ghost", (c["ny"], c["nx"])),
                     field(c, "eps_wall_ghost", (c["ny"], c["nx"])),
                     c["beta_p"], c["beta_d"])
    assert relerr(k, field(c, "k_out")) <= 1e-12
    assert relerr(e, field(c, "eps_out")) <= 1e-12
    assert relerr(nu, field(c, "nu_t_out")) <= 1e-12


@pytest.mark.parametrize("c", cases("support"), ids=case_ids("support"))
def test_support_sponge_soil_gas_energy(c):
    sh = (c["nz"], c["ny"], c["nx"])
    # sponge
    u = field(c, "sponge_u_in")
    T.apply_outflow_sponge(u, field(c, "sponge_u_target", (c["nz"], c["ny"])),
                           arr(c, "sponge_sigma_x"), field(c, "sponge_Y_F"),
                           c["sponge_Y_F_skip"], c["sponge_dt"])
    assert np.array_equal(u, field(c, "sponge_u_out"))
    # soil
    s_dz, s_da, s_db, depth = T.build_soil_grid()
    assert np.array_equal(s_dz, arr(c, "soil_dz"))
    assert np.array_equal(s_da, arr(c, "soil_d_above"))
    assert np.array_equal(s_db, arr(c, "soil_d_below"))
    assert depth == c["soil_depth"]
    Tsoil = field(c, "soil_T_in", (c["n_soil"], c["ny"], c["nx"]))
    T.step_soil_conduction(Tsoil, field(c, "soil_q_in", (c["ny"], c["nx"])),
                           c["soil_dt"], s_dz, s_da, s_db, T_amb=300.0)
    assert relerr(Tsoil, field(c, "soil_T_out", Tsoil.shape)) <= 1e-15
    # gas energy
    Tg = field(c, "Tg_in")
    T.adv_gas_energy(Tg, field(c, "u"), field(c, "v"), field(c, "w"),
                     c["gas_dt"], c["dx"], c["dy"], arr(c, "dz_arr"),
                     arr(c, "d_face_above"), arr(c, "d_face_below"),
                     alpha_th=2.0e-5, T_amb=300.0)
    assert np.array_equal(Tg, field(c, "Tg_out"))
