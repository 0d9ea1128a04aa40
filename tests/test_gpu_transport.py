"""GPU parity: transport stencils through the C ABI vs the golden vectors.

FP64 kernels are held to the golden vectors at the reference's own JS-port
standard (docs/kerneltest.mjs): bit-exact for pure stencils, 1e-12 relative
where sqrt/tanh enter.  FP32 kernels are held to the FP64 oracle at 2e-5
relative (max-norm over the field), the FP32 budget of a single stencil
application (a few ulps of float32 on each of ~20 operations).
"""
import numpy as np
import pytest

from tests.golden_io import arr, case_ids, cases, field, relerr

pytestmark = pytest.mark.gpu

F32_TOL = 2e-5


def _p3():
    from firewebapp_amd import physics_3d  # noqa: F401
    from firewebapp_amd.physics_3d import (drag_3d, momentum_3d, muscl_3d,
                                           soil_3d, solid_conduction_3d,
                                           species_3d, turbulence_3d)
    return dict(drag=drag_3d, mom=momentum_3d, muscl=muscl_3d, soil=soil_3d,
                cond=solid_conduction_3d, spec=species_3d, turb=turbulence_3d)


def _f32(a):
    import torch
    return torch.as_tensor(np.asarray(a, np.float32), device="cuda")


def _np(t):
    return t.double().cpu().numpy()


@pytest.mark.parametrize("c", cases("muscl"), ids=case_ids("muscl"))
def test_muscl(c):
    m = _p3()["muscl"]
    rhs = field(c, "rhs_in")
    args = [field(c, k) for k in ("phi", "u", "v", "w")]
    m.advect_3d_scalar_muscl(*args, c["dx"], c["dy"], arr(c, "d_face_above"),
                             arr(c, "d_face_below"), rhs, c["phi_inlet"])
    assert np.array_equal(rhs, field(c, "rhs_out"))
    r32 = _f32(field(c, "rhs_in"))
    m.advect_3d_scalar_muscl(*[_f32(a) for a in args], c["dx"], c["dy"],
                             arr(c, "d_face_above"), arr(c, "d_face_below"),
                             r32, c["phi_inlet"])
    assert relerr(_np(r32), field(c, "rhs_out")) <= F32_TOL


@pytest.mark.parametrize("c", cases("species"), ids=case_ids("species"))
def test_species(c):
    s = _p3()["spec"]
    Y = field(c, "Y_in")
    rest = [field(c, k) for k in ("rho", "u", "v", "w", "S")]
    geo = (c["dt"], c["dx"], c["dy"], arr(c, "dz_arr"), arr(c, "d_face_above"),
           arr(c, "d_face_below"), c["D"], c["Y_inlet"])
    s.step_species_transport(Y, *rest, *geo)
    assert np.array_equal(Y, field(c, "Y_out"))
    Y32 = _f32(field(c, "Y_in"))
    s.step_species_transport(Y32, *[_f32(a) for a in rest], *geo)
    assert relerr(_np(Y32), field(c, "Y_out")) <= F32_TOL


@pytest.mark.parametrize("c", cases("momentum"), ids=case_ids("momentum"))
def test_momentum(c):
    mm = _p3()["mom"]
    sh2 = (c["nz"], c["ny"])
    u, v, w = field(c, "u_in"), field(c, "v_in"), field(c, "w_in")
    rest = [field(c, k) for k in ("rho", "T_g", "Fx", "Fy", "Fz")]
    inl = [field(c, k, sh2) for k in ("u_inlet", "v_inlet", "w_inlet")]
    geo = (c["dt"], c["dx"], c["dy"], arr(c, "dz_arr"), arr(c, "d_face_above"),
           arr(c, "d_face_below"), c["T_amb"])
    mm.step_tentative_velocity(u, v, w, *rest, *geo, *inl)
    for got, key in ((u, "u_out"), (v, "v_out"), (w, "w_out")):
        assert np.array_equal(got, field(c, key)), key
    U = [_f32(field(c, k)) for k in ("u_in", "v_in", "w_in")]
    mm.step_tentative_velocity(*U, *[_f32(a) for a in rest], *geo,
                               *[_f32(a) for a in inl])
    for got, key in zip(U, ("u_out", "v_out", "w_out")):
        assert relerr(_np(got), field(c, key)) <= F32_TOL, key


@pytest.mark.parametrize("c", cases("drag"), ids=case_ids("drag"))
def test_drag(c):
    d = _p3()["drag"]
    sh = (c["nz"], c["ny"], c["nx"])
    F = [np.full(sh, -7.0) for _ in range(3)]
    ins = [field(c, k) for k in ("u", "v", "w", "rho", "alpha_s")]
    d.step_drag_force(*ins, c["sigma_sav"], *F, c["C_D"])
    for key, f in zip(("Fx", "Fy", "Fz"), F):
        assert relerr(f, field(c, key)) <= 1e-15, key


@pytest.mark.parametrize("c", cases("solid_conduction"),
                         ids=case_ids("solid_conduction"))
def test_solid_conduction(c):
    sc = _p3()["cond"]
    Ts = field(c, "T_s_in")
    sc.step_solid_conduction_vertical(Ts, field(c, "alpha_s"), arr(c, "dz_arr"),
                                      arr(c, "d_face_above"),
                                      arr(c, "d_face_below"), c["k_solid"],
                                      c["rho_solid"], c["cp_solid"], c["dt"])
    assert np.array_equal(Ts, field(c, "T_s_out"))


@pytest.mark.parametrize("c", cases("turb_diff"), ids=case_ids("turb_diff"))
def test_turb_diff(c):
    t = _p3()["turb"]
    f = field(c, "field_in")
    t.apply_turbulent_diffusion(f, field(c, "nu_t"), c["sc_t"], c["dt"], c["dx"],
                                c["dy"], arr(c, "dz_arr"), arr(c, "d_face_above"),
                                arr(c, "d_face_below"))
    assert np.array_equal(f, field(c, "field_out"))


@pytest.mark.parametrize("c", cases("kepsilon"), ids=case_ids("kepsilon"))
def test_kepsilon(c):
    t = _p3()["turb"]
    sh = (c["nz"], c["ny"], c["nx"])
    k, e, nu = field(c, "k_in"), field(c, "eps_in"), np.zeros(sh)
    t.step_k_epsilon(k, e, nu, field(c, "u"), field(c, "v"), field(c, "w"),
                     field(c, "T_g"), field(c, "rho"), field(c, "alpha_s"),
                     c["sigma_sav"], c["dt"], c["dx"], c["dy"], arr(c, "dz_arr"),
                     arr(c, "d_face_above"), arr(c, "d_face_below"), c["T_amb"],
                     None, None, field(c, "u_inlet", (c["nz"], c["ny"])),
                     field(c, "k_wall_ghost", (c["ny"], c["nx"])),
                     field(c, "eps_wall_ghost", (c["ny"], c["nx"])),
                     c["beta_p"], c["beta_d"])
    assert relerr(k, field(c, "k_out")) <= 1e-12
    assert relerr(e, field(c, "eps_out")) <= 1e-12
    assert relerr(nu, field(c, "nu_t_out")) <= 1e-12


@pytest.mark.parametrize("c", cases("support"), ids=case_ids("support"))
def test_sponge_soil_gas_energy(c):
    p = _p3()
    from firewebapp_amd import spread_3d
    u = field(c, "sponge_u_in")
    p["mom"].apply_outflow_sponge(u, field(c, "sponge_u_target", (c["nz"], c["ny"])),
                                  arr(c, "sponge_sigma_x"), field(c, "sponge_Y_F"),
                                  c["sponge_Y_F_skip"], c["sponge_dt"])
    assert np.array_equal(u, field(c, "sponge_u_out"))
    s = p["soil"]
    s_dz, s_da, s_db, _ = s.build_soil_grid()
    Tsoil = field(c, "soil_T_in", (c["n_soil"], c["ny"], c["nx"]))
    s.step_soil_conduction(Tsoil, field(c, "soil_q_in", (c["ny"], c["nx"])),
                           c["soil_dt"], s_dz, s_da, s_db, 0.0, s.K_SOIL_DEFAULT,
                           s.RHO_SOIL_DEFAULT, s.CP_SOIL_DEFAULT,
                           s.EPS_SOIL_DEFAULT, 300.0)
    assert relerr(Tsoil, field(c, "soil_T_out", Tsoil.shape)) <= 1e-15
    Tg = field(c, "Tg_in")
    spread_3d._adv_gas_energy(Tg, field(c, "u"), field(c, "v"), field(c, "w"),
                              c["gas_dt"], c["dx"], c["dy"], arr(c, "dz_arr"),
                              arr(c, "d_face_above"), arr(c, "d_face_below"),
                              alpha_th=2.0e-5, T_amb=300.0)
    assert np.array_equal(Tg, field(c, "Tg_out"))
