"""Host-side parts of the flame_front_3d, lagrangian_particles_3d,
boundary_conditions and lagrangian_bed_3d drop-ins, against the reference's
golden vectors and its documented behaviour (no GPU)."""
import numpy as np
import pytest

from tests.golden_io import arr, case_ids, cases, field


@pytest.mark.parametrize("c", cases("flame_front"), ids=case_ids("flame_front"))
def test_flame_tilt_band_golden(c):
    from firewebapp_amd.physics_3d.flame_front_3d import flame_tilt_band_m
    for u, want in c["tilt_probe"]:
        assert flame_tilt_band_m(u) == want
    assert flame_tilt_band_m(5.0) == c["band_m"]


def test_kinematic_buffers_and_scalar_locator():
    from firewebapp_amd.physics_3d import lagrangian_particles_3d as lp
    buf = lp.allocate_kinematic_buffers(7)
    assert sorted(buf) == sorted(["x", "y", "z", "u", "v", "w", "alive", "age"])
    assert buf["alive"].dtype == np.int8 and buf["x"].dtype == np.float64
    assert all(not a.any() for a in buf.values())
    with pytest.raises(ValueError):
        lp.allocate_kinematic_buffers(-1)
    zf = np.array([0.0, 0.1, 0.25, 0.5])
    assert lp.locate_k_from_z(0.0, zf, 3) == 0
    assert lp.locate_k_from_z(0.1, zf, 3) == 1        # a face belongs to the cell above
    assert lp.locate_k_from_z(0.4999, zf, 3) == 2
    assert lp.locate_k_from_z(0.5, zf, 3) == -1       # top face is outside
    assert lp.locate_k_from_z(-1e-12, zf, 3) == -1
    # x < 0 is outside before the truncation (int(-0.05 / 0.1) would be 0)
    assert lp.locate_cell(-0.05, 0.05, 0.05, 0.1, 0.1, zf, 3, 10, 2) == (-1, 0, 0)
    assert lp.locate_cell(0.95, 0.19, 0.3, 0.1, 0.1, zf, 3, 10, 2) == (9, 1, 2)
    assert lp.locate_cell(1.0, 0.2, 0.3, 0.1, 0.1, zf, 3, 10, 2) == (-1, -1, 2)
    alive = np.array([1, 1, 0, 1, 0], np.int8)
    assert lp.alloc_dead_slot(alive) == 2
    assert lp.alloc_dead_slot(np.ones(3, np.int8)) == -1


def test_boundary_condition_registry():
    from firewebapp_amd import boundary_conditions as bc
    from firewebapp_amd.grid import Grid3D, build_u_inlet
    assert bc.available() == ("outdoor_wind",)
    cls = bc.get_bc_class("outdoor_wind")
    assert cls is bc.OutdoorWindBC and cls.kind == "outdoor_wind"
    with pytest.raises(ValueError, match="Unknown boundary_condition_kind"):
        bc.get_bc_class("cup_burner")
    with pytest.raises(ValueError):
        bc.OutdoorWindBC(4.0, wind_profile_type="ramp")
    with pytest.raises(ValueError, match="canopy_bl"):
        bc.OutdoorWindBC(4.0, wind_profile_type="canopy_bl")
    g = Grid3D.build(Lx=3.0, Ly=0.4, Lz=2.0, dx=0.1, h_bed=0.2, n_z_bed=4)
    b = bc.OutdoorWindBC(4.0)
    u_in = b.build_u_inlet(g)
    assert u_in.shape == (g.Nz, g.Ny)
    np.testing.assert_array_equal(u_in, build_u_inlet(g, 4.0))

    class Proj:
        def set_inlet_BC(self, u):
            self.u_inlet = u

    class State:
        u = np.zeros((g.Nz, g.Ny, g.Nx))

    p, s = Proj(), State()
    b.configure(p, g, s)
    np.testing.assert_array_equal(p.u_inlet, u_in)
    np.testing.assert_array_equal(s.u, np.broadcast_to(u_in[:, :, None], s.u.shape))


@pytest.mark.parametrize("c", cases("lagrangian_bed"), ids=case_ids("lagrangian_bed"))
def test_product_particle_init_bit_exact(c):
    """The drop-in's initialize_bed_particles_from_alpha_s (the one the GPU
    step's set-up calls) against the golden init: slot count, positions,
    masses and alive flags bit for bit (lagrangian_bed_3d.py:221-329)."""
    from firewebapp_amd.physics_3d import lagrangian_bed_3d as lb
    buf = lb.allocate_bed_particle_buffers(c["n_max"])
    n = lb.initialize_bed_particles_from_alpha_s(
        buf, field(c, "alpha_s"), c["rho_b_dry"], c["moisture_frac"], c["T_amb"],
        c["dx"], c["dy"], arr(c, "dz_arr"), c["n_z_bed"], c["n_per_cell"], sav=c["sav"])
    assert n == c["n_alloc"]
    for k in ("x", "y", "z", "m_solid", "m_water"):
        assert np.array_equal(buf[k], arr(c, "init_" + k)), k
    assert np.array_equal(buf["alive"], np.asarray(c["init_alive"], np.int8))
    assert np.all(buf["T_s"][:n] == c["T_amb"])
