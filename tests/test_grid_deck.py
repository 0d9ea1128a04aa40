"""Host setup: grid/mesh against the golden mesh vectors (bit-exact), deck
parsing rules, and the benchmark mesh recipes (Nz of the 128x32x32 and
256x64x48 configurations)."""
import numpy as np
import pytest

from firewebapp_amd import deck, grid
from tests.golden_io import case_ids, cases


@pytest.mark.parametrize("c", cases("mesh"), ids=case_ids("mesh"))
def test_mesh_golden(c):
    g = grid.Grid3D.build(**c["args"])
    assert (g.Nx, g.Ny, g.Nz, g.n_z_bed) == (c["nx"], c["ny"], c["nz"], c["n_z_bed"])
    assert g.dx == c["dx"] and g.dy == c["dy"] and g.dz == c["dz"]
    assert g.Lz == c["Lz_actual"]
    for k in ("dz_arr", "z_face", "z_mid", "x_mid", "y_mid", "d_face_above",
              "d_face_below"):
        assert np.array_equal(getattr(g, k), np.asarray(c[k])), k


def test_deck_parse_rules():
    ri = deck.parse_deck_text(
        "# comment\n"
        "outdoor.Lx = 3.0   # trailing\n"
        "outdoor.wall_bl_N = 1\n"
        "outdoor.flag = no\n"
        "outdoor.level_set_passive = true\n"
        "outdoor.thing = none\n"
        "outdoor.turbulence_model = k_epsilon\n"
        "outdoor.sigma_x = 5\n"
        "t_end = 0.6\n")
    o = ri.outdoor_overrides
    assert o["lx"] == 3.0
    assert o["wall_bl_n"] is True          # "1" parses as a boolean
    assert o["flag"] is False and o["level_set_passive"] is True
    assert o["thing"] is None and o["turbulence_model"] == "k_epsilon"
    assert "sigma_x" not in o
    cfg = deck.outdoor_env_from_dict({"Fuel_Depth_M": 0.21, "bulk_density_kg_m3": 1.4762})
    assert cfg.fuel_depth_m == 0.21 and cfg.bulk_density_kg_m3 == 1.4762


def test_integration_deck_grid():
    ri = deck.load_text_input("tests/golden/integration_case.txt")
    o = ri.outdoor_overrides
    g = grid.Grid3D.build(Lx=o["lx"], Ly=o["ly"], Lz=o["lz"], dx=o["dx"], dy=o["dy"],
                          h_bed=o["fuel_depth_m"], n_z_bed=int(o["n_z_bed"]),
                          atm_growth=o["atm_growth"], atm_max_dz=o["atm_max_dz"])
    # integration_reference.json: Nz x Ny x Nx = 24 x 1 x 30, n_z_bed 4
    assert (g.Nz, g.Ny, g.Nx, g.n_z_bed) == (24, 1, 30, 4)


@pytest.mark.parametrize("L,expect", [((12.8, 3.2, 3.2, 2, 0.105), (128, 32, 32)),
                                      ((25.6, 6.4, 4.51, 5, 0.042), (256, 64, 48))])
def test_baseline_meshes(L, expect):
    Lx, Ly, Lz, nzb, first = L
    g = grid.Grid3D.build(Lx=Lx, Ly=Ly, Lz=Lz, dx=0.1, dy=0.1, h_bed=0.21,
                          n_z_bed=nzb, wall_bl_N=1, wall_bl_first_dz=first,
                          wall_bl_growth=1.0, atm_uniform_dz=0.1)
    assert (g.Nx, g.Ny, g.Nz) == expect
    assert g.n_z_bed == nzb
    assert abs(g.z_face[nzb] - 0.21) < 1e-12
