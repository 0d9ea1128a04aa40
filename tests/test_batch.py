"""The batch.* deck section and its batch table (host side)."""
import os

import numpy as np
import pytest

from firewebapp_amd.batch import batch_table
from firewebapp_amd.deck import parse_deck_text

BASE = "outdoor.dx = 0.1\noutdoor.wind_speed_m_s = 2.5\n"


def test_no_batch_section():
    assert batch_table(parse_deck_text(BASE)) == []


def test_cycle_lists_grids_and_draws():
    ri = parse_deck_text(BASE + "batch.n_cases = 10\nbatch.seed = 3\n"
                         "batch.lx = 6.4, 12.8\nbatch.wind_speed_m_s = 1:7:2\n"
                         "batch.initial_moisture_frac = 0.02~0.12\nbatch.n_steps = 30\n"
                         "batch.terrain = open\n")
    assert ri.outdoor_overrides == {"dx": 0.1, "wind_speed_m_s": 2.5}
    t = batch_table(ri)
    assert [c.index for c in t] == list(range(10))
    assert [c.overrides["lx"] for c in t] == [6.4, 12.8] * 5
    assert [c.wind_speed_m_s for c in t] == [1.0, 3.0, 5.0, 7.0] * 2 + [1.0, 3.0]
    m = np.array([c.overrides["initial_moisture_frac"] for c in t])
    assert np.array_equal(m, np.random.default_rng(3).uniform(0.02, 0.12, 10))
    assert all(c.overrides["n_steps"] == 30 and isinstance(c.overrides["n_steps"], int)
               for c in t)
    assert all(c.overrides["terrain"] == "open" for c in t)
    # the case deck: base keys kept, case keys applied, no batch section
    ri_c = t[1].inputs(ri)
    assert ri_c.outdoor_overrides["lx"] == 12.8 and ri_c.outdoor_overrides["dx"] == 0.1
    assert ri_c.batch_overrides == {} and ri.batch_overrides


def test_product_mode_and_deck_wind():
    ri = parse_deck_text(BASE + "batch.mode = product\nbatch.ly = 3.2, 6.4\n"
                         "batch.bulk_density_kg_m3 = 1.476, 1.238, 1.0\n")
    t = batch_table(ri)
    assert len(t) == 6
    got = {(c.overrides["bulk_density_kg_m3"], c.overrides["ly"]) for c in t}
    assert got == {(b, y) for b in (1.476, 1.238, 1.0) for y in (3.2, 6.4)}
    assert all(c.wind_speed_m_s == 2.5 for c in t)


def test_bad_batch_sections():
    with pytest.raises(ValueError):
        batch_table(parse_deck_text(BASE + "batch.lx = 1, 2\n"))          # no n_cases
    with pytest.raises(ValueError):
        batch_table(parse_deck_text(BASE + "batch.mode = zigzag\nbatch.lx = 1\n"))
    with pytest.raises(ValueError):
        batch_table(parse_deck_text(BASE + "batch.n_cases = 2\nbatch.lx = 1:5:0\n"))


def test_c4_sweep_deck():
    """The shipped 48-case sweep spans the stated ranges."""
    from firewebapp_amd.deck import load_text_input
    from firewebapp_amd.grid import Grid3D
    from firewebapp_amd.spread_3d import Spread3DConfig
    path = os.path.join(os.path.dirname(__file__), "..", "firewebapp_amd", "data",
                        "sweep_c4.txt")
    ri = load_text_input(path)
    t = batch_table(ri)
    assert len(t) == 48
    dims = set()
    for c in t:
        cfg = Spread3DConfig.resolve(c.inputs(ri), wind_speed_m_s=c.wind_speed_m_s)
        g = Grid3D.build(Lx=cfg.Lx, Ly=cfg.Ly, Lz=cfg.Lz, dx=cfg.dx, dy=cfg.dy,
                         h_bed=cfg.h_bed, n_z_bed=cfg.n_z_bed, **cfg.mesh_kwargs)
        dims.add((g.Nx, g.Ny, g.Nz))
        assert 0.02 <= cfg.moisture <= 0.12 and cfg.n_steps == 300
        assert cfg.dom_quadrature == 8 and cfg.level_set_passive
    assert {d[0] for d in dims} == {64, 128, 192, 256}
    assert {d[1] for d in dims} == {32, 64}
    assert {d[2] for d in dims} == {32, 40, 48}
    assert {c.wind_speed_m_s for c in t} == {1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0}
