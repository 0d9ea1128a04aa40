"""The oracle's active level set (level_set_passive = false) on the
reference's active integration deck (tests/golden/integration_case_active.txt,
scripts/integration_case_active.txt of the reference).

The reference ships no run of this deck.  Its kernels are pinned by the
golden vectors (compute_v_n_3d, the Godunov evolve and the reinitialisation,
bit-exact); the composition is checked here: with the empirical blend at
weight 1 the front speed is the empirical ROS in every cell and the front
advances at that speed; with the resolved speed, v_n is compute_v_n_3d of
the DOM forward flux in the band of the step's level set (cell by cell,
from the oracle's intensities), non-negative and zero outside the band."""
import os

import numpy as np
import pytest

from tests.golden_io import GOLDEN_DIR

DECK = os.path.join(GOLDEN_DIR, "integration_case_active.txt")


def _oracle(n_steps, **extra):
    from firewebapp_amd.deck import load_text_input
    from firewebapp_amd.grid import Grid3D
    from firewebapp_amd.spread_3d import Spread3DConfig
    from oracle.step import OracleSpread3D
    ri = load_text_input(DECK)
    ri.outdoor_overrides.update(extra)
    cfg = Spread3DConfig.resolve(ri, wind_speed_m_s=4.0)
    g = Grid3D.build(Lx=cfg.Lx, Ly=cfg.Ly, Lz=cfg.Lz, dx=cfg.dx, dy=cfg.dy,
                     h_bed=cfg.h_bed, n_z_bed=cfg.n_z_bed, **cfg.mesh_kwargs)
    o = OracleSpread3D(cfg, g)
    hist = []
    for _ in range(n_steps):
        phi0 = o.phi.copy()
        I0 = o.dom.I_set.copy()
        M0 = o.M_local.copy()
        r = o.step()
        hist.append((phi0, I0, M0, o.v_n.copy()))
        if r is not None:
            break
    return o, cfg, hist


def test_active_deck_is_active():
    from firewebapp_amd.spread_3d import Spread3DConfig
    cfg = Spread3DConfig.resolve(DECK, wind_speed_m_s=4.0)
    assert not cfg.level_set_passive and not cfg.empirical_ros_enable


def test_empirical_front_moves_at_ros():
    o, cfg, hist = _oracle(40, empirical_ros_enable=True, empirical_ros_u_threshold_m_s=10.0)
    ros = cfg.empirical_ros_m_s
    assert cfg.empirical_blend_w == 1.0 and ros > 0.1
    for *_, v in hist:
        assert np.all(v == ros)
    # with v_n = ROS the zero crossing of a unit-slope level set moves ROS dt
    # per step; the source patch's kinks (slope 3 at the bed front, a step in
    # z at the bed top) and the reinitialisations bend that locally, so the
    # fitted front speed matches ROS to a few per cent
    t, xf = np.asarray(o.front_t), np.asarray(o.front_x)
    slope = np.polyfit(t[1:], xf[1:], 1)[0]
    print("front slope", slope, "ROS", ros, flush=True)
    assert abs(slope / ros - 1.0) < 0.05


def test_resolved_speed_from_dom_forward_flux():
    from oracle import levelset as LS
    from oracle import snapshot as S
    o, cfg, hist = _oracle(12)
    assert all(np.all(v >= 0.0) for *_, v in hist)
    assert any(v.any() for *_, v in hist)
    # the last step by hand: the band of the level set it began with, the
    # intensities and M_local it left (its DOM solve precedes the level set)
    g = o.g
    phi0, _, _, v = hist[-1]
    mask = S.ahead_band_mask(phi0, max(g.dx, 0.20), g.n_z_bed)
    q = S.q_dom_fwd_at_band(o.dom.Omega, o.dom.weights, o.dom.I_set, mask)
    want = LS.compute_v_n_3d(np.where(mask, q, 0.0), cfg.rho_b, 1300.0, cfg.h_bed, o.T_ign,
                             cfg.T_amb, M_local=o.M_local)
    assert np.array_equal(v, want)
    assert not v[~mask].any() and v[mask].any()
