"""Oracle of the Eulerian bed (oracle/eulerian.py).

* gas_solid_coupling: the golden `coupling` vectors (the reference's own
  step_gas_solid_coupling on seeded inputs, slab2d and small3d);
* the pyrolysis kernels have no golden vectors; they are pinned by what the
  mathematics fixes: the exponential decay composes (two half steps at fixed
  T equal one full step), mass is conserved between the pools and the
  emitted source (drying: water + Q dt / L; MD2004: solid + char + S dt),
  the moisture gate ramps linearly from 1 (dry) to 0 at 1 % of the initial
  water, the oxidation caps (heat <= Q_max, at most half the pool per step)
  and the heat equals HOC times the mass burnt; every skip branch leaves the
  cell untouched;
* the oracle's Eulerian whole step runs on the reference's integration deck
  with the particles switched off and keeps the energy and mass bookkeeping
  finite and the pools non-negative.
"""
import os

import numpy as np
import pytest

from oracle import eulerian as E
from tests.golden_io import GOLDEN_DIR, arr, case_ids, cases, field

SEC = "coupling"


@pytest.mark.parametrize("c", cases(SEC), ids=case_ids(SEC))
def test_coupling_golden(c):
    sh = (c["nz"], c["ny"], c["nx"])
    Tg, Ts = field(c, "T_g_in"), field(c, "T_s_in")
    mw = field(c, "m_water_in")
    E.gas_solid_coupling(Tg, Ts, field(c, "rho"), field(c, "u"), field(c, "v"), field(c, "w"),
                         field(c, "alpha_s"), c["sigma_sav"], field(c, "q_rad_in"),
                         field(c, "Q_pyro"), field(c, "Q_comb"), mw, c["L_v"], c["dt"],
                         arr(c, "dz_arr"), c["T_amb"])
    np.testing.assert_allclose(Tg, field(c, "T_g_out"), rtol=1e-14, atol=0)
    np.testing.assert_allclose(Ts, field(c, "T_s_out"), rtol=1e-14, atol=0)
    np.testing.assert_allclose(mw, field(c, "m_water_out"), rtol=1e-14, atol=1e-300)
    # cells whose water the step removed completely
    assert int((field(c, "m_water_in") > 0).sum() - (mw > 0).sum()) == c["n_dried"]


def _cells(n=400, seed=1):
    rng = np.random.default_rng(seed)
    T = 280.0 + 900.0 * rng.random(n)
    a = np.where(rng.random(n) < 0.85, 0.003 * rng.random(n), 0.0)
    y = np.where(rng.random(n) < 0.9, 0.232 * rng.random(n), 0.0)
    return T, a, y, rng


def test_drying_semigroup_and_mass():
    T, a, _, rng = _cells()
    mw = np.where(rng.random(T.size) < 0.8, 0.1 * rng.random(T.size), 0.0)
    one, q1 = mw.copy(), np.zeros_like(mw)
    E.drying(T, one, a, 0.02, q1)
    two, qh = mw.copy(), np.zeros_like(mw)
    E.drying(T, two, a, 0.01, qh)
    E.drying(T, two, a, 0.01, qh)
    np.testing.assert_allclose(two, one, rtol=1e-13)
    # water removed = Q dt / L_v, and nothing where there is no solid
    np.testing.assert_allclose(mw - one, q1 * 0.02 / E.L_VAP_WATER, rtol=1e-12, atol=1e-18)
    assert np.all(one[a <= 0.0] == mw[a <= 0.0]) and np.all(q1[a <= 0.0] == 0.0)
    assert np.all(one <= mw) and np.all(one >= 0.0)


def test_md2004_mass_balance_and_gate():
    T, a, y, rng = _cells(seed=2)
    n = T.size
    m = 0.5 * rng.random(n)
    mc = 0.01 * rng.random(n)
    mw_init = 0.05
    mw = np.where(rng.random(n) < 0.5, 0.0, mw_init * 0.02 * rng.random(n))
    S, Q = np.zeros(n), np.zeros(n)
    m1, mc1 = m.copy(), mc.copy()
    dt = 0.01
    E.pyrolysis_md2004(T, m1, mc1, y, a, mw, mw_init, dt, S, Q)
    dm = m - m1
    # differences of float64 pools: to the rounding of the pools (1e-16 of 0.5)
    np.testing.assert_allclose(mc1 - mc, E.CHAR_YIELD_MD2004 * dm, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(S * dt, E.ETA_MD2004 * dm, rtol=1e-12, atol=1e-16)
    # the gate: no loss at >= 1 % of the initial water, full loss when dry
    wet = mw / mw_init >= 0.01
    assert np.all(dm[wet] == 0.0) and np.all(S[wet] == 0.0)
    dry = (mw == 0.0) & (a > 0.0)
    k = E.A_MD2004 * np.exp(-E.E_MD2004 / (E.R_GAS * T)) + np.where(
        y > E.Y_O2_MIN_OP, E.A_OP * np.exp(-E.E_OP / (E.R_GAS * T)) * y, 0.0)
    np.testing.assert_allclose(dm[dry], (m * (1.0 - np.exp(-k * dt)))[dry], rtol=1e-9,
                               atol=1e-16)
    # half the gate at half a per cent of the initial water
    m2, mc2, S2, Q2 = np.array([0.4]), np.array([0.0]), np.zeros(1), np.zeros(1)
    E.pyrolysis_md2004(np.array([900.0]), m2, mc2, np.array([0.0]), np.array([0.002]),
                       np.array([0.005 * mw_init]), mw_init, dt, S2, Q2)
    m3, mc3, S3, Q3 = np.array([0.4]), np.array([0.0]), np.zeros(1), np.zeros(1)
    E.pyrolysis_md2004(np.array([900.0]), m3, mc3, np.array([0.0]), np.array([0.002]),
                       np.array([0.0]), mw_init, dt, S3, Q3)
    np.testing.assert_allclose(0.4 - m2, 0.5 * (0.4 - m3), rtol=1e-12)
    # thermal only (no O2): Q = rate * heat of pyrolysis
    np.testing.assert_allclose(Q3, S3 / E.ETA_MD2004 * E.HEAT_OF_PYROLYSIS, rtol=1e-12)


@pytest.mark.parametrize("kind", ["char", "smold"])
def test_oxidation_caps_and_heat(kind):
    T, a, y, rng = _cells(seed=3)
    m = np.where(rng.random(T.size) < 0.1, 1e-7, 0.2 * rng.random(T.size))
    fn, HOC, QMAX, Ton = ((E.char_oxidation, E.HOC_CHAR, E.Q_CHAR_MAX, E.T_CHAR_ONSET)
                          if kind == "char" else
                          (E.smoldering, E.HOC_SMOLD, E.Q_SMOLD_MAX, E.T_SMOLD_ONSET))
    for dt in (0.001, 1.0):
        m1, Q = m.copy(), np.zeros_like(m)
        fn(T, m1, y, a, dt, Q)
        burnt = m - m1
        # burnt = m - (m - consumed) carries the rounding of m (~1e-17) times HOC
        np.testing.assert_allclose(Q * dt, burnt * HOC, rtol=1e-12, atol=HOC * 1e-16)
        assert np.all(burnt <= 0.5 * m * (1 + 1e-15))
        assert np.all(Q <= QMAX * (1 + 1e-12))
        skip = (a <= 0.0) | (m <= 1e-6) | (T < Ton) | (y < 1e-3)
        assert np.all(burnt[skip] == 0.0) and np.all(Q[skip] == 0.0)
        assert np.any(burnt > 0.0)


def test_oracle_eulerian_step_runs():
    from firewebapp_amd.deck import load_text_input
    from firewebapp_amd.grid import Grid3D
    from firewebapp_amd.spread_3d import Spread3DConfig
    from oracle.step import OracleSpread3D
    ri = load_text_input(os.path.join(GOLDEN_DIR, "integration_case.txt"))
    ri.outdoor_overrides["lagrangian_bed_enable"] = False
    cfg = Spread3DConfig.resolve(ri, wind_speed_m_s=4.0)
    assert not cfg.lagrangian_bed
    g = Grid3D.build(Lx=cfg.Lx, Ly=cfg.Ly, Lz=cfg.Lz, dx=cfg.dx, dy=cfg.dy, h_bed=cfg.h_bed,
                     n_z_bed=cfg.n_z_bed, **cfg.mesh_kwargs)
    o = OracleSpread3D(cfg, g)
    m0 = o.eu_m_solid.sum()
    o.run(40)
    assert o.n_alloc == 0 and o.step_no == 40
    for a in (o.T_g, o.T_s, o.eu_m_solid, o.eu_m_char, o.eu_m_water):
        assert np.all(np.isfinite(a))
    assert o.eu_m_solid.min() >= 0.0 and o.eu_m_water.min() >= 0.0
    assert o.eu_m_solid.sum() <= m0
    # the drip torch heats the source cells' solid through the Q_pyro sink
    s0, s1 = o.i_src
    assert o.T_s[:g.n_z_bed, :, s0:s1].max() > cfg.T_amb + 50.0
