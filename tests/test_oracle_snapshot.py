"""Oracle of the snapshot fields, and the result file round trip (CPU).

* phi_flame: the golden flame_front vectors (the reference's own
  compute_phi_flame_from_state, both clauses of the active criterion, a
  non-uniform dz) and a literal all-pairs minimum on random tiny grids,
  including the all-active / none-active sentinels;
* ahead band mask: the golden ahead_mask from the golden reinitialised level
  set;
* q_in_at_front_2d: the golden q_in (which adds a burst term to the top band
  cell of each column) with that term taken back out, and numpy's own z sum;
* bed aggregates: invariants (the live, located particles are counted once,
  float32 mass sums within float32 rounding of the float64 sums, the weighted
  T_s inside each cell's range) and a hand-worked two-particle cell;
* tau_mix: its two floors by hand;
* Spread3DResult.save / load: the reference's key set and JSON metadata, a
  lossless round trip, and the defaults for files without Y_H2O, Y_CO, m_char.
"""
import json
import math

import numpy as np
import pytest

from oracle import snapshot as S
from tests.golden_io import arr, case_ids, cases, field

SEC = "flame_front"


@pytest.mark.parametrize("c", cases(SEC), ids=case_ids(SEC))
def test_phi_flame_golden(c):
    dz = arr(c, "dz_arr")
    got = S.phi_flame(field(c, "omega"), field(c, "T_g"), field(c, "Y_fuel"), c["dx"],
                      c["dy"], dz)
    want = field(c, "phi_flame")
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)
    brute = S.phi_flame_brute(field(c, "omega"), field(c, "T_g"), field(c, "Y_fuel"),
                              c["dx"], c["dy"], dz)
    np.testing.assert_allclose(brute, want, rtol=1e-12, atol=0)


def _random_state(rng, shape, frac):
    om = np.where(rng.random(shape) < frac, 5e-2, 0.0)
    Tg = np.where(rng.random(shape) < 0.1, 1400.0, 300.0)
    Yf = np.where(rng.random(shape) < 0.5, 0.02, 0.0)
    return om, Tg, Yf


@pytest.mark.parametrize("shape,frac", [((1, 1, 1), 0.5), ((3, 1, 7), 0.2), ((4, 3, 5), 0.1),
                                        ((5, 4, 3), 0.6), ((2, 6, 2), 0.0), ((3, 2, 4), 1.0)])
def test_phi_flame_brute_force(shape, frac):
    rng = np.random.default_rng(int(np.prod(shape)) + int(10 * frac))
    om, Tg, Yf = _random_state(rng, shape, frac)
    if frac == 0.0:
        Tg[:] = 300.0
    if frac == 1.0:
        om[:] = 1.0
    dz = 0.03 + 0.1 * rng.random(shape[0])
    got = S.phi_flame(om, Tg, Yf, 0.1, 0.07, dz)
    want = S.phi_flame_brute(om, Tg, Yf, 0.1, 0.07, dz)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    act = S.flame_active(om, Tg, Yf)
    if act.all():
        assert np.all(got == -1.0e6)
    elif not act.any():
        assert np.all(got == 1.0e6)
    else:
        assert np.array_equal(got <= 0.0, act)


@pytest.mark.parametrize("c", cases(SEC), ids=case_ids(SEC))
def test_ahead_band_golden(c):
    m = S.ahead_band_mask(field(c, "phi_reinit"), c["band_m"], c["n_z_bed"])
    assert np.array_equal(m.astype(np.uint8), field(c, "ahead_mask").astype(np.uint8))


@pytest.mark.parametrize("c", cases(SEC), ids=case_ids(SEC))
def test_q_in_2d_golden(c):
    mask = field(c, "ahead_mask") != 0
    q_dom = field(c, "q_dom_fwd")
    burst = arr(c, "q_burst").reshape(c["ny"], c["nx"])
    got = S.q_in_at_front_2d(q_dom, mask)
    # the golden q_in carries the burst term in the top band cell of each
    # column that has one
    want = field(c, "q_in").sum(axis=0) - np.where(mask.any(axis=0), burst, 0.0)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-9)
    assert np.array_equal(got, np.where(mask, q_dom, 0.0).sum(axis=0))


@pytest.mark.parametrize("N", [4, 8])
def test_q_dom_fwd_hemisphere_moments(N):
    """Closed forms of the forward flux on the S4 / S8 sets: an intensity
    equal to the direction cosine xi gives sum_{xi>0} w xi^2 = 2 pi / 3 (half
    the even moment, exact for both sets); an isotropic intensity I0 gives
    sum_{xi>0} w xi I0 = pi I0 (S4; the LQ8 half-range first moment is 1.7 %
    above pi); an intensity on backward ordinates only gives 0."""
    from oracle import dom as D
    Om, w = D.sn_ordinates(N)
    sh = (3, 2, 4)
    mask = np.zeros(sh, bool)
    mask[0, :, 1:3] = True
    mask[2, 1, 0] = True
    I0 = 7.5
    iso = np.full((len(w),) + sh, I0)
    q = S.q_dom_fwd_at_band(Om, w, iso, mask)
    np.testing.assert_allclose(q[mask], math.pi * I0, rtol=1e-5 if N == 4 else 0.02)
    assert np.all(q[~mask] == 0.0)
    lin = np.broadcast_to(Om[:, 0][:, None, None, None], (len(w),) + sh)
    q = S.q_dom_fwd_at_band(Om, w, lin, mask)
    np.testing.assert_allclose(q[mask], 2.0 * math.pi / 3.0, rtol=1e-5)
    back = np.broadcast_to((Om[:, 0] < 0.0)[:, None, None, None] * 1.0, (len(w),) + sh)
    assert not S.q_dom_fwd_at_band(Om, w, back, mask).any()


def test_tau_mix_floors():
    tb = S.tau_buoy(0.1, 0.1, 0.042)
    assert tb == math.sqrt(2.0 * 0.1 / 9.81)
    k = np.array([1.0, 1.0, 1e-6, 0.0])
    eps = np.array([2.0, 0.0, 1.0, 1.0])
    got = S.tau_mix(k, eps, tb)
    assert got[0] == 0.5
    assert got[1] == 1.0e12           # eps floored at 1e-12
    assert got[2] == tb and got[3] == tb


def _particles(rng, n, shape, dx, dy, z_face):
    Nz, Ny, Nx = shape
    x = rng.uniform(-0.05, (Nx + 0.5) * dx, n)
    y = rng.uniform(0.0, Ny * dy * 1.05, n)
    z = rng.uniform(0.0, z_face[-1] * 1.1, n)
    alive = (rng.random(n) < 0.8).astype(np.int8)
    ms, mw, mc = rng.random(n) * 1e-3, rng.random(n) * 1e-4, rng.random(n) * 1e-5
    Ts = 300.0 + 900.0 * rng.random(n)
    return x, y, z, alive, ms, mw, mc, Ts


def test_bed_snapshot_invariants():
    rng = np.random.default_rng(5)
    shape, dx, dy = (3, 4, 6), 0.1, 0.1
    z_face = np.array([0.0, 0.04, 0.08, 0.2])
    x, y, z, alive, ms, mw, mc, Ts = _particles(rng, 400, shape, dx, dy, z_face)
    cnt, s, w, ch, tavg = S.bed_snapshot(x, y, z, alive, ms, mw, mc, Ts, dx, dy, z_face,
                                         shape)
    for a in (cnt, s, w, ch, tavg):
        assert a.dtype == np.float32 and a.shape == shape
    inside = ((x >= 0) & (x < shape[2] * dx) & (y < shape[1] * dy) & (z < z_face[-1]))
    # int() truncates toward zero: -0.05 < x < 0 still lands in column 0
    inside |= (x > -dx) & (x < 0) & (y < shape[1] * dy) & (z < z_face[-1])
    live = inside & (alive != 0)
    assert int(cnt.sum()) == int(live.sum())
    assert abs(float(s.sum(dtype=np.float64)) - ms[live].sum()) <= 1e-6 * ms[live].sum()
    occ = cnt > 0
    assert np.all(tavg[~occ] == 0.0)
    kk = np.searchsorted(z_face, z, side="right") - 1
    ii, jj = (x / dx).astype(int), (y / dy).astype(int)
    for c in zip(*np.nonzero(occ)):
        sel = live & (kk == c[0]) & (jj == c[1]) & (ii == c[2])
        assert Ts[sel].min() * (1 - 1e-6) <= tavg[c] <= Ts[sel].max() * (1 + 1e-6)


def test_bed_snapshot_two_particles():
    x = np.array([0.05, 0.06])
    y = np.array([0.05, 0.05])
    z = np.array([0.01, 0.02])
    alive = np.array([1, 1], np.int8)
    ms = np.array([2.0, 1.0])
    mw = np.array([1.0, 0.0])
    mc = np.array([0.0, 1.0])
    Ts = np.array([400.0, 1000.0])
    cnt, s, w, ch, tavg = S.bed_snapshot(x, y, z, alive, ms, mw, mc, Ts, 0.1, 0.1,
                                         np.array([0.0, 0.1]), (1, 1, 1))
    assert cnt[0, 0, 0] == 2.0 and s[0, 0, 0] == 3.0 and w[0, 0, 0] == 1.0
    assert ch[0, 0, 0] == 1.0
    # (400 * 3 + 1000 * 2) / 5
    assert tavg[0, 0, 0] == np.float32(640.0)


# ---- Spread3DResult.save / load ------------------------------------------------

def _result(rng):
    from firewebapp_amd.grid import Grid3D
    from firewebapp_amd.spread_3d import CellState3D, Spread3DResult
    g = Grid3D.build(Lx=0.8, Ly=0.3, Lz=0.5, dx=0.1, dy=0.1, h_bed=0.1, n_z_bed=2)
    sh = (g.Nz, g.Ny, g.Nx)
    st = CellState3D(**{k: rng.random(sh) for k in (
        "rho", "u", "v", "w", "p_dyn", "T_g", "Y_fuel", "Y_O2", "Y_H2O", "Y_CO", "alpha_s",
        "T_s", "m_hemi", "m_cell", "m_lign", "m_char")})
    return Spread3DResult(
        ros_m_s=0.0123, n_cells_ignited=17, front_t=[0.0, 0.5], front_x=[0.3, 0.4], grid=g,
        state_final=st, diag_t=[0.0, 0.1], diag_Tg_max=[300.0, 900.0],
        diag_Ts_max=[300.0, 700.0], diag_Sp_max=[0.0, 1e-3], diag_Qc_max=[0.0, 5e5],
        diag_omega_max=[0.0, 1e-2], diag_Y_max=[0.0, 3e-2], diag_n_ign=[0, 4],
        diag_proj_div_max=[0.0, 1e-4], diag_proj_n_iter=[0, 3])


def test_result_save_load_round_trip(tmp_path):
    from firewebapp_amd.spread_3d import Spread3DResult
    r = _result(np.random.default_rng(9))
    r.save(tmp_path / "res", metadata={"wind": 4.0, "deck": "x.txt"})
    path = tmp_path / "res.npz"
    with np.load(path, allow_pickle=False) as z:
        assert sorted(z.files) == sorted(S.RESULT_KEYS)
        meta = json.loads(str(z["_meta_json"]))
    assert list(meta) == list(S.RESULT_META_KEYS)
    assert meta["user"] == {"wind": 4.0, "deck": "x.txt"} and meta["dz"] == r.grid.dz
    back = Spread3DResult.load(tmp_path / "res")
    assert back.ros_m_s == r.ros_m_s and back.n_cells_ignited == r.n_cells_ignited
    assert back.front_t == r.front_t and back.front_x == r.front_x
    for k in ("diag_t", "diag_Tg_max", "diag_Ts_max", "diag_Sp_max", "diag_Qc_max",
              "diag_omega_max", "diag_Y_max", "diag_n_ign"):
        assert getattr(back, k) == getattr(r, k), k
    # the reference's file has no projection diagnostics
    assert back.diag_proj_div_max == [] and back.diag_proj_n_iter == []
    for k in ("x_mid", "y_mid", "z_mid", "dz_arr", "inv_dz_arr", "z_face", "d_face_above",
              "d_face_below"):
        assert np.array_equal(getattr(back.grid, k), getattr(r.grid, k)), k
    for k in ("Nx", "Ny", "Nz", "dx", "dy", "dz", "Lx", "Ly", "Lz", "n_z_bed"):
        assert getattr(back.grid, k) == getattr(r.grid, k), k
    for k in ("rho", "u", "v", "w", "p_dyn", "T_g", "Y_fuel", "Y_O2", "Y_H2O", "Y_CO",
              "alpha_s", "T_s", "m_hemi", "m_cell", "m_lign", "m_char"):
        assert np.array_equal(getattr(back.state_final, k), getattr(r.state_final, k)), k


def test_result_load_defaults(tmp_path):
    from firewebapp_amd.spread_3d import Spread3DResult
    r = _result(np.random.default_rng(10))
    r.save(tmp_path / "full.npz")
    with np.load(tmp_path / "full.npz", allow_pickle=False) as z:
        d = {k: z[k] for k in z.files if k not in ("Y_H2O", "Y_CO", "m_char")}
    np.savez_compressed(tmp_path / "old.npz", **d)
    back = Spread3DResult.load(tmp_path / "old.npz")
    for k in ("Y_H2O", "Y_CO", "m_char"):
        a = getattr(back.state_final, k)
        assert a.shape == r.state_final.T_g.shape and not a.any(), k
