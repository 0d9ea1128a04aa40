"""Oracle level set and flame-body predicate vs the golden flame_front
vectors (bit-exact: Godunov gradients, evolve, reinit, masks, front_x), plus
the sign property that lets the step replace the two distance transforms of
compute_phi_flame_from_state by the `active` predicate."""
import math

import numpy as np
import pytest

from oracle import levelset as LS
from tests.golden_io import arr, case_ids, cases, field

SEC = "flame_front"


@pytest.mark.parametrize("c", cases(SEC), ids=case_ids(SEC))
def test_init_grad_evolve_reinit(c):
    sh = (c["nz"], c["ny"], c["nx"])
    dz = arr(c, "dz_arr")
    x_mid = arr(c, "x_mid")
    nzb = c["n_z_bed"]
    phi = LS.initialize_source_patch(sh, 1, 4, nzb - 1, x_mid, c["dx"])
    assert np.array_equal(phi, field(c, "phi_init"))
    g = LS.godunov_grad_norm(phi, c["dx"], c["dy"], dz)
    assert np.array_equal(g, field(c, "grad_out"))
    gr = LS.reinit_godunov_grad(field(c, "phi_dist"), phi, c["dx"], c["dy"], dz)
    assert np.array_equal(gr, field(c, "grad_reinit"))
    ev = LS.evolve(phi, c["dt"], field(c, "v_n_in"), c["dx"], c["dy"], dz)
    assert np.array_equal(ev, field(c, "phi_evolved"))
    re = LS.reinitialize(ev, c["dx"], c["dy"], dz)
    assert np.array_equal(re, field(c, "phi_reinit"))
    fb = (re >= -LS.L_BURNOUT_M) & (re <= 0.0)
    assert np.array_equal(fb.astype(np.uint8), field(c, "flame_body_mask").astype(np.uint8))
    assert np.array_equal((re <= 0.0).astype(np.uint8),
                          field(c, "burned_mask").astype(np.uint8))
    fx = LS.front_x(re, 1, c["ny"] // 2, c["dx"])
    assert fx == c["front_x"]


@pytest.mark.parametrize("c", cases(SEC), ids=case_ids(SEC))
def test_flame_active_equals_phi_flame_sign(c):
    om, tg, yf = field(c, "omega"), field(c, "T_g"), field(c, "Y_fuel")
    act = LS.flame_active(om, tg, yf)
    assert np.array_equal(act, field(c, "phi_flame") <= 0.0)
    assert np.array_equal(act.astype(np.uint8), field(c, "fb_from_phi").astype(np.uint8))
    assert int(act.sum()) == c["n_active"]
    # both clauses are exercised: reacting core and hot fuel-bearing tail
    assert ((om > LS.OMEGA_MIN_FLAME) & act).any()
    assert ((om <= LS.OMEGA_MIN_FLAME) & act).any()


def test_evolve_with_zero_speed_is_identity():
    rng = np.random.default_rng(0)
    phi = rng.normal(size=(5, 3, 7))
    dz = 0.02 + 0.05 * rng.random(5)
    assert np.array_equal(LS.evolve(phi, 0.01, np.zeros_like(phi), 0.1, 0.1, dz), phi)


def test_front_x_linear_crossing():
    phi = np.zeros((2, 1, 6))
    phi[:, 0] = [-0.3, -0.1, 0.1, 0.3, 0.5, 0.7]
    # zero crossing half-way between cells 1 and 2 -> x = (1 + 0.5 + 0.5) dx
    assert LS.front_x(phi, 1, 0, 0.1) == pytest.approx(0.2, abs=1e-15)
    assert LS.front_x(np.ones((1, 1, 4)), 0, 0, 0.1) == math.inf


@pytest.mark.parametrize("c", cases(SEC), ids=case_ids(SEC))
def test_v_n_golden(c):
    """compute_v_n_3d, dry and moisture-aware, on the golden q_in."""
    q = field(c, "q_in")
    dry = LS.compute_v_n_3d(q, 1.07, 1850.0, 0.25, 600.0, 300.0)
    wet = LS.compute_v_n_3d(q, 1.07, 1850.0, 0.25, 600.0, 300.0, M_local=field(c, "M_local"))
    assert np.array_equal(dry, field(c, "v_n_dry"))
    assert np.array_equal(wet, field(c, "v_n_wet"))
