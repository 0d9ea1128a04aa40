"""Oracle EDC vs the golden vectors and vs a literal sequential loop.

Standard (docs/kerneltest.mjs:289-298): omega / Y at 1e-9 relative, T_g at
1e-2 relative, because the discontinuous wet-bulb gate turns ulps into
kelvins.  The oracle here reaches ~1e-12 on every field.
"""
import numpy as np
import pytest

from oracle import edc as E
from tests.golden_io import case_ids, cases, field, relerr


def _run(fn, c):
    T, Yf, YO = field(c, "T_g_in"), field(c, "Y_fuel_in"), field(c, "Y_O2_in")
    om = np.zeros_like(T)
    fn(field(c, "rho"), T, Yf, YO, field(c, "k_turb"), field(c, "eps_turb"),
       c["chi_rad"], c["cp_g"], c["dt"], c["n_substeps"], om, field(c, "Y_H2O"),
       extinction_enable=c["extinction_enable"], s_stoich=c["s_stoich"],
       hoc_J=c["hoc_J"])
    return T, Yf, YO, om


@pytest.mark.parametrize("c", cases("edc"), ids=case_ids("edc"))
def test_edc_golden(c):
    T, Yf, YO, om = _run(E.step_chemistry_ode_edc, c)
    assert relerr(Yf, field(c, "Y_fuel_out")) <= 1e-12
    assert relerr(YO, field(c, "Y_O2_out")) <= 1e-12
    assert relerr(om, field(c, "omega_out")) <= 1e-12
    assert relerr(T, field(c, "T_g_out")) <= 1e-12


@pytest.mark.parametrize("c", cases("edc"), ids=case_ids("edc"))
def test_edc_scan_equals_sequential_loop(c):
    a = _run(E.step_chemistry_ode_edc, c)
    b = _run(E.step_chemistry_ode_edc_sequential, c)
    for x, y in zip(a, b):
        assert relerr(x, y) <= 1e-14


def test_flag_leak_is_exercised():
    """A seeded case where a fuel-free cell inherits a True flag from the
    preceding cell and is wet-bulb cooled; a dropped carry changes T_g."""
    rng = np.random.default_rng(7)
    sh = (3, 2, 9)
    rho = 0.5 + 0.5 * rng.random(sh)
    T = 900.0 + 600.0 * rng.random(sh)
    Yf = np.where(rng.random(sh) < 0.5, 0.0, 0.05 * rng.random(sh))
    YO = 0.232 * rng.random(sh)
    Yw = 0.15 * rng.random(sh)
    k = 0.5 * rng.random(sh)
    e = 0.5 * rng.random(sh)
    inc = E.carried_flags(Yf, YO, Yw)
    no_reset = ~((Yf > 1e-9) & (YO > 1e-9))
    assert (inc & no_reset).any()
    outs = []
    for fn in (E.step_chemistry_ode_edc, E.step_chemistry_ode_edc_sequential):
        Tt, Yft, YOt, om = T.copy(), Yf.copy(), YO.copy(), np.zeros(sh)
        fn(rho, Tt, Yft, YOt, k, e, 0.34, 1100.0, 0.002, 1, om, Yw)
        outs.append(Tt)
    assert np.array_equal(outs[0], outs[1])
    # without the carry (flag forced False on fuel-free cells) T_g differs
    Tt, Yft, YOt, om = T.copy(), Yf.copy(), YO.copy(), np.zeros(sh)
    orig = E.carried_flags
    try:
        E.carried_flags = lambda a, b, c: np.zeros(a.shape, bool)
        E.step_chemistry_ode_edc(rho, Tt, Yft, YOt, k, e, 0.34, 1100.0, 0.002, 1, om, Yw)
    finally:
        E.carried_flags = orig
    assert not np.array_equal(Tt, outs[0])
