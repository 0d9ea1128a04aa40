"""GPU parity of the EDC closure through the C ABI.

FP64: golden vectors at 1e-12 relative on every field (the reference's own
standard is 1e-9 for omega/Y and 1e-2 for T_g, docs/kerneltest.mjs:289-298).
A seeded 8x16x64 case (32 scan segments, long fuel-free runs, wet cells)
pins the carried wet-bulb flag across segment boundaries against the
literal sequential oracle loop.  FP32: the golden vectors at the reference's
cross-implementation standard (T_g 1e-2, Y/omega 1e-4 relative).
"""
import numpy as np
import pytest

from oracle import edc as E
from tests.golden_io import case_ids, cases, field, relerr

pytestmark = pytest.mark.gpu


def _gpu_edc(args, c, to=None):
    from firewebapp_amd.physics_3d import chemistry_closures as cc
    rho, T, Yf, YO, k, e, Yw = args
    om = np.zeros_like(T) if to is None else to(np.zeros_like(T))
    if to is not None:
        rho, T, Yf, YO, k, e, Yw = (to(a) for a in (rho, T, Yf, YO, k, e, Yw))
    cc.run("edc", rho=rho, T_g=T, Y_fuel=Yf, Y_O2=YO, k_turb=k, eps_turb=e,
           chi_rad=c["chi_rad"], cp_g=c["cp_g"], dt=c["dt"],
           n_substeps=c["n_substeps"], omega_out=om, Y_H2O=Yw,
           extinction_enable=c["extinction_enable"], tau_mix=None)
    return T, Yf, YO, om


def _args(c):
    return [field(c, k) for k in ("rho", "T_g_in", "Y_fuel_in", "Y_O2_in",
                                  "k_turb", "eps_turb", "Y_H2O")]


@pytest.mark.parametrize("c", cases("edc"), ids=case_ids("edc"))
def test_edc_fp64_golden(c):
    T, Yf, YO, om = _gpu_edc(_args(c), c)
    assert relerr(Yf, field(c, "Y_fuel_out")) <= 1e-12
    assert relerr(YO, field(c, "Y_O2_out")) <= 1e-12
    assert relerr(om, field(c, "omega_out")) <= 1e-12
    assert relerr(T, field(c, "T_g_out")) <= 1e-12


@pytest.mark.parametrize("c", cases("edc"), ids=case_ids("edc"))
def test_edc_fp32_golden(c):
    import torch
    to = lambda a: torch.as_tensor(np.asarray(a, np.float32), device="cuda")
    T, Yf, YO, om = (x.double().cpu().numpy() for x in _gpu_edc(_args(c), c, to))
    assert relerr(T, field(c, "T_g_out")) <= 1e-2
    assert relerr(Yf, field(c, "Y_fuel_out")) <= 1e-4
    assert relerr(YO, field(c, "Y_O2_out")) <= 1e-4


def test_edc_flag_carry_across_segments():
    rng = np.random.default_rng(2024)
    sh = (8, 16, 64)
    rho = 0.4 + 0.8 * rng.random(sh)
    T = 350.0 + 1500.0 * rng.random(sh)
    Yf = np.where(rng.random(sh) < 0.97, 0.0, 0.08 * rng.random(sh))
    YO = 0.232 * rng.random(sh)
    Yw = 0.2 * rng.random(sh)
    k = 0.8 * rng.random(sh)
    e = 0.8 * rng.random(sh)
    c = dict(chi_rad=0.34, cp_g=1100.0, dt=0.002, n_substeps=1,
             extinction_enable=False)
    want = [a.copy() for a in (T, Yf, YO)] + [np.zeros(sh)]
    E.step_chemistry_ode_edc_sequential(rho, want[0], want[1], want[2], k, e,
                                        0.34, 1100.0, 0.002, 1, want[3], Yw)
    got = _gpu_edc([rho, T.copy(), Yf.copy(), YO.copy(), k, e, Yw], c)
    for x, y in zip(got, want):
        assert relerr(x, y) <= 1e-12


@pytest.mark.parametrize("fuel_cells", [5, 0])
def test_edc_flag_carry_long_lookback(fuel_cells):
    """Resetting cells only at the very start (or none): segments more than
    256 segments later inherit their flag through several look-back rounds
    (or keep the initial false flag)."""
    rng = np.random.default_rng(77)
    sh = (20, 16, 256)          # 320 segments of 256 cells
    rho = 0.4 + 0.8 * rng.random(sh)
    T = 350.0 + 1500.0 * rng.random(sh)
    Yf = np.zeros(sh)
    Yf.reshape(-1)[:fuel_cells] = 0.05
    YO = 0.232 * rng.random(sh)
    Yw = 0.2 * rng.random(sh)
    Yw.reshape(-1)[fuel_cells - 1 if fuel_cells else 0] = 0.17   # quenching flag set
    k = 0.8 * rng.random(sh)
    e = 0.8 * rng.random(sh)
    c = dict(chi_rad=0.34, cp_g=1100.0, dt=0.002, n_substeps=1,
             extinction_enable=False)
    want = [a.copy() for a in (T, Yf, YO)] + [np.zeros(sh)]
    E.step_chemistry_ode_edc_sequential(rho, want[0], want[1], want[2], k, e,
                                        0.34, 1100.0, 0.002, 1, want[3], Yw)
    got = _gpu_edc([rho, T.copy(), Yf.copy(), YO.copy(), k, e, Yw], c)
    for x, y in zip(got, want):
        assert relerr(x, y) <= 1e-12
