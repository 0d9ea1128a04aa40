"""FP64 oracle of the EDC closure (test infrastructure only).

chemistry_closures/edc.py:133-334 (step_chemistry_ode_edc) with constants
from chemistry_closures/_constants.py:15-36.

The reference's ``_h2o_quench_substantial`` flag is assigned only when a
cell's fuel and O2 are both above 1e-9 (edc.py:222-244), so a cell without
fuel or O2 inherits the flag left by the previous cell in (k, j, i) loop
order (edc.py:223; docs/js/physics/edc.js:162-178 documents and reproduces
it).  A cell's own flag depends only on its Y_H2O, so the inherited value is
the flag of the nearest preceding cell whose entry Yf and YO2 both exceed
1e-9 (False if none).  This module computes that carry with a prefix scan
and checks it against a literal sequential loop in the tests.
"""
from __future__ import annotations

import math

import numpy as np

C_GAMMA_EDC = 2.1377
C_TAU_EDC = 0.4083
NU_GAS_EDC = 1.5e-5
K_TURB_FLOOR_EDC = 1.0e-4
EPS_TURB_FLOOR_EDC = 1.0e-6
CP_VAPOR_EDC = 2000.0
EXTINCTION_F_SAFETY = 1.5
Y_INERT_CRIT = 0.88
T_IGNITION_MIN = 1200.0
Y_H2O_QUENCH = 0.18
T_BOIL_WATER_EDC = 373.15
Y_H2O_SAT_THRESH = 0.10
S_STOICH = 1.3
HOC_J = 17_000_000.0


def own_quench_flag(Y_H2O):
    """Flag a resetting cell leaves behind (edc.py:222-244)."""
    f = 1.0 - Y_H2O / Y_H2O_QUENCH
    f = np.where(f < 0.0, 0.0, f)
    return (Y_H2O > 0.0) & (f < 0.5)


def carried_flags(Y_fuel, Y_O2, Y_H2O):
    """Incoming flag of every cell in flat (k, j, i) order."""
    reset = ((Y_fuel > 1e-9) & (Y_O2 > 1e-9)).ravel()
    own = own_quench_flag(Y_H2O).ravel()
    idx = np.where(reset, np.arange(reset.size), -1)
    last = np.maximum.accumulate(idx)
    prev = np.concatenate([[-1], last[:-1]])
    inc = np.where(prev >= 0, own[np.maximum(prev, 0)], False)
    return inc.reshape(Y_fuel.shape)


def step_chemistry_ode_edc(rho, T_g, Y_fuel, Y_O2, k_turb, eps_turb, chi_rad,
                           cp_g, dt, n_substeps, omega_int_out, Y_H2O,
                           extinction_enable=False, s_stoich=S_STOICH,
                           hoc_J=HOC_J):
    """In place on T_g, Y_fuel, Y_O2; writes omega_int_out."""
    n = max(int(n_substeps), 1)
    h = dt / n
    HoC_eff = hoc_J * (1.0 - chi_rad)
    flag_in = carried_flags(Y_fuel, Y_O2, Y_H2O)
    Yf = Y_fuel.copy()
    YO2 = Y_O2.copy()
    Tg = T_g.copy()
    k_t = np.where(k_turb < K_TURB_FLOOR_EDC, K_TURB_FLOOR_EDC, k_turb)
    e_t = np.where(eps_turb < EPS_TURB_FLOOR_EDC, EPS_TURB_FLOOR_EDC, eps_turb)
    ratio = NU_GAS_EDC * e_t / (k_t * k_t)
    ratio = np.where(ratio < 1e-30, 1e-30, ratio)
    gb = C_GAMMA_EDC * np.power(ratio, 0.25)
    gamma = gb * (gb * gb)
    gamma = np.where(gamma > 1.0, 1.0, gamma)
    tau = C_TAU_EDC * np.sqrt(NU_GAS_EDC / e_t)
    flag = flag_in.copy()
    acc = np.zeros_like(Tg)
    own = own_quench_flag(Y_H2O)
    h2o_f = np.where(1.0 - Y_H2O / Y_H2O_QUENCH < 0.0, 0.0, 1.0 - Y_H2O / Y_H2O_QUENCH)
    relax = 1.0 - math.exp(-h / 0.5)
    for _ in range(n):
        active = (Yf > 1e-9) & (YO2 > 1e-9)
        Y_lim = np.where(Yf < YO2 / s_stoich, Yf, YO2 / s_stoich)
        omega = gamma * (rho * Y_lim / tau)
        omega = np.where(Y_H2O > 0.0, omega * h2o_f, omega)
        omega = np.where(active, omega, 0.0)
        flag = np.where(active, own, flag)
        ext = np.zeros(Tg.shape, dtype=bool)
        if extinction_enable:
            on = omega > 0.0
            Yib = 1.0 - Yf - YO2
            supp = (1.0 - Yib) / (1.0 - Y_INERT_CRIT)
            supp = np.where(supp < 0.0, 0.0, supp)
            hit = on & (Yib > Y_INERT_CRIT)
            omega = np.where(hit, omega * supp, omega)
            ext |= hit & (supp < 0.5)
            g2 = on & (omega > 0.0) & (Tg < T_IGNITION_MIN)
            omega = np.where(g2, 0.0, omega)
            ext |= g2
            g3 = on & (omega > 0.0) & (omega * HoC_eff < EXTINCTION_F_SAFETY * 5.0e4)
            omega = np.where(g3, 0.0, omega)
            ext |= g3
        wb_on = (ext | flag) & (Tg > T_BOIL_WATER_EDC) & (Y_H2O > 0.0)
        wb = np.where(Y_H2O / Y_H2O_SAT_THRESH > 1.0, 1.0, Y_H2O / Y_H2O_SAT_THRESH)
        dT_wb = (Tg - T_BOIL_WATER_EDC) * relax * wb
        Tg = np.where(wb_on, Tg - dT_wb, Tg)
        dY = -omega * h / rho
        Yf = Yf + dY
        Yf = np.where(Yf < 0.0, 0.0, Yf)
        YO2 = YO2 + s_stoich * dY
        YO2 = np.where(YO2 < 0.0, 0.0, YO2)
        cp_mix = (1.0 - Y_H2O) * cp_g + Y_H2O * CP_VAPOR_EDC
        Tg = Tg + omega * HoC_eff * h / (rho * cp_mix)
        Tg = np.where(Tg > 2400.0, 2400.0, Tg)
        acc = acc + omega * h
    Y_fuel[...] = Yf
    Y_O2[...] = YO2
    T_g[...] = Tg
    omega_int_out[...] = acc / dt if dt > 0.0 else 0.0


def step_chemistry_ode_edc_sequential(rho, T_g, Y_fuel, Y_O2, k_turb,
                                      eps_turb, chi_rad, cp_g, dt, n_substeps,
                                      omega_int_out, Y_H2O,
                                      extinction_enable=False,
                                      s_stoich=S_STOICH, hoc_J=HOC_J):
    """Literal scalar-loop restatement (small grids only): pins the
    prefix-scan formulation of the carried flag above."""
    shp = rho.shape
    R, Tg_, Yf_, YO_, kt, et, Yw, om = (a.reshape(-1) for a in
                                         (rho, T_g, Y_fuel, Y_O2, k_turb,
                                          eps_turb, Y_H2O, omega_int_out))
    n = max(int(n_substeps), 1)
    h = dt / n
    HoC_eff = hoc_J * (1.0 - chi_rad)
    flag = False
    for c in range(R.size):
        Yf, YO2, Tg, rho_i = Yf_[c], YO_[c], Tg_[c], R[c]
        k_t = kt[c] if not kt[c] < K_TURB_FLOOR_EDC else K_TURB_FLOOR_EDC
        e_t = et[c] if not et[c] < EPS_TURB_FLOOR_EDC else EPS_TURB_FLOOR_EDC
        ratio = NU_GAS_EDC * e_t / (k_t * k_t)
        if ratio < 1e-30:
            ratio = 1e-30
        gb = C_GAMMA_EDC * ratio ** 0.25
        gamma = gb * (gb * gb)
        if gamma > 1.0:
            gamma = 1.0
        tau = C_TAU_EDC * math.sqrt(NU_GAS_EDC / e_t)
        acc = 0.0
        for _ in range(n):
            if Yf <= 1e-9 or YO2 <= 1e-9:
                omega = 0.0
            else:
                Y_lim = Yf if Yf < (YO2 / s_stoich) else (YO2 / s_stoich)
                omega = gamma * (rho_i * Y_lim / tau)
                y = Yw[c]
                flag = False
                if y > 0.0:
                    f = 1.0 - y / Y_H2O_QUENCH
                    if f < 0.0:
                        f = 0.0
                    omega = omega * f
                    if f < 0.5:
                        flag = True
            ext = False
            if extinction_enable and omega > 0.0:
                Yib = 1.0 - Yf - YO2
                if Yib > Y_INERT_CRIT:
                    s = (1.0 - Yib) / (1.0 - Y_INERT_CRIT)
                    if s < 0.0:
                        s = 0.0
                    omega = omega * s
                    if s < 0.5:
                        ext = True
                if omega > 0.0 and Tg < T_IGNITION_MIN:
                    omega = 0.0
                    ext = True
                if omega > 0.0 and omega * HoC_eff < EXTINCTION_F_SAFETY * 5.0e4:
                    omega = 0.0
                    ext = True
            if (ext or flag) and Tg > T_BOIL_WATER_EDC:
                y = Yw[c]
                if y > 0.0:
                    wb = y / Y_H2O_SAT_THRESH
                    if wb > 1.0:
                        wb = 1.0
                    Tg = Tg - (Tg - T_BOIL_WATER_EDC) * (1.0 - math.exp(-h / 0.5)) * wb
            dY = -omega * h / rho_i
            Yf = Yf + dY
            if Yf < 0.0:
                Yf = 0.0
            YO2 = YO2 + s_stoich * dY
            if YO2 < 0.0:
                YO2 = 0.0
            y = Yw[c]
            cp_mix = (1.0 - y) * cp_g + y * CP_VAPOR_EDC
            Tg = Tg + omega * HoC_eff * h / (rho_i * cp_mix)
            if Tg > 2400.0:
                Tg = 2400.0
            acc += omega * h
        Yf_[c], YO_[c], Tg_[c] = Yf, YO2, Tg
        om[c] = acc / dt if dt > 0.0 else 0.0
    for dst, src in ((T_g, Tg_), (Y_fuel, Yf_), (Y_O2, YO_), (omega_int_out, om)):
        dst[...] = src.reshape(shp)
