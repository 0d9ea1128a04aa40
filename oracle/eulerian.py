"""Eulerian fuel bed (oracle; test infrastructure only).

Restates, vectorised over cells (every kernel is cell-local, so the order
of operations per cell is the reference's):
* ``drying``              pyrolysis_3d.step_drying (pyrolysis_3d.py:270-309)
* ``pyrolysis_md2004``    pyrolysis_3d.step_pyrolysis_md2004 (:311-462)
* ``char_oxidation``      pyrolysis_3d.step_char_oxidation (:465-564)
* ``smoldering``          pyrolysis_3d.step_smoldering_oxidation (:567-614)
* ``gas_solid_coupling``  coupling_3d.step_gas_solid_coupling
                          (coupling_3d.py:76-192)
Arrays are float64; functions update their in/out arrays in place like the
reference kernels.
"""
from __future__ import annotations

import numpy as np

R_GAS = 8.314
A_DRY, E_DRY, L_VAP_WATER = 4.29e3, 43_800.0, 2.26e6
A_MD2004, E_MD2004 = 36_280.0, 58_600.0
HEAT_OF_PYROLYSIS = 4.0e5
CHAR_YIELD_MD2004 = 0.15
ETA_MD2004 = 1.0 - CHAR_YIELD_MD2004
A_OP, E_OP, HOR_OP, Y_O2_MIN_OP, N_O2_OP = 7.26e4, 55_000.0, -1.0e6, 1.0e-3, 1.0
A_CHAR, E_CHAR, HOC_CHAR = 7.09e5, 94_000.0, 32_000_000.0
T_CHAR_ONSET, Y_O2_MIN_CHAR, Q_CHAR_MAX = 600.0, 1.0e-3, 5.0e5
A_SMOLD, E_SMOLD, HOC_SMOLD = 1.0e6, 80_000.0, 28_000_000.0
T_SMOLD_ONSET, Y_O2_MIN_SMOLD, Q_SMOLD_MAX = 473.0, 1.0e-3, 2.0e5
MU_GAS, K_GAS, PR_GAS = 1.8e-5, 0.026, 0.7
RHO_SOLID, CP_SOLID, EPS_SOLID, SIGMA_SB, H_GROUND = 500.0, 1300.0, 0.9, 5.67e-8, 5.0


def drying(T_s, m_water, alpha_s, dt, Q_dry_out):
    live = (alpha_s > 0.0) & (m_water > 0.0) & (T_s > 0.0)
    Ts = np.where(live, T_s, 1.0)
    k = A_DRY * np.exp(-E_DRY / (R_GAS * Ts))
    mw_new = m_water * np.exp(-k * dt)
    dm = m_water - mw_new
    Q_dry_out[...] = np.where(live, dm * L_VAP_WATER / dt if dt > 0.0 else 0.0, 0.0)
    m_water[...] = np.where(live, mw_new, m_water)


def pyrolysis_md2004(T_s, m_solid, m_char, Y_O2, alpha_s, m_water, m_water_init, dt,
                     S_out, Q_out, A_p=A_MD2004, E_p=E_MD2004):
    inv = 1.0 / m_water_init if m_water_init > 0.0 else 0.0
    gate = 1.0 - 100.0 * (m_water * inv) if inv > 0.0 else np.ones_like(T_s)
    live = (alpha_s > 0.0) & (T_s > 0.0) & (gate > 0.0) & (m_solid > 0.0)
    Ts = np.where(live, T_s, 1.0)
    k_th = A_p * np.exp(-E_p / (R_GAS * Ts))
    k_ox = np.where(Y_O2 > Y_O2_MIN_OP,
                    A_OP * np.exp(-E_OP / (R_GAS * Ts)) * Y_O2 ** N_O2_OP, 0.0)
    k_tot = k_th + k_ox
    m_new = m_solid * np.exp(-k_tot * dt)
    dm = (m_solid - m_new) * gate
    rate = dm / dt if dt > 0.0 else np.zeros_like(dm)
    pos = k_tot > 0.0
    kt = np.where(pos, k_tot, 1.0)
    f_th = np.where(pos, k_th / kt, 1.0)
    f_ox = np.where(pos, k_ox / kt, 0.0)
    S_out[...] = np.where(live, ETA_MD2004 * rate, 0.0)
    Q_out[...] = np.where(live, rate * (f_th * HEAT_OF_PYROLYSIS + f_ox * HOR_OP), 0.0)
    m_char[...] = np.where(live, m_char + CHAR_YIELD_MD2004 * dm, m_char)
    m_solid[...] = np.where(live, m_solid - dm, m_solid)


def _oxidation(T_s, m, Y_O2, alpha_s, dt, Q_out, A, E, HOC, T_on, y_min, Q_max):
    live = (alpha_s > 0.0) & (m > 1.0e-6) & (T_s >= T_on) & (Y_O2 >= y_min)
    k = A * np.exp(-E / (R_GAS * np.where(live, T_s, 1.0)))
    m_dot = k * m * Y_O2
    m_dot = np.where(m_dot * HOC > Q_max, Q_max / HOC, m_dot)
    mc = m_dot * dt
    mc = np.where(mc > 0.5 * m, 0.5 * m, mc)
    Q_out[...] = np.where(live, mc * HOC / dt if dt > 0.0 else 0.0, 0.0)
    m[...] = np.where(live, m - mc, m)


def char_oxidation(T_s, m_solid, Y_O2, alpha_s, dt, Q_char_out):
    _oxidation(T_s, m_solid, Y_O2, alpha_s, dt, Q_char_out, A_CHAR, E_CHAR, HOC_CHAR,
               T_CHAR_ONSET, Y_O2_MIN_CHAR, Q_CHAR_MAX)


def smoldering(T_s, m_solid, Y_O2, alpha_s, dt, Q_smold_out):
    _oxidation(T_s, m_solid, Y_O2, alpha_s, dt, Q_smold_out, A_SMOLD, E_SMOLD, HOC_SMOLD,
               T_SMOLD_ONSET, Y_O2_MIN_SMOLD, Q_SMOLD_MAX)


def gas_solid_coupling(T_g, T_s, rho, u, v, w, alpha_s, sigma_sav, q_rad_in, Q_pyro, Q_comb,
                       m_water, L_v, dt, dz_arr, T_amb, q_loss_enable=True, h_conv_mult=1.0):
    if sigma_sav <= 0.0:
        return
    d_p = 4.0 / sigma_sav
    cp_g = 1100.0
    dz = np.asarray(dz_arr, np.float64)[:, None, None]
    Tg_new = T_g + Q_comb / (rho * cp_g) * dt
    solid = alpha_s > 0.0
    speed = np.sqrt(u ** 2 + v ** 2 + w ** 2)
    Re = np.maximum(rho * speed * d_p / MU_GAS, 0.1)
    Nu = 2.0 + 0.6 * (Re ** 0.5) * (PR_GAS ** (1.0 / 3.0))
    h_p = Nu * K_GAS / d_p * h_conv_mult
    a_v = sigma_sav * alpha_s
    q_conv = h_p * a_v * (Tg_new - T_s)
    q_loss = (EPS_SOLID * SIGMA_SB * (T_s ** 4 - T_amb ** 4) * a_v if q_loss_enable
              else np.zeros_like(T_s))
    q_rad_vol = q_rad_in / dz
    q_ground = np.zeros_like(T_s)
    q_ground[0] = H_GROUND * (T_s[0] - T_amb) / dz[0]
    q_in = q_rad_vol + q_conv - q_loss - q_ground
    evap = solid & (m_water > 0.0) & (L_v > 0.0) & (dt > 0.0) & (q_in > 0.0)
    q_max = m_water * L_v / dt if (L_v > 0.0 and dt > 0.0) else np.zeros_like(T_s)
    q_use = np.where(q_in < q_max, q_in, q_max)
    mw_new = np.maximum(m_water - q_use * dt / L_v if L_v > 0.0 else m_water, 0.0)
    q_res = np.where(evap, q_in - q_use, q_in)
    C_s = RHO_SOLID * CP_SOLID * alpha_s
    Cs = np.where(C_s > 0.0, C_s, 1.0)
    Ts_new = np.where(C_s > 0.0, T_s + (q_res - Q_pyro) / Cs * dt, T_s)
    Tg_solid = Tg_new - q_conv / (rho * cp_g) * dt
    m_water[...] = np.where(evap, mw_new, m_water)
    T_s[...] = np.where(solid, Ts_new, T_s)
    T_g[...] = np.where(solid, Tg_solid, Tg_new)
