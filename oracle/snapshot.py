"""Snapshot and result file fields (oracle; test infrastructure only).

Restates what the reference writes into snap_%04d.npz (spread_3d.py:3537-3670)
and Spread3DResult.save (spread_3d.py:460-511):

* ``phi_flame``          compute_phi_flame_from_state (flame_front_3d.py:756-800):
                         scipy.ndimage.distance_transform_edt, called as the
                         reference calls it (sampling (mean dz, dy, dx)).
* ``phi_flame_brute``    the same distances as a literal minimum over every
                         pair of cells (tiny grids only).
* ``ahead_band_mask``    LevelSetFront3D.ahead_band_mask (flame_front_3d.py:
                         387-389) cleared above the bed (spread_3d.py:2565).
* ``q_dom_fwd_at_band``  flame_front_3d.py:594-630.
* ``q_in_at_front_2d``   spread_3d.py:3227-3233: z sum of
                         compute_q_in_at_front_3d (flame_front_3d.py:440-471)
                         with a zero Frankman flux and no burst term.
* ``tau_mix``            spread_3d.py:1559-1562 and :2454-2455.
* ``bed_snapshot``       the per-particle aggregation loop, spread_3d.py:3569-3610.
* ``SNAPSHOT_KEYS``, ``RESULT_KEYS``, ``RESULT_META_KEYS``: the key sets of
  the two files (spread_3d.py:3611-3668, :479-511).
"""
from __future__ import annotations

import math

import numpy as np

OMEGA_MIN_FLAME = 1.0e-3
T_PLUME_MIN = 1000.0
Y_F_MIN_PLUME = 1.0e-3
G = 9.81

SNAPSHOT_KEYS = (
    "t", "step", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "u", "v", "w", "rho", "omega",
    "S_pyro", "k_turb", "eps_turb", "nu_t", "tau_mix", "phi_bed", "phi_flame", "front_x",
    "q_dom_fwd_3d", "q_frankman_3d", "q_in_at_front_2d", "v_n_2d", "ft_sink_M",
    "ft_sink_E", "ft_dep_M", "ft_dep_E", "ft_sink_t_rem", "ft_dep_t_rem", "bp_count",
    "bp_m_solid", "bp_m_water", "bp_m_char", "bp_T_s_avg", "x_mid", "z_mid", "dx", "dy",
    "dz_arr", "wind_U", "closure")
# (Nz, Ny, Nx) float32 fields of the snapshot
SNAPSHOT_3D = (
    "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "u", "v", "w", "rho", "omega", "S_pyro",
    "k_turb", "eps_turb", "nu_t", "tau_mix", "phi_bed", "phi_flame", "q_dom_fwd_3d",
    "q_frankman_3d", "ft_sink_M", "ft_sink_E", "ft_dep_M", "ft_dep_E", "ft_sink_t_rem",
    "ft_dep_t_rem", "bp_count", "bp_m_solid", "bp_m_water", "bp_m_char", "bp_T_s_avg")

RESULT_META_KEYS = ("ros_m_s", "n_cells_ignited", "Nx", "Ny", "Nz", "dx", "dy", "dz", "Lx",
                    "Ly", "Lz", "n_z_bed", "user")
RESULT_KEYS = (
    "_meta_json", "x_mid", "y_mid", "z_mid", "dz_arr", "inv_dz_arr", "z_face",
    "d_face_above", "d_face_below", "rho", "u", "v", "w", "p_dyn", "T_g", "Y_fuel", "Y_O2",
    "Y_H2O", "Y_CO", "alpha_s", "T_s", "m_hemi", "m_cell", "m_lign", "m_char", "front_t",
    "front_x", "diag_t", "diag_Tg_max", "diag_Ts_max", "diag_Sp_max", "diag_Qc_max",
    "diag_omega_max", "diag_Y_max", "diag_n_ign")


def flame_active(omega, T_g, Y_fuel):
    return (omega > OMEGA_MIN_FLAME) | ((T_g > T_PLUME_MIN) & (Y_fuel > Y_F_MIN_PLUME))


def phi_flame(omega, T_g, Y_fuel, dx, dy, dz_arr):
    from scipy import ndimage
    active = flame_active(omega, T_g, Y_fuel)
    sampling = (float(np.asarray(dz_arr).mean()), dy, dx)
    if active.any() and (~active).any():
        out = ndimage.distance_transform_edt(~active, sampling=sampling)
        ins = ndimage.distance_transform_edt(active, sampling=sampling)
        return np.where(active, -ins, out).astype(np.float64)
    return np.full(active.shape, -1.0e6 if active.all() else 1.0e6)


def phi_flame_brute(omega, T_g, Y_fuel, dx, dy, dz_arr):
    """Signed distance by a literal minimum over all cell pairs."""
    active = flame_active(omega, T_g, Y_fuel)
    h = (float(np.asarray(dz_arr).mean()), dy, dx)
    idx = np.argwhere(np.ones(active.shape, bool)).astype(np.float64) * np.asarray(h)
    flat = active.ravel()
    out = np.empty(flat.size)
    for c in range(flat.size):
        other = idx[flat != flat[c]]
        if other.size == 0:
            out[c] = -1.0e6 if flat[c] else 1.0e6
            continue
        d = math.sqrt(float(np.min(np.sum((other - idx[c]) ** 2, axis=1))))
        out[c] = -d if flat[c] else d
    return out.reshape(active.shape)


def ahead_band_mask(phi, band_m, n_z_bed):
    m = (phi > 0.0) & (phi <= band_m)
    m[n_z_bed:] = False
    return m


def q_dom_fwd_at_band(Omega, weights, I_set, mask):
    out = np.zeros(I_set.shape[1:])
    for n in range(I_set.shape[0]):
        xi = float(Omega[n, 0])
        if xi <= 0.0:
            continue
        out += np.where(mask, weights[n] * abs(xi) * I_set[n], 0.0)
    return out


def q_in_at_front_2d(q_dom_fwd, mask):
    q3 = np.where(mask, 0.0 + q_dom_fwd, 0.0)
    out = np.zeros(q3.shape[1:])
    for k in range(q3.shape[0]):
        out = out + q3[k]
    return out


def tau_buoy(dx, dy, dz):
    return math.sqrt(2.0 * max(dx, dy, dz) / G)


def tau_mix(k, eps, tau_b):
    return np.maximum(k / np.maximum(eps, 1.0e-12), tau_b)


def bed_snapshot(x, y, z, alive, m_solid, m_water, m_char, T_s, dx, dy, z_face, shape):
    """count, m_solid, m_water, m_char (float32) and the mass-weighted T_s
    (float32) per cell, accumulated particle by particle as the reference's
    loop does."""
    Nz, Ny, Nx = shape
    cnt = np.zeros(shape, np.float32)
    ms = np.zeros(shape, np.float32)
    mw = np.zeros(shape, np.float32)
    mc = np.zeros(shape, np.float32)
    tw = np.zeros(shape)
    mt = np.zeros(shape)
    for p in range(alive.shape[0]):
        if alive[p] == 0:
            continue
        xi = int(x[p] / dx)
        yj = int(y[p] / dy)
        if not (0 <= xi < Nx and 0 <= yj < Ny):
            continue
        zk = -1
        for k in range(Nz):
            if z[p] < z_face[k + 1]:
                zk = k
                break
        if zk < 0:
            continue
        a, b, d, T = m_solid[p], m_water[p], m_char[p], T_s[p]
        tot = a + b + d
        c = (zk, yj, xi)
        cnt[c] += 1.0
        ms[c] += a
        mw[c] += b
        mc[c] += d
        tw[c] += T * tot
        mt[c] += tot
    tavg = np.where(mt > 0.0, tw / np.maximum(mt, 1.0e-30), 0.0).astype(np.float32)
    return cnt, ms, mw, mc, tavg
