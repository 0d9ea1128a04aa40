"""Level-set front and flame-body predicate (oracle; test infrastructure only).

Restates physics_3d/flame_front_3d.py of the reference:
* ``initialize_source_patch``  flame_front_3d.py:304-330
* ``godunov_grad_norm``        flame_front_3d.py:113-188
* ``reinit_godunov_grad``      flame_front_3d.py:191-263
* ``reinitialize``             flame_front_3d.py:339-355 (5 sub-steps,
                               dtau = 0.5 min(dx, dy, dz_min))
* ``front_x``                  flame_front_3d.py:391-402
* ``compute_v_n_3d``           flame_front_3d.py:474-514 (front speed from the
                               heat flux into the band, moisture-aware)
* ``flame_active``             the `active` predicate of
                               compute_phi_flame_from_state (:756-800); the
                               flame-body mask with band 0 (:803-813) equals it
                               cell for cell, because the distance transform
                               only scales magnitudes and both branches keep
                               the sign (+-1e6 sentinels when all/none active).
"""
import math

import numpy as np

L_BURNOUT_M = 0.50
PHI_INIT_UNBURNED = 100.0
LEVELSET_REINIT_INTERVAL = 10
REINIT_SUBSTEPS = 5
DT_REINIT_FRAC = 0.5
L_VAP_WATER = 2.26e6
OMEGA_MIN_FLAME = 1.0e-3
T_PLUME_MIN = 1000.0
Y_F_MIN_PLUME = 1.0e-3


def initialize_source_patch(shape, i_start, i_end, k_top_bed, x_mid, dx,
                            L_burnout=L_BURNOUT_M):
    Nz, Ny, Nx = shape
    phi = np.full(shape, PHI_INIT_UNBURNED)
    x_front = x_mid[i_end - 1] + 0.5 * dx
    d = x_mid - x_front
    row = np.where(d > 0.0, d, np.maximum(d, -L_burnout / 2.0))
    phi[:] = row[None, None, :]
    phi[:k_top_bed + 1, :, i_start:i_end] = -L_burnout / 2.0
    return phi


def _inv_dz_pm(dz_arr):
    Nz = dz_arr.shape[0]
    im = np.empty(Nz)
    ip = np.empty(Nz)
    for k in range(Nz):
        if k == 0:
            im[k] = 1.0 / dz_arr[0]
            ip[k] = 2.0 / (dz_arr[0] + dz_arr[1]) if Nz > 1 else 1.0 / dz_arr[0]
        elif k == Nz - 1:
            ip[k] = 1.0 / dz_arr[Nz - 1]
            im[k] = 2.0 / (dz_arr[Nz - 1] + dz_arr[Nz - 2])
        else:
            im[k] = 2.0 / (dz_arr[k] + dz_arr[k - 1])
            ip[k] = 2.0 / (dz_arr[k] + dz_arr[k + 1])
    return im, ip


def _one_sided(phi, dx, dy, dz_arr):
    Nz, Ny, Nx = phi.shape
    im, ip = _inv_dz_pm(dz_arr)
    dmx = np.zeros_like(phi)
    dpx = np.zeros_like(phi)
    dmx[:, :, 1:] = (phi[:, :, 1:] - phi[:, :, :-1]) * (1.0 / dx)
    dpx[:, :, :-1] = (phi[:, :, 1:] - phi[:, :, :-1]) * (1.0 / dx)
    dmy = (phi - np.roll(phi, 1, axis=1)) * (1.0 / dy)
    dpy = (np.roll(phi, -1, axis=1) - phi) * (1.0 / dy)
    dmz = np.zeros_like(phi)
    dpz = np.zeros_like(phi)
    dmz[1:] = (phi[1:] - phi[:-1]) * im[1:, None, None]
    dpz[:-1] = (phi[1:] - phi[:-1]) * ip[:-1, None, None]
    return dmx, dpx, dmy, dpy, dmz, dpz


def _norm(a, b, c, d, e, f):
    return np.sqrt(a * a + b * b + c * c + d * d + e * e + f * f)


def godunov_grad_norm(phi, dx, dy, dz_arr):
    dmx, dpx, dmy, dpy, dmz, dpz = _one_sided(phi, dx, dy, dz_arr)
    return _norm(np.maximum(dmx, 0.0), np.minimum(dpx, 0.0), np.maximum(dmy, 0.0),
                 np.minimum(dpy, 0.0), np.maximum(dmz, 0.0), np.minimum(dpz, 0.0))


def reinit_godunov_grad(phi, phi_0, dx, dy, dz_arr):
    dmx, dpx, dmy, dpy, dmz, dpz = _one_sided(phi, dx, dy, dz_arr)
    pos = phi_0 > 0.0
    g_pos = _norm(np.maximum(dmx, 0.0), np.minimum(dpx, 0.0), np.maximum(dmy, 0.0),
                  np.minimum(dpy, 0.0), np.maximum(dmz, 0.0), np.minimum(dpz, 0.0))
    g_neg = _norm(np.minimum(dmx, 0.0), np.maximum(dpx, 0.0), np.minimum(dmy, 0.0),
                  np.maximum(dpy, 0.0), np.minimum(dmz, 0.0), np.maximum(dpz, 0.0))
    return np.where(pos, g_pos, g_neg)


def evolve(phi, dt, v_n, dx, dy, dz_arr):
    """phi - dt v_n |grad phi| (flame_front_3d.py:332-337)."""
    return phi - dt * v_n * godunov_grad_norm(phi, dx, dy, dz_arr)


def reinitialize(phi, dx, dy, dz_arr):
    phi_0 = phi.copy()
    dtau = DT_REINIT_FRAC * min(dx, dy, float(np.min(dz_arr)))
    s = np.where(phi_0 > 0.0, 1.0, np.where(phi_0 < 0.0, -1.0, 0.0))
    for _ in range(REINIT_SUBSTEPS):
        g = reinit_godunov_grad(phi, phi_0, dx, dy, dz_arr)
        phi = phi - dtau * s * (g - 1.0)
    return phi


def front_x(phi, k, j, dx):
    row = phi[k, j]
    for i in range(row.shape[0] - 1):
        p0, p1 = row[i], row[i + 1]
        if p0 < 0.0 < p1 or p1 < 0.0 < p0:
            return (i + (-p0 / (p1 - p0)) + 0.5) * dx
    return math.inf


def flame_active(omega, T_g, Y_fuel):
    return (omega > OMEGA_MIN_FLAME) | ((T_g > T_PLUME_MIN) & (Y_fuel > Y_F_MIN_PLUME))


def compute_v_n_3d(q_in, rho_b, cp_s, h_bed, T_ign, T_amb, M_local=None, L_vap=L_VAP_WATER,
                   f_dry_to_ignite=1.0):
    E_sens = rho_b * cp_s * h_bed * (T_ign - T_amb)
    if M_local is None:
        E = np.full_like(q_in, E_sens)
    else:
        E = E_sens + rho_b * h_bed * L_vap * f_dry_to_ignite * M_local
    E = np.maximum(E, 1.0)
    return np.maximum(q_in / E, 0.0)
