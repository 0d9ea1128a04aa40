"""FP64 oracle for the transport / stencil kernels (test infrastructure only).

Vectorised NumPy restatements.  Every elementwise expression keeps the
reference's operation order so that the IEEE results equal the reference's
scalar loops.  Arrays are (Nz, Ny, Nx) float64, C order.
"""
from __future__ import annotations

import math

import numpy as np

# ---------------------------------------------------------------------------
# neighbour helpers
# ---------------------------------------------------------------------------


def _sx(a, s):
    """a[..., i+s] with wrap (callers mask the wrapped cells)."""
    return np.roll(a, -s, axis=2)


def _sy(a, s):
    """a[:, j+s, :] periodic."""
    return np.roll(a, -s, axis=1)


def _sz(a, s):
    return np.roll(a, -s, axis=0)


def _kcol(v):
    return np.asarray(v, dtype=np.float64).reshape(-1, 1, 1)


def minmod(a, b):
    """muscl_3d.py:36-45."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    pos = np.where(a < b, a, b)
    neg = np.where(a > b, a, b)
    out = np.where(a > 0.0, pos, neg)
    return np.where(a * b <= 0.0, 0.0, out)


def muscl_face_value(pm1, p0, pp1, pp2, uf):
    """muscl_3d.py:48-64 (upwind side from the sign of uf)."""
    s_up = minmod(p0 - pm1, pp1 - p0)
    s_dn = minmod(pp1 - p0, pp2 - pp1)
    return np.where(uf >= 0.0, p0 + 0.5 * s_up, pp1 - 0.5 * s_dn)


def _masks(shape):
    Nz, Ny, Nx = shape
    i = np.arange(Nx).reshape(1, 1, -1)
    k = np.arange(Nz).reshape(-1, 1, 1)
    mx = (i >= 2) & (i <= Nx - 3)
    mz = (k >= 2) & (k <= Nz - 3)
    return i, k, mx, mz


def _x_ghosts(f, left_ghost):
    """f[i-1] with ghost at i=0, f[i+1] with zero-grad at i=Nx-1."""
    fl = _sx(f, -1).copy()
    fl[:, :, 0] = left_ghost
    fr = _sx(f, 1).copy()
    fr[:, :, -1] = f[:, :, -1]
    return fl, fr


def _z_ghosts(f, bottom_ghost):
    fl = _sz(f, -1).copy()
    fl[0] = bottom_ghost
    fr = _sz(f, 1).copy()
    fr[-1] = f[-1]
    return fl, fr


def _muscl_diffs(f, u, v, w):
    """Face differences (f_{+1/2} - f_{-1/2}) in x, y, z from MUSCL."""
    fxp = muscl_face_value(_sx(f, -1), f, _sx(f, 1), _sx(f, 2), u)
    fxm = muscl_face_value(_sx(f, -2), _sx(f, -1), f, _sx(f, 1), u)
    fyp = muscl_face_value(_sy(f, -1), f, _sy(f, 1), _sy(f, 2), v)
    fym = muscl_face_value(_sy(f, -2), _sy(f, -1), f, _sy(f, 1), v)
    fzp = muscl_face_value(_sz(f, -1), f, _sz(f, 1), _sz(f, 2), w)
    fzm = muscl_face_value(_sz(f, -2), _sz(f, -1), f, _sz(f, 1), w)
    return fxp - fxm, fyp - fym, fzp - fzm


# ---------------------------------------------------------------------------
# MUSCL scalar advection  (muscl_3d.py:67-171)
# ---------------------------------------------------------------------------


def advect_3d_scalar_muscl(phi, u, v, w, dx, dy, d_face_above, d_face_below,
                           rhs, phi_inlet):
    """rhs -= u.grad(phi), in place.  muscl_3d.py:67-171."""
    i, k, mx, mz = _masks(phi.shape)
    da = _kcol(d_face_above)
    db = _kcol(d_face_below)
    inv_db = 1.0 / db
    inv_da = 1.0 / da
    xL, xR = _x_ghosts(phi, phi_inlet)
    zL, zR = _z_ghosts(phi, phi[0])
    ddx, ddy, ddz = _muscl_diffs(phi, u, v, w)
    fx_m = u * ddx / dx
    fx_u = np.where(u >= 0.0, u * (phi - xL) / dx, u * (xR - phi) / dx)
    flux_x = np.where(mx, fx_m, fx_u)
    flux_y = v * ddy / dy
    fz_m = w * ddz / (0.5 * (da + db))
    fz_u = np.where(w >= 0.0, w * (phi - zL) * inv_db, w * (zR - phi) * inv_da)
    flux_z = np.where(mz, fz_m, fz_u)
    rhs -= flux_x + flux_y + flux_z


# ---------------------------------------------------------------------------
# species transport  (species_3d.py:40-175)
# ---------------------------------------------------------------------------


def step_species_transport(Y, rho, u, v, w, S, dt, dx, dy, dz_arr,
                           d_face_above, d_face_below, D, Y_inlet):
    """In-place species update; returns the number of clipped cells.

    species_3d.py:40-175 (no z-min inlet: the outdoor path).
    """
    i, k, mx, mz = _masks(Y.shape)
    inv_dx = 1.0 / dx
    inv_dy = 1.0 / dy
    inv_dx2 = inv_dx * inv_dx
    inv_dy2 = inv_dy * inv_dy
    da = _kcol(d_face_above)
    db = _kcol(d_face_below)
    inv_da = 1.0 / da
    inv_db = 1.0 / db
    inv_dz = 1.0 / _kcol(dz_arr)
    xL, xR = _x_ghosts(Y, Y_inlet)
    zL, zR = _z_ghosts(Y, Y[0])
    ddx, ddy, ddz = _muscl_diffs(Y, u, v, w)
    flux_x = np.where(mx, u * ddx * inv_dx,
                      np.where(u >= 0.0, u * (Y - xL) * inv_dx,
                               u * (xR - Y) * inv_dx))
    flux_y = v * ddy * inv_dy
    flux_z = np.where(mz, w * ddz / (0.5 * (da + db)),
                      np.where(w >= 0.0, w * (Y - zL) * inv_db,
                               w * (zR - Y) * inv_da))
    adv = -(flux_x + flux_y + flux_z)
    d2x = (xR - 2.0 * Y + xL) * inv_dx2
    d2y = (_sy(Y, 1) - 2.0 * Y + _sy(Y, -1)) * inv_dy2
    d2z = ((zR - Y) * inv_da - (Y - zL) * inv_db) * inv_dz
    diff = D * (d2x + d2y + d2z)
    src = S / rho
    dY = (adv + diff + src) * dt
    new = Y + dY
    n_clip = int(np.count_nonzero((new < 0.0) | (new > 1.0)))
    Y[...] = np.clip(new, 0.0, 1.0)
    return n_clip


# ---------------------------------------------------------------------------
# gas energy advection-diffusion  (spread_3d.py:3827-3897)
# ---------------------------------------------------------------------------


def adv_gas_energy(T_g, u, v, w, dt, dx, dy, dz_arr, d_face_above,
                   d_face_below, alpha_th, T_amb):
    """In place.  spread_3d.py:3827-3897 (outdoor: no z-min inlet)."""
    inv_d_below = (1.0 / d_face_below[1:-1]).reshape(-1, 1, 1)
    inv_d_above = (1.0 / d_face_above[1:-1]).reshape(-1, 1, 1)
    inv_dz_int = (1.0 / dz_arr[1:-1]).reshape(-1, 1, 1)
    dT = np.zeros_like(T_g)
    advect_3d_scalar_muscl(T_g, u, v, w, dx, dy, d_face_above, d_face_below,
                           dT, T_amb)
    inv_dx2 = 1.0 / (dx * dx)
    dT[:, :, 1:-1] += alpha_th * (T_g[:, :, 2:] - 2.0 * T_g[:, :, 1:-1]
                                  + T_g[:, :, :-2]) * inv_dx2
    dT[:, :, 0] += alpha_th * (T_g[:, :, 1] - 2.0 * T_g[:, :, 0] + T_amb) * inv_dx2
    dT[:, :, -1] += alpha_th * (T_g[:, :, -2] - T_g[:, :, -1]) * inv_dx2
    dT += alpha_th * (np.roll(T_g, -1, axis=1) - 2.0 * T_g
                      + np.roll(T_g, 1, axis=1)) / (dy * dy)
    dT[1:-1] += alpha_th * ((T_g[2:] - T_g[1:-1]) * inv_d_above
                            - (T_g[1:-1] - T_g[:-2]) * inv_d_below) * inv_dz_int
    dT[0] += alpha_th * ((T_g[1] - T_g[0]) / d_face_above[0]) / dz_arr[0]
    dT[-1] += alpha_th * (-(T_g[-1] - T_g[-2]) / d_face_below[-1]) / dz_arr[-1]
    T_g += dT * dt
    np.maximum(T_g, T_amb, out=T_g)
    np.minimum(T_g, 10_000.0, out=T_g)


# ---------------------------------------------------------------------------
# turbulent diffusion  (turbulence_3d.py:861-949)
# ---------------------------------------------------------------------------


def turb_diff_nsub(nu_t, sc_t, dt, dx, dy, dz_arr):
    """Sub-step count of apply_turbulent_diffusion; 0 means "skip".

    turbulence_3d.py:879-918.  nu_t max is over interior cells only.
    """
    Nz, Ny, Nx = nu_t.shape
    dz_min = float(np.min(dz_arr))
    h_min2 = min(dx * dx, min(dy * dy, dz_min * dz_min))
    inner = nu_t[1:Nz - 1, :, 1:Nx - 1]
    # the reference scans with `if v > nu_t_max` from 0.0: NaN never wins
    finite = inner[inner > 0.0]
    nu_t_max = float(finite.max()) if finite.size else 0.0
    D_t_max = nu_t_max / sc_t
    if D_t_max <= 1.0e-12:
        return 0, 0.0
    if not math.isfinite(D_t_max):
        return 0, 0.0
    dt_sub_max = 0.4 * h_min2 / D_t_max
    if dt_sub_max <= 0.0:
        return 0, 0.0
    n_sub_target = dt / dt_sub_max
    if n_sub_target > 1000:
        return 0, 0.0
    n_sub = max(1, int(math.ceil(n_sub_target)))
    return n_sub, dt / n_sub


def apply_turbulent_diffusion(field, nu_t, sc_t, dt, dx, dy, dz_arr,
                              d_face_above, d_face_below):
    """In place; returns n_sub (0 = skipped).  turbulence_3d.py:861-949."""
    n_sub, dt_sub = turb_diff_nsub(nu_t, sc_t, dt, dx, dy, dz_arr)
    if n_sub == 0:
        return 0
    inv_dx2 = 1.0 / (dx * dx)
    inv_dy2 = 1.0 / (dy * dy)
    inv_dz = 1.0 / _kcol(dz_arr)
    inv_da = 1.0 / _kcol(d_face_above)
    inv_db = 1.0 / _kcol(d_face_below)
    D_t = nu_t / sc_t
    for _ in range(n_sub):
        f = field
        xL, xR = _x_ghosts(f, f[:, :, 0])
        zL, zR = _z_ghosts(f, f[0])
        d2x = (xR - 2.0 * f + xL) * inv_dx2
        d2y = (_sy(f, 1) - 2.0 * f + _sy(f, -1)) * inv_dy2
        d2z = ((zR - f) * inv_da - (f - zL) * inv_db) * inv_dz
        df = D_t * (d2x + d2y + d2z) * dt_sub
        field += df
    return n_sub


# ---------------------------------------------------------------------------
# solid vertical conduction  (solid_conduction_3d.py:56-127)
# ---------------------------------------------------------------------------


def step_solid_conduction_vertical(T_s, alpha_s, dz_arr, d_face_above,
                                   d_face_below, k_solid, rho_solid,
                                   cp_solid, dt):
    inv_rho_cp = 1.0 / (rho_solid * cp_solid)
    a = alpha_s
    Ta = np.roll(T_s, -1, axis=0)
    Tb = np.roll(T_s, 1, axis=0)
    aa = np.roll(a, -1, axis=0)
    ab = np.roll(a, 1, axis=0)
    Nz = T_s.shape[0]
    kk = np.arange(Nz).reshape(-1, 1, 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        a_eff_up = 2.0 * a * aa / (a + aa)
        flux_up = k_solid * a_eff_up * (Ta - T_s) / _kcol(d_face_above)
        flux_up = np.where((kk < Nz - 1) & (aa > 0.0), flux_up, 0.0)
        a_eff_dn = 2.0 * a * ab / (a + ab)
        flux_dn = k_solid * a_eff_dn * (Tb - T_s) / _kcol(d_face_below)
        flux_dn = np.where((kk > 0) & (ab > 0.0), flux_dn, 0.0)
        src = (flux_up + flux_dn) / _kcol(dz_arr)
        dT = dt * src * inv_rho_cp / a
    T_s[...] = np.where(a > 0.0, T_s + dT, T_s)


# ---------------------------------------------------------------------------
# soil  (soil_3d.py:73-153)
# ---------------------------------------------------------------------------

RHO_SOIL_DEFAULT = 1500.0
CP_SOIL_DEFAULT = 800.0
K_SOIL_DEFAULT = 0.48
EPS_SOIL_DEFAULT = 0.85
SIGMA_SB = 5.67e-8


def build_soil_grid(n_soil=6, dz_first=0.001, growth=1.5):
    """soil_3d.py:73-102."""
    soil_dz = np.array([dz_first * growth ** k for k in range(n_soil)],
                       dtype=np.float64)
    d_above = np.empty(n_soil)
    d_below = np.empty(n_soil)
    d_above[0] = soil_dz[0] / 2.0
    for k in range(1, n_soil):
        d_above[k] = 0.5 * (soil_dz[k - 1] + soil_dz[k])
    for k in range(n_soil - 1):
        d_below[k] = 0.5 * (soil_dz[k] + soil_dz[k + 1])
    d_below[-1] = soil_dz[-1] / 2.0
    return soil_dz, d_above, d_below, float(soil_dz.sum())


def step_soil_conduction(T_soil, q_in_surface, dt, soil_dz, d_above, d_below,
                         k_s=K_SOIL_DEFAULT, rho_s=RHO_SOIL_DEFAULT,
                         cp_s=CP_SOIL_DEFAULT, eps_s=EPS_SOIL_DEFAULT,
                         T_amb=300.0):
    """In-place Gauss-Seidel sweep in k (soil_3d.py:105-153)."""
    n = T_soil.shape[0]
    rho_cp_inv = 1.0 / (rho_s * cp_s)
    inv_da = 1.0 / d_above
    inv_db = 1.0 / d_below
    for k in range(n):
        T_k = T_soil[k].copy()
        if k == 0:
            T4 = (T_k * T_k) * (T_k * T_k)
            q_top = q_in_surface - eps_s * SIGMA_SB * T4
        else:
            q_top = k_s * (T_soil[k - 1] - T_k) * inv_da[k]
        if k == n - 1:
            q_bot = k_s * (T_k - T_amb) * inv_db[k]
        else:
            q_bot = k_s * (T_k - T_soil[k + 1]) * inv_db[k]
        T_soil[k] = T_k + dt * (q_top - q_bot) * rho_cp_inv / soil_dz[k]


# ---------------------------------------------------------------------------
# drag  (drag_3d.py:99-144)
# ---------------------------------------------------------------------------

ERGUN_VISC_K = 50.0
MU_GAS_DRAG = 3.0e-5


def step_drag_force(u, v, w, rho, alpha_s, sigma_sav, Fx, Fy, Fz, C_D):
    sigma2 = sigma_sav * sigma_sav
    a = alpha_s
    speed = np.sqrt(u * u + v * v + w * w)
    one_m_a = 1.0 - a
    with np.errstate(divide="ignore", invalid="ignore"):
        K_visc = (ERGUN_VISC_K * MU_GAS_DRAG * sigma2 * a * a
                  / (one_m_a * one_m_a * one_m_a))
    K_quad = C_D * sigma_sav * a * 0.5 * rho * speed
    coeff = -(K_visc + K_quad)
    on = a > 0.0
    Fx[...] = np.where(on, coeff * u, 0.0)
    Fy[...] = np.where(on, coeff * v, 0.0)
    Fz[...] = np.where(on, coeff * w, 0.0)


# ---------------------------------------------------------------------------
# tentative velocity  (momentum_3d.py:41-276) and sponge (:292-317)
# ---------------------------------------------------------------------------

_G = 9.81
_MU_GAS_MOM = 1.8e-5


def step_tentative_velocity(u, v, w, rho, T_g, Fx, Fy, Fz, dt, dx, dy,
                            dz_arr, d_face_above, d_face_below, T_amb,
                            u_inlet, v_inlet, w_inlet):
    i, k, mx, mz = _masks(u.shape)
    inv_dx = 1.0 / dx
    inv_dy = 1.0 / dy
    inv_dx2 = inv_dx * inv_dx
    inv_dy2 = inv_dy * inv_dy
    da = _kcol(d_face_above)
    db = _kcol(d_face_below)
    nu = _MU_GAS_MOM / rho
    inv_dz_eff = 1.0 / (0.5 * (da + db))
    inv_db = 1.0 / db
    inv_da = 1.0 / da
    inv_dz = 1.0 / _kcol(dz_arr)
    out = []
    for f, f_in in ((u, u_inlet), (v, v_inlet), (w, w_inlet)):
        xL, xR = _x_ghosts(f, np.asarray(f_in)[:, :])
        zL, zR = _z_ghosts(f, 0.0)
        ddx, ddy, ddz = _muscl_diffs(f, u, v, w)
        dfdx = np.where(mx, ddx * inv_dx,
                        np.where(u >= 0.0, (f - xL) * inv_dx, (xR - f) * inv_dx))
        dfdy = ddy * inv_dy
        dfdz = np.where(mz, ddz * inv_dz_eff,
                        np.where(w >= 0.0, (f - zL) * inv_db, (zR - f) * inv_da))
        adv = -(u * dfdx + v * dfdy + w * dfdz)
        d2x = (xR - 2.0 * f + xL) * inv_dx2
        d2y = (_sy(f, 1) - 2.0 * f + _sy(f, -1)) * inv_dy2
        d2z = ((zR - f) * inv_da - (f - zL) * inv_db) * inv_dz
        visc = nu * (d2x + d2y + d2z)
        out.append((adv, visc))
    buoy_w = _G * (T_g - T_amb) / T_amb
    du = (out[0][0] + out[0][1] + Fx / rho) * dt
    dv = (out[1][0] + out[1][1] + Fy / rho) * dt
    dw = (out[2][0] + out[2][1] + buoy_w + Fz / rho) * dt
    u += du
    v += dv
    w += dw


def apply_outflow_sponge(u, u_target_2d, sigma_x, Y_F, Y_F_skip, dt):
    sig = sigma_x.reshape(1, 1, -1)
    ut = u_target_2d[:, :, None]
    on = (sig > 0.0) & (Y_F < Y_F_skip)
    u[...] = np.where(on, u + sig * dt * (ut - u), u)


def sponge_sigma_profile(Nx, dx):
    """spread_3d.py:1474-1484."""
    raise NotImplementedError  # filled in by oracle.step


# ---------------------------------------------------------------------------
# strain / vorticity and k-epsilon  (turbulence_3d.py:391-858, with the
# one-sided dT/dz at k=0 and k=Nz-1 of the fixed upstream kernel,
# docs/js/physics/kepsilon.js:24-31,233-236)
# ---------------------------------------------------------------------------

_NU_GAS = 1.5e-5
C_MU = 0.09
C_1EPS = 1.44
C_2EPS = 1.92
SIGMA_K = 1.0
SIGMA_EPS = 1.3
PR_T = 0.85
SC_T = 0.7
C_D_DRAG = 1.0
C_EPS4_CANOPY = 0.9
C_EPS5_CANOPY = 0.9
ETA0 = 4.38
BETA_RNG = 0.012
K_MIN = 1.0e-8
EPS_MIN = 1.0e-8
NU_T_MIN = 0.0
A_0_REAL = 4.04
A_S_REAL = 4.5
U_TINY_HENKES = 0.01
C_LIM_P = 10.0


def _inv_dz_central(d_face_above, d_face_below):
    Nz = len(d_face_above)
    out = np.empty(Nz)
    for k in range(Nz):
        if k == 0:
            out[k] = 1.0 / d_face_above[0]
        elif k == Nz - 1:
            out[k] = 1.0 / d_face_below[Nz - 1]
        else:
            out[k] = 1.0 / (d_face_below[k] + d_face_above[k])
    return out.reshape(-1, 1, 1)


def strain_and_vorticity_squared(u, v, w, dx, dy, d_face_above, d_face_below,
                                 u_inlet):
    inv_dx = 1.0 / dx
    inv_dy = 1.0 / dy
    idzc = _inv_dz_central(d_face_above, d_face_below)
    uxL, uxR = _x_ghosts(u, u_inlet)
    vxL, vxR = _x_ghosts(v, 0.0)
    wxL, wxR = _x_ghosts(w, 0.0)
    uzL, uzR = _z_ghosts(u, 0.0)
    vzL, vzR = _z_ghosts(v, 0.0)
    wzL, wzR = _z_ghosts(w, 0.0)
    dudx = (uxR - uxL) * 0.5 * inv_dx
    dudy = (_sy(u, 1) - _sy(u, -1)) * 0.5 * inv_dy
    dudz = (uzR - uzL) * idzc
    dvdx = (vxR - vxL) * 0.5 * inv_dx
    dvdy = (_sy(v, 1) - _sy(v, -1)) * 0.5 * inv_dy
    dvdz = (vzR - vzL) * idzc
    dwdx = (wxR - wxL) * 0.5 * inv_dx
    dwdy = (_sy(w, 1) - _sy(w, -1)) * 0.5 * inv_dy
    dwdz = (wzR - wzL) * idzc
    S11, S22, S33 = dudx, dvdy, dwdz
    S12 = 0.5 * (dudy + dvdx)
    S13 = 0.5 * (dudz + dwdx)
    S23 = 0.5 * (dvdz + dwdy)
    S2 = (2.0 * (S11 * S11 + S22 * S22 + S33 * S33)
          + 4.0 * (S12 * S12 + S13 * S13 + S23 * S23))
    O12 = 0.5 * (dudy - dvdx)
    O13 = 0.5 * (dudz - dwdx)
    O23 = 0.5 * (dvdz - dwdy)
    O2 = 4.0 * (O12 * O12 + O13 * O13 + O23 * O23)
    return S2, O2


def step_k_epsilon(k_turb, eps_turb, nu_t_out, u, v, w, T_g, rho, alpha_s,
                   sigma_sav, dt, dx, dy, dz_arr, d_face_above, d_face_below,
                   T_amb, u_inlet, k_wall_ghost, eps_wall_ghost,
                   beta_p_canopy, beta_d_canopy):
    """In place on k_turb, eps_turb; writes nu_t_out.  bvg_factor = 0 and the
    L-min / Durbin caps off (the production call, spread_3d.py:2409-2426)."""
    S2, O2 = strain_and_vorticity_squared(u, v, w, dx, dy, d_face_above,
                                          d_face_below, u_inlet)
    i, kk, mx, mz = _masks(u.shape)
    inv_dx = 1.0 / dx
    inv_dy = 1.0 / dy
    inv_dx2 = 1.0 / (dx * dx)
    inv_dy2 = 1.0 / (dy * dy)
    da = _kcol(d_face_above)
    db = _kcol(d_face_below)
    inv_dz = 1.0 / _kcol(dz_arr)
    inv_da = 1.0 / da
    inv_db = 1.0 / db
    idzc = _inv_dz_central(d_face_above, d_face_below)

    k_loc = np.where(k_turb < K_MIN, K_MIN, k_turb)
    e_loc = np.where(eps_turb < EPS_MIN, EPS_MIN, eps_turb)

    def xg(f, loc, left):
        fl = _sx(f, -1).copy()
        fl[:, :, 0] = left
        fr = _sx(f, 1).copy()
        fr[:, :, -1] = loc[:, :, -1]
        return fl, fr

    def zg(f, loc, bottom):
        fl = _sz(f, -1).copy()
        fl[0] = bottom
        fr = _sz(f, 1).copy()
        fr[-1] = loc[-1]
        return fl, fr

    kxL, kxR = xg(k_turb, k_loc, K_MIN)
    exL, exR = xg(eps_turb, e_loc, EPS_MIN)
    kzL, kzR = zg(k_turb, k_loc, k_wall_ghost)
    ezL, ezR = zg(eps_turb, e_loc, eps_wall_ghost)

    U_star = np.sqrt(0.5 * (S2 + O2))
    C_mu_real = 1.0 / (A_0_REAL + A_S_REAL * U_star * k_loc / e_loc)
    nu_t = C_mu_real * k_loc * k_loc / e_loc
    nu_t = np.where(nu_t < NU_T_MIN, NU_T_MIN, nu_t)
    P_k = nu_t * S2
    P_k_lim = np.where(P_k > C_LIM_P * e_loc, C_LIM_P * e_loc, P_k)

    Tp = _sz(T_g, 1)
    Tm = _sz(T_g, -1)
    dTdz = (Tp - Tm) * idzc
    dTdz[0] = (T_g[1] - T_g[0]) * idzc[0]
    dTdz[-1] = (T_g[-1] - T_g[-2]) * idzc[-1]
    T_for_buoy = np.where(T_g < T_amb, T_amb, T_g)
    G_k = (nu_t / PR_T) * (_G / T_for_buoy) * dTdz
    G_k = np.where(G_k < 0.0, 0.0, G_k)
    G_k_for_k = np.where(G_k > P_k_lim, P_k_lim, G_k)

    u_h = np.sqrt(u * u + v * v)
    u_h = np.where(u_h < U_TINY_HENKES, U_TINY_HENKES, u_h)
    C_3eps = np.tanh(np.abs(w) / u_h)

    a_s = alpha_s
    a_v = sigma_sav * a_s
    speed = np.sqrt(u * u + v * v + w * w)
    cd_av_speed = C_D_DRAG * a_v * speed
    P_k_canopy = np.where(a_s > 0.0, beta_p_canopy * cd_av_speed * speed * speed, 0.0)
    D_k_canopy = np.where(a_s > 0.0, beta_d_canopy * cd_av_speed * k_loc, 0.0)
    G_B = 0.0  # bvg_factor = 0

    def adv(f, loc, fL, fR):
        # MUSCL neighbours read the raw field; the centre uses the floored value.
        fxp = muscl_face_value(_sx(f, -1), loc, _sx(f, 1), _sx(f, 2), u)
        fxm = muscl_face_value(_sx(f, -2), _sx(f, -1), loc, _sx(f, 1), u)
        fyp = muscl_face_value(_sy(f, -1), loc, _sy(f, 1), _sy(f, 2), v)
        fym = muscl_face_value(_sy(f, -2), _sy(f, -1), loc, _sy(f, 1), v)
        fzp = muscl_face_value(_sz(f, -1), loc, _sz(f, 1), _sz(f, 2), w)
        fzm = muscl_face_value(_sz(f, -2), _sz(f, -1), loc, _sz(f, 1), w)
        ax = np.where(mx, u * (fxp - fxm) * inv_dx,
                      np.where(u >= 0.0, u * (loc - fL[0]) * inv_dx,
                               u * (fR[0] - loc) * inv_dx))
        ay = v * (fyp - fym) * inv_dy
        az = np.where(mz, w * (fzp - fzm) / (0.5 * (da + db)),
                      np.where(w >= 0.0, w * (loc - fL[1]) * inv_db,
                               w * (fR[1] - loc) * inv_da))
        return ax + ay + az

    def lap(f, loc, fL, fR):
        d2x = (fR[0] - 2.0 * loc + fL[0]) * inv_dx2
        d2y = (_sy(f, 1) - 2.0 * loc + _sy(f, -1)) * inv_dy2
        d2z = ((fR[1] - loc) * inv_da - (loc - fL[1]) * inv_db) * inv_dz
        return d2x + d2y + d2z

    adv_k = adv(k_turb, k_loc, (kxL, kzL), (kxR, kzR))
    alpha_k = _NU_GAS + nu_t / SIGMA_K
    diff_k = alpha_k * lap(k_turb, k_loc, (kxL, kzL), (kxR, kzR))
    S_pos_k = P_k_lim + G_k_for_k + P_k_canopy + G_B
    S_neg_k = e_loc + D_k_canopy
    k_new = (k_loc + (-adv_k + diff_k + S_pos_k) * dt) / (1.0 + S_neg_k * dt / k_loc)
    k_new = np.where(k_new < K_MIN, K_MIN, k_new)

    adv_e = adv(eps_turb, e_loc, (exL, ezL), (exR, ezR))
    alpha_eps = _NU_GAS + nu_t / SIGMA_EPS
    diff_e = alpha_eps * lap(eps_turb, e_loc, (exL, ezL), (exR, ezR))
    S_mag = np.sqrt(S2)
    eta = S_mag * k_loc / e_loc
    eta3 = eta * (eta * eta)
    rng_corr = C_MU * eta3 * (1.0 - eta / ETA0) / (1.0 + BETA_RNG * eta3)
    rng_corr = np.where(rng_corr < 0.0, 0.0, rng_corr)
    C_2_eff = C_2EPS + rng_corr
    ek_ratio = e_loc / k_loc
    S_pos_eps = (C_1EPS * (P_k_lim + C_3eps * G_k) * ek_ratio
                 + C_EPS4_CANOPY * P_k_canopy * ek_ratio)
    S_neg_eps = (C_2_eff * e_loc * ek_ratio
                 + C_EPS5_CANOPY * D_k_canopy * ek_ratio)
    e_new = (e_loc + (-adv_e + diff_e + S_pos_eps) * dt) / (1.0 + S_neg_eps * dt / e_loc)
    e_new = np.where(e_new < EPS_MIN, EPS_MIN, e_new)

    C_mu_new = 1.0 / (A_0_REAL + A_S_REAL * U_star * k_new / e_new)
    nu_t_new = C_mu_new * k_new * k_new / e_new
    nu_t_new = np.where(nu_t_new < NU_T_MIN, NU_T_MIN, nu_t_new)
    k_turb[...] = k_new
    eps_turb[...] = e_new
    nu_t_out[...] = nu_t_new


# ---------------------------------------------------------------------------
# equation of state  (spread_3d.py:2042, :3440)
# ---------------------------------------------------------------------------

P0 = 101325.0
R_AIR = 8.314 / 0.029


def eos(T_g, T_amb):
    return P0 / (R_AIR * np.maximum(T_g, T_amb))
