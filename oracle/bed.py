"""FP64 oracle of the Lagrangian bed particles (test infrastructure only).

lagrangian_bed_3d.py of the reference: buffers (:193-218), initialisation
(:221-329), step_bed_particles (:593-1070), aggregate_particles_to_T_s_grid
(:500-546), aggregate_particles_to_M_local_grid (:549-590),
step_horizontal_solid_conduction_scatter (:427-497); cell location from
lagrangian_particles_3d.py:94-138; kinetic constants from pyrolysis_3d.py
(:45, :117, :217-267, :446-562).

Particles are processed vectorised; the per-cell sources are accumulated
with np.add.at in slot order, which performs the reference's sequential
``cell += contribution`` sums in the same order.
"""
from __future__ import annotations

import math

import numpy as np

R_GAS = 8.314
HEAT_OF_PYROLYSIS = 4.0e5
A_MD2004 = 36_280.0
E_MD2004 = 58_600.0
CHAR_YIELD_MD2004 = 0.15
ETA_MD2004 = 1.0 - CHAR_YIELD_MD2004
A_OP_MD2004 = 7.26e4
E_OP_MD2004 = 55_000.0
N_O2_OP = 1.0
HOR_OP_MD2004 = -1.0e6
Y_O2_MIN_OP = 1.0e-3
A_DRY = 4.29e3
E_DRY = 43_800.0
L_VAP_WATER = 2.26e6
A_CHAR = 7.09e5
E_CHAR = 94_000.0
HOC_CHAR = 32_000_000.0
T_CHAR_ONSET = 600.0
Y_O2_MIN_CHAR = 1.0e-3
A_SMOLD = 1.0e6
E_SMOLD = 80_000.0
HOC_SMOLD = 28_000_000.0
T_SMOLD_ONSET = 473.0
Y_O2_MIN_SMOLD = 1.0e-3

RHO_SOLID_TRUE_GRASS = 380.0
CP_SOLID_GRASS = 1500.0
H_CONV_DEFAULT = 25.0
SAV_GRASS_DEFAULT = 2000.0
M_PARTICLE_BURNOUT = 1.0e-8
SIGMA_SB = 5.67e-8
EPS_SOLID_DEFAULT = 0.9
T_BOIL_WATER = 373.15
A_DRY_GRASS = 4.29e3
E_DRY_GRASS = 30_000.0
DRY_MODE_ARRHENIUS = 0
DRY_MODE_EQUILIBRIUM = 1
DRY_MODE_COMBINED = 2

FIELDS = ("x", "y", "z", "u", "v", "w", "age", "m_solid", "m_water", "m_char",
          "T_s", "m_solid_0", "m_water_0", "sav", "m_char_max")


def allocate_bed_particle_buffers(N_max):
    if N_max < 0:
        raise ValueError(f"N_max must be >= 0; got {N_max}")
    buf = {k: np.zeros(N_max) for k in FIELDS}
    buf["alive"] = np.zeros(N_max, dtype=np.int8)
    return buf


def initialize_bed_particles_from_alpha_s(buf, alpha_s, rho_b_dry, moisture_frac,
                                          T_amb, dx, dy, dz_arr, n_z_bed,
                                          n_per_cell, sav=SAV_GRASS_DEFAULT,
                                          m_scale=None):
    """lagrangian_bed_3d.py:221-329.  Optional `m_scale` (per-slot factor on
    the dry and water mass, new) supports seeded per-particle load jitter;
    None keeps the reference exactly."""
    Nz, Ny, Nx = alpha_s.shape
    N_max = buf["alive"].shape[0]
    nb = int(np.count_nonzero(alpha_s[:min(n_z_bed, Nz)] > 0.0))
    if nb * n_per_cell > N_max:
        raise ValueError(f"Buffer too small: need {nb * n_per_cell} slots "
                         f"({nb} bed cells x {n_per_cell} particles); have {N_max}")
    slot = 0
    for k in range(min(n_z_bed, Nz)):
        V_cell = dx * dy * dz_arr[k]
        z_face_k = float(np.sum(dz_arr[:k]))
        ks = np.argwhere(alpha_s[k] > 0.0)          # (j, i) in C order
        if ks.size == 0:
            continue
        m_solid_cell = rho_b_dry * V_cell
        m_solid_per_p = m_solid_cell / n_per_cell
        m_water_per_p = moisture_frac * m_solid_per_p
        nc = ks.shape[0]
        p = np.arange(n_per_cell)
        fx = ((p * 13) % n_per_cell + 0.5) / n_per_cell
        fy = ((p * 7) % n_per_cell + 0.5) / n_per_cell
        fz = (p + 0.5) / n_per_cell
        jj = ks[:, 0][:, None]
        ii = ks[:, 1][:, None]
        sl = slice(slot, slot + nc * n_per_cell)
        buf["x"][sl] = ((ii + fx[None, :]) * dx).ravel()
        buf["y"][sl] = ((jj + fy[None, :]) * dy).ravel()
        buf["z"][sl] = np.broadcast_to(z_face_k + fz * dz_arr[k], (nc, n_per_cell)).ravel()
        for f in ("u", "v", "w", "age", "m_char"):
            buf[f][sl] = 0.0
        buf["alive"][sl] = 1
        if m_scale is None:
            buf["m_solid"][sl] = m_solid_per_p
            buf["m_water"][sl] = m_water_per_p
        else:
            sc = np.asarray(m_scale)[sl]
            buf["m_solid"][sl] = m_solid_per_p * sc
            buf["m_water"][sl] = m_water_per_p * sc
        buf["T_s"][sl] = T_amb
        buf["m_solid_0"][sl] = buf["m_solid"][sl]
        buf["m_water_0"][sl] = buf["m_water"][sl]
        buf["sav"][sl] = sav
        slot += nc * n_per_cell
    return slot


def locate_k_from_z(z, z_face, Nz):
    """lagrangian_particles_3d.py:94-108, vectorised."""
    z = np.asarray(z)
    k = np.searchsorted(z_face[1:Nz + 1], z, side="right")
    bad = (z < z_face[0]) | (z >= z_face[Nz])
    return np.where(bad, -1, np.minimum(k, Nz - 1))


def locate_cell(x, y, z, dx, dy, z_face, Nz, Nx, Ny):
    """lagrangian_particles_3d.py:111-138 (x < 0 -> -1 before truncation)."""
    with np.errstate(invalid="ignore"):
        i = np.where(x < 0.0, -1, np.trunc(x / dx)).astype(np.int64)
        i = np.where(i >= Nx, -1, i)
        j = np.where(y < 0.0, -1, np.trunc(y / dy)).astype(np.int64)
        j = np.where(j >= Ny, -1, j)
    k = locate_k_from_z(z, z_face, Nz)
    return i, j, k


def agg_cell(x, y, z, dx, dy, z_face, Nz, Nx, Ny):
    """Index used by the aggregators and the conduction scatter: int() first,
    then the range test (lagrangian_bed_3d.py:490-497, :528-534, :575-581)."""
    i = np.trunc(x / dx).astype(np.int64)
    j = np.trunc(y / dy).astype(np.int64)
    ok = (i >= 0) & (i < Nx) & (j >= 0) & (j < Ny)
    k = locate_k_from_z(z, z_face, Nz)
    return np.where(ok, i, -1), np.where(ok, j, -1), k


def step_bed_particles(buf, T_g, Y_O2, Q_ext, n_per_cell, dx, dy, dz_arr, z_face,
                       h_conv, rho_solid_true, cp_solid, eps_solid, T_amb,
                       view_factor, view_factor_geometric, h_bed, kappa_bed_eff,
                       dt, do_drying, do_pyrolysis, do_char_ox, do_smolder,
                       drying_mode, char_ox_flux_cap, char_ox_ash_exp):
    """Returns (sources dict of 8 fields, n_alive, n_burned, diag[16])."""
    Nz, Ny, Nx = T_g.shape
    names = ("S_pyro", "S_drying", "Q_pyro", "Q_drying", "Y_F_source", "Q_char",
             "Q_smold", "Q_g_conv")
    src = {k: np.zeros((Nz, Ny, Nx)) for k in names}
    alive = buf["alive"].astype(bool)
    i, j, k = locate_cell(buf["x"], buf["y"], buf["z"], dx, dy, z_face, Nz, Nx, Ny)
    lost = alive & ((i < 0) | (j < 0) | (k < 0))
    buf["alive"][lost] = 0
    act = alive & ~lost
    idx = np.nonzero(act)[0]
    i, j, k = i[idx], j[idx], k[idx]
    cell = (k * Ny + j) * Nx + i
    inv_dt = 1.0 / dt if dt > 0.0 else 0.0
    V_cell = dx * dy * np.asarray(dz_arr)[k]
    T_s = buf["T_s"][idx].copy()
    Tg = T_g.reshape(-1)[cell]
    YO = Y_O2.reshape(-1)[cell]
    ms = buf["m_solid"][idx].copy()
    mw = buf["m_water"][idx].copy()
    mc = buf["m_char"][idx].copy()
    mw0 = buf["m_water_0"][idx]
    sav = buf["sav"][idx]
    mcmax = buf["m_char_max"][idx].copy()
    z = buf["z"][idx]
    # the fuel top above each particle: the buffer's optional per-particle
    # h_top (a seeded depth field or depth jitter; no reference counterpart),
    # else the scalar h_bed of the reference's uniform bed
    h_top = buf["h_top"][idx] if "h_top" in buf else h_bed
    n = idx.size
    # drying (Arrhenius / combined)
    dm_evap = np.zeros(n)
    if do_drying:
        on = (mw > 0.0) & (T_s > 0.0)
        if drying_mode == DRY_MODE_ARRHENIUS:
            A, E = A_DRY, E_DRY
        else:
            A, E = A_DRY_GRASS, E_DRY_GRASS
        if drying_mode in (DRY_MODE_ARRHENIUS, DRY_MODE_COMBINED):
            with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
                kd = A * np.exp(-E / (R_GAS * T_s))
                mw_new = mw * np.exp(-kd * dt)
            dm_evap = np.where(on, mw - mw_new, 0.0)
            mw = np.where(on, mw_new, mw)
    # pyrolysis
    rate = np.zeros(n)
    f_th = np.ones(n)
    f_op = np.zeros(n)
    if do_pyrolysis:
        on = (ms > 0.0) & (T_s > 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            wet = np.where(mw0 > 0.0, mw / np.where(mw0 > 0.0, mw0, 1.0), 0.0)
        gate = np.where(mw0 > 0.0, np.where(1.0 - wet < 0.0, 0.0, 1.0 - wet), 1.0)
        on = on & (gate > 0.0)
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            kth = A_MD2004 * np.exp(-E_MD2004 / (R_GAS * T_s))
            kop = np.where(YO > Y_O2_MIN_OP,
                           A_OP_MD2004 * np.exp(-E_OP_MD2004 / (R_GAS * T_s)) * YO ** N_O2_OP,
                           0.0)
            ktot = kth + kop
            on = on & (ktot > 0.0)
            ms_new = ms * np.exp(-ktot * dt)
            dm_full = ms - ms_new
            dm_pyro = np.where(on, dm_full * gate, 0.0)
            rate = np.where(on, dm_pyro * inv_dt, 0.0)
            f_th = np.where(on, kth / ktot, 1.0)
            f_op = np.where(on, kop / ktot, 0.0)
        mc = np.where(on, mc + CHAR_YIELD_MD2004 * dm_pyro, mc)
        ms = np.where(on, ms - dm_pyro, ms)
        mcmax = np.where(on & (mc > mcmax), mc, mcmax)
    A_now = sav * (ms + mc) / rho_solid_true
    # char oxidation
    dm_ch = np.zeros(n)
    Q_ch = np.zeros(n)
    if do_char_ox:
        on = (mc > 1.0e-12) & (T_s >= T_CHAR_ONSET) & (YO >= Y_O2_MIN_CHAR)
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            kch = A_CHAR * np.exp(-E_CHAR / (R_GAS * T_s))
            mdot = kch * mc * YO
            mcons = mdot * dt
            mcons = np.where(mcons > 0.5 * mc, 0.5 * mc, mcons)
            Ar = A_now.copy()
            if char_ox_ash_exp > 0.0:
                pen = mcmax > 1.0e-15
                bf = mc / np.where(pen, mcmax, 1.0)
                bf = np.where(bf < 0.0, 0.0, bf)
                Ar = np.where(pen, Ar * bf ** char_ox_ash_exp, Ar)
            Qcap = char_ox_flux_cap * Ar
            Qarr = (mcons * HOC_CHAR) * inv_dt
            mcons = np.where(Qarr > Qcap, Qcap / (HOC_CHAR * inv_dt), mcons)
        dm_ch = np.where(on, mcons, 0.0)
        Q_ch = np.where(on, (dm_ch * HOC_CHAR) * inv_dt, 0.0)
        mc = np.where(on, mc - dm_ch, mc)
    # smoulder
    Q_sm = np.zeros(n)
    if do_smolder:
        on = (T_s >= T_SMOLD_ONSET) & (YO >= Y_O2_MIN_SMOLD)
        avail = ms + mc
        on = on & (avail > 1.0e-12)
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            ksm = A_SMOLD * np.exp(-E_SMOLD / (R_GAS * T_s))
            mdot = ksm * avail * YO
            mcons = mdot * dt
            mcons = np.where(mcons > 0.5 * avail, 0.5 * avail, mcons)
            Qcap = 4.0e4 * A_now
            Qarr = (mcons * HOC_SMOLD) * inv_dt
            mcons = np.where(Qarr > Qcap, Qcap / (HOC_SMOLD * inv_dt), mcons)
            fs = ms / np.where(avail > 0.0, avail, 1.0)
        d_s = np.where(on, fs * mcons, 0.0)
        d_c = np.where(on, (1.0 - fs) * mcons, 0.0)
        ms = ms - d_s
        mc = mc - d_c
        Q_sm = np.where(on, (mcons * HOC_SMOLD) * inv_dt, 0.0)
    inv_V = 1.0 / V_cell
    c_Sp = ETA_MD2004 * rate * inv_V
    c_Sd1 = dm_evap * inv_dt * inv_V
    c_Qp = (rate * (f_th * HEAT_OF_PYROLYSIS + f_op * HOR_OP_MD2004)) * inv_V
    c_Qd1 = dm_evap * L_VAP_WATER * inv_dt * inv_V
    c_YF = ETA_MD2004 * rate * inv_V
    c_Qch = Q_ch * inv_V
    c_Qsm = Q_sm * inv_V
    # energy / Newton for T_s
    mt = ms + mw + mc
    has = mt > 0.0
    A_p = sav * (ms + mc) / rho_solid_true
    Q_conv = np.where(has, h_conv * A_p * (Tg - T_s), 0.0)
    Q_rxn = -rate * (f_th * HEAT_OF_PYROLYSIS + f_op * HOR_OP_MD2004)
    Q_dry = -dm_evap * L_VAP_WATER * inv_dt
    Q_ext_p = (Q_ext.reshape(-1)[cell] * V_cell / n_per_cell) if n_per_cell > 0 else np.zeros(n)
    if view_factor_geometric:
        h_above = h_top - z
        h_above = np.where(h_above < 0.0, 0.0, h_above)
        fgeom = np.exp(-kappa_bed_eff * h_above)
    else:
        fgeom = np.ones(n)
    C_rad = eps_solid * SIGMA_SB * A_p * view_factor * fgeom
    mcp = mt * cp_solid
    Tamb4 = T_amb * T_amb * T_amb * T_amb
    Q_other = Q_conv + Q_rxn + Q_dry + Q_ch + Q_sm + Q_ext_p
    mc_dt = mcp / dt
    T_old = T_s.copy()
    Ti = T_s.copy()
    F = np.zeros(n)
    for _ in range(5):
        T2 = Ti * Ti
        T3 = T2 * Ti
        T4 = T3 * Ti
        F = mc_dt * (Ti - T_old) - Q_other + C_rad * (T4 - Tamb4)
        Fp = mc_dt + 4.0 * C_rad * T3
        Ti = Ti - F / Fp
        Ti = np.where(Ti < T_amb, T_amb, Ti)
    T_new = np.where(has, Ti, T_s)
    T_newton = T_new.copy()          # the diagnostics see T_s before the clamp
    absF = np.where(has, np.abs(F), 0.0)
    T_new = np.where(has & (T_new > 1.0e4), 1.0e4, T_new)
    # equilibrium drying above boiling
    c_Sd2 = np.zeros(n)
    c_Qd2 = np.zeros(n)
    if do_drying and drying_mode in (DRY_MODE_EQUILIBRIUM, DRY_MODE_COMBINED):
        on = has & (mw > 0.0) & (T_new > T_BOIL_WATER)
        excess = mcp * (T_new - T_BOIL_WATER)
        dm_eq = excess / L_VAP_WATER
        full = dm_eq >= mw
        resid = excess - mw * L_VAP_WATER
        with np.errstate(divide="ignore", invalid="ignore"):
            T_full = T_BOIL_WATER + np.where(mcp > 0.0, resid / np.where(mcp > 0.0, mcp, 1.0), 0.0)
        dm_eq = np.where(full, mw, dm_eq)
        T_new = np.where(on, np.where(full, T_full, T_BOIL_WATER), T_new)
        mw = np.where(on, np.where(full, 0.0, mw - dm_eq), mw)
        c_Sd2 = np.where(on, dm_eq * inv_dt * inv_V, 0.0)
        c_Qd2 = np.where(on, dm_eq * L_VAP_WATER * inv_dt * inv_V, 0.0)
    c_Qgc = Q_conv * inv_V
    # slot-ordered scatter
    for name, parts in (("S_pyro", [c_Sp]), ("S_drying", [c_Sd1, c_Sd2]),
                        ("Q_pyro", [c_Qp]), ("Q_drying", [c_Qd1, c_Qd2]),
                        ("Y_F_source", [c_YF]), ("Q_char", [c_Qch]),
                        ("Q_smold", [c_Qsm]), ("Q_g_conv", [c_Qgc])):
        vals = np.stack(parts, axis=1).ravel()
        cells = np.repeat(cell, len(parts))
        np.add.at(src[name].reshape(-1), cells, vals)
    buf["T_s"][idx] = T_new
    buf["m_solid"][idx] = ms
    buf["m_water"][idx] = mw
    buf["m_char"][idx] = mc
    buf["m_char_max"][idx] = mcmax
    burned = mt < M_PARTICLE_BURNOUT
    dead_idx = idx[burned]
    buf["alive"][dead_idx] = 0
    n_burned = int(burned.sum())
    n_alive = int((~burned).sum())
    diag = np.zeros(16)
    if np.any(has):
        # first particle (slot order) attaining the running maximum of T_s
        Tm = np.where(has, T_newton, -np.inf)
        q = int(np.argmax(Tm))
        if Tm[q] > 0.0:
            Tq = T_newton[q]
            T_rad_loss = C_rad[q] * ((Tq * Tq) * (Tq * Tq) - Tamb4)
            diag[:13] = [Tq, Q_conv[q], Q_rxn[q], Q_dry[q], Q_ch[q], Q_sm[q],
                         Q_ext_p[q], T_rad_loss, Q_other[q], Tg[q], mc_dt[q],
                         C_rad[q], absF[q]]
            diag[14] = A_p[q]
            diag[15] = mt[q]
        diag[13] = float(absF.max()) if n else 0.0
    return src, n_alive, n_burned, diag


def aggregate_T_s_grid(buf, dx, dy, z_face, T_s_grid):
    Nz, Ny, Nx = T_s_grid.shape
    alive = buf["alive"] != 0
    i, j, k = agg_cell(buf["x"], buf["y"], buf["z"], dx, dy, z_face, Nz, Nx, Ny)
    mt = buf["m_solid"] + buf["m_water"] + buf["m_char"]
    ok = alive & (i >= 0) & (k >= 0) & (mt > 0.0)
    cell = ((k * Ny + j) * Nx + i)[ok]
    num = np.zeros(Nz * Ny * Nx)
    den = np.zeros(Nz * Ny * Nx)
    np.add.at(num, cell, (buf["T_s"] * mt)[ok])
    np.add.at(den, cell, mt[ok])
    flat = T_s_grid.reshape(-1)
    has = den > 0.0
    flat[has] = num[has] / den[has]


def aggregate_M_local_grid(buf, dx, dy, z_face, M_grid):
    Nz, Ny, Nx = M_grid.shape
    alive = buf["alive"] != 0
    i, j, k = agg_cell(buf["x"], buf["y"], buf["z"], dx, dy, z_face, Nz, Nx, Ny)
    ok = alive & (i >= 0) & (k >= 0)
    cell = ((k * Ny + j) * Nx + i)[ok]
    num = np.zeros(Nz * Ny * Nx)
    den = np.zeros(Nz * Ny * Nx)
    np.add.at(num, cell, buf["m_water"][ok])
    np.add.at(den, cell, buf["m_solid"][ok])
    has = den > 0.0
    out = np.zeros(Nz * Ny * Nx)
    out[has] = num[has] / den[has]
    M_grid[...] = out.reshape(M_grid.shape)


def horizontal_conduction(buf, T_s_grid, alpha_s, dx, dy, z_face, k_solid,
                          rho_solid_true, cp_solid, n_z_bed, dt):
    Nz, Ny, Nx = T_s_grid.shape
    diff = k_solid / (rho_solid_true * cp_solid)
    inv_dx2 = 1.0 / (dx * dx)
    inv_dy2 = 1.0 / (dy * dy)
    To = T_s_grid.copy()
    dT_cell = np.zeros((Nz, Ny, Nx))
    nb = min(n_z_bed, Nz)
    T = To[:nb]
    a = alpha_s[:nb]
    Txp = np.where(np.pad(a[:, :, 1:] > 0.0, ((0, 0), (0, 0), (0, 1))),
                   np.pad(T[:, :, 1:], ((0, 0), (0, 0), (0, 1))), T)
    Txm = np.where(np.pad(a[:, :, :-1] > 0.0, ((0, 0), (0, 0), (1, 0))),
                   np.pad(T[:, :, :-1], ((0, 0), (0, 0), (1, 0))), T)
    Typ = np.where(np.roll(a, -1, axis=1) > 0.0, np.roll(T, -1, axis=1), T)
    Tym = np.where(np.roll(a, 1, axis=1) > 0.0, np.roll(T, 1, axis=1), T)
    lap = ((Txp - 2.0 * T + Txm) * inv_dx2 + (Typ - 2.0 * T + Tym) * inv_dy2)
    dT = diff * lap * dt
    on = a > 0.0
    dT_cell[:nb] = np.where(on, dT, 0.0)
    T_s_grid[:nb] = np.where(on, T + dT, T_s_grid[:nb])
    alive = buf["alive"] != 0
    i, j, k = agg_cell(buf["x"], buf["y"], buf["z"], dx, dy, z_face, Nz, Nx, Ny)
    ok = alive & (i >= 0) & (k >= 0) & (k < n_z_bed)
    cell = ((k * Ny + j) * Nx + i)[ok]
    buf["T_s"][ok] += dT_cell.reshape(-1)[cell]
