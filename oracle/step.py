"""FP64 NumPy restatement of one Tier-3 outer step (oracle; test
infrastructure only).

Follows run_3d_spread of the reference (spread_3d.py:1126-1896 setup,
:1897-3683 loop body, :3743-3824 front tracking and ROS) for the
configuration the GPU driver implements: k-epsilon (any turbulence_model
other than smagorinsky), EDC, DOM every ``dom_subcycle_every`` steps,
Lagrangian bed particles or the Eulerian bed, passive or active level set,
outdoor log-law inflow,
periodic y.  The kernels are the oracle's own restatements (``oracle.bed``,
``oracle.dom``, ``oracle.edc``, ``oracle.levelset``, ``oracle.projection``,
``oracle.transport``).

Dead work of the reference (outputs nothing downstream consumes on this
path) is skipped here exactly as in the GPU driver: the O2-supply limiter
and omega_max_T (discarded by the EDC closure), tau_mix, burning_mask,
Q_pyro, q_dom_fwd / q_in_at_front, update_cell_age, and the distance
transforms of compute_phi_flame_from_state (only the sign is consumed).

``grid`` is any object with the Grid3D attributes (Nx, Ny, Nz, dx, dy, Lx,
dz_arr, d_face_above, d_face_below, z_face, x_mid, z_mid, n_z_bed); ``cfg``
any object with the resolved run parameters (names of the deck keys).
"""
from __future__ import annotations

import math

import numpy as np

from . import bed as Bd
from . import dom as Dm
from . import edc as Ed
from . import eulerian as Eu
from . import levelset as LS
from . import snapshot as Sn
from . import projection as Pj
from . import transport as Tr

P0 = 101_325.0
R_AIR = 8.314 / 0.029
G = 9.81
RHO_PARTICLE = 500.0
RHO_SOLID = 500.0          # coupling_3d.py:50
CP_SOLID = 1300.0          # coupling_3d.py:51
K_SOLID_GRASS = 0.20       # solid_conduction_3d.py:52
SC_T = 0.7
PR_T = 0.85
C_MU = 0.09
K_MIN = 1.0e-8
EPS_MIN = 1.0e-8
T_BOOTSTRAP_S = 2.0        # flame_front_3d.py
N_SUB = 10
T_CAP = 10_000.0
CP_SOLID_VN = 1300.0       # spread_3d.py:80, cp_s of compute_v_n_3d
Sn_DX_VN_BAND_M = 0.20     # flame_front_3d.py:64
STALL_WINDOW_S = 3.0       # spread_3d.py:1800


def midflame_wind_speed(U_10m, terrain="open"):
    """boundary.py:49-65."""
    return U_10m * {"open": 0.9, "shrub": 0.6}.get(terrain.lower(), 0.9)


def log_law(z, U_ref, z_ref=10.0, z_0=0.01):
    """boundary.py:68-111."""
    if z <= 0.0 or U_ref <= 0.0 or z_ref <= 0.0:
        return 0.0
    u_tk = U_ref / math.log((z_ref + z_0) / z_0)
    return max(0.0, u_tk * math.log((z + z_0) / z_0))


class OracleSpread3D:
    def __init__(self, cfg, grid, bed=None):
        """bed: (fill, m_scale, h_top) of a seeded bed variation (no reference
        counterpart; None parts keep the reference's uniform bed): fill
        multiplies alpha_s and the Eulerian pools per cell, m_scale the
        particles' masses, h_top is the per-particle fuel top."""
        fill, particle_mass_scale, h_top = bed if bed is not None else (None, None, None)
        self.cfg, self.g = cfg, grid
        g = grid
        sh = (g.Nz, g.Ny, g.Nx)
        self.shape = sh
        T_amb = cfg.T_amb
        z = lambda: np.zeros(sh)
        # CellState3D.allocate (spread_3d.py:88-155)
        self.rho = np.full(sh, P0 / (R_AIR * T_amb))
        self.u, self.v, self.w = z(), z(), z()
        self.T_g = np.full(sh, T_amb)
        self.T_s = np.full(sh, T_amb)
        self.Y_fuel, self.Y_H2O = z(), z()
        self.Y_O2 = np.full(sh, 0.232)
        self.alpha_s = z()
        i0 = int(np.searchsorted(g.x_mid, cfg.bed_x_start, side="left"))
        i1 = int(np.searchsorted(g.x_mid, cfg.bed_x_end, side="right"))
        if i0 >= i1:
            raise ValueError("empty fuel bed")
        self.i_bed = (i0, i1)
        self.alpha_s[:g.n_z_bed, :, i0:i1] = cfg.rho_b / RHO_PARTICLE
        if fill is not None:
            self.alpha_s *= fill
        self.eulerian = not getattr(cfg, "lagrangian_bed", True)
        # Eulerian pools (spread_3d.py:1186-1195, :1274-1278)
        self.eu_m_solid, self.eu_m_char, self.eu_m_water = z(), z(), z()
        self.eu_Q_pyro = z()
        self.eu_m_water_init = cfg.rho_b * cfg.moisture
        if self.eulerian:
            self.eu_m_solid[:g.n_z_bed, :, i0:i1] = cfg.rho_b * 1.0
            self.eu_m_water[:g.n_z_bed, :, i0:i1] = cfg.rho_b * cfg.moisture
            if fill is not None:
                self.eu_m_solid *= fill
                self.eu_m_water *= fill
        # bed particles (spread_3d.py:1300-1330)
        nb = 0 if self.eulerian else int((self.alpha_s[:g.n_z_bed, :, i0:i1] > 0).sum())
        self.buf = Bd.allocate_bed_particle_buffers(nb * cfg.n_per_cell)
        self.n_alloc = 0 if self.eulerian else Bd.initialize_bed_particles_from_alpha_s(
            self.buf, self.alpha_s, cfg.rho_b, cfg.moisture, T_amb, g.dx, g.dy,
            g.dz_arr, g.n_z_bed, cfg.n_per_cell, sav=cfg.sigma_sav,
            m_scale=particle_mass_scale)
        if h_top is not None and self.n_alloc > 0:
            self.buf["h_top"] = np.asarray(h_top, np.float64)[:self.n_alloc].copy()
        # ignition patch + solid-phase seeding (:1486-1526)
        n_src = max(1, int(round(0.5 * cfg.ignition_width_mult / g.dx)))
        s0, s1 = i0, min(i1, i0 + n_src)
        self.i_src = (s0, s1)
        if cfg.solid_phase_ignition and not self.eulerian:
            x_lo = g.x_mid[s0] - 0.5 * g.dx
            x_hi = g.x_mid[s1 - 1] + 0.5 * g.dx
            b = self.buf
            sel = (b["alive"] != 0) & (x_lo <= b["x"]) & (b["x"] < x_hi)
            b["T_s"][sel] = cfg.solid_phase_T_s_K
        # inlet, sponge, turbulence (:1391-1580)
        self.U_mf = midflame_wind_speed(cfg.wind_speed_m_s, cfg.terrain)
        self.u_inlet = np.zeros((g.Nz, g.Ny))
        for k in range(g.Nz):
            self.u_inlet[k, :] = log_law(g.z_mid[k], cfg.wind_speed_m_s)
        # seeded inlet gusts (no reference counterpart): the log law times
        # g(t), g piecewise half-cosine between nodes 1 + a_m at t = m T,
        # a_0 = 0, a_m = A (2 U_m - 1) with U_m the m-th PCG64(seed) draw
        self.u_inlet_base = self.u_inlet.copy()
        self.gust_A = float(getattr(cfg, "gust_amplitude", 0.0))
        self.gust_T = float(getattr(cfg, "gust_period_s", 5.0))
        self._gust_rng = np.random.default_rng(int(getattr(cfg, "gust_seed", 0)))
        self._gust_nodes = [0.0]
        self.u[:] = self.u_inlet[:, :, None]
        self.v_inlet = np.zeros((g.Nz, g.Ny))
        self.w_inlet = np.zeros((g.Nz, g.Ny))
        Ns = max(3, min(int(round(0.5 / g.dx)), g.Nx // 4))
        self.sigma_x = np.zeros(g.Nx)
        for i in range(g.Nx - Ns, g.Nx):
            fr = (i - (g.Nx - Ns)) / Ns
            self.sigma_x[i] = 5.0 * fr * fr
        k_init = max(1.5 * (0.10 * max(self.U_mf, 0.5)) ** 2, 1.0e-4)
        eps_init = C_MU * k_init * k_init / 0.01
        self.k = np.full(sh, k_init)
        self.eps = np.full(sh, eps_init)
        self.nu_t = np.full(sh, 0.01)
        self.k_wall = np.full((g.Ny, g.Nx), K_MIN)
        self.e_wall = np.full((g.Ny, g.Nx), EPS_MIN)
        # work fields
        self.omega, self.S_pyro, self.Q_comb = z(), z(), z()
        self.q_rad, self.q_rad_gas = z(), z()
        self.Fx, self.Fy, self.Fz = z(), z(), z()
        self.M_local = z()
        self.src = {k: z() for k in ("S_pyro", "S_drying", "Q_pyro", "Q_drying",
                                     "Y_F_source", "Q_char", "Q_smold", "Q_g_conv")}
        # soil (:1545-1551)
        self.soil_dz, self.soil_da, self.soil_db, _ = Tr.build_soil_grid(6, 0.001, 1.5)
        self.T_soil = np.full((6, g.Ny, g.Nx), T_amb)
        self.q_in_soil = np.zeros((g.Ny, g.Nx))
        # level set (:1591-1625)
        self.phi = np.full(sh, LS.PHI_INIT_UNBURNED)
        if g.n_z_bed > 0 and s1 > s0:
            self.phi = LS.initialize_source_patch(sh, s0, s1, g.n_z_bed - 1, g.x_mid, g.dx)
        self.lset_step = 0
        self.seed_count = 0
        self.v_n = np.zeros(sh)
        # solvers (:1719-1780)
        self.proj = Pj.ProjectionSolver3D(g.Nz, g.Ny, g.Nx, g.dy, g.dx, g.dz_arr,
                                          g.d_face_above, g.d_face_below,
                                          cg_rtol=cfg.cg_rtol)
        self.proj.set_inlet_BC(self.u_inlet)
        self.dom = Dm.DOMRadiationSolver(g.Nz, g.Ny, g.Nx, g.dy, g.dx, g.dz_arr,
                                         N_quadrature=cfg.dom_quadrature)
        # dt and histories (:1782-1830)
        self.T_ign = T_amb + 300.0
        self.dz_min = float(g.dz_arr.min())
        U_char = max(self.U_mf, math.sqrt(G * cfg.h_bed * 0.5), 0.1)
        dt_adv = cfg.cfl_factor * min(g.dx, g.dy, self.dz_min) / U_char
        dt_diff = 0.25 * self.dz_min * self.dz_min / max(float(self.nu_t.max()), 1.0e-3)
        self.dt = min(dt_adv, dt_diff, 0.05)
        self.rho_prev = self.rho.copy()
        x0 = float(g.x_mid[s1 - 1] + 0.5 * g.dx) if (g.n_z_bed > 0 and s1 > s0) else 0.0
        self.front_t, self.front_x = [0.0], [x0]
        self.diag = {"t": [0.0], "Tg": [float(self.T_g.max())], "Ts": [float(self.T_s.max())],
                     "Sp": [0.0], "Qc": [0.0], "om": [0.0], "Y": [float(self.Y_fuel.max())],
                     "n_ign": [0], "div": [0.0], "iters": [0]}
        self.t = 0.0
        self.step_no = 0
        self.extinct = 0
        self.last_front = 0.0
        self.last_advance_t = 0.0
        self.stop_reason = None

    # -----------------------------------------------------------------------
    def gust(self, t):
        """g(t) of the inlet gusts (1 without them)."""
        if not (self.gust_A > 0.0 and self.gust_T > 0.0):
            return 1.0
        m = int(math.floor(t / self.gust_T))
        while len(self._gust_nodes) <= m + 1:
            self._gust_nodes.append(self.gust_A * (2.0 * float(self._gust_rng.random()) - 1.0))
        a0, a1 = self._gust_nodes[m], self._gust_nodes[m + 1]
        f = t / self.gust_T - m
        return 1.0 + a0 + (a1 - a0) * 0.5 * (1.0 - math.cos(math.pi * f))

    def step(self):
        """One pass of the loop body; returns a stop reason or None."""
        cfg, g = self.cfg, self.g
        if self.gust_A > 0.0:
            self.u_inlet[...] = self.u_inlet_base * self.gust(self.t)
        T_amb = cfg.T_amb
        dx, dy, dz, da, db = g.dx, g.dy, g.dz_arr, g.d_face_above, g.d_face_below
        self.step_no += 1
        t = self.t
        # adaptive dt (:1906-1941)
        u_now = max(float(np.abs(self.u).max()), float(np.abs(self.v).max()),
                    float(np.abs(self.w).max()), self.U_mf,
                    math.sqrt(G * cfg.h_bed * 0.5), 0.1)
        nu_now = max(float(self.nu_t.max()), 1.0e-3)
        dt_new = min(cfg.cfl_factor * min(dx, dy, self.dz_min) / u_now,
                     0.25 * self.dz_min * self.dz_min / nu_now, 0.05)
        dt_new = min(dt_new, self.dt * 1.5)
        if dt_new < cfg.min_dt_s:
            dt_new = cfg.min_dt_s
        if not np.isfinite(float(self.T_g.max())):
            raise RuntimeError(f"State NaN/Inf at t={t:.3f}s -- simulation exploded.")
        dt = self.dt = dt_new
        combustion = t >= cfg.spin_up_s
        ign_on = t < cfg.ignition_duration_s
        s0, s1 = self.i_src
        nzb = g.n_z_bed
        # rho_prev, EoS, inlet O2 (:1984-2050)
        np.copyto(self.rho_prev, self.rho)
        self.rho[...] = P0 / (R_AIR * np.maximum(self.T_g, T_amb))
        self.Y_O2[:, :, 0] = 0.232
        # bed (:2064-2215)
        if self.eulerian:
            if combustion:
                mw_init = self.eu_m_water_init
                Q_dry = np.zeros(self.shape)
                if mw_init > 0.0:
                    Eu.drying(self.T_s, self.eu_m_water, self.alpha_s, dt, Q_dry)
                Eu.pyrolysis_md2004(self.T_s, self.eu_m_solid, self.eu_m_char, self.Y_O2,
                                    self.alpha_s, self.eu_m_water, mw_init, dt, self.S_pyro,
                                    self.eu_Q_pyro, cfg.pyro_A_p, cfg.pyro_E_p)
                self.eu_Q_pyro += Q_dry
                Q_char = np.zeros(self.shape)
                Eu.char_oxidation(self.T_s, self.eu_m_char, self.Y_O2, self.alpha_s, dt, Q_char)
                self.eu_Q_pyro -= Q_char
                Q_sm = np.zeros(self.shape)
                Eu.smoldering(self.T_s, self.eu_m_char, self.Y_O2, self.alpha_s, dt, Q_sm)
                self.eu_Q_pyro -= Q_sm
            else:
                self.S_pyro.fill(0.0)
                self.eu_Q_pyro.fill(0.0)
            self.src = {k: np.zeros(self.shape) for k in self.src}
            self.src["S_pyro"] = self.S_pyro
        elif combustion and self.n_alloc > 0:
            Qsx = np.zeros(self.shape)
            if ign_on:
                Qsx[:nzb, :, s0:s1] += 0.80 * (30_000.0 / max(cfg.h_bed, 1e-3))
            for k in range(nzb):
                Qsx[k] += np.clip(self.q_rad[k] / dz[k], -1.0e5, 1.0e5)
            self.src, _, _, _ = Bd.step_bed_particles(
                self.buf, self.T_g, self.Y_O2, Qsx, cfg.n_per_cell, dx, dy, dz, g.z_face,
                cfg.h_conv, cfg.rho_solid_true, cfg.cp_solid, cfg.eps_solid, T_amb,
                cfg.view_factor, cfg.view_factor_geometric, cfg.h_bed,
                cfg.sigma_sav * cfg.rho_b / cfg.rho_solid_true, dt, True, True, True,
                True, cfg.drying_mode, cfg.char_ox_flux_cap, cfg.char_ox_ash_exp)
            np.copyto(self.S_pyro, self.src["S_pyro"])
            self.S_pyro += self.src["S_drying"]
            Bd.aggregate_T_s_grid(self.buf, dx, dy, g.z_face, self.T_s)
            Bd.horizontal_conduction(self.buf, self.T_s, self.alpha_s, dx, dy, g.z_face,
                                     K_SOLID_GRASS, cfg.rho_solid_true, cfg.cp_solid,
                                     nzb, dt)
        else:
            self.S_pyro.fill(0.0)
        # drag + tentative velocity (:2217-2277)
        Tr.step_drag_force(self.u, self.v, self.w, self.rho, self.alpha_s, cfg.sigma_sav,
                           self.Fx, self.Fy, self.Fz, cfg.canopy_C_d)
        Tr.step_tentative_velocity(self.u, self.v, self.w, self.rho, self.T_g, self.Fx,
                                   self.Fy, self.Fz, dt, dx, dy, dz, da, db, T_amb,
                                   self.u_inlet, self.v_inlet, self.w_inlet)
        # projection loop (:2293-2360)
        drho_dt = (self.rho - self.rho_prev) / dt
        div_target = (self.S_pyro - drho_dt) / np.maximum(self.rho, 0.1)
        self.proj.rebuild_for_rho(self.rho)
        proj_n, proj_div = 0, 0.0
        for it in range(cfg.proj_max_iter):
            self.proj.project(self.u, self.v, self.w, self.rho, dt, div_target=div_target)
            dv = self.proj.divergence(self.u, self.v, self.w) - div_target
            proj_div = float(np.max(np.abs(dv)))
            proj_n = it + 1
            if proj_div < cfg.proj_div_tol:
                break
        # sponge (:2368-2372)
        Tr.apply_outflow_sponge(self.u, self.u_inlet, self.sigma_x, self.Y_fuel, 1.0e-3, dt)
        # k-epsilon (:2408-2433)
        Tr.step_k_epsilon(self.k, self.eps, self.nu_t, self.u, self.v, self.w, self.T_g,
                          self.rho, self.alpha_s, cfg.sigma_sav, dt, dx, dy, dz, da, db,
                          T_amb, self.u_inlet, self.k_wall, self.e_wall,
                          cfg.canopy_beta_p, cfg.canopy_beta_d)
        # DOM every dom_subcycle_every steps, soil every step (:2455-2510)
        if self.step_no % max(cfg.dom_subcycle_every, 1) == 0:
            if not self.eulerian:
                Bd.aggregate_M_local_grid(self.buf, dx, dy, g.z_face, self.M_local)
            self.dom.solve(self.T_s, self.T_g, self.alpha_s, self.omega, cfg.sigma_sav,
                           T_amb, self.q_rad, self.q_rad_gas, T_soil_surface=self.T_soil[0],
                           q_in_soil_out=self.q_in_soil, Y_H2O=self.Y_H2O, rho=self.rho,
                           bed_moisture_per_cell=None if self.eulerian else self.M_local)
        Tr.step_soil_conduction(self.T_soil, self.q_in_soil, dt, self.soil_dz,
                                self.soil_da, self.soil_db, T_amb=T_amb)
        # ignition pulse (:2528-2532)
        if ign_on:
            self.q_rad[nzb - 1, :, s0:s1] += 240_000.0 * cfg.ignition_q_mult
        # flame-body mask from the pre-sub-step state (:2551-2566)
        flame_n = int(LS.flame_active(self.omega, self.T_g, self.Y_fuel).sum())
        # gas-energy advection (:2613-2622)
        Tr.adv_gas_energy(self.T_g, self.u, self.v, self.w, dt, dx, dy, dz, da, db,
                          2.0e-5, T_amb)
        # N_SUB sub-steps (:2678-3200)
        dt_sub = dt / N_SUB
        Sp, Sd, Qgc = self.src["S_pyro"], self.src["S_drying"], self.src["Q_g_conv"]
        drip_gas = (1.0 - 0.80) * (30_000.0 / max(cfg.h_bed, 1e-3))
        for _ in range(N_SUB):
            Ed.step_chemistry_ode_edc(self.rho, self.T_g, self.Y_fuel, self.Y_O2, self.k,
                                      self.eps, 0.34, 1100.0, dt_sub, 1, self.omega,
                                      self.Y_H2O, extinction_enable=False)
            self.Q_comb.fill(0.0)
            if combustion and ign_on:
                if self.eulerian:
                    self.eu_Q_pyro[:nzb, :, s0:s1] -= 0.80 * (30_000.0 / max(cfg.h_bed, 1e-3))
                self.Q_comb[:nzb, :, s0:s1] += drip_gas
            S_tot = Sp + Sd
            S_F = Sp - self.Y_fuel * S_tot
            S_O = -self.Y_O2 * S_tot
            S_W = Sd - self.Y_H2O * S_tot
            Tr.step_species_transport(self.Y_fuel, self.rho, self.u, self.v, self.w, S_F,
                                      dt_sub, dx, dy, dz, da, db, 1.0e-5, 0.0)
            Tr.step_species_transport(self.Y_O2, self.rho, self.u, self.v, self.w, S_O,
                                      dt_sub, dx, dy, dz, da, db, 1.0e-5, 0.232)
            Tr.step_species_transport(self.Y_H2O, self.rho, self.u, self.v, self.w, S_W,
                                      dt_sub, dx, dy, dz, da, db, 1.0e-5, 0.0)
            np.clip(self.Y_O2, 0.0, 0.232, out=self.Y_O2)
            np.clip(self.Y_H2O, 0.0, 1.0, out=self.Y_H2O)
            for f in (self.Y_fuel, self.Y_O2, self.Y_H2O):
                Tr.apply_turbulent_diffusion(f, self.nu_t, SC_T, dt_sub, dx, dy, dz, da, db)
            np.clip(self.Y_O2, 0.0, 0.232, out=self.Y_O2)
            np.clip(self.Y_fuel, 0.0, 1.0, out=self.Y_fuel)
            np.clip(self.Y_H2O, 0.0, 1.0, out=self.Y_H2O)
            for k in range(g.Nz):
                self.Q_comb[k] += self.q_rad_gas[k] / dz[k]
            if self.eulerian:
                # coupling_3d with q_rad (+ a zero Frankman flux) into the
                # solid, no emission loss (:3094, :3148-3156)
                Eu.gas_solid_coupling(self.T_g, self.T_s, self.rho, self.u, self.v, self.w,
                                      self.alpha_s, cfg.sigma_sav, self.q_rad + 0.0,
                                      self.eu_Q_pyro, self.Q_comb, self.eu_m_water,
                                      2_257_000.0 if self.eu_m_water_init > 0.0 else 0.0,
                                      dt_sub, dz, T_amb, q_loss_enable=False,
                                      h_conv_mult=cfg.h_conv_mult)
            else:
                cp_mix = (1.0 - self.Y_H2O) * 1100.0 + self.Y_H2O * 2000.0
                debit = (Sp * np.maximum(self.T_g - self.T_s, 0.0)
                         + Sd * np.maximum(self.T_g - 373.15, 0.0)) * 2000.0
                inv = dt_sub / (np.maximum(self.rho, 1.0e-3) * cp_mix)
                self.T_g += (self.Q_comb - Qgc - debit) * inv
                np.maximum(self.T_g, T_amb, out=self.T_g)
            Tr.step_solid_conduction_vertical(self.T_s, self.alpha_s, dz, da, db,
                                              K_SOLID_GRASS, RHO_SOLID, CP_SOLID, dt_sub)
            Tr.apply_turbulent_diffusion(self.T_g, self.nu_t, PR_T, dt_sub, dx, dy, dz, da, db)
            np.minimum(self.T_g, T_CAP, out=self.T_g)
            np.maximum(self.T_g, T_amb, out=self.T_g)
            np.minimum(self.T_s, T_CAP, out=self.T_s)
            np.maximum(self.T_s, T_amb, out=self.T_s)
        # level set (:3204-3300; passive: the v_n = 0 evolve is the identity),
        # exits (:3413-3440)
        if nzb > 0:
            self.lset_step += 1
            if not cfg.level_set_passive:
                # band of the step's level set and its DOM forward flux
                # (:2555-2565, :2591), front speed (:3227-3256), evolve (:3260)
                mask = Sn.ahead_band_mask(self.phi, max(dx, Sn_DX_VN_BAND_M), nzb)
                q = Sn.q_dom_fwd_at_band(self.dom.Omega, self.dom.weights, self.dom.I_set,
                                         mask)
                q_in = np.where(mask, 0.0 + q, 0.0)
                v = LS.compute_v_n_3d(q_in, cfg.rho_b, CP_SOLID_VN, cfg.h_bed, self.T_ign,
                                      T_amb, M_local=self.M_local)
                w = cfg.empirical_blend_w if cfg.empirical_ros_enable else 0.0
                if w > 0.0:
                    v *= (1.0 - w)
                    v += w * cfg.empirical_ros_m_s
                self.v_n = v
                self.phi = LS.evolve(self.phi, dt, v, dx, dy, dz)
            if self.lset_step % LS.LEVELSET_REINIT_INTERVAL == 0:
                self.phi = LS.reinitialize(self.phi, dx, dy, dz)
            fx_exit = LS.front_x(self.phi, 1, g.Ny // 2, dx)
            # dynamic seeding behind the empirical-ROS front (:3271-3300)
            if (not cfg.level_set_passive and cfg.empirical_ros_enable
                    and cfg.empirical_blend_w > 0.0 and not self.eulerian):
                self.seed_count += 1
                if self.seed_count >= 50:
                    self.seed_count = 0
                    b = self.buf
                    x_lo = g.x_mid[self.i_bed[0]] - 0.5 * dx
                    sel = ((b["alive"] == 1) & (b["x"] >= x_lo) & (b["x"] <= fx_exit)
                           & (b["T_s"] < self.T_ign) & (b["m_solid"] > 1.0e-9))
                    b["T_s"][sel] = cfg.solid_phase_T_s_K
            if flame_n == 0 and t > cfg.ignition_duration_s + T_BOOTSTRAP_S:
                self.extinct += 1
                if self.extinct > 50:
                    self.stop_reason = "extinct"
                    return self.stop_reason
            else:
                self.extinct = 0
            if fx_exit > cfg.bed_x_end - dx:
                self.stop_reason = "front_exit"
                return self.stop_reason
        # EoS (:3440) and front tracking (:3455-3475)
        self.rho[...] = P0 / (R_AIR * np.maximum(self.T_g, T_amb))
        front_ts = self._front_ts()
        if nzb > 0:
            fl = LS.front_x(self.phi, nzb // 2, g.Ny // 2, dx)
            if math.isfinite(fl):
                if fl > (self.front_x[-1] if self.front_x else 0.0):
                    self.front_t.append(t)
                    self.front_x.append(fl)
                new_front = fl
            else:
                new_front = front_ts
        else:
            new_front = front_ts
        if new_front > self.last_front:
            self.last_front = new_front
            self.last_advance_t = t
        self.t = t + dt
        d = self.diag
        d["t"].append(self.t)
        d["Tg"].append(float(self.T_g.max()))
        d["Ts"].append(float(self.T_s.max()))
        d["Sp"].append(float(self.S_pyro.max()))
        d["Qc"].append(float(self.Q_comb.max()))
        d["om"].append(float(self.omega.max()))
        d["Y"].append(float(self.Y_fuel.max()))
        d["n_ign"].append(int(np.sum(self.T_s >= self.T_ign)))
        d["div"].append(proj_div)
        d["iters"].append(proj_n)
        if self.front_x[-1] > g.Lx * 0.85:
            self.stop_reason = "domain_end"
            return self.stop_reason
        # stall of an active level set (:3675-3683)
        if (not cfg.level_set_passive and self.t - self.last_advance_t > STALL_WINDOW_S
                and self.t > cfg.ignition_duration_s):
            self.stop_reason = "stalled"
            return self.stop_reason
        return None

    def _front_ts(self):
        """_update_front_tracking with append=False (spread_3d.py:3743-3780)."""
        g = self.g
        nzb = g.n_z_bed
        burning = (self.T_s[:nzb] >= self.T_ign) & (self.alpha_s[:nzb] > 0.0)
        if not burning.any():
            return self.front_x[-1] if self.front_x else 0.0
        cols = np.where(burning.any(axis=(0, 1)))[0]
        return float(g.x_mid[int(cols[-1])])

    def run(self, n_steps):
        for _ in range(n_steps):
            if self.step() is not None:
                break
        return self

    def ros(self):
        """_compute_steady_ros (spread_3d.py:3796-3824)."""
        ft, fx = self.front_t, self.front_x
        if len(ft) < 2:
            return 0.0
        n_src = self.i_src[1] - self.i_src[0]
        if fx[-1] - n_src * self.g.dx < 0.1:
            return 0.0
        if self.t - ft[-1] > 30.0 and not fx[-1] > self.g.Lx * 0.7:
            return 0.0
        return (fx[-1] - fx[0]) / (self.t - ft[0]) if self.t > ft[0] else 0.0
