"""GPU driver of the Tier-3 spread time step: ``run_3d_spread`` of the
reference (spread_3d.py:563-3729) for the k-epsilon / EDC / DOM
configuration with a Lagrangian (particle) or Eulerian fuel bed
(``lagrangian_bed_enable``) and a passive or an active level set
(``level_set_passive``; both default false as in the reference), every
field resident in device memory.

Precision: the grid fields are float32 by default (deck key
``outdoor.dtype = float64`` selects the float64 parity build); bed particles,
the level set, the pressure warm start and every reduction stay float64.

Each operation is one call into libfirewave.  The host keeps only the
reference's scalar decisions (dt, the projection loop, the turbulent
diffusion sub-step count, exits), read back from one fused reduction per
step plus the projection and nu_t scalars.

Work the reference computes but never consumes on this path is skipped; the
skip is exact because nothing downstream reads it: the O2-supply limiter and
omega_max_T (ignored by the EDC closure), tau_mix, burning_mask, Q_pyro,
q_dom_fwd / q_in_at_front with a passive level set, update_cell_age (read
only by the ebu_bootstrap closure), the passive level set's evolve with
v_n = 0 (the identity), and the distance transforms of
compute_phi_flame_from_state: the flame-body mask with band 0 is exactly the
`active` predicate (flame_front_3d.py:756-813), and its count for step n is
the count of the state the previous step ended with (omega, T_g and Y_fuel
are not written between the end of a step and that point of the next one).
The snapshot writer computes the skipped fields it writes when a snapshot is
due.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np
import torch

from . import _lib
from .deck import load_text_input, outdoor_env_from_dict
from .device import DeviceGrid, Marshal, device, grid_for, ptr, stream_ptr
from .bed_variation import GustSchedule, bed_variation, fuel_depth_field
from .grid import Grid3D, build_u_inlet, midflame_wind_speed
from .physics_3d import dom_3d as _dom
from .physics_3d import lagrangian_bed_3d as _lb
from .physics_3d import turbulence_3d as _turb
from .physics_3d.fft_poisson_3d import create_handle as _proj_handle
from .physics_3d.soil_3d import (CP_SOIL_DEFAULT, EPS_SOIL_DEFAULT, K_SOIL_DEFAULT,
                                 RHO_SOIL_DEFAULT, build_soil_grid)

_P0 = 101_325.0
_R_AIR = 8.314 / 0.029
_G = 9.81
_RHO_PARTICLE = 500.0
L_BURNOUT_M = 0.50          # flame_front_3d.py:62
DX_VN_BAND_M = 0.20         # flame_front_3d.py:64
L_VAP_WATER = 2.26e6        # flame_front_3d.py:437
_CP_SOLID_VN = 1300.0       # spread_3d.py:80, the cp_s of compute_v_n_3d
_STALL_WINDOW_S = 3.0       # spread_3d.py:1800


class FwSpreadFields(C.Structure):
    """fw_spread_fields (include/firewave.h)."""
    _fields_ = ([(n, C.c_void_p) for n in (
        "rho", "rho_prev", "u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O",
        "alpha_s", "Fx", "Fy", "Fz", "k", "eps", "nu_t", "omega", "S_pyro", "div_target",
        "Q_comb", "q_rad", "q_rad_gas", "Q_ext", "M_local", "T_s_old", "tmp1", "tmp2",
        "tmp3")] + [("src", C.c_void_p * 8)] + [(n, C.c_void_p) for n in (
            "u_inlet", "v_inlet", "w_inlet", "sigma_x", "k_wall", "eps_wall", "T_soil",
            "q_in_soil", "soil_dz", "soil_d_above", "soil_d_below", "phi", "phi_tmp",
            "phi0", "red_work", "red_out", "red_host", "nu_work", "contrib", "dT_work",
            "counters", "edc_work", "lset_mask", "q_dom_fwd", "v_n", "eu_m_solid",
            "eu_m_char", "eu_m_water", "Q_pyro", "graphs", "step_scalars")])


class FwSpreadParams(C.Structure):
    """fw_spread_params (include/firewave.h)."""
    _fields_ = [(n, C.c_double) for n in (
        "T_amb", "h_bed", "sigma_sav", "canopy_C_d", "canopy_beta_p", "canopy_beta_d",
        "cfl_factor", "min_dt", "U_mf", "dz_min", "T_ign", "ignition_duration", "spin_up",
        "ignition_q_mult", "bed_x_end", "Lx", "x_step", "proj_div_tol", "rho_solid_true",
        "cp_solid", "soil_k", "soil_rho", "soil_cp", "soil_eps")] + [
        (n, C.c_int32) for n in ("n_soil", "i_src0", "i_src1", "proj_max_iter",
                                 "dom_every", "reserved")] + [
        (n, C.c_double) for n in ("band_m", "E_sens", "E_lat_coef", "emp_w", "emp_ros",
                                  "seed_T_s", "x_lo_bed", "stall_window")] + [
        (n, C.c_int32) for n in ("lset_active", "reserved2")] + [
        (n, C.c_double) for n in ("eu_m_water_init", "eu_A_p", "eu_E_p", "eu_L_v",
                                  "h_conv_mult")] + [
        (n, C.c_int32) for n in ("eulerian", "reserved3")]


class FwSpreadState(C.Structure):
    """fw_spread_state (include/firewave.h)."""
    _fields_ = [(n, C.c_double) for n in (
        "t", "dt", "front_exit_x", "front_mid_x", "last_front", "last_advance_t",
        "front_x_last", "front_t_new", "front_x_new", "proj_div")] + [
        ("red", C.c_double * 14)] + [(n, C.c_int32) for n in (
            "step_no", "lset_step", "extinct_count", "stop", "front_appended", "proj_n",
            "pcg_iters", "reserved", "seed_count", "reserved2")]


# grid-type fields the native step reads, writes and swaps
_GRID_SLOTS = ("rho", "rho_prev", "u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O",
               "alpha_s", "Fx", "Fy", "Fz", "k", "eps", "nu_t", "omega", "S_pyro",
               "div_target", "Q_comb", "q_rad", "q_rad_gas", "Q_ext", "M_local", "T_s_old",
               "tmp1", "tmp2", "tmp3")
_SRC_NAMES = ("Sp", "Sd", "Qp", "Qd", "YF", "Qch", "Qsm", "Qgc")
_STOPS = {1: "extinct", 2: "front_exit", 3: "domain_end", 4: "stalled"}


def _adv_gas_energy(T_g, u, v, w, dt, dx, dy, dz_arr, d_face_above,
                    d_face_below, alpha_th, T_amb, T_inlet_zmin=None,
                    z_min_inlet_active=False):
    """MUSCL advection + alpha_th diffusion of T_g, clamped to [T_amb, 1e4].

    spread_3d.py:3827-3897 of the reference; in place.
    """
    if z_min_inlet_active:
        raise ValueError("z-min inlet (cup burner) is not part of the outdoor hot path")
    m = Marshal(T_g)
    g = grid_for(m, tuple(T_g.shape), dx, dy, dz_arr, d_face_above, d_face_below)
    Tg = m.out(T_g)
    To = torch.empty_like(Tg)
    _lib.call("fw_gas_energy", g.ref, ptr(Tg), ptr(To), ptr(m.dev(u)), ptr(m.dev(v)),
              ptr(m.dev(w)), float(dt), float(alpha_th), float(T_amb), stream_ptr())
    Tg.copy_(To)
    m.writeback()


# ---------------------------------------------------------------------------
# configuration (spread_3d.py:959-1125)
# ---------------------------------------------------------------------------

# deck switches whose other values select a branch this driver does not
# implement: (deck key, the one accepted value, reference default)
_REQUIRED = (
    ("combustion_closure", "edc", "edc"),
    ("radiation_solver", "dom", "dom"),
    ("boundary_condition_kind", "outdoor_wind", "outdoor_wind"),
    ("chemistry_family", "biomass", "biomass"),
    ("ignition_T_pin_enable", False, False),
    ("edc_extinction_enable", False, False),
    ("finney_tendril_enable", False, False),
    ("moisture_jump_enable", False, False),
)
_OUTDOOR_OFF = ("sem_enable", "fuel_pert_enable", "volume_weighted_projection",
                "rothermel_packing_enable", "free_burning_mode")


@dataclass
class Spread3DConfig:
    """Resolved parameters of one run.  Mesh keys resolve keyword -> deck ->
    default (`_pick`, spread_3d.py:972-1000); case-spec keys deck -> keyword
    -> default (`_deck_first`, :1034-1085)."""
    wind_speed_m_s: float
    Lx: float
    Ly: float
    Lz: float
    dx: float
    dy: float
    n_z_bed: int
    mesh_kwargs: dict
    bed_x_start: float
    bed_x_end: float
    cfl_factor: float
    proj_max_iter: int
    proj_div_tol: float
    h_bed: float
    rho_b: float
    T_amb: float
    sigma_sav: float
    moisture: float
    ignition_duration_s: float
    terrain: str
    canopy_C_d: float
    canopy_beta_p: float
    canopy_beta_d: float
    dom_subcycle_every: int
    dom_quadrature: int
    n_per_cell: int
    drying_mode: int
    h_conv: float
    rho_solid_true: float
    cp_solid: float
    eps_solid: float
    view_factor: float
    view_factor_geometric: bool
    char_ox_flux_cap: float
    char_ox_ash_exp: float
    solid_phase_ignition: bool
    solid_phase_T_s_K: float
    ignition_q_mult: float
    ignition_width_mult: float
    min_dt_s: float
    max_wall_time_s: float
    cg_rtol: float
    n_steps: Optional[int]
    dtype: torch.dtype
    particle_jitter: float
    particle_jitter_seed: int
    spin_up_s: float
    level_set_passive: bool = True
    lagrangian_bed: bool = True
    h_conv_mult: float = 1.0
    pyro_A_p: float = 36280.0
    pyro_E_p: float = 58600.0
    empirical_ros_enable: bool = False
    empirical_ros_m_s: float = 0.0
    empirical_blend_w: float = 0.0
    # seeded bed and inflow variation (no reference counterpart)
    fuel_depth_cv: float = 0.0
    fuel_depth_corr_m: float = 1.0
    fuel_depth_seed: int = 0
    bed_zone_m: float = 0.0          # mesh height of the bed layers (0: h_bed)
    gust_amplitude: float = 0.0
    gust_period_s: float = 5.0
    gust_seed: int = 0

    @classmethod
    def resolve(cls, ri_or_path, wind_speed_m_s=0.0, **kw):
        ri = load_text_input(ri_or_path) if isinstance(ri_or_path, (str, Path)) else ri_or_path
        ovr = ri.outdoor_overrides
        ocfg = outdoor_env_from_dict(ovr)
        ocfg.wind_speed_m_s = wind_speed_m_s

        def deck(key):
            for k in (key, key.lower()):
                if k in ovr and ovr[k] is not None:
                    return ovr[k]
            return None

        def pick(name, default):
            v = kw.pop(name, None)
            if v is not None:
                return v
            d = deck(name)
            return default if d is None else d

        def deck_first(name, default):
            v = kw.pop(name, None)
            d = deck(name)
            if d is not None:
                return d
            return default if v is None else v

        Lx = float(pick("Lx", 10.0))
        Ly = float(pick("Ly", 2.0))
        Lz = float(pick("Lz", 1.5))
        dx = float(pick("dx", 0.05))
        n_z_bed = int(pick("n_z_bed", 4))
        mesh = dict(
            dz_expansion=float(pick("dz_expansion", 1.0)),
            bl_growth=float(pick("bl_growth", 1.0)),
            wall_bl_N=int(pick("wall_bl_N", 0)),
            wall_bl_first_dz=float(pick("wall_bl_first_dz", 0.0)),
            wall_bl_growth=float(pick("wall_bl_growth", 1.3)),
            bed_top_inner_bl_N=int(pick("bed_top_inner_bl_N", 0)),
            bed_top_inner_bl_first_dz=float(pick("bed_top_inner_bl_first_dz", 0.0)),
            bed_top_inner_bl_growth=float(pick("bed_top_inner_bl_growth", 1.3)),
            bed_top_outer_bl_N=int(pick("bed_top_outer_bl_N", 0)),
            bed_top_outer_bl_first_dz=float(pick("bed_top_outer_bl_first_dz", 0.0)),
            bed_top_outer_bl_growth=float(pick("bed_top_outer_bl_growth", 1.3)),
            atm_growth=float(pick("atm_growth", 1.3)),
        )
        for opt in ("dz_first", "atm_max_dz", "atm_uniform_dz"):
            v = pick(opt, None)
            mesh[opt] = float(v) if v is not None else None
        cfl = float(pick("cfl_factor", 0.4))
        if bool(pick("wall_function", False)):
            raise ValueError("wall_function=true is not implemented on the GPU step")
        pmi, pdt = deck("proj_max_iter"), deck("proj_div_tol")
        proj_max_iter = int(pmi) if pmi is not None else 50
        proj_div_tol = float(pdt) if pdt is not None else 1.0e-3
        bed_x_start = float(pick("bed_x_start", 0.0))
        bed_x_end = float(pick("bed_x_end", Lx))
        dy_kw = kw.pop("dy", None)
        dy_d = deck("dy")
        dy = float(dy_d if dy_d is not None else (dy_kw if dy_kw is not None else dx))
        n_z_bed = int(deck_first("n_z_bed", n_z_bed))
        for key, want, default in _REQUIRED:
            got = deck_first(key, default)
            got = bool(got) if type(want) is bool else str(got)
            if got != want:
                raise ValueError(f"{key}={got!r}: only {want!r} is implemented on the "
                                 f"GPU step")
        if str(deck_first("turbulence_model", "laminar")) == "smagorinsky":
            raise ValueError("turbulence_model='smagorinsky' is not implemented on the "
                             "GPU step (every other value runs k-epsilon, "
                             "spread_3d.py:2394-2433)")
        # every projection_method solves the same linear system; the GPU solve
        # is PCG to projection_cg_rtol
        deck_first("projection_method", "pardiso")
        if float(deck_first("nu_t_multiplier", 1.0)) != 1.0:
            raise ValueError("nu_t_multiplier != 1 is not implemented on the GPU step")
        for key in _OUTDOOR_OFF:
            if bool(getattr(ocfg, key, False)):
                raise ValueError(f"outdoor.{key}=true is not implemented on the GPU step")
        dmode = str(deck_first("lagrangian_bed_drying_mode", "arrhenius"))
        dmap = {"arrhenius": 0, "equilibrium": 1, "combined": 2}
        if dmode not in dmap:
            raise ValueError(f"lagrangian_bed_drying_mode must be one of {sorted(dmap)} "
                             f"(got {dmode!r})")
        dt_key = str(deck_first("dtype", "float32"))
        if dt_key not in ("float32", "float64"):
            raise ValueError("dtype must be float32 or float64")
        moisture = float(getattr(ocfg, "initial_moisture_frac", 0.0) or 0.0)
        mmin, mmax = deck("moisture_min"), deck("moisture_max")
        if mmin is not None and mmax is not None:
            seed = int(deck("moisture_seed") or 0)
            moisture = float(np.random.default_rng(seed).uniform(float(mmin), float(mmax)))
        n_steps = deck_first("n_steps", None)
        quad = int(deck_first("dom_quadrature", 4))
        if quad not in (4, 8):
            raise ValueError("dom_quadrature must be 4 or 8")
        sigma_sav = float(getattr(ocfg, "sav_ratio_1_m", 6562.0))
        sav_bed = kw.pop("lagrangian_bed_sav_1_m", None)
        if sav_bed is not None and float(sav_bed) != sigma_sav:
            raise ValueError("lagrangian_bed_sav_1_m different from the deck SAV is not "
                             "implemented on the GPU step")
        cg = kw.pop("projection_cg_rtol", None)
        if cg is None:
            cg = deck("projection_cg_rtol")
        y_bc = kw.pop("y_bc", "periodic")
        if y_bc != "periodic":
            raise ValueError("only y_bc='periodic' is implemented on the GPU step")
        # level-set mode and the empirical-ROS blend (spread_3d.py:1075-1090,
        # :1246-1272); the ROS and its weight are fixed for the run
        passive = bool(deck_first("level_set_passive", False))
        lagrangian = bool(deck_first("lagrangian_bed_enable", False))
        # run_3d_spread keywords of the Eulerian bed (spread_3d.py:703, :712-713,
        # :1845-1858)
        h_conv_mult = float(kw.pop("h_conv_mult", 1.0))
        a_ov = kw.pop("pyrolysis_A_p_override", None)
        e_ov = kw.pop("pyrolysis_E_p_override", None)
        if (a_ov is None) != (e_ov is None):
            raise ValueError("pyrolysis_A_p_override and pyrolysis_E_p_override must be "
                             "set together")
        pyro_A_p = float(a_ov) if a_ov is not None else 36280.0
        pyro_E_p = float(e_ov) if e_ov is not None else 58600.0
        emp_on = bool(deck_first("empirical_ros_enable", False))
        emp_ros, emp_w = 0.0, 0.0
        if emp_on:
            from . import empirical_ros as _er
            model = str(deck_first("empirical_ros_model", "cheney_eq6"))
            a_ch = float(deck_first("empirical_ros_a_ch", 0.406))
            u_thr = float(deck_first("empirical_ros_u_threshold_m_s", 1.4))
            width = float(deck_first("empirical_ros_blend_width_m_s", 0.5))
            age = float(deck_first("empirical_ros_age_yr", 10.0))
            emp_ros = _er.evaluate_empirical_ros(model, float(wind_speed_m_s), moisture, a_ch,
                                                 age_yr=age)
            emp_w = _er.blend_resolved_empirical(float(wind_speed_m_s), u_thr, width)
        out = cls(
            wind_speed_m_s=float(wind_speed_m_s), Lx=Lx, Ly=Ly, Lz=Lz, dx=dx, dy=dy,
            n_z_bed=n_z_bed, mesh_kwargs=mesh, bed_x_start=bed_x_start,
            bed_x_end=bed_x_end, cfl_factor=cfl, proj_max_iter=proj_max_iter,
            proj_div_tol=proj_div_tol, h_bed=float(ocfg.fuel_depth_m),
            rho_b=float(ocfg.bulk_density_kg_m3), T_amb=float(ocfg.ambient_T_K),
            sigma_sav=sigma_sav, moisture=moisture,
            ignition_duration_s=float(ocfg.ignition_duration_s), terrain=str(ocfg.terrain),
            canopy_C_d=float(ocfg.canopy_C_d), canopy_beta_p=float(ocfg.canopy_beta_p),
            canopy_beta_d=float(ocfg.canopy_beta_d),
            dom_subcycle_every=int(deck_first("dom_subcycle_every", 1)),
            dom_quadrature=quad,
            n_per_cell=int(deck_first("lagrangian_bed_N_per_cell", 50)),
            drying_mode=dmap[dmode],
            h_conv=float(deck_first("lagrangian_bed_h_conv", 25.0)),
            rho_solid_true=float(kw.pop("lagrangian_bed_rho_solid_true", 380.0)),
            cp_solid=float(kw.pop("lagrangian_bed_cp_solid", 1500.0)),
            eps_solid=float(kw.pop("lagrangian_bed_eps_solid", 0.9)),
            view_factor=float(kw.pop("lagrangian_bed_view_factor", 1.0)),
            view_factor_geometric=bool(kw.pop("lagrangian_bed_view_factor_geometric", True)),
            char_ox_flux_cap=float(deck_first("char_ox_flux_cap_W_m2", 1.0e5)),
            char_ox_ash_exp=float(deck_first("char_ox_ash_exp", 0.0)),
            solid_phase_ignition=bool(deck_first("solid_phase_ignition_enable", False)),
            solid_phase_T_s_K=float(deck_first("solid_phase_ignition_T_s_K", 800.0)),
            ignition_q_mult=float(deck_first("ignition_q_mult", 1.0)),
            ignition_width_mult=float(deck_first("ignition_width_mult", 1.0)),
            min_dt_s=float(deck_first("min_dt_s", 1.0e-4)),
            max_wall_time_s=float(deck_first("max_wall_time_s", 60.0)),
            cg_rtol=float(cg) if cg is not None else 1.0e-6,
            n_steps=int(n_steps) if n_steps is not None else None,
            dtype=torch.float64 if dt_key == "float64" else torch.float32,
            particle_jitter=float(deck_first("particle_jitter", 0.0)),
            particle_jitter_seed=int(deck_first("particle_jitter_seed", 0)),
            spin_up_s=float(ocfg.spin_up_s),
            level_set_passive=passive, lagrangian_bed=lagrangian, h_conv_mult=h_conv_mult,
            pyro_A_p=pyro_A_p, pyro_E_p=pyro_E_p, empirical_ros_enable=emp_on,
            empirical_ros_m_s=float(emp_ros), empirical_blend_w=float(emp_w),
            fuel_depth_cv=float(deck_first("fuel_depth_cv", 0.0)),
            fuel_depth_corr_m=float(deck_first("fuel_depth_corr_length_m", 1.0)),
            fuel_depth_seed=int(deck_first("fuel_depth_seed", 0)),
            bed_zone_m=float(deck_first("fuel_depth_max_m", 0.0)),
            gust_amplitude=float(deck_first("wind_gust_amplitude", 0.0)),
            gust_period_s=float(deck_first("wind_gust_period_s", 5.0)),
            gust_seed=int(deck_first("wind_gust_seed", 0)),
        )
        if out.fuel_depth_cv > 0.0 and out.bed_zone_m < out.h_bed:
            raise ValueError("fuel_depth_cv > 0 needs fuel_depth_max_m >= fuel_depth_m "
                             "(the mesh height of the bed layers; deeper columns are clipped)")
        if kw:
            raise TypeError(f"unsupported keyword arguments: {sorted(kw)}")
        return out


def build_grid(cfg: Spread3DConfig) -> Grid3D:
    """The configuration's mesh; with a fuel depth field its bed layers span
    fuel_depth_max_m instead of the mean depth."""
    zone = cfg.bed_zone_m if cfg.bed_zone_m > 0.0 else cfg.h_bed
    return Grid3D.build(Lx=cfg.Lx, Ly=cfg.Ly, Lz=cfg.Lz, dx=cfg.dx, dy=cfg.dy, h_bed=zone,
                        n_z_bed=cfg.n_z_bed, **cfg.mesh_kwargs)


def bed_inputs(cfg: Spread3DConfig, g: Grid3D):
    """(fill, m_scale, h_top) of the configuration's seeded bed variation
    (bed_variation.bed_variation; all None for the reference's uniform
    bed): the lognormal fuel depth field clipped to the bed layers' height,
    and the per-particle load and depth jitter."""
    i0 = int(np.searchsorted(g.x_mid, cfg.bed_x_start, side="left"))
    i1 = int(np.searchsorted(g.x_mid, cfg.bed_x_end, side="right"))
    mask = np.zeros((g.Nz, g.Ny, g.Nx), bool)
    mask[:g.n_z_bed, :, i0:i1] = True
    depth = None
    if cfg.fuel_depth_cv > 0.0:
        depth = fuel_depth_field(g.Nx, g.Ny, g.dx, g.dy, cfg.h_bed, cfg.fuel_depth_cv,
                                 cfg.fuel_depth_corr_m, cfg.fuel_depth_seed)
        depth = np.minimum(depth, float(g.z_face[g.n_z_bed]))
    jit = cfg.particle_jitter if cfg.lagrangian_bed else 0.0
    return bed_variation(mask, g.z_face, g.n_z_bed, cfg.n_per_cell, cfg.h_bed, depth, jit,
                         cfg.particle_jitter_seed)


@dataclass
class CellState3D:
    """The state arrays of the result (spread_3d.py:88-158), float64
    (Nz, Ny, Nx).  On this configuration p_dyn and Y_CO stay zero (no step
    writes them) and the Eulerian solid pools m_hemi, m_cell, m_lign, m_char
    are zero (the particles own the bed, spread_3d.py:1318-1327)."""
    rho: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    p_dyn: np.ndarray
    T_g: np.ndarray
    Y_fuel: np.ndarray
    Y_O2: np.ndarray
    Y_H2O: np.ndarray
    Y_CO: np.ndarray
    alpha_s: np.ndarray
    T_s: np.ndarray
    m_hemi: np.ndarray
    m_cell: np.ndarray
    m_lign: np.ndarray
    m_char: np.ndarray


_STATE_KEYS = ("rho", "u", "v", "w", "p_dyn", "T_g", "Y_fuel", "Y_O2", "Y_H2O", "Y_CO",
               "alpha_s", "T_s", "m_hemi", "m_cell", "m_lign", "m_char")
_GRID_KEYS = ("x_mid", "y_mid", "z_mid", "dz_arr", "inv_dz_arr", "z_face", "d_face_above",
              "d_face_below")
_DIAG_KEYS = ("diag_t", "diag_Tg_max", "diag_Ts_max", "diag_Sp_max", "diag_Qc_max",
              "diag_omega_max", "diag_Y_max", "diag_n_ign")


def _npz_path(path):
    path = Path(path)
    return path if str(path).endswith(".npz") else path.with_suffix(".npz")


@dataclass
class Spread3DResult:
    """The Spread3DResult of run_3d_spread (spread_3d.py:428-559)."""
    ros_m_s: float
    n_cells_ignited: int
    front_t: list
    front_x: list
    grid: Grid3D
    state_final: CellState3D
    diag_t: list = field(default_factory=list)
    diag_Tg_max: list = field(default_factory=list)
    diag_Ts_max: list = field(default_factory=list)
    diag_Sp_max: list = field(default_factory=list)
    diag_Qc_max: list = field(default_factory=list)
    diag_omega_max: list = field(default_factory=list)
    diag_Y_max: list = field(default_factory=list)
    diag_n_ign: list = field(default_factory=list)
    diag_proj_div_max: list = field(default_factory=list)
    diag_proj_n_iter: list = field(default_factory=list)
    k_turb_final: Optional[np.ndarray] = None
    eps_turb_final: Optional[np.ndarray] = None
    nu_t_final: Optional[np.ndarray] = None
    n_steps: int = 0
    stop_reason: Optional[str] = None

    def save(self, path, metadata: Optional[dict] = None) -> None:
        """One compressed .npz (``.npz`` appended when missing) with the keys
        of Spread3DResult.save (spread_3d.py:460-511): ``_meta_json`` (the
        scalars and ``user`` = metadata), the grid axes, the 16 state arrays,
        the front history and the eight per-step diagnostics."""
        import json
        path = _npz_path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        g, s = self.grid, self.state_final
        meta = {"ros_m_s": float(self.ros_m_s), "n_cells_ignited": int(self.n_cells_ignited),
                "Nx": int(g.Nx), "Ny": int(g.Ny), "Nz": int(g.Nz), "dx": float(g.dx),
                "dy": float(g.dy), "dz": float(g.dz), "Lx": float(g.Lx), "Ly": float(g.Ly),
                "Lz": float(g.Lz), "n_z_bed": int(g.n_z_bed), "user": metadata or {}}
        arrays = {"_meta_json": np.array(json.dumps(meta))}
        arrays.update({k: getattr(g, k) for k in _GRID_KEYS})
        arrays.update({k: getattr(s, k) for k in _STATE_KEYS})
        arrays["front_t"] = np.array(self.front_t)
        arrays["front_x"] = np.array(self.front_x)
        arrays.update({k: np.array(getattr(self, k)) for k in _DIAG_KEYS})
        np.savez_compressed(path, **arrays)

    @classmethod
    def load(cls, path) -> "Spread3DResult":
        """Read a file of ``save`` (spread_3d.py:513-559), with no pickles;
        Y_H2O, Y_CO and m_char default to zeros when absent."""
        import json
        z = np.load(_npz_path(path), allow_pickle=False)
        meta = json.loads(str(z["_meta_json"]))
        grid = Grid3D(Nx=int(meta["Nx"]), Ny=int(meta["Ny"]), Nz=int(meta["Nz"]),
                      dx=float(meta["dx"]), dy=float(meta["dy"]), dz=float(meta["dz"]),
                      Lx=float(meta["Lx"]), Ly=float(meta["Ly"]), Lz=float(meta["Lz"]),
                      n_z_bed=int(meta["n_z_bed"]), **{k: z[k] for k in _GRID_KEYS})
        st = {}
        for k in _STATE_KEYS:
            if k in z.files:
                st[k] = z[k]
            elif k in ("Y_H2O", "Y_CO"):
                st[k] = np.zeros_like(z["Y_O2"])
            elif k == "m_char":
                st[k] = np.zeros_like(z["m_hemi"])
            else:
                raise KeyError(f"{path}: missing state array {k!r}")
        return cls(ros_m_s=float(meta["ros_m_s"]),
                   n_cells_ignited=int(meta["n_cells_ignited"]),
                   front_t=z["front_t"].tolist(), front_x=z["front_x"].tolist(),
                   grid=grid, state_final=CellState3D(**st),
                   **{k: z[k].tolist() for k in _DIAG_KEYS})


# ---------------------------------------------------------------------------
# solver
# ---------------------------------------------------------------------------

class Spread3DSolver:
    """Device-resident state and one outer step of the reference loop."""

    FIELDS = ("rho", "rho_prev", "u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2",
              "Y_H2O", "alpha_s", "Fx", "Fy", "Fz", "k", "eps", "nu_t", "omega",
              "S_pyro", "div_target", "Q_comb", "q_rad", "q_rad_gas", "Q_ext", "Sp",
              "Sd", "Qp", "Qd", "YF", "Qch", "Qsm", "Qgc", "M_local", "T_s_old",
              "tmp1", "tmp2", "tmp3")

    def __init__(self, cfg: Spread3DConfig, grid: Optional[Grid3D] = None,
                 graphs: Optional[bool] = None):
        """graphs: replay the step's sub-step loop from captured hipGraphs
        (default: off unless FIREWAVE_GRAPHS=1); bit-identical to direct
        launches, and on MI355X no faster (256 x 64 x 48: 1.972 against
        1.970 ms per step; the 48-case sweep: 10.9 against 10.0 s, the
        recordings of every case's launch sequences costing more than the
        dispatches saved).  With graphs on and the caller on the default
        stream, the steps run on a stream of the solver's own, ordered
        against the caller's stream on entry and exit."""
        self.cfg = cfg
        self.proj = None
        self._graphs = None
        self.dev = device()
        self.dtype = cfg.dtype
        g = grid or build_grid(cfg)
        if g.n_z_bed <= 0:
            raise ValueError("the GPU step needs a fuel bed (n_z_bed > 0)")
        self.grid = g
        self.dg = DeviceGrid(g.Nx, g.Ny, g.Nz, g.dx, g.dy, g.dz_arr, g.d_face_above,
                             g.d_face_below, z_face=g.z_face, n_z_bed=g.n_z_bed,
                             dtype=self.dtype)
        self.dg64 = (self.dg if self.dtype == torch.float64
                     else self.dg.with_dtype(torch.float64))
        self.shape = (g.Nz, g.Ny, g.Nx)
        self.N = g.Nx * g.Ny * g.Nz
        self.f = {n: torch.zeros(self.shape, dtype=self.dtype, device=self.dev)
                  for n in self.FIELDS}
        self._setup(graphs)

    # -- setup (spread_3d.py:1126-1896) --------------------------------------
    def _setup(self, graphs):
        cfg, g, f, dev = self.cfg, self.grid, self.f, self.dev
        T_amb = cfg.T_amb
        f64 = dict(dtype=torch.float64, device=dev)
        f["rho"].fill_(_P0 / (_R_AIR * T_amb))
        f["T_g"].fill_(T_amb)
        f["T_s"].fill_(T_amb)
        f["Y_O2"].fill_(0.232)
        i0 = int(np.searchsorted(g.x_mid, cfg.bed_x_start, side="left"))
        i1 = int(np.searchsorted(g.x_mid, cfg.bed_x_end, side="right"))
        if i0 >= i1:
            raise ValueError(f"bed_x_start={cfg.bed_x_start} >= bed_x_end={cfg.bed_x_end} "
                             f"after grid snapping (i_bed_start={i0}, i_bed_end={i1})")
        self.i_bed = (i0, i1)
        alpha = np.zeros(self.shape)
        alpha[:g.n_z_bed, :, i0:i1] = cfg.rho_b / _RHO_PARTICLE
        self.bed_fill, self.m_scale, self.h_top = bed_inputs(cfg, g)
        if self.bed_fill is not None:
            alpha *= self.bed_fill
        f["alpha_s"].copy_(torch.as_tensor(alpha))
        n_src = max(1, int(round(0.5 * cfg.ignition_width_mult / g.dx)))
        s0, s1 = i0, min(i1, i0 + n_src)
        self.i_src = (s0, s1)
        self.counters = torch.zeros(2, dtype=torch.int32, device=dev)
        self.dT_work = torch.zeros(self.N, **f64)
        self.src_arr = (C.c_void_p * 8)(*[f[k].data_ptr() for k in
                                          ("Sp", "Sd", "Qp", "Qd", "YF", "Qch", "Qsm", "Qgc")])
        if cfg.lagrangian_bed:
            self._setup_particles(alpha, i0, i1)
        else:
            self._setup_eulerian_bed(i0, i1)
        # inflow, sponge, turbulence (:1391-1580)
        self.U_mf = midflame_wind_speed(cfg.wind_speed_m_s, cfg.terrain)
        u_in = build_u_inlet(g, cfg.wind_speed_m_s)
        self.u_inlet = torch.as_tensor(u_in, dtype=self.dtype, device=dev)
        self.v_inlet = torch.zeros_like(self.u_inlet)
        self.w_inlet = torch.zeros_like(self.u_inlet)
        f["u"].copy_(self.u_inlet[:, :, None].expand(self.shape))
        self.gust = GustSchedule(cfg.gust_amplitude, cfg.gust_period_s, cfg.gust_seed)
        self.u_inlet_base = torch.as_tensor(u_in, dtype=torch.float64, device=dev)
        Ns = max(3, min(int(round(0.5 / g.dx)), g.Nx // 4))
        sig = np.zeros(g.Nx)
        for i in range(g.Nx - Ns, g.Nx):
            fr = (i - (g.Nx - Ns)) / Ns
            sig[i] = 5.0 * fr * fr
        self.sigma_x = torch.as_tensor(sig, dtype=self.dtype, device=dev)
        k_init = max(1.5 * (0.10 * max(self.U_mf, 0.5)) ** 2, 1.0e-4)
        eps_init = _turb.C_MU * k_init * k_init / 0.01
        f["k"].fill_(k_init)
        f["eps"].fill_(eps_init)
        f["nu_t"].fill_(0.01)
        self.k_wall = torch.full((g.Ny, g.Nx), _turb.K_MIN, dtype=self.dtype, device=dev)
        self.e_wall = torch.full((g.Ny, g.Nx), _turb.EPS_MIN, dtype=self.dtype, device=dev)
        # soil (:1545-1551)
        sdz, sda, sdb, _ = build_soil_grid(6, 0.001, 1.5)
        self.soil_dz = torch.as_tensor(sdz, **f64)
        self.soil_da = torch.as_tensor(sda, **f64)
        self.soil_db = torch.as_tensor(sdb, **f64)
        self.T_soil = torch.full((6, g.Ny, g.Nx), T_amb, dtype=self.dtype, device=dev)
        self.q_in_soil = torch.zeros((g.Ny, g.Nx), dtype=self.dtype, device=dev)
        # level set, float64 (flame_front_3d.py:304-330)
        x_front = g.x_mid[s1 - 1] + 0.5 * g.dx
        d = g.x_mid - x_front
        row = np.where(d > 0.0, d, np.maximum(d, -L_BURNOUT_M / 2.0))
        phi = np.broadcast_to(row[None, None, :], self.shape).copy()
        phi[:g.n_z_bed, :, s0:s1] = -L_BURNOUT_M / 2.0
        self.phi = torch.as_tensor(phi, **f64)
        self.phi_tmp = torch.empty_like(self.phi)
        self.phi0 = torch.empty_like(self.phi)
        self.lset_step = 0
        # projection and DOM handles (:1719-1780)
        self.proj = _proj_handle(self.dg, 1.0e-6, cfg.cg_rtol)
        self.dom = _dom.DOMRadiationSolver(g.Nz, g.Ny, g.Nx, g.dy, g.dx, g.dz_arr,
                                           g.d_face_above, g.d_face_below, "periodic",
                                           cfg.dom_quadrature)
        self.dom_h = self.dom._handle(self.dtype)
        # time step and histories (:1782-1830)
        self.T_ign = T_amb + 300.0
        self.dz_min = float(g.dz_arr.min())
        U_char = max(self.U_mf, math.sqrt(_G * cfg.h_bed * 0.5), 0.1)
        dt_adv = cfg.cfl_factor * min(g.dx, g.dy, self.dz_min) / U_char
        dt_diff = 0.25 * self.dz_min * self.dz_min / max(0.01, 1.0e-3)
        self.red_work = torch.empty(14 * _lib.FW_REDUCE_BLOCKS, **f64)
        self.red_out = torch.zeros(14, **f64)
        self.red_host = torch.zeros(14, dtype=torch.float64).pin_memory()
        self.nu_work = torch.empty(_lib.FW_REDUCE_BLOCKS, **f64)
        self.edc_work = torch.empty(max(_lib.lib().fw_edc_workspace_ints(self.dg.ref), 1),
                                    dtype=torch.int32, device=dev)
        # the native step's views of all of the above
        self.params = FwSpreadParams(
            T_amb, cfg.h_bed, cfg.sigma_sav, cfg.canopy_C_d, cfg.canopy_beta_p,
            cfg.canopy_beta_d, cfg.cfl_factor, cfg.min_dt_s, self.U_mf, self.dz_min,
            self.T_ign, cfg.ignition_duration_s, cfg.spin_up_s, cfg.ignition_q_mult,
            cfg.bed_x_end, g.Lx, g.Lx / g.Nx, cfg.proj_div_tol, cfg.rho_solid_true,
            cfg.cp_solid, K_SOIL_DEFAULT, RHO_SOIL_DEFAULT, CP_SOIL_DEFAULT, EPS_SOIL_DEFAULT,
            6, s0, s1, cfg.proj_max_iter, max(cfg.dom_subcycle_every, 1), 0)
        self.fields = FwSpreadFields()
        fs = self.fields
        for name in _GRID_SLOTS:
            setattr(fs, name, f[name].data_ptr())
        for q, name in enumerate(_SRC_NAMES):
            fs.src[q] = f[name].data_ptr()
        for name, t in (("u_inlet", self.u_inlet), ("v_inlet", self.v_inlet),
                        ("w_inlet", self.w_inlet), ("sigma_x", self.sigma_x),
                        ("k_wall", self.k_wall), ("eps_wall", self.e_wall),
                        ("T_soil", self.T_soil), ("q_in_soil", self.q_in_soil),
                        ("soil_dz", self.soil_dz), ("soil_d_above", self.soil_da),
                        ("soil_d_below", self.soil_db), ("phi", self.phi),
                        ("phi_tmp", self.phi_tmp), ("phi0", self.phi0),
                        ("red_work", self.red_work), ("red_out", self.red_out),
                        ("red_host", self.red_host), ("nu_work", self.nu_work),
                        ("contrib", self.contrib), ("dT_work", self.dT_work),
                        ("counters", self.counters), ("edc_work", self.edc_work)):
            setattr(fs, name, t.data_ptr())
        if graphs is None:
            graphs = os.environ.get("FIREWAVE_GRAPHS", "0") == "1"
        self._stream = None
        if graphs:
            h = C.c_void_p()
            _lib.call("fw_graphs_create", C.byref(h))
            self._graphs = h
            self.step_scalars = torch.zeros(_lib.FW_STEP_SCALARS, **f64)
            fs.graphs = h.value
            fs.step_scalars = self.step_scalars.data_ptr()
            self._stream = torch.cuda.Stream(device=dev)
        # active level set: the band, its forward flux and the front speed,
        # kept across the step for the snapshot writer
        cfg_active = not cfg.level_set_passive
        if cfg_active:
            self.lset_mask = torch.zeros(self.shape, dtype=torch.uint8, device=dev)
            self.q_dom_fwd = torch.zeros(self.shape, dtype=self.dtype, device=dev)
            self.v_n = torch.zeros(self.shape, **f64)
            for name, t in (("lset_mask", self.lset_mask), ("q_dom_fwd", self.q_dom_fwd),
                            ("v_n", self.v_n)):
                setattr(fs, name, t.data_ptr())
        pr = self.params
        pr.band_m = max(g.dx, DX_VN_BAND_M)
        pr.E_sens = cfg.rho_b * _CP_SOLID_VN * cfg.h_bed * (self.T_ign - T_amb)
        pr.E_lat_coef = cfg.rho_b * cfg.h_bed * L_VAP_WATER * 1.0
        pr.emp_w = cfg.empirical_blend_w if cfg.empirical_ros_enable else 0.0
        pr.emp_ros = cfg.empirical_ros_m_s if cfg.empirical_ros_enable else 0.0
        pr.seed_T_s = cfg.solid_phase_T_s_K
        pr.x_lo_bed = float(g.x_mid[i0] - 0.5 * g.dx)
        pr.stall_window = _STALL_WINDOW_S
        pr.lset_active = 1 if cfg_active else 0
        # Eulerian bed: pools, the coupling's latent heat (spread_3d.py:3152:
        # L_v only with moisture) and the MD2004 pair
        pr.h_conv_mult = cfg.h_conv_mult
        if self.eu is not None:
            for name in ("m_solid", "m_char", "m_water"):
                setattr(fs, "eu_" + name, self.eu[name].data_ptr())
            fs.Q_pyro = self.eu["Q_pyro"].data_ptr()
            pr.eulerian = 1
            pr.eu_m_water_init = self.eu_m_water_init
            pr.eu_A_p, pr.eu_E_p = cfg.pyro_A_p, cfg.pyro_E_p
            pr.eu_L_v = 2_257_000.0 if self.eu_m_water_init > 0.0 else 0.0
        self._by_ptr = {t.data_ptr(): t for t in list(f.values()) +
                        [self.phi, self.phi_tmp, self.phi0]}
        self.state = FwSpreadState()
        st = self.state
        st.t = 0.0
        st.dt = min(dt_adv, dt_diff, 0.05)
        st.front_x_last = float(x_front)
        self.front_t = [0.0]
        self.front_x = [float(x_front)]
        s = stream_ptr()
        self._c("fw_spread_front", self.dg64.ref, C.byref(fs), C.byref(st), s)
        self._c("fw_spread_reduce", self.dg.ref, C.byref(fs), C.byref(self.params), 0,
                C.byref(st), s)
        self.diag = {k: [] for k in ("t", "Tg", "Ts", "Sp", "Qc", "om", "Y", "n_ign",
                                     "div", "iters")}
        self._append_diag(0.0, st.red, 0.0, 0, sources=False)
        self.stop_reason = None

    def _setup_particles(self, alpha, i0, i1):
        """Bed particles and solid-phase seeding (spread_3d.py:1296-1330,
        :1486-1526)."""
        cfg, g, dev = self.cfg, self.grid, self.dev
        f64 = dict(dtype=torch.float64, device=dev)
        T_amb = cfg.T_amb
        n_cells = int((alpha[:g.n_z_bed, :, i0:i1] > 0).sum())
        buf = _lb.allocate_bed_particle_buffers(n_cells * cfg.n_per_cell)
        n_alloc = _lb.initialize_bed_particles_from_alpha_s(
            buf, alpha, cfg.rho_b, cfg.moisture, T_amb, g.dx, g.dy, g.dz_arr,
            g.n_z_bed, cfg.n_per_cell, sav=cfg.sigma_sav, i_lo=i0, i_hi=i1,
            mass_scale=self.m_scale)
        if self.h_top is not None:
            buf["h_top"] = np.asarray(self.h_top, np.float64)[:n_alloc].copy()
        if n_alloc <= 0:
            raise ValueError("no bed particles were allocated")
        self.n_particles = n_alloc
        s0, s1 = self.i_src
        if cfg.solid_phase_ignition:
            x_lo = g.x_mid[s0] - 0.5 * g.dx
            x_hi = g.x_mid[s1 - 1] + 0.5 * g.dx
            x = buf["x"][:n_alloc]
            sel = (buf["alive"][:n_alloc] != 0) & (x_lo <= x) & (x < x_hi)
            buf["T_s"][:n_alloc][sel] = cfg.solid_phase_T_s_K
        self.P = _lb.DeviceParticles(buf, n_alloc)
        cs, ca = _lb._locate(self.dg, self.P)
        if not torch.equal(cs, ca):
            raise RuntimeError("particle cell maps of the step and the aggregates differ")
        if int(cs.min()) < 0:
            raise RuntimeError("bed particle outside the grid")
        self.ix = _lb.DeviceIndex(cs)
        self.contrib = torch.empty(10 * n_alloc, **f64)
        self.bed_params = _lb.FwBedParams(
            0.0, cfg.h_conv, cfg.rho_solid_true, cfg.cp_solid, cfg.eps_solid, T_amb,
            cfg.view_factor, cfg.h_bed, cfg.sigma_sav * cfg.rho_b / cfg.rho_solid_true,
            cfg.char_ox_flux_cap, cfg.char_ox_ash_exp, cfg.n_per_cell,
            int(cfg.view_factor_geometric), 1, 1, 1, 1, cfg.drying_mode, 0)
        self.eu = None

    def _setup_eulerian_bed(self, i0, i1):
        """Eulerian pools (spread_3d.py:1186-1195, :1236, :1274-1278): the
        reactive mass rho_b (uniform profile) and the water rho_b M in the bed
        cells, no char; no particles and no solid-phase seeding (:1508)."""
        cfg, g, dev = self.cfg, self.grid, self.dev
        f64 = dict(dtype=torch.float64, device=dev)
        m = np.zeros(self.shape)
        m[:g.n_z_bed, :, i0:i1] = cfg.rho_b * 1.0
        w = np.zeros(self.shape)
        w[:g.n_z_bed, :, i0:i1] = cfg.rho_b * cfg.moisture
        if self.bed_fill is not None:
            m *= self.bed_fill
            w *= self.bed_fill
        self.eu = {"m_solid": torch.as_tensor(m, **f64), "m_char": torch.zeros(self.shape, **f64),
                   "m_water": torch.as_tensor(w, **f64),
                   "Q_pyro": torch.zeros(self.shape, dtype=self.dtype, device=dev)}
        self.eu_m_water_init = cfg.rho_b * cfg.moisture
        self.n_particles = 0
        empty = _lb.allocate_bed_particle_buffers(0)
        self.P = _lb.DeviceParticles(empty, 0)
        self.ix = None
        self.contrib = torch.zeros(1, **f64)
        self.bed_params = None

    # -- helpers ---------------------------------------------------------------
    @staticmethod
    def _c(name, *args):
        _lib.call(name, *args)

    @property
    def t(self):
        return self.state.t

    @property
    def dt(self):
        return self.state.dt

    @property
    def step_no(self):
        return self.state.step_no

    @property
    def last_pcg_iters(self):
        return self.state.pcg_iters

    def _append_diag(self, t, r, div, iters, sources=True):
        d = self.diag
        d["t"].append(t)
        d["Tg"].append(float(r[4]))
        d["Ts"].append(float(r[6]))
        d["Sp"].append(float(r[7]) if sources else 0.0)
        d["Qc"].append(float(r[8]) if sources else 0.0)
        d["om"].append(float(r[9]) if sources else 0.0)
        d["Y"].append(float(r[10]))
        d["n_ign"].append(int(r[11]) if sources else 0)
        d["div"].append(div)
        d["iters"].append(iters)

    def _remap(self):
        """Re-read which buffer holds which field after the step's swaps."""
        fs = self.fields
        for name in _GRID_SLOTS:
            self.f[name] = self._by_ptr[getattr(fs, name)]
        self.phi = self._by_ptr[fs.phi]
        self.phi_tmp = self._by_ptr[fs.phi_tmp]

    # -- one outer step (spread_3d.py:1897-3683) --------------------------------
    def step(self):
        """Advance one outer step in the native driver; returns the stop reason
        (``extinct``, ``front_exit``, ``domain_end``, ``stalled``) or None."""
        st = self.state
        if self.gust.enabled:
            # the inlet of a step starting at t: the log law times g(t)
            self.u_inlet.copy_(self.u_inlet_base * self.gust.factor(st.t))
        rc = self._native("fw_spread_step")
        if rc == _lib.FW_E_NONFINITE:
            raise RuntimeError(_lib.lib().fw_last_error().decode())
        _lib.check(rc, "fw_spread_step")
        return self._after_step(st)

    def advance(self, n):
        """Up to n outer steps in one native call (fw_spread_run): the
        reference's driver loop without a return to Python between steps.
        Diagnostics, front points and exits are those of n calls of step();
        returns the stop reason or None.  With gusts the inlet of each step is
        set from Python, so the steps go one by one."""
        if n <= 0:
            return None
        if self.gust.enabled:
            for _ in range(n):
                r = self.step()
                if r is not None:
                    return r
            return None
        hist = (FwSpreadState * n)()
        done = C.c_int(0)
        rc = self._native("fw_spread_run", C.c_int(n), hist, C.byref(done))
        reason = None
        for i in range(done.value):
            reason = self._after_step(hist[i])
            if reason is not None:
                break
        if rc == _lib.FW_E_NONFINITE:
            raise RuntimeError(_lib.lib().fw_last_error().decode())
        _lib.check(rc, "fw_spread_run")
        return reason

    def _native(self, name, *extra):
        """Call fw_spread_step / fw_spread_run on the step stream, then re-read
        which buffer holds which field."""
        cur = torch.cuda.current_stream()
        run = cur
        if self._stream is not None and cur.cuda_stream == 0:
            # graph capture needs a stream other than the legacy default one
            run = self._stream
            run.wait_stream(cur)
        rc = getattr(_lib.lib(), name)(
            self.dg.ref, self.dg64.ref, C.byref(self.fields), C.byref(self.params),
            C.byref(self.state), self.proj, self.dom_h, C.byref(self.P.c),
            C.byref(self.ix.c) if self.ix is not None else None,
            C.byref(self.bed_params) if self.bed_params is not None else None,
            C.c_void_p(run.cuda_stream), *extra)
        if run is not cur:
            cur.wait_stream(run)
        self._remap()
        return rc

    def _after_step(self, st):
        """The exits, front points and diagnostics of one step's state."""
        if st.stop in (1, 2):
            self.stop_reason = _STOPS[st.stop]
            return self.stop_reason
        if st.front_appended:
            self.front_t.append(st.front_t_new)
            self.front_x.append(st.front_x_new)
        self._append_diag(st.t, st.red, st.proj_div, st.proj_n)
        if st.stop in (3, 4):
            self.stop_reason = _STOPS[st.stop]
            return self.stop_reason
        return None

    # -- whole run -------------------------------------------------------------
    def run(self, n_steps=None, snapshot_dir=None, snapshot_interval_s=1.0):
        """Step until n_steps (or the deck's n_steps, else max_wall_time_s)
        or an exit.  With snapshot_dir, write snap_%04d.npz after each step
        that ends >= snapshot_interval_s after the last one
        (spread_3d.py:3537-3670); the extinction and front exits leave before
        the snapshot, the domain-end exit after it, as in the reference."""
        cfg = self.cfg
        n_target = n_steps if n_steps is not None else cfg.n_steps
        writer = None
        if snapshot_dir is not None:
            from .snapshot import SnapshotWriter, snapshot_path
            Path(snapshot_dir).mkdir(parents=True, exist_ok=True)
            writer = SnapshotWriter(self, cfg.wind_speed_m_s)
            last_t, idx = -1.0e30, 0
        try:
            while writer is None and n_target is not None and self.step_no < n_target:
                # no snapshots: the steps run in native calls of at most 1000
                # steps, back to Python only at an exit
                if self.advance(min(n_target - self.step_no, 1000)) is not None:
                    break
            while writer is not None or n_target is None:
                if n_target is not None:
                    if self.step_no >= n_target:
                        break
                elif self.t >= cfg.max_wall_time_s:
                    break
                r = self.step()
                if r in ("extinct", "front_exit"):
                    break
                if writer is not None and self.t - last_t >= snapshot_interval_s:
                    last_t = self.t
                    writer.submit(snapshot_path(snapshot_dir, idx))
                    idx += 1
                if r is not None:
                    break
        finally:
            if writer is not None:
                writer.close()
        return self

    def state_numpy(self):
        torch.cuda.current_stream().synchronize()
        out = {k: self.f[k].double().cpu().numpy() for k in
               ("rho", "u", "v", "w", "T_g", "Y_fuel", "Y_O2", "Y_H2O", "T_s", "alpha_s",
                "k", "eps", "nu_t", "omega", "q_rad", "q_rad_gas", "S_pyro", "Q_comb")}
        out["phi"] = self.phi.cpu().numpy()
        out["T_soil"] = self.T_soil.double().cpu().numpy()
        for k in ("x", "y", "z", "T_s", "m_solid", "m_water", "m_char", "alive"):
            out["p_" + k] = self.P.t[k].cpu().numpy().copy()
        if self.eu is not None:
            for k in ("m_solid", "m_char", "m_water", "Q_pyro"):
                out["eu_" + k] = self.eu[k].double().cpu().numpy()
        return out

    def ros(self):
        """_compute_steady_ros (spread_3d.py:3796-3824)."""
        g = self.grid
        ft, fx = self.front_t, self.front_x
        if len(ft) < 2:
            return 0.0
        n_src = self.i_src[1] - self.i_src[0]
        if fx[-1] - n_src * g.dx < 0.1:
            return 0.0
        if self.t - ft[-1] > 30.0 and not fx[-1] > g.Lx * 0.7:
            return 0.0
        return (fx[-1] - fx[0]) / (self.t - ft[0]) if self.t > ft[0] else 0.0

    def result(self):
        """Spread3DResult of the current state (spread_3d.py:3708-3729)."""
        torch.cuda.current_stream().synchronize()
        f = self.f

        def h64(t):
            return t.double().cpu().numpy()

        st = {k: h64(f[k]) for k in ("rho", "u", "v", "w", "T_g", "Y_fuel", "Y_O2", "Y_H2O",
                                     "alpha_s", "T_s")}
        zero = np.zeros(self.shape)
        for k in ("p_dyn", "Y_CO", "m_hemi", "m_cell", "m_lign", "m_char"):
            st[k] = zero.copy()
        if self.eu is not None:
            st["m_hemi"] = self.eu["m_solid"].cpu().numpy()
            st["m_char"] = self.eu["m_char"].cpu().numpy()
        d = self.diag
        return Spread3DResult(
            ros_m_s=self.ros(), n_cells_ignited=int(np.sum(st["T_s"] >= self.T_ign)),
            front_t=list(self.front_t), front_x=list(self.front_x), grid=self.grid,
            state_final=CellState3D(**st), diag_t=d["t"], diag_Tg_max=d["Tg"],
            diag_Ts_max=d["Ts"], diag_Sp_max=d["Sp"], diag_Qc_max=d["Qc"],
            diag_omega_max=d["om"], diag_Y_max=d["Y"], diag_n_ign=d["n_ign"],
            diag_proj_div_max=d["div"], diag_proj_n_iter=d["iters"],
            k_turb_final=h64(f["k"]), eps_turb_final=h64(f["eps"]),
            nu_t_final=h64(f["nu_t"]), n_steps=self.step_no, stop_reason=self.stop_reason)

    def graph_stats(self):
        """(replays, captures) of the step's graph cache, or None without one."""
        if self._graphs is None:
            return None
        r, c = C.c_longlong(), C.c_longlong()
        _lib.call("fw_graphs_stats", self._graphs, C.byref(r), C.byref(c))
        return r.value, c.value

    def close(self):
        if self.proj is not None:
            _lib.lib().fw_proj_destroy(self.proj)
            self.proj = None
        if self._graphs is not None:
            if self._stream is not None:
                self._stream.synchronize()
            _lib.lib().fw_graphs_destroy(self._graphs)
            self._graphs = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_3d_spread(ri_or_path, *, wind_speed_m_s=0.0, snapshot_dir=None,
                  snapshot_interval_s=1.0, **kwargs) -> Spread3DResult:
    """Drop-in for run_3d_spread (spread_3d.py:563-3729): deck + keyword
    overrides, run until max_wall_time_s (or the new deck key n_steps),
    optional periodic snapshots (snapshot_dir, snapshot_interval_s), and
    return the result summary."""
    cfg = Spread3DConfig.resolve(ri_or_path, wind_speed_m_s, **kwargs)
    solver = Spread3DSolver(cfg)
    try:
        solver.run(snapshot_dir=snapshot_dir, snapshot_interval_s=snapshot_interval_s)
        return solver.result()
    finally:
        solver.close()
