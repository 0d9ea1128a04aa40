"""turbulence_3d drop-in (turbulence_3d.py:481-949 of the reference).

Production subset: realizable k-epsilon and the turbulent-diffusion helper.
Smagorinsky and the wall function are not on the hot path.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .. import _lib
from ..device import device
from ._common import Marshal, call, grid_for, ptr, stream_ptr

C_MU = 0.09
C_1EPS = 1.44
C_2EPS = 1.92
SIGMA_K = 1.0
SIGMA_EPS = 1.3
PR_T = 0.85
SC_T = 0.7
K_MIN = 1.0e-8
EPS_MIN = 1.0e-8
BETA_P_CANOPY_DEFAULT = 1.0
BETA_D_CANOPY_DEFAULT = 4.0


def step_k_epsilon(k_turb, eps_turb, nu_t_out, u, v, w, T_g, rho, alpha_s,
                   sigma_sav, dt, dx, dy, dz_arr, d_face_above, d_face_below,
                   T_amb, S_mag2_work, Omega_mag2_work, u_inlet, k_wall_ghost,
                   eps_wall_ghost, beta_p_canopy, beta_d_canopy,
                   bvg_factor=0.0, eps_realiz_L_min_m=0.0,
                   eps_realiz_durbin_alpha=0.0, eps_cap_count_out=None,
                   eps_durbin_count_out=None):
    """Realizable k-eps step, in place on k_turb/eps_turb, writes nu_t_out.

    The strain/vorticity work arrays are not written (the kernel computes
    them in registers); the BVG term and the L-min / Durbin caps are not on
    the production path and are rejected when enabled.
    """
    if bvg_factor != 0.0 or eps_realiz_L_min_m > 0.0 or eps_realiz_durbin_alpha > 0.0:
        raise ValueError("BVG / L-min / Durbin options are not on the hot path")
    m = Marshal(k_turb)
    g = grid_for(m, tuple(k_turb.shape), dx, dy, dz_arr, d_face_above, d_face_below)
    K, E, NU = m.out(k_turb), m.out(eps_turb), m.out(nu_t_out)
    Ko, Eo = torch.empty_like(K), torch.empty_like(E)
    call("fw_k_epsilon", g.ref, ptr(K), ptr(E), ptr(Ko), ptr(Eo), ptr(NU),
         ptr(m.dev(u)), ptr(m.dev(v)), ptr(m.dev(w)), ptr(m.dev(T_g)),
         ptr(m.dev(alpha_s)), ptr(m.dev(u_inlet)), ptr(m.dev(k_wall_ghost)),
         ptr(m.dev(eps_wall_ghost)), float(sigma_sav), float(dt), float(T_amb),
         float(beta_p_canopy), float(beta_d_canopy), stream_ptr())
    K.copy_(Ko)
    E.copy_(Eo)
    m.writeback()


def turbulent_diffusion_substeps(nu_t_max, sc_t, dt, dx, dy, dz_min):
    """(n_sub, dt_sub) of apply_turbulent_diffusion; n_sub = 0 means skip.

    turbulence_3d.py:879-918: stability-limited explicit sub-stepping with a
    Fourier target of 0.4, skipped for non-finite, negligible or too stiff nu_t.
    """
    h_min2 = min(dx * dx, min(dy * dy, dz_min * dz_min))
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
    """Explicit turbulent diffusion with zero-flux ghosts; in place."""
    m = Marshal(field)
    g = grid_for(m, tuple(field.shape), dx, dy, dz_arr, d_face_above, d_face_below)
    F = m.out(field)
    NU = m.dev(nu_t)
    work = torch.empty(_lib.FW_REDUCE_BLOCKS, dtype=torch.float64, device=device())
    import ctypes as C
    mx = C.c_double(0.0)
    call("fw_nu_t_max_interior", g.ref, ptr(NU), ptr(work), C.byref(mx), stream_ptr())
    n_sub, dt_sub = turbulent_diffusion_substeps(
        mx.value, sc_t, dt, dx, dy, float(np.min(np.asarray(dz_arr, np.float64))))
    if n_sub == 0:
        m.writeback()
        return
    a, b = F, torch.empty_like(F)
    for _ in range(n_sub):
        call("fw_turbulent_diffusion_pass", g.ref, ptr(a), ptr(b), ptr(NU),
             float(sc_t), float(dt_sub), 1.0, 0.0, stream_ptr())
        a, b = b, a
    if a is not F:
        F.copy_(a)
    m.writeback()
