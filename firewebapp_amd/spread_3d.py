"""GPU time-step driver of the Tier-3 spread (spread_3d.py of the reference).

Helpers with reference names live here too (``_adv_gas_energy``).
"""
from __future__ import annotations

import torch

from .physics_3d._common import Marshal, call, grid_for, ptr, stream_ptr


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
    call("fw_gas_energy", g.ref, ptr(Tg), ptr(To), ptr(m.dev(u)), ptr(m.dev(v)),
         ptr(m.dev(w)), float(dt), float(alpha_th), float(T_amb), stream_ptr())
    Tg.copy_(To)
    m.writeback()
