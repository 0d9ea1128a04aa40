"""solid_conduction_3d drop-in (solid_conduction_3d.py:56-127 of the reference)."""
from __future__ import annotations

import torch

from ._common import Marshal, call, grid_for, ptr, stream_ptr

K_SOLID_GRASS = 0.20


def step_solid_conduction_vertical(T_s, alpha_s, dz_arr, d_face_above,
                                   d_face_below, k_solid, rho_solid, cp_solid,
                                   dt):
    """Harmonic-face fin conduction in k on alpha_s > 0 cells; in place."""
    m = Marshal(T_s)
    g = grid_for(m, tuple(T_s.shape), 1.0, 1.0, dz_arr, d_face_above, d_face_below)
    Ts = m.out(T_s)
    To = torch.empty_like(Ts)
    call("fw_solid_conduction_vertical", g.ref, ptr(Ts), ptr(To),
         ptr(m.dev(alpha_s)), float(k_solid), float(rho_solid), float(cp_solid),
         float(dt), 1.0, 0.0, stream_ptr())
    Ts.copy_(To)
    m.writeback()
