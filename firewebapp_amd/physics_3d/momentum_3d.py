"""momentum_3d drop-in (momentum_3d.py:41-317 of the reference)."""
from __future__ import annotations

import torch

from ._common import Marshal, call, grid_for, ptr, stream_ptr


def step_tentative_velocity(u, v, w, rho, T_g, Fx_ext, Fy_ext, Fz_ext, dt, dx,
                            dy, dz_arr, d_face_above, d_face_below, T_amb,
                            u_inlet, v_inlet, w_inlet):
    """MUSCL advection + viscosity + Boussinesq buoyancy + body force; in place."""
    m = Marshal(u)
    g = grid_for(m, tuple(u.shape), dx, dy, dz_arr, d_face_above, d_face_below)
    U, V, W = m.out(u), m.out(v), m.out(w)
    Uo, Vo, Wo = torch.empty_like(U), torch.empty_like(V), torch.empty_like(W)
    call("fw_tentative_velocity", g.ref, ptr(U), ptr(V), ptr(W), ptr(Uo),
         ptr(Vo), ptr(Wo), ptr(m.dev(rho)), ptr(m.dev(T_g)), ptr(m.dev(Fx_ext)),
         ptr(m.dev(Fy_ext)), ptr(m.dev(Fz_ext)), ptr(m.dev(u_inlet)),
         ptr(m.dev(v_inlet)), ptr(m.dev(w_inlet)), float(dt), float(T_amb),
         stream_ptr())
    U.copy_(Uo)
    V.copy_(Vo)
    W.copy_(Wo)
    m.writeback()


def apply_outflow_sponge(u, u_target_2d, sigma_x, Y_F, Y_F_skip, dt):
    """u += sigma dt (u_target - u) in the sponge where Y_F < Y_F_skip."""
    m = Marshal(u)
    nz, ny, nx = u.shape
    g = grid_for(m, (nz, ny, nx), 1.0, 1.0, [1.0] * nz, [1.0] * nz, [1.0] * nz)
    U = m.out(u)
    call("fw_outflow_sponge", g.ref, ptr(U), ptr(m.dev(u_target_2d)),
         ptr(m.dev(sigma_x)), ptr(m.dev(Y_F)), float(Y_F_skip), float(dt),
         stream_ptr())
    m.writeback()
