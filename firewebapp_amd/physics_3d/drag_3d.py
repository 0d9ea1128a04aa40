"""drag_3d drop-in (drag_3d.py:99-144 of the reference)."""
from __future__ import annotations

from ._common import Marshal, call, grid_for, ptr, stream_ptr

C_D_DEFAULT = 0.30
C_D = C_D_DEFAULT
MU_GAS = 3.0e-5
ERGUN_VISC_K = 50.0


def step_drag_force(u, v, w, rho, alpha_s, sigma_sav, Fx_out, Fy_out, Fz_out,
                    C_D):
    """Ergun two-term porous drag on alpha_s > 0 cells (zero elsewhere)."""
    m = Marshal(u)
    nz, ny, nx = u.shape
    g = grid_for(m, (nz, ny, nx), 1.0, 1.0, [1.0] * nz, [1.0] * nz, [1.0] * nz)
    fx, fy, fz = m.out(Fx_out), m.out(Fy_out), m.out(Fz_out)
    call("fw_drag_force", g.ref, ptr(m.dev(u)), ptr(m.dev(v)), ptr(m.dev(w)),
         ptr(m.dev(rho)), ptr(m.dev(alpha_s)), float(sigma_sav), float(C_D),
         ptr(fx), ptr(fy), ptr(fz), stream_ptr())
    m.writeback()
