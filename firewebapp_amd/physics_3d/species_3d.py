"""species_3d drop-in (species_3d.py:40-175 of the reference)."""
from __future__ import annotations

import torch

from ._common import Marshal, call, grid_for, ptr, stream_ptr

D_LAMINAR = 1.0e-4


def step_species_transport(Y, rho, u, v, w, S_per_volume, dt, dx, dy, dz_arr,
                           d_face_above, d_face_below, D, Y_inlet,
                           Y_inlet_zmin=None, z_min_inlet_active=False):
    """MUSCL advection + central diffusion + S/rho, clip [0, 1]; in place."""
    if z_min_inlet_active:
        raise ValueError("z-min inlet (cup burner) is not part of the outdoor hot path")
    m = Marshal(Y)
    g = grid_for(m, tuple(Y.shape), dx, dy, dz_arr, d_face_above, d_face_below)
    y = m.out(Y)
    y_new = torch.empty_like(y)
    call("fw_species_transport", g.ref, ptr(y), ptr(y_new), ptr(m.dev(rho)),
         ptr(m.dev(u)), ptr(m.dev(v)), ptr(m.dev(w)), ptr(m.dev(S_per_volume)),
         float(dt), float(D), float(Y_inlet), 1.0, stream_ptr())
    y.copy_(y_new)
    m.writeback()
