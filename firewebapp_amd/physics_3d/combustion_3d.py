"""combustion_3d drop-in: step_o2_supply_rate (combustion_3d.py:185-260 of the
reference), the only function of the module the outer step calls
(spread_3d.py:2722-2726).

The step fills omega_O2 with 1e30 and calls this limiter every sub-step, but
the EDC closure (chemistry_closures/edc.py:133-376) never reads omega_O2, so
the native step driver skips the call; the function is here for callers of
the module surface.
"""
from __future__ import annotations

import numpy as np

from ..grid import Grid3D
from ._common import Marshal, call, grid_for, ptr, stream_ptr
from .chemistry_closures._constants import S_STOICH


def step_o2_supply_rate(rho, u, v, w, Y_O2, dx, dy, dz_arr, omega_O2_out,
                        s_stoich=S_STOICH):
    """omega_O2_out = upwind O2 mass inflow over the six faces / s_stoich on
    interior cells (1 <= k <= Nz-2, 1 <= j <= Ny-2, 1 <= i <= Nx-2); the
    other cells keep the caller's values.  In place on omega_O2_out."""
    m = Marshal(rho)
    sh = tuple(rho.shape)
    dz = np.asarray(dz_arr, np.float64)
    da, db = Grid3D._faces(dz)
    g = grid_for(m, sh, dx, dy, dz, da, db)
    out = m.out(omega_O2_out, shape=sh)
    call("fw_o2_supply_rate", g.ref, ptr(m.dev(rho, shape=sh)), ptr(m.dev(u, shape=sh)),
         ptr(m.dev(v, shape=sh)), ptr(m.dev(w, shape=sh)), ptr(m.dev(Y_O2, shape=sh)),
         ptr(out), float(s_stoich), stream_ptr())
    m.writeback()
