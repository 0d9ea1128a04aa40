"""muscl_3d drop-in (muscl_3d.py:67-171 of the reference)."""
from __future__ import annotations

import numpy as np

from ._common import Marshal, call, dz_from_faces, grid_for, ptr, stream_ptr


def advect_3d_scalar_muscl(phi, u, v, w, dx, dy, d_face_above, d_face_below,
                           rhs, phi_inlet, phi_inlet_zmin=None,
                           z_min_inlet_active=False):
    """rhs -= u.grad(phi) (minmod MUSCL), in place.  muscl_3d.py:67-171."""
    if z_min_inlet_active:
        raise ValueError("z-min inlet (cup burner) is not part of the outdoor hot path")
    m = Marshal(phi)
    g = grid_for(m, tuple(phi.shape), dx, dy, dz_from_faces(d_face_above, d_face_below),
                 d_face_above, d_face_below)
    r = m.out(rhs)
    call("fw_advect_scalar_muscl", g.ref, ptr(m.dev(phi)), ptr(m.dev(u)),
         ptr(m.dev(v)), ptr(m.dev(w)), ptr(r), float(phi_inlet), stream_ptr())
    m.writeback()
