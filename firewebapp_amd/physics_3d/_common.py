"""Shared helpers of the drop-in modules."""
from __future__ import annotations

import numpy as np

from .. import _lib
from ..device import Marshal, grid_for, ptr, stream_ptr  # noqa: F401


def dz_from_faces(d_face_above, d_face_below):
    """Cell heights recovered from the face distances, for reference
    signatures that pass only d_face_above / d_face_below.

    The grid's faces are d_face_above[k] = (dz[k] + dz[k+1]) / 2 and the top
    half cell d_face_above[Nz-1] = dz[Nz-1] / 2 (spread_3d.py:281-288), so
    dz[Nz-1] = 2 d_face_above[Nz-1] and dz[k] = 2 d_face_above[k] - dz[k+1]
    downwards.  d_face_below is implied by the same heights.
    """
    da = np.asarray(d_face_above, dtype=np.float64)
    dz = np.empty_like(da)
    if da.size:
        dz[-1] = 2.0 * da[-1]
        for k in range(da.size - 2, -1, -1):
            dz[k] = 2.0 * da[k] - dz[k + 1]
    return dz


call = _lib.call
