"""Shared helpers of the drop-in modules."""
from __future__ import annotations

import numpy as np

from .. import _lib
from ..device import Marshal, grid_for, ptr, stream_ptr  # noqa: F401


def dz_from_faces(d_face_above, d_face_below):
    """Placeholder dz for kernels whose reference signature has no dz_arr."""
    return np.asarray(d_face_above, dtype=np.float64)


call = _lib.call
