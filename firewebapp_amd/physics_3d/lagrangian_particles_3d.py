"""lagrangian_particles_3d drop-in: the locator and buffer helpers the bed
particles use (lagrangian_particles_3d.py:78-186 of the reference).

locate_k_from_z and locate_cell are the reference's scalar helpers (one
position per call), kept as host functions with the same semantics: numba
callers use them inside their own loops.  The GPU step never calls them; its
particle -> cell map is fw_bed_locate, the same rule applied to every
particle in one launch (locate_cells below), computed once because the bed
particles do not move.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from ..device import DeviceGrid, device, ptr, stream_ptr
from ..grid import Grid3D

# lagrangian_particles_3d.py:83-88
GRAV = 9.81
CP_GAS = 1100.0
ALIVE_FALSE = 0
ALIVE_TRUE = 1


def locate_k_from_z(z, z_face, Nz):
    """k of the cell containing z (z_face[k] <= z < z_face[k+1]), -1 outside
    [z_face[0], z_face[Nz]) (lagrangian_particles_3d.py:94-108)."""
    if z < z_face[0] or z >= z_face[Nz]:
        return -1
    for k in range(Nz):
        if z < z_face[k + 1]:
            return k
    return Nz - 1


def locate_cell(x, y, z, dx, dy, z_face, Nz, Nx, Ny):
    """(i, j, k) of the cell containing (x, y, z); -1 on an axis where the
    point is outside (negative coordinates before the truncation)
    (lagrangian_particles_3d.py:111-138)."""
    if x < 0.0:
        i = -1
    else:
        i = int(x / dx)
        if i >= Nx:
            i = -1
    if y < 0.0:
        j = -1
    else:
        j = int(y / dy)
        if j >= Ny:
            j = -1
    return i, j, locate_k_from_z(z, z_face, Nz)


def locate_cells(x, y, z, dx, dy, z_face, Nz, Nx, Ny):
    """Flat cell index (k Ny + j) Nx + i of every particle by the rule of
    locate_cell, -1 where any axis is outside: fw_bed_locate on the device
    (the map the GPU step builds its per-cell segments from).  x, y, z:
    float64 arrays or CUDA tensors; returns an int32 array of the same kind."""
    dev = device()
    zf = np.asarray(z_face, np.float64)
    dz = np.diff(zf)
    if dz.size != Nz:
        raise ValueError(f"z_face has {zf.size} values, Nz + 1 = {Nz + 1} are required")
    da, db = Grid3D._faces(dz)
    g = DeviceGrid(Nx, Ny, Nz, dx, dy, dz, da, db, z_face=zf, dtype=torch.float64)
    t = [torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous() for a in (x, y, z)]
    n = t[0].numel()
    if t[1].numel() != n or t[2].numel() != n:
        raise ValueError("x, y and z must have the same length")
    cs = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ca = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    _lib.call("fw_bed_locate", g.ref, int(n), ptr(t[0]), ptr(t[1]), ptr(t[2]), ptr(cs),
              ptr(ca), stream_ptr())
    out = cs[:n]
    if isinstance(x, torch.Tensor):
        return out
    torch.cuda.current_stream().synchronize()
    return out.cpu().numpy()


def alloc_dead_slot(part_alive):
    """First slot with alive == ALIVE_FALSE, or -1 (lagrangian_particles_3d.py:144-157)."""
    dead = np.flatnonzero(np.asarray(part_alive) == ALIVE_FALSE)
    return int(dead[0]) if dead.size else -1


def allocate_kinematic_buffers(N_max):
    """Zeroed kinematic arrays x, y, z, u, v, w, age (float64) and alive
    (int8) of N_max slots (lagrangian_particles_3d.py:160-186)."""
    if N_max < 0:
        raise ValueError(f"N_max must be ≥ 0; got {N_max}")
    buf = {k: np.zeros(N_max, dtype=np.float64) for k in ("x", "y", "z", "u", "v", "w")}
    buf["alive"] = np.zeros(N_max, dtype=np.int8)
    buf["age"] = np.zeros(N_max, dtype=np.float64)
    return buf
