"""fft_poisson_3d drop-in (fft_poisson_3d.py:45-228 of the reference).

The eigenbases are built once on the host, like the reference's constructor
(fft_poisson_3d.py:70-125):

* x: the Neumann-left / Dirichlet-right 1-D Laplacian has the closed-form
  DCT-IV eigenbasis cos(pi (m+1/2)(i+1/2)/Nx), lambda_m =
  -(4/dx^2) sin^2(pi (2m+1)/(4 Nx));
* y: the real orthonormal periodic basis (1, cos, sin, (-1)^j) with
  lambda = (2/dy^2)(cos(2 pi m/Ny) - 1), the same operator the reference
  applies through the FFT;
* z: eigh of sqrt(dz) L_z sqrt(dz)^-1 (non-uniform dz has no closed form).

Applying the inverse runs on the GPU (libfirewave: batched dense transforms).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from ..device import Marshal, grid_for, ptr, stream_ptr


def build_L_z(Nz, dz_arr, d_above, d_below):
    """1-D z Laplacian, Neumann at k=0, Dirichlet face at the top
    (fft_poisson_3d.py:150-181)."""
    L = np.zeros((Nz, Nz))
    for k in range(Nz):
        inv_dz_k = 1.0 / dz_arr[k]
        if 1 <= k <= Nz - 2:
            ca = inv_dz_k / d_above[k]
            cb = inv_dz_k / d_below[k]
            L[k, k] = -ca - cb
            L[k, k + 1] = ca
            L[k, k - 1] = cb
        elif k == 0:
            ca = inv_dz_k / d_above[0]
            L[0, 0] = -ca
            L[0, 1] = ca
        else:
            cb = inv_dz_k / d_below[Nz - 1]
            L[Nz - 1, Nz - 1] = -2.0 * inv_dz_k * inv_dz_k - cb
            L[Nz - 1, Nz - 2] = cb
    return L


def x_basis(Nx, dx):
    i = np.arange(Nx)
    U = np.sqrt(2.0 / Nx) * np.cos(np.pi * (i[None, :] + 0.5) * (i[:, None] + 0.5) / Nx)
    lam = -(4.0 / (dx * dx)) * np.sin(np.pi * (2 * i + 1) / (4.0 * Nx)) ** 2
    return lam, U


def y_basis(Ny, dy):
    j = np.arange(Ny)
    lam = lambda m: (2.0 / (dy * dy)) * (np.cos(2.0 * np.pi * m / Ny) - 1.0)
    cols, lams = [np.full(Ny, 1.0 / np.sqrt(Ny))], [lam(0)]
    for m in range(1, (Ny - 1) // 2 + 1):
        cols += [np.sqrt(2.0 / Ny) * np.cos(2.0 * np.pi * m * j / Ny),
                 np.sqrt(2.0 / Ny) * np.sin(2.0 * np.pi * m * j / Ny)]
        lams += [lam(m), lam(m)]
    if Ny % 2 == 0 and Ny > 1:
        cols.append(((-1.0) ** j) / np.sqrt(Ny))
        lams.append(lam(Ny // 2))
    return np.array(lams, dtype=np.float64), np.ascontiguousarray(np.stack(cols, axis=1))


def z_basis(Nz, dz_arr, d_above, d_below):
    L = build_L_z(Nz, dz_arr, d_above, d_below)
    D = np.sqrt(np.asarray(dz_arr, np.float64))
    Ls = (D[:, None] * L) * (1.0 / D)[None, :]
    Ls = 0.5 * (Ls + Ls.T)
    lam, U = np.linalg.eigh(Ls)
    return lam, np.ascontiguousarray(U)


def bases(Nz, Ny, Nx, dx, dy, dz_arr, d_above, d_below):
    lx, Ux = x_basis(Nx, dx)
    ly, Uy = y_basis(Ny, dy)
    lz, Uz = z_basis(Nz, dz_arr, d_above, d_below)
    return (lx, Ux), (ly, Uy), (lz, Uz)


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.c_void_p)


def create_handle(grid, eps_reg=1.0e-6, cg_rtol=1.0e-6, maxiter=400):
    """fw_proj_create for a DeviceGrid; returns the opaque handle."""
    dz = grid.dz.cpu().numpy()
    da = grid.d_above.cpu().numpy()
    db = grid.d_below.cpu().numpy()
    (lx, Ux), (ly, Uy), (lz, Uz) = bases(grid.nz, grid.ny, grid.nx, grid.dx,
                                         grid.dy, dz, da, db)
    keep = [_dp(a) for a in (Ux, Uy, Uz, lx, ly, lz)]
    h = C.c_void_p()
    _lib.call("fw_proj_create", grid.ref, *[k[1] for k in keep], float(eps_reg),
              float(cg_rtol), int(maxiter), C.byref(h))
    return h


class SeparableLaplacian3D:
    """Constant-coefficient inverse of the projection operator (GPU)."""

    def __init__(self, Nz, Ny, Nx, dx, dy, dz_arr, d_face_above, d_face_below,
                 eps_reg=1.0e-6):
        self.Nz, self.Ny, self.Nx = Nz, Ny, Nx
        self.dx, self.dy = dx, dy
        self.dz_arr = np.asarray(dz_arr, np.float64)
        self.d_above = np.asarray(d_face_above, np.float64)
        self.d_below = np.asarray(d_face_below, np.float64)
        self.eps_reg = eps_reg
        self._h = {}
        self._grids = {}

    def _handle(self, m):
        if m.dtype not in self._h:
            g = grid_for(m, (self.Nz, self.Ny, self.Nx), self.dx, self.dy,
                         self.dz_arr, self.d_above, self.d_below)
            self._grids[m.dtype] = g
            self._h[m.dtype] = create_handle(g, self.eps_reg)
        return self._h[m.dtype]

    def solve(self, rhs):
        if tuple(rhs.shape) != (self.Nz, self.Ny, self.Nx):
            raise ValueError(f"rhs shape {tuple(rhs.shape)} != expected "
                             f"({self.Nz}, {self.Ny}, {self.Nx})")
        m = Marshal(rhs)
        h = self._handle(m)
        r = m.dev(rhs)
        import torch
        z = torch.empty_like(r)
        _lib.call("fw_proj_precondition", h, ptr(r), ptr(z), stream_ptr())
        torch.cuda.current_stream().synchronize()
        if isinstance(rhs, np.ndarray):
            return z.cpu().numpy().astype(rhs.dtype)
        return z

    def __del__(self):
        try:
            for h in self._h.values():
                _lib.lib().fw_proj_destroy(h)
        except Exception:
            pass
