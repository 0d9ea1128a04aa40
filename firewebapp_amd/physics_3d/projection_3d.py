"""projection_3d drop-in: ProjectionSolver3D, fft_pcg method
(projection_3d.py:214-948 of the reference).

Differences that are part of the design, not of the call surface:
* the Krylov method is PCG on the dz-symmetrised operator (the reference's
  BiCGSTAB with the same preconditioner and the same stopping rule
  ||b - A p||_2 <= cg_rtol ||b||_2); both return a p inside that ball;
* with float32 fields the solve is mixed precision: float64 right-hand side,
  pressure and refinement residual, float32 inner PCG, float64 dot products.
Other methods ('pardiso', 'amg_cg') are not on the hot path and raise.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..device import Marshal, grid_for, ptr, stream_ptr
from .fft_poisson_3d import create_handle


class ProjectionSolver3D:
    def __init__(self, Nz, Ny, Nx, dy, dx, dz_arr, d_face_above, d_face_below,
                 y_bc="periodic", method="fft_pcg", cg_rtol=1.0e-6,
                 amg_rebuild_every=100):
        if method != "fft_pcg":
            raise ValueError(f"projection method {method!r} is not on the hot path "
                             "(only 'fft_pcg' is implemented)")
        if y_bc != "periodic":
            raise ValueError("only y_bc='periodic' is on the hot path")
        self.Nz, self.Ny, self.Nx = Nz, Ny, Nx
        self.dx, self.dy = dx, dy
        self.dz_arr = np.asarray(dz_arr, np.float64)
        self.d_face_above = np.asarray(d_face_above, np.float64)
        self.d_face_below = np.asarray(d_face_below, np.float64)
        self.dz = float(self.dz_arr[0])
        self.y_bc = y_bc
        self.method = method
        self.cg_rtol = cg_rtol
        self._u_inlet = np.zeros((Nz, Ny))
        self._alpha_g = None
        self._h = {}
        self._grids = {}
        self._rho = None
        self._last_fft_pcg_iters = 0
        self.last_rel_resid = 0.0
        self.last_div_max = 0.0

    # -- state ---------------------------------------------------------------
    def _handle(self, dtype):
        if dtype not in self._h:
            m = Marshal(torch.empty(0, dtype=dtype))
            g = grid_for(m, (self.Nz, self.Ny, self.Nx), self.dx, self.dy,
                         self.dz_arr, self.d_face_above, self.d_face_below)
            self._grids[dtype] = g
            self._h[dtype] = create_handle(g, 1.0e-6, self.cg_rtol)
        return self._h[dtype]

    def set_alpha_g(self, alpha_g):
        if alpha_g is not None:
            raise ValueError("volume-weighted projection (alpha_g) is not on the hot path")
        self._alpha_g = None

    def set_inlet_BC(self, u_inlet):
        self._u_inlet = u_inlet

    def rebuild_for_rho(self, rho):
        """The operator is applied matrix-free from rho; only rho is kept."""
        self._rho = rho

    # -- operations ----------------------------------------------------------
    def project(self, u, v, w, rho, dt, div_target=None):
        m = Marshal(u)
        h = self._handle(m.dtype)
        sh = (self.Nz, self.Ny, self.Nx)
        U, V, W = m.out(u, shape=sh), m.out(v, shape=sh), m.out(w, shape=sh)
        R = m.dev(rho, shape=sh)
        T = m.dev(div_target, shape=sh) if div_target is not None else None
        UI = m.dev(self._u_inlet, shape=(self.Nz, self.Ny))
        _lib.call("fw_proj_project", h, ptr(U), ptr(V), ptr(W), ptr(R), float(dt),
                  ptr(T), ptr(UI), stream_ptr())
        it, rel, dmax = C.c_int(), C.c_double(), C.c_double()
        _lib.call("fw_proj_stats", h, C.byref(it), C.byref(rel), C.byref(dmax))
        self._last_fft_pcg_iters = it.value
        self.last_rel_resid = rel.value
        self.last_div_max = dmax.value
        p = torch.empty((self.Nz, self.Ny, self.Nx), dtype=torch.float64, device=U.device)
        _lib.call("fw_proj_pressure", h, ptr(p), stream_ptr())
        m.writeback()
        return p.cpu().numpy() if isinstance(u, np.ndarray) else p

    def divergence(self, u, v, w):
        m = Marshal(u)
        h = self._handle(m.dtype)
        sh = (self.Nz, self.Ny, self.Nx)
        U, V, W = m.dev(u, shape=sh), m.dev(v, shape=sh), m.dev(w, shape=sh)
        out = torch.empty_like(U)
        _lib.call("fw_proj_divergence", h, ptr(U), ptr(V), ptr(W),
                  ptr(m.dev(self._u_inlet, shape=(self.Nz, self.Ny))), ptr(out),
                  stream_ptr())
        torch.cuda.current_stream().synchronize()
        return out.cpu().numpy() if isinstance(u, np.ndarray) else out

    def matvec(self, x, rho=None):
        """A @ x for the current rho (the reference's ``self._A @ x``)."""
        rho = self._rho if rho is None else rho
        m = Marshal(rho)
        h = self._handle(m.dtype)
        X = m.dev(np.asarray(x).reshape(self.Nz, self.Ny, self.Nx)
                  if isinstance(x, np.ndarray) else x.reshape(self.Nz, self.Ny, self.Nx))
        Y = torch.empty_like(X)
        _lib.call("fw_proj_matvec", h, ptr(m.dev(rho, shape=(self.Nz, self.Ny, self.Nx))),
                  ptr(X), ptr(Y), stream_ptr())
        torch.cuda.current_stream().synchronize()
        return Y.cpu().numpy().reshape(-1) if isinstance(x, np.ndarray) else Y

    def __del__(self):
        try:
            for h in self._h.values():
                _lib.lib().fw_proj_destroy(h)
        except Exception:
            pass
