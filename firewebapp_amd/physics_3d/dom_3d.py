"""dom_3d drop-in: DOMRadiationSolver (dom_3d.py:227-473 of the reference).

N_quadrature 4 uses the reference's S4 set verbatim (6-decimal cosines,
dom_3d.py:107-141).  N_quadrature 8 (absent from the reference, which raises)
is the level-symmetric LQ8 set: mu_i^2 = (1 + 6 (i-1))/21 and the three
published weight classes 0.1209877 / 0.0907407 / 0.0925926 per octant.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .. import _lib
from ..device import Marshal, device_view, grid_for, ptr, stream_ptr

SIGMA_SB = 5.67e-8
KAPPA_SOOT_HOT = 0.5
OMEGA_COMB_THRESH = 1.0e-3
T_GAS_RAD_REF = 600.0
T_GAS_RAD_DT = 400.0
EPS_W_GROUND = 0.85

_OCTANTS = [(1, 1, 1), (1, 1, -1), (1, -1, 1), (-1, 1, 1),
            (1, -1, -1), (-1, 1, -1), (-1, -1, 1), (-1, -1, -1)]


def _generate_sn_ordinates(N):
    """Ordinates and weights (dom_3d.py:98-142 for S4; LQ8 for S8)."""
    if N == 4:
        mu1, mu2 = 0.295876, 0.908248
        per = [(mu1, mu1, mu2), (mu1, mu2, mu1), (mu2, mu1, mu1)]
        wper = None
    elif N == 8:
        mu = [math.sqrt((1.0 + 6.0 * i) / 21.0) for i in range(4)]
        cls_w = {(0, 0, 3): 0.1209877, (0, 1, 2): 0.0907407, (1, 1, 1): 0.0925926}
        per, wper = [], []
        for a in range(4):
            for b in range(4):
                c = 3 - a - b
                if 0 <= c <= 3:
                    per.append((mu[a], mu[b], mu[c]))
                    wper.append(cls_w[tuple(sorted((a, b, c)))])
        tot = sum(wper)
        wper = [x / tot for x in wper]
    else:
        raise NotImplementedError(f"S{N} not implemented (S4 and S8 are)")
    Om, w = [], []
    for sx, sy, sz in _OCTANTS:
        for q, o in enumerate(per):
            Om.append((sx * o[0], sy * o[1], sz * o[2]))
            if wper is not None:
                w.append(wper[q] * 4.0 * math.pi / 8.0)
    Om = np.array(Om, dtype=np.float64)
    if wper is None:
        w = np.full(Om.shape[0], 4.0 * math.pi / Om.shape[0])
    return Om, np.asarray(w, dtype=np.float64)


class DOMRadiationSolver:
    def __init__(self, Nz, Ny, Nx, dy, dx, dz_arr, d_face_above, d_face_below,
                 y_bc="periodic", N_quadrature=4, eps_w_ground=EPS_W_GROUND,
                 max_source_iter=30, tol_source_iter=1.0e-3, omega_relax=0.7,
                 kappa_gas_max=KAPPA_SOOT_HOT, **_unused):
        if y_bc != "periodic":
            raise ValueError("only y_bc='periodic' is on the hot path")
        self.Nz, self.Ny, self.Nx = Nz, Ny, Nx
        self.dx, self.dy = dx, dy
        self.dz_arr = np.asarray(dz_arr, np.float64).copy()
        self.d_face_above = np.asarray(d_face_above, np.float64)
        self.d_face_below = np.asarray(d_face_below, np.float64)
        self.y_periodic = True
        self.eps_w_ground = float(eps_w_ground)
        self.max_iter = int(max_source_iter)
        self.tol = float(tol_source_iter)
        self.omega = float(omega_relax)
        self.kappa_gas_max = float(kappa_gas_max)
        self.Omega, self.weights = _generate_sn_ordinates(N_quadrature)
        self.M = self.Omega.shape[0]
        self._h = {}
        self._grids = {}
        self.last_iters = 0
        self._dtype = None

    def _handle(self, dtype):
        if dtype not in self._h:
            m = Marshal(torch.empty(0, dtype=dtype))
            g = grid_for(m, (self.Nz, self.Ny, self.Nx), self.dx, self.dy,
                         self.dz_arr, self.d_face_above, self.d_face_below)
            self._grids[dtype] = g
            h = C.c_void_p()
            Om = np.ascontiguousarray(self.Omega)
            w = np.ascontiguousarray(self.weights)
            _lib.call("fw_dom_create", g.ref, int(self.M),
                      Om.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                      self.eps_w_ground, self.max_iter, self.tol, self.omega,
                      self.kappa_gas_max, C.byref(h))
            self._h[dtype] = h
        self._dtype = dtype
        return self._h[dtype]

    @property
    def I_set(self):
        """(M, Nz, Ny, Nx) intensities (zero-copy device tensor; owned by the
        handle and overwritten by the next solve)."""
        h = self._h[self._dtype]
        p = C.c_void_p()
        _lib.call("fw_dom_intensity", h, C.byref(p))
        return device_view(p.value, (self.M, self.Nz, self.Ny, self.Nx), self._dtype)

    def solve(self, T_s, T_g, alpha_s, omega_comb, sigma_sav, T_amb,
              q_rad_solid_out, q_rad_gas_out, T_soil_surface=None,
              q_in_soil_out=None, Y_H2O=None, rho=None,
              bed_moisture_per_cell=None):
        m = Marshal(T_s)
        h = self._handle(m.dtype)
        sh = (self.Nz, self.Ny, self.Nx)
        sh2 = (self.Ny, self.Nx)
        qs = m.out(q_rad_solid_out, shape=sh)
        qg = m.out(q_rad_gas_out, shape=sh)
        qsoil = m.out(q_in_soil_out, shape=sh2) if q_in_soil_out is not None else None
        opt = lambda a, s: m.dev(a, shape=s) if a is not None else None
        _lib.call("fw_dom_solve", h, ptr(m.dev(T_s, shape=sh)), ptr(m.dev(T_g, shape=sh)),
                  ptr(m.dev(alpha_s, shape=sh)), ptr(m.dev(omega_comb, shape=sh)),
                  float(sigma_sav), float(T_amb), ptr(qs), ptr(qg),
                  ptr(opt(T_soil_surface, sh2)), ptr(qsoil), ptr(opt(Y_H2O, sh)),
                  ptr(opt(rho, sh)), ptr(opt(bed_moisture_per_cell, sh)), stream_ptr())
        it = C.c_int()
        _lib.call("fw_dom_stats", h, C.byref(it))
        self.last_iters = it.value
        m.writeback()

    def __del__(self):
        try:
            for h in self._h.values():
                _lib.lib().fw_dom_destroy(h)
        except Exception:
            pass
