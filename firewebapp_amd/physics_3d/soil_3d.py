"""soil_3d drop-in (soil_3d.py:66-153 of the reference)."""
from __future__ import annotations

import numpy as np
import torch

from ..device import device
from ._common import Marshal, call, grid_for, ptr, stream_ptr

RHO_SOIL_DEFAULT = 1500.0
CP_SOIL_DEFAULT = 800.0
K_SOIL_DEFAULT = 0.48
EPS_SOIL_DEFAULT = 0.85
SIGMA_SB = 5.67e-8


def build_soil_grid(n_soil=6, dz_first=0.001, growth=1.5):
    """Geometric soil layers (soil_3d.py:73-102); host-side, one-shot."""
    soil_dz = np.array([dz_first * growth ** k for k in range(n_soil)], dtype=np.float64)
    d_above = np.empty(n_soil)
    d_below = np.empty(n_soil)
    d_above[0] = soil_dz[0] / 2.0
    for k in range(1, n_soil):
        d_above[k] = 0.5 * (soil_dz[k - 1] + soil_dz[k])
    for k in range(n_soil - 1):
        d_below[k] = 0.5 * (soil_dz[k] + soil_dz[k + 1])
    d_below[-1] = soil_dz[-1] / 2.0
    return soil_dz, d_above, d_below, float(soil_dz.sum())


def step_soil_conduction(T_soil, q_in_surface, dt, soil_dz, d_above, d_below,
                         alpha_s, k_s, rho_s, cp_s, eps_s, T_amb):
    """Explicit FV soil column, Gauss-Seidel order in k; in place."""
    m = Marshal(T_soil)
    n_soil, ny, nx = T_soil.shape
    g = grid_for(m, (1, ny, nx), 1.0, 1.0, [1.0], [1.0], [1.0])
    Ts = m.out(T_soil)
    f64 = dict(dtype=torch.float64, device=device())
    sdz = torch.as_tensor(np.asarray(soil_dz, np.float64), **f64)
    sda = torch.as_tensor(np.asarray(d_above, np.float64), **f64)
    sdb = torch.as_tensor(np.asarray(d_below, np.float64), **f64)
    call("fw_soil_conduction", g.ref, int(n_soil), ptr(Ts), ptr(m.dev(q_in_surface)),
         ptr(sdz), ptr(sda), ptr(sdb), float(dt), float(k_s), float(rho_s),
         float(cp_s), float(eps_s), float(T_amb), stream_ptr())
    m.writeback()
