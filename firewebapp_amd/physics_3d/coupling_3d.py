"""coupling_3d drop-in (coupling_3d.py:76-192 of the reference): one gas-solid
coupling step through fw_gas_solid_coupling, in place on T_g, T_s and
m_water."""
from __future__ import annotations

import numpy as np
import torch

from ._common import Marshal, call, grid_for, ptr, stream_ptr

_RHO_SOLID = 500.0
_CP_SOLID = 1300.0


def step_gas_solid_coupling(T_g, T_s, rho, u, v, w, alpha_s, sigma_sav, q_rad_in, Q_pyro,
                            Q_comb, m_water, L_v, dt, dz_arr, T_amb, q_loss_enable=True,
                            h_conv_mult=1.0):
    m = Marshal(T_g)
    sh = tuple(T_g.shape)
    dz = np.asarray(dz_arr, np.float64)
    g = grid_for(m, sh, 1.0, 1.0, dz, dz, dz)
    Tg, Ts = m.out(T_g, shape=sh), m.out(T_s, shape=sh)
    mw = m.out(m_water, dtype=torch.float64, shape=sh)
    d = lambda a: ptr(m.dev(a, shape=sh))
    call("fw_gas_solid_coupling", g.ref, ptr(Tg), ptr(Ts), d(rho), d(u), d(v), d(w),
         d(alpha_s), float(sigma_sav), d(q_rad_in), d(Q_pyro), d(Q_comb), ptr(mw), float(L_v),
         float(dt), float(T_amb), int(bool(q_loss_enable)), float(h_conv_mult), stream_ptr())
    m.writeback()
