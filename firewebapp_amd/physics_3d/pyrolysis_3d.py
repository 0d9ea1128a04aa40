"""pyrolysis_3d drop-in: the Eulerian bed kernels of the reference
(pyrolysis_3d.py:270-614) -- drying, single-pool MD2004 pyrolysis with char
yield, char oxidation and smouldering oxidation -- through fw_pyro_step.
In-place semantics as in the reference: the mass pools and the outputs are
updated in the caller's arrays.  Mass pools run in float64 on the device."""
from __future__ import annotations

import numpy as np
import torch

from ._common import Marshal, call, grid_for, ptr, stream_ptr

A_MD2004 = 36_280.0
E_MD2004 = 58_600.0
CHAR_YIELD_MD2004 = 0.15
ETA_MD2004 = 1.0 - CHAR_YIELD_MD2004
A_DRY = 4.29e3
E_DRY = 43_800.0
L_VAP_WATER = 2.26e6

_DRY, _MD, _CHAR, _SMOLD = 1, 2, 4, 8


def _run(process, T_s, alpha_s, dt, Y_O2=None, m_solid=None, m_char=None, m_water=None,
         m_water_init=0.0, A_p=A_MD2004, E_p=E_MD2004, outs=()):
    m = Marshal(T_s)
    sh = tuple(T_s.shape)
    g = grid_for(m, sh, 1.0, 1.0, np.ones(sh[0]), np.ones(sh[0]), np.ones(sh[0]))
    f64 = lambda a: m.out(a, dtype=torch.float64, shape=sh) if a is not None else None
    ms, mc, mw = f64(m_solid), f64(m_char), f64(m_water)
    Yo = m.dev(Y_O2, shape=sh) if Y_O2 is not None else None
    slots = {"S": None, "Q": None, "dry": None, "char": None, "smold": None}
    for name, arr in outs:
        slots[name] = m.out(arr, shape=sh)
    P = lambda t: ptr(t) if t is not None else None
    call("fw_pyro_step", g.ref, ptr(m.dev(T_s, shape=sh)), ptr(m.dev(alpha_s, shape=sh)), P(Yo),
         P(ms), P(mc), P(mw), 1, float(m_water_init), float(A_p), float(E_p), float(dt),
         process, P(slots["S"]), P(slots["Q"]), P(slots["dry"]), P(slots["char"]),
         P(slots["smold"]), stream_ptr())
    m.writeback()


def step_drying(T_s, m_water, alpha_s, dt, Q_dry_out):
    """pyrolysis_3d.py:270-309: m_water *= exp(-A exp(-E/RT) dt), Q_dry the
    latent heat rate of the removed water."""
    _run(_DRY, T_s, alpha_s, dt, m_water=m_water, outs=(("dry", Q_dry_out),))


def step_pyrolysis_md2004(T_s, m_solid, m_initial, m_char, Y_O2, alpha_s, m_water,
                          m_water_init, dt, S_pyro_out, Q_pyro_out, A_p=A_MD2004,
                          E_p=E_MD2004):
    """pyrolysis_3d.py:311-462 (m_initial is not read, as in the reference)."""
    _run(_MD, T_s, alpha_s, dt, Y_O2=Y_O2, m_solid=m_solid, m_char=m_char, m_water=m_water,
         m_water_init=m_water_init, A_p=A_p, E_p=E_p,
         outs=(("S", S_pyro_out), ("Q", Q_pyro_out)))


def step_char_oxidation(T_s, m_solid, Y_O2, alpha_s, dt, Q_char_out):
    """pyrolysis_3d.py:465-564: consumes the pool passed as m_solid (the
    step passes m_char)."""
    _run(_CHAR, T_s, alpha_s, dt, Y_O2=Y_O2, m_char=m_solid, outs=(("char", Q_char_out),))


def step_smoldering_oxidation(T_s, m_solid, Y_O2, alpha_s, dt, Q_smold_out):
    """pyrolysis_3d.py:567-614."""
    _run(_SMOLD, T_s, alpha_s, dt, Y_O2=Y_O2, m_char=m_solid, outs=(("smold", Q_smold_out),))
