"""EDC closure drop-in (chemistry_closures/edc.py:133-376 of the reference)."""
from __future__ import annotations

import torch

from ... import _lib
from ...device import device
from .._common import Marshal, call, grid_for, ptr, stream_ptr
from ._constants import HOC_J, S_STOICH


def step_chemistry_ode_edc(rho, T_g, Y_fuel, Y_O2, k_turb, eps_turb, chi_rad,
                           cp_g, dt, n_substeps, omega_int_out, Y_H2O,
                           extinction_enable=False, s_stoich=S_STOICH,
                           hoc_J=HOC_J):
    """Magnussen EDC with Y_H2O quench and wet-bulb relaxation; in place."""
    m = Marshal(T_g)
    nz, ny, nx = T_g.shape
    g = grid_for(m, (nz, ny, nx), 1.0, 1.0, [1.0] * nz, [1.0] * nz, [1.0] * nz)
    T, Yf, YO, om = m.out(T_g), m.out(Y_fuel), m.out(Y_O2), m.out(omega_int_out)
    nwork = _lib.lib().fw_edc_workspace_ints(g.ref)
    work = torch.empty(max(nwork, 1), dtype=torch.int32, device=device())
    call("fw_edc", g.ref, ptr(m.dev(rho)), ptr(T), ptr(Yf), ptr(YO),
         ptr(m.dev(k_turb)), ptr(m.dev(eps_turb)), ptr(m.dev(Y_H2O)), ptr(om),
         float(chi_rad), float(cp_g), float(dt), int(n_substeps),
         int(bool(extinction_enable)), float(s_stoich), float(hoc_J), ptr(work),
         stream_ptr())
    m.writeback()


def run(*, rho, T_g, Y_fuel, Y_O2, k_turb, eps_turb, chi_rad, cp_g, dt,
        n_substeps, omega_out, Y_H2O=None, extinction_enable=False,
        s_stoich=S_STOICH, hoc_J=HOC_J, **_unused):
    """Registry entry: unused keyword arguments are ignored, as in the reference."""
    if Y_H2O is None:
        Y_H2O = (torch.zeros_like(T_g) if isinstance(T_g, torch.Tensor)
                 else T_g * 0.0)
    step_chemistry_ode_edc(rho, T_g, Y_fuel, Y_O2, k_turb, eps_turb, chi_rad,
                           cp_g, dt, n_substeps, omega_out, Y_H2O,
                           bool(extinction_enable), s_stoich=s_stoich,
                           hoc_J=hoc_J)
