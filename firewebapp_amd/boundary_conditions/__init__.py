"""boundary_conditions drop-in: the BC registry (boundary_conditions/
__init__.py:32-51 of the reference) and OutdoorWindBC
(boundary_conditions/outdoor_wind.py:34-130), the wind-inlet pattern of the
spread step: x-min log-law inlet, x-max outflow, periodic y, no-slip bed base,
free top.

Only "outdoor_wind" is registered: the cup burner (boundary_conditions/
cup_burner.py) is not on this path, and asking for it raises the registry's
ValueError.  Of OutdoorWindBC's two inlet profiles the log law is built; the
Raupach canopy profile ("canopy_bl") raises ValueError.
"""
from __future__ import annotations

import numpy as np

from ..grid import build_u_inlet


class BoundaryCondition:
    """BC contract (boundary_conditions/base.py:37-80): configure() once at
    set-up, apply_per_step() before momentum each step."""

    kind: str = ""

    def configure(self, proj_solver, grid, state) -> None:
        raise NotImplementedError

    def apply_per_step(self, state, grid, t: float) -> None:
        pass


class OutdoorWindBC(BoundaryCondition):
    """x-face wind inlet, outflow, periodic y, wall bed base, free top."""

    kind = "outdoor_wind"

    def __init__(self, wind_speed_m_s, wind_profile_type="log_law", h_bed=0.0,
                 sigma_sav=2000.0, alpha_s_avg=0.0):
        if wind_profile_type not in ("log_law", "canopy_bl"):
            raise ValueError(f"wind_profile_type={wind_profile_type!r} not in "
                             "{'log_law', 'canopy_bl'}")
        if wind_profile_type != "log_law":
            raise ValueError("wind_profile_type='canopy_bl' (the Raupach canopy inlet) is "
                             "not built here; the spread step uses 'log_law'")
        self.wind_speed_m_s = float(wind_speed_m_s)
        self.wind_profile_type = str(wind_profile_type)
        self.h_bed = float(h_bed)
        self.sigma_sav = float(sigma_sav)
        self.alpha_s_avg = float(alpha_s_avg)
        self.u_inlet = None

    def build_u_inlet(self, grid):
        """(Nz, Ny) inlet u: the log law U10 ln((z+z0)/z0)/ln((10+z0)/z0) at
        z_mid, z0 = 0.01 m (outdoor_wind.py:80-110)."""
        self.u_inlet = build_u_inlet(grid, self.wind_speed_m_s)
        return self.u_inlet

    def configure(self, proj_solver, grid, state) -> None:
        """Install the inlet in the projection solver and start the interior
        from the inlet profile (outdoor_wind.py:112-118)."""
        u_inlet = self.build_u_inlet(grid)
        proj_solver.set_inlet_BC(u_inlet)
        u = state.u
        if isinstance(u, np.ndarray):
            u[:, :, :] = u_inlet[:, :, np.newaxis]
        else:
            import torch
            u.copy_(torch.as_tensor(u_inlet, dtype=u.dtype, device=u.device)[:, :, None]
                    .expand(u.shape))


_REGISTRY = {OutdoorWindBC.kind: OutdoorWindBC}


def available():
    """Sorted tuple of registered BC kinds."""
    return tuple(sorted(_REGISTRY.keys()))


def get_bc_class(kind):
    """The BC class of a kind; ValueError on an unknown kind."""
    if kind not in _REGISTRY:
        raise ValueError(f"Unknown boundary_condition_kind={kind!r}.  "
                         f"Registered: {available()}")
    return _REGISTRY[kind]
