"""flame_front_3d drop-in (flame_front_3d.py of the reference): the level-set
front tracker LevelSetFront3D and the flame-front helpers the outer step
calls (flame_front_3d.py:70-111, 113-402, 440-514, 594-630, 756-835).

phi is float64 on the device, as the reference keeps it float64.  Every
array operation runs in libfirewave (fw_lset_init_patch, fw_levelset_update,
fw_phi_window_mask, fw_lset_front_x, fw_phi_z_mean, fw_phi_flame,
fw_update_cell_age, fw_q_in_at_front_3d, fw_lset_speed,
fw_dom_forward_flux).

Array conventions follow the rest of the drop-in package: NumPy arguments are
copied to the device and the results come back as NumPy arrays; CUDA tensors
are used in place.  LevelSetFront3D itself has no array argument to infer
that from, so it takes the additive keyword ``device_arrays``: False (the
default, the reference's behaviour) makes ``phi`` and the masks NumPy arrays;
True keeps them as CUDA tensors.  ``phi`` is a property: assigning to it
replaces the device field, and with NumPy arrays a read returns a copy, so
edit phi by assignment (the reference's loop only reads it).

Scalar helpers (flame_tilt_band_m, compute_v_n on 2-D input sums) are host
arithmetic as in the reference.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .. import _lib
from ..device import DeviceGrid, Marshal, device, grid_for, ptr, stream_ptr
from ..grid import Grid3D

# constants of flame_front_3d.py:61-67, :437, :522-523
T_FLAME_K = 1500.0
L_BURNOUT_M = 0.50
H_FLAME_FRANKMAN = 200.0
DX_VN_BAND_M = 0.20
G_ACCEL = 9.81
WIND_MIDFLAME_FRAC = 0.723
L_VAP_WATER = 2.26e6
Q_BOOTSTRAP_W_M3 = 500_000.0
T_BOOTSTRAP_S = 2.0
# LevelSetFront3D scheduling and initial value (flame_front_3d.py:107-110)
PHI_INIT_UNBURNED = 100.0
LEVELSET_REINIT_INTERVAL = 10
REINIT_SUBSTEPS = 5
DT_REINIT_FRAC = 0.5
# active-flame criterion of compute_phi_flame_from_state (flame_front_3d.py:751-753)
OMEGA_MIN_FLAME = 1.0e-3
T_PLUME_MIN = 1000.0
Y_F_MIN_PLUME = 1.0e-3


def flame_tilt_band_m(u_10_m_s, L_flame=L_BURNOUT_M, g=G_ACCEL,
                      midflame_frac=WIND_MIDFLAME_FRAC):
    """Albini flame-tilt band length L sin(theta), tan(theta) = U_mid /
    sqrt(2 g L) (flame_front_3d.py:70-111); host scalar arithmetic."""
    u_buoy = (2.0 * g * L_flame) ** 0.5
    u_mid = max(u_10_m_s, 0.0) * midflame_frac
    tan_th = u_mid / u_buoy
    sin_th = tan_th / (1.0 + tan_th * tan_th) ** 0.5
    return L_flame * sin_th


def _grid64(shape, dx, dy, dz_arr):
    nz, ny, nx = shape
    dz = np.asarray(dz_arr, np.float64)
    da, db = Grid3D._faces(dz)
    return DeviceGrid(nx, ny, nz, dx, dy, dz, da, db, dtype=torch.float64)


def _bool(mask_u8, numpy_out):
    b = mask_u8.bool()
    return b.cpu().numpy() if numpy_out else b


def _mask_u8(mask, m: Marshal):
    """uint8 device copy of a boolean / integer mask (NumPy or tensor)."""
    if isinstance(mask, torch.Tensor):
        t = mask.to(device=device(), dtype=torch.uint8).contiguous()
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(mask), dtype=np.uint8),
                            device=device())
    m._keep.append(t)
    return t


def godunov_grad_norm(phi, dx, dy, dz_arr, grad_out):
    """|grad phi| with first-order Godunov upwinding (flame_front_3d.py:113-188),
    into grad_out."""
    m = Marshal(phi)
    if m.dtype != torch.float64:
        raise TypeError("the level set is float64")
    g = _grid64(tuple(phi.shape), dx, dy, dz_arr)
    out = m.out(grad_out)
    _lib.call("fw_levelset_update", g.ref, ptr(m.dev(phi)), None, None, ptr(out), 0.0, 2,
              stream_ptr())
    m.writeback()


def reinit_godunov_grad(phi, phi_0, dx, dy, dz_arr, grad_out):
    """|grad phi| with the sign(phi_0)-aware upwinding of the Sussman
    reinitialisation (flame_front_3d.py:191-263), into grad_out."""
    m = Marshal(phi)
    if m.dtype != torch.float64:
        raise TypeError("the level set is float64")
    g = _grid64(tuple(phi.shape), dx, dy, dz_arr)
    out = m.out(grad_out)
    _lib.call("fw_levelset_update", g.ref, ptr(m.dev(phi)), ptr(m.dev(phi_0)), None,
              ptr(out), 0.0, 3, stream_ptr())
    m.writeback()


class LevelSetFront3D:
    """Level-set front tracker (flame_front_3d.py:266-402): phi < 0 burned,
    phi > 0 unburned, phi = 0 the front."""

    def __init__(self, Nz, Ny, Nx, dx, dy, dz_arr, L_burnout=L_BURNOUT_M, *,
                 device_arrays=False):
        self.Nz, self.Ny, self.Nx = int(Nz), int(Ny), int(Nx)
        self.dx, self.dy = float(dx), float(dy)
        self.dz_arr = np.asarray(dz_arr, np.float64).copy()
        self.L_burnout = float(L_burnout)
        self.device_arrays = bool(device_arrays)
        self._g = _grid64((self.Nz, self.Ny, self.Nx), self.dx, self.dy, self.dz_arr)
        sh = (self.Nz, self.Ny, self.Nx)
        f64 = dict(dtype=torch.float64, device=device())
        self._phi = torch.full(sh, PHI_INIT_UNBURNED, **f64)
        self._tmp = torch.empty(sh, **f64)
        self._phi_0 = torch.zeros(sh, **f64)
        self._mask = torch.empty(sh, dtype=torch.uint8, device=device())
        self._step = 0

    # -- phi --------------------------------------------------------------------
    @property
    def phi(self):
        """The float64 level set: a NumPy copy, or the device tensor itself
        with device_arrays=True."""
        if self.device_arrays:
            return self._phi
        torch.cuda.current_stream().synchronize()
        return self._phi.cpu().numpy()

    @phi.setter
    def phi(self, value):
        t = torch.as_tensor(np.asarray(value) if not isinstance(value, torch.Tensor) else value)
        if tuple(t.shape) != tuple(self._phi.shape):
            raise ValueError(f"phi of shape {tuple(t.shape)} where "
                             f"{tuple(self._phi.shape)} is required")
        self._phi.copy_(t.to(dtype=torch.float64))

    # -- updates ----------------------------------------------------------------
    def initialize_source_patch(self, i_start, i_end, k_top_bed, x_mid):
        """Signed distance to the right edge of the source strip, -L_burnout/2
        inside the patch (flame_front_3d.py:304-330)."""
        xm = torch.as_tensor(np.ascontiguousarray(np.asarray(x_mid, np.float64))
                             if not isinstance(x_mid, torch.Tensor) else x_mid,
                             dtype=torch.float64, device=device()).contiguous()
        if xm.numel() != self.Nx:
            raise ValueError(f"x_mid has {xm.numel()} values, the grid has Nx={self.Nx}")
        _lib.call("fw_lset_init_patch", self._g.ref, ptr(self._phi), ptr(xm), int(i_start),
                  int(i_end), int(k_top_bed), self.L_burnout, stream_ptr())

    def _swap(self):
        self._phi, self._tmp = self._tmp, self._phi

    def evolve(self, dt, v_n_field):
        """phi -= dt v_n |grad phi|_Godunov (flame_front_3d.py:332-337)."""
        m = Marshal(self._phi)
        vn = m.dev(v_n_field, dtype=torch.float64, shape=self._phi.shape)
        _lib.call("fw_levelset_update", self._g.ref, ptr(self._phi), None, ptr(vn),
                  ptr(self._tmp), float(dt), 0, stream_ptr())
        self._swap()
        self._step += 1
        torch.cuda.current_stream().synchronize()

    def reinitialize(self):
        """REINIT_SUBSTEPS Sussman sub-steps with dtau = 0.5 min(dx, dy, dz_min)
        and the sign of the phi before them (flame_front_3d.py:339-355)."""
        self._phi_0.copy_(self._phi)
        dtau = DT_REINIT_FRAC * min(self.dx, self.dy, float(self.dz_arr.min()))
        for _ in range(REINIT_SUBSTEPS):
            _lib.call("fw_levelset_update", self._g.ref, ptr(self._phi), ptr(self._phi_0),
                      None, ptr(self._tmp), dtau, 1, stream_ptr())
            self._swap()
        torch.cuda.current_stream().synchronize()

    def maybe_reinitialize(self):
        """Reinitialise every LEVELSET_REINIT_INTERVAL evolves
        (flame_front_3d.py:357-360)."""
        if self._step % LEVELSET_REINIT_INTERVAL == 0 and self._step > 0:
            self.reinitialize()

    def enforce_z_uniformity(self):
        """phi <- its z mean in every column (flame_front_3d.py:362-377)."""
        _lib.call("fw_phi_z_mean", self._g.ref, ptr(self._phi), stream_ptr())

    # -- masks ------------------------------------------------------------------
    def _window(self, lo, lo_strict, hi):
        _lib.call("fw_phi_window_mask", self._g.ref, ptr(self._phi), float(lo),
                  int(lo_strict), float(hi), self.Nz, ptr(self._mask), stream_ptr())
        return _bool(self._mask, not self.device_arrays)

    def burned_mask(self):
        """phi <= 0 (flame_front_3d.py:379-381)."""
        return self._window(-math.inf, 0, 0.0)

    def flame_body_mask(self):
        """-L_burnout <= phi <= 0 (flame_front_3d.py:383-385)."""
        return self._window(-self.L_burnout, 0, 0.0)

    def ahead_band_mask(self, band_m=DX_VN_BAND_M):
        """0 < phi <= band_m (flame_front_3d.py:387-389)."""
        return self._window(0.0, 1, float(band_m))

    def front_x(self, k=0, j=0):
        """Linear zero crossing of row (k, j), +inf if none
        (flame_front_3d.py:391-402)."""
        out = C.c_double()
        _lib.call("fw_lset_front_x", self._g.ref, ptr(self._phi), int(k), int(j),
                  C.byref(out), stream_ptr())
        return out.value


# -- front speed and fluxes ------------------------------------------------------

def compute_v_n(q_into_unburned_band, rho_b, cp_s, h_bed, T_ign, T_amb):
    """Legacy 2-D front speed max(q / E_ign, 0) (flame_front_3d.py:405-434)."""
    E = rho_b * cp_s * h_bed * (T_ign - T_amb)
    if E <= 0.0:
        return (torch.zeros_like(q_into_unburned_band)
                if isinstance(q_into_unburned_band, torch.Tensor)
                else np.zeros_like(q_into_unburned_band))
    shape = tuple(q_into_unburned_band.shape)
    q3 = q_into_unburned_band.reshape((1,) + shape)
    return compute_v_n_3d(q3, rho_b, cp_s, h_bed, T_ign, T_amb).reshape(shape)


def compute_q_in_at_front_3d(q_frankman, q_dom_fwd, ahead_band_mask, q_burst_conv_2d=None):
    """Per-cell forward flux into the ahead band, plus the burst flux on each
    column's highest band cell (flame_front_3d.py:440-471); returns a new
    array."""
    m = Marshal(q_dom_fwd)
    sh = tuple(q_dom_fwd.shape)
    g = grid_for(m, sh, 1.0, 1.0, np.ones(sh[0]), np.ones(sh[0]), np.ones(sh[0]))
    out = torch.empty(sh, dtype=m.dtype, device=device())
    qb = m.dev(q_burst_conv_2d, shape=sh[1:]) if q_burst_conv_2d is not None else None
    qf = m.dev(q_frankman, shape=sh) if q_frankman is not None else None
    _lib.call("fw_q_in_at_front_3d", g.ref, ptr(qf),
              ptr(m.dev(q_dom_fwd, shape=sh)), ptr(_mask_u8(ahead_band_mask, m)), ptr(qb),
              ptr(out), stream_ptr())
    return _result(out, q_dom_fwd, m)


def compute_v_n_3d(q_in_3d, rho_b, cp_s, h_bed, T_ign, T_amb, M_local=None,
                   L_vap=L_VAP_WATER, f_dry_to_ignite=1.0):
    """v_n = max(q / max(E_sens + E_lat M, 1), 0) per cell
    (flame_front_3d.py:474-514); returns a new float64 array."""
    m = Marshal(q_in_3d)
    sh = tuple(q_in_3d.shape)
    g = grid_for(m, sh, 1.0, 1.0, np.ones(sh[0]), np.ones(sh[0]), np.ones(sh[0]))
    E_sens = rho_b * cp_s * h_bed * (T_ign - T_amb)
    if M_local is None:
        M = torch.zeros(sh, dtype=m.dtype, device=device())
        E_lat = 0.0
    else:
        M = m.dev(M_local, shape=sh)
        E_lat = rho_b * h_bed * L_vap * f_dry_to_ignite
    v = torch.empty(sh, dtype=torch.float64, device=device())
    _lib.call("fw_lset_speed", g.ref, ptr(m.dev(q_in_3d, shape=sh)), ptr(M), float(E_sens),
              float(E_lat), 0.0, 0.0, ptr(v), stream_ptr())
    return _result(v, q_in_3d, m)


def compute_q_dom_fwd_at_band(rad_solver, ahead_band_mask, q_dom_fwd_out):
    """q_dom_fwd_out = sum over ordinates with xi > 0 of w |xi| I_n inside the
    ahead band, 0 outside (flame_front_3d.py:594-630).  rad_solver: the
    drop-in dom_3d.DOMRadiationSolver after a solve."""
    m = Marshal(q_dom_fwd_out)
    h = rad_solver._handle(m.dtype)
    out = m.out(q_dom_fwd_out, shape=(rad_solver.Nz, rad_solver.Ny, rad_solver.Nx))
    _lib.call("fw_dom_forward_flux", h, ptr(_mask_u8(ahead_band_mask, m)), ptr(out),
              stream_ptr())
    m.writeback()


# -- flame geometry ----------------------------------------------------------------

def compute_phi_flame_from_state(omega, T_g, Y_fuel, dx, dy, dz_arr):
    """Signed exact Euclidean distance to the active-flame cells, sampling
    (mean dz, dy, dx); +-1e6 when no / every cell is active
    (flame_front_3d.py:756-800).  Returns a new float64 array."""
    m = Marshal(omega)
    sh = tuple(omega.shape)
    dz = np.asarray(dz_arr, np.float64)
    da, db = Grid3D._faces(dz)
    g = grid_for(m, sh, dx, dy, dz, da, db)
    phi = torch.empty(sh, dtype=torch.float64, device=device())
    work = torch.empty(sh, dtype=torch.float64, device=device())
    _lib.call("fw_phi_flame", g.ref, ptr(m.dev(omega, shape=sh)), ptr(m.dev(T_g, shape=sh)),
              ptr(m.dev(Y_fuel, shape=sh)), float(dz.mean()), ptr(phi), ptr(work),
              stream_ptr())
    return _result(phi, omega, m)


def flame_body_mask_from_phi_flame(phi_flame, band_m=0.0):
    """phi_flame <= band_m (flame_front_3d.py:803-813)."""
    m = Marshal(phi_flame)
    sh = tuple(phi_flame.shape)
    g = _grid64(sh, 1.0, 1.0, np.ones(sh[0]))
    mask = torch.empty(sh, dtype=torch.uint8, device=device())
    _lib.call("fw_phi_window_mask", g.ref, ptr(m.dev(phi_flame, dtype=torch.float64)),
              -math.inf, 0, float(band_m), sh[0], ptr(mask), stream_ptr())
    return _bool(mask, not isinstance(phi_flame, torch.Tensor))


def update_cell_age(cell_age, flame_body_mask, dt):
    """In place: flame-body cells age by dt (from 0 when inf), the others
    become inf (flame_front_3d.py:816-835)."""
    m = Marshal(cell_age)
    sh = tuple(cell_age.shape)
    g = grid_for(m, sh, 1.0, 1.0, np.ones(sh[0]), np.ones(sh[0]), np.ones(sh[0]))
    age = m.out(cell_age)
    _lib.call("fw_update_cell_age", g.ref, ptr(_mask_u8(flame_body_mask, m)), ptr(age),
              float(dt), stream_ptr())
    m.writeback()


def _result(t, like, m: Marshal):
    """A new result array in the caller's convention (NumPy for NumPy input)."""
    if isinstance(like, torch.Tensor):
        return t
    torch.cuda.current_stream().synchronize()
    m._keep.clear()
    return t.cpu().numpy()
