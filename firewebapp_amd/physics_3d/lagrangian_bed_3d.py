"""lagrangian_bed_3d drop-in (lagrangian_bed_3d.py:147-1070 of the reference).

Buffer allocation and initialisation are one-shot host code (as in the
reference); the per-step kernels run on the GPU through libfirewave.
``DeviceBed`` keeps the particles resident in HBM for the time-step driver,
with the particle -> cell map computed once (particles never move).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..device import Marshal, device, grid_for, ptr, stream_ptr
from ..grid import Grid3D

RHO_SOLID_TRUE_GRASS = 380.0
CP_SOLID_GRASS = 1500.0
H_CONV_DEFAULT = 25.0
SAV_GRASS_DEFAULT = 2000.0
M_PARTICLE_BURNOUT = 1.0e-8
SIGMA_SB = 5.67e-8
EPS_SOLID_DEFAULT = 0.9
T_BOIL_WATER = 373.15
A_DRY_GRASS = 4.29e3
E_DRY_GRASS = 30_000.0
DRY_MODE_ARRHENIUS = 0
DRY_MODE_EQUILIBRIUM = 1
DRY_MODE_COMBINED = 2
ALIVE_FALSE = 0
ALIVE_TRUE = 1

_F64 = ("x", "y", "z", "u", "v", "w", "age", "m_solid", "m_water", "m_char",
        "T_s", "m_solid_0", "m_water_0", "sav", "m_char_max")
_STATE = ("x", "y", "z", "m_solid", "m_water", "m_char", "T_s", "m_water_0",
          "sav", "m_char_max")


class FwBedParticles(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in _STATE] + [("alive", C.c_void_p),
                                                     ("n", C.c_int32),
                                                     ("reserved", C.c_int32),
                                                     ("h_top", C.c_void_p)]


class FwBedIndex(C.Structure):
    _fields_ = [("cell", C.c_void_p), ("seg_ptr", C.c_void_p),
                ("seg_idx", C.c_void_p), ("n_seg_cells", C.c_int32),
                ("reserved", C.c_int32)]


class FwBedParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "dt", "h_conv", "rho_solid_true", "cp_solid", "eps_solid", "T_amb",
        "view_factor", "h_bed", "kappa_bed_eff", "char_ox_flux_cap",
        "char_ox_ash_exp")] + [(k, C.c_int32) for k in (
            "n_per_cell_for_split", "view_factor_geometric", "do_drying",
            "do_pyrolysis", "do_char_ox", "do_smolder", "drying_mode", "reserved")]


def allocate_bed_particle_buffers(N_max: int) -> dict:
    """Host particle buffers (lagrangian_bed_3d.py:193-218)."""
    if N_max < 0:
        raise ValueError(f"N_max must be >= 0; got {N_max}")
    buf = {k: np.zeros(N_max, dtype=np.float64) for k in _F64}
    buf["alive"] = np.zeros(N_max, dtype=np.int8)
    return buf


def initialize_bed_particles_from_alpha_s(buf, alpha_s, rho_b_dry, moisture_frac,
                                          T_amb, dx, dy, dz_arr, n_z_bed,
                                          n_per_cell, sav=SAV_GRASS_DEFAULT,
                                          i_lo=None, i_hi=None, j_lo=None,
                                          j_hi=None, mass_scale=None):
    """Slots in (k, j, i, p) order (lagrangian_bed_3d.py:221-329).

    Sub-cell positions fx = ((13 p) % n + 0.5)/n, fy = ((7 p) % n + 0.5)/n,
    fz = (p + 0.5)/n.  ``mass_scale`` (new, optional): per-slot factor on the
    dry and water mass for seeded load jitter; None reproduces the reference.
    """
    alpha_s = np.asarray(alpha_s)
    Nz, Ny, Nx = alpha_s.shape
    N_max = buf["alive"].shape[0]
    i_lo = 0 if i_lo is None else i_lo
    i_hi = Nx if i_hi is None else i_hi
    j_lo = 0 if j_lo is None else j_lo
    j_hi = Ny if j_hi is None else j_hi
    kb = min(n_z_bed, Nz)
    win = alpha_s[:kb, j_lo:min(j_hi, Ny), i_lo:min(i_hi, Nx)] > 0.0
    n_req = int(win.sum()) * n_per_cell
    if n_req > N_max:
        raise ValueError(f"Buffer too small: need {n_req} slots "
                         f"({int(win.sum())} bed cells x {n_per_cell} particles); "
                         f"have {N_max}")
    p = np.arange(n_per_cell)
    fx = ((p * 13) % n_per_cell + 0.5) / n_per_cell
    fy = ((p * 7) % n_per_cell + 0.5) / n_per_cell
    fz = (p + 0.5) / n_per_cell
    slot = 0
    for k in range(kb):
        V_cell = dx * dy * dz_arr[k]
        z_face_k = float(np.sum(dz_arr[:k]))
        jj, ii = np.nonzero(win[k])
        nc = jj.size
        if nc == 0:
            continue
        jj = (jj + j_lo)[:, None]
        ii = (ii + i_lo)[:, None]
        m_s = (rho_b_dry * V_cell) / n_per_cell
        m_w = moisture_frac * m_s
        sl = slice(slot, slot + nc * n_per_cell)
        buf["x"][sl] = ((ii + fx[None, :]) * dx).ravel()
        buf["y"][sl] = ((jj + fy[None, :]) * dy).ravel()
        buf["z"][sl] = np.broadcast_to(z_face_k + fz * dz_arr[k], (nc, n_per_cell)).ravel()
        for f in ("u", "v", "w", "age", "m_char"):
            buf[f][sl] = 0.0
        buf["alive"][sl] = ALIVE_TRUE
        if mass_scale is None:
            buf["m_solid"][sl] = m_s
            buf["m_water"][sl] = m_w
        else:
            sc = np.asarray(mass_scale)[sl]
            buf["m_solid"][sl] = m_s * sc
            buf["m_water"][sl] = m_w * sc
        buf["T_s"][sl] = T_amb
        buf["m_solid_0"][sl] = buf["m_solid"][sl]
        buf["m_water_0"][sl] = buf["m_water"][sl]
        buf["sav"][sl] = sav
        slot += nc * n_per_cell
    return slot


def build_index(cell: np.ndarray):
    """CSR inverse of a particle -> cell map, particles in slot order."""
    cell = np.asarray(cell, dtype=np.int64)
    valid = cell >= 0
    n_seg = int(cell.max()) + 1 if valid.any() else 0
    counts = np.bincount(cell[valid], minlength=n_seg)
    seg_ptr = np.zeros(n_seg + 1, dtype=np.int32)
    seg_ptr[1:] = np.cumsum(counts)
    order = np.argsort(np.where(valid, cell, np.iinfo(np.int64).max), kind="stable")
    seg_idx = order[:int(valid.sum())].astype(np.int32)
    return seg_ptr, seg_idx, n_seg


class DeviceIndex:
    def __init__(self, cell_dev: torch.Tensor):
        seg_ptr, seg_idx, n_seg = build_index(cell_dev.cpu().numpy())
        dev = device()
        self.cell = cell_dev
        self.seg_ptr = torch.as_tensor(seg_ptr, device=dev)
        self.seg_idx = torch.as_tensor(seg_idx if seg_idx.size else np.zeros(1, np.int32),
                                       device=dev)
        self.n_seg = n_seg
        self.c = FwBedIndex(cell_dev.data_ptr(), self.seg_ptr.data_ptr(),
                            self.seg_idx.data_ptr(), n_seg, 0)


class DeviceParticles:
    """Device float64 SoA of a particle buffer dict."""

    def __init__(self, buf, n=None):
        dev = device()
        self.n = int(n if n is not None else buf["alive"].shape[0])
        self.t = {}
        for k in _STATE:
            a = buf[k]
            self.t[k] = (a[:self.n] if isinstance(a, torch.Tensor) and a.is_cuda
                         else torch.as_tensor(np.ascontiguousarray(np.asarray(a)[:self.n],
                                                                   dtype=np.float64),
                                              device=dev))
        a = buf["alive"]
        self.t["alive"] = (a[:self.n] if isinstance(a, torch.Tensor) and a.is_cuda
                           else torch.as_tensor(np.ascontiguousarray(np.asarray(a)[:self.n],
                                                                     dtype=np.int8),
                                                device=dev))
        h = buf.get("h_top")
        if h is not None:
            self.t["h_top"] = (h[:self.n] if isinstance(h, torch.Tensor) and h.is_cuda
                               else torch.as_tensor(np.ascontiguousarray(
                                   np.asarray(h)[:self.n], dtype=np.float64), device=dev))
        self.c = FwBedParticles(*[self.t[k].data_ptr() for k in _STATE],
                                self.t["alive"].data_ptr(), self.n, 0,
                                self.t["h_top"].data_ptr() if h is not None else None)

    def write_back(self, buf):
        torch.cuda.current_stream().synchronize()
        for k in _STATE + ("alive",):
            a = buf[k]
            if isinstance(a, np.ndarray):
                a[:self.n] = self.t[k].cpu().numpy()


def _locate(g, P: DeviceParticles):
    dev = device()
    cs = torch.empty(max(P.n, 1), dtype=torch.int32, device=dev)
    ca = torch.empty(max(P.n, 1), dtype=torch.int32, device=dev)
    _lib.call("fw_bed_locate", g.ref, P.n, ptr(P.t["x"]), ptr(P.t["y"]), ptr(P.t["z"]),
              ptr(cs), ptr(ca), stream_ptr())
    return cs[:P.n], ca[:P.n]


def _grid(m, shape, dx, dy, dz_arr, z_face, n_z_bed=0):
    dz = np.asarray(dz_arr, np.float64)
    da, db = Grid3D._faces(dz)
    return grid_for(m, shape, dx, dy, dz, da, db, z_face=z_face, n_z_bed=n_z_bed)


def step_bed_particles(part_x, part_y, part_z, part_alive, part_m_solid,
                       part_m_water, part_m_char, part_T_s, part_m_water_0,
                       part_sav, T_g_grid, Y_O2_grid, Q_solid_ext_per_cell,
                       n_per_cell_for_split, S_pyro_per_cell, S_drying_per_cell,
                       Q_pyro_per_cell, Q_drying_per_cell, Y_F_source_per_cell,
                       Q_char_per_cell, Q_smold_per_cell, Q_g_conv_per_cell,
                       dx, dy, dz_arr, z_face, h_conv, rho_solid_true, cp_solid,
                       eps_solid, T_amb, view_factor, view_factor_geometric, h_bed,
                       kappa_bed_eff, dt, do_drying, do_pyrolysis, do_char_ox,
                       do_smolder, drying_mode, char_ox_flux_cap_W_m2,
                       char_ox_ash_exp, part_m_char_max, n_alive_out, n_burned_out,
                       diag_max_out):
    """One bed step for every particle; in place on the particle arrays and
    the 8 per-cell source fields (zeroed over the full grid first)."""
    buf = {"x": part_x, "y": part_y, "z": part_z, "alive": part_alive,
           "m_solid": part_m_solid, "m_water": part_m_water, "m_char": part_m_char,
           "T_s": part_T_s, "m_water_0": part_m_water_0, "sav": part_sav,
           "m_char_max": part_m_char_max}
    m = Marshal(T_g_grid)
    sh = tuple(T_g_grid.shape)
    g = _grid(m, sh, dx, dy, dz_arr, z_face)
    P = DeviceParticles(buf)
    cs, _ = _locate(g, P)
    ix = DeviceIndex(cs)
    srcs = [m.out(a, shape=sh) for a in (S_pyro_per_cell, S_drying_per_cell,
                                         Q_pyro_per_cell, Q_drying_per_cell,
                                         Y_F_source_per_cell, Q_char_per_cell,
                                         Q_smold_per_cell, Q_g_conv_per_cell)]
    for s_ in srcs:
        s_.zero_()
    arr = (C.c_void_p * 8)(*[s_.data_ptr() for s_ in srcs])
    dev = device()
    contrib = torch.empty(10 * max(P.n, 1), dtype=torch.float64, device=dev)
    counters = torch.zeros(2, dtype=torch.int32, device=dev)
    diag_w = torch.empty(15 * max(P.n, 1), dtype=torch.float64, device=dev)
    diag = torch.zeros(16, dtype=torch.float64, device=dev)
    prm = FwBedParams(float(dt), float(h_conv), float(rho_solid_true), float(cp_solid),
                      float(eps_solid), float(T_amb), float(view_factor), float(h_bed),
                      float(kappa_bed_eff), float(char_ox_flux_cap_W_m2),
                      float(char_ox_ash_exp), int(n_per_cell_for_split),
                      int(bool(view_factor_geometric)), int(bool(do_drying)),
                      int(bool(do_pyrolysis)), int(bool(do_char_ox)),
                      int(bool(do_smolder)), int(drying_mode), 0)
    _lib.call("fw_bed_step", g.ref, C.byref(P.c), C.byref(ix.c), C.byref(prm),
              ptr(m.dev(T_g_grid, shape=sh)), ptr(m.dev(Y_O2_grid, shape=sh)),
              ptr(m.dev(Q_solid_ext_per_cell, shape=sh)), arr, ptr(contrib),
              ptr(counters), ptr(diag_w), ptr(diag), stream_ptr())
    P.write_back(buf)
    m.writeback()
    cnt = counters.cpu().numpy()
    n_alive_out[0] = int(cnt[0])
    n_burned_out[0] = int(cnt[1])
    d = diag.cpu().numpy()
    if isinstance(diag_max_out, np.ndarray):
        diag_max_out[:16] = d[:diag_max_out.shape[0]] if diag_max_out.shape[0] < 16 else d
    else:
        diag_max_out.copy_(torch.as_tensor(d))


def _agg(part_x, part_y, part_z, part_alive, part_m_solid, part_m_water,
         part_m_char, part_T_s, dx, dy, z_face, grid_like, want_T, want_M):
    buf = {"x": part_x, "y": part_y, "z": part_z, "alive": part_alive,
           "m_solid": part_m_solid, "m_water": part_m_water,
           "m_char": part_m_char if part_m_char is not None else np.zeros_like(part_x),
           "T_s": part_T_s if part_T_s is not None else np.zeros_like(part_x),
           "m_water_0": np.zeros_like(part_x), "sav": np.zeros_like(part_x),
           "m_char_max": np.zeros_like(part_x)}
    m = Marshal(grid_like)
    sh = tuple(grid_like.shape)
    zf = np.asarray(z_face, np.float64)
    g = _grid(m, sh, dx, dy, np.diff(zf), zf)
    P = DeviceParticles(buf)
    _, ca = _locate(g, P)
    ix = DeviceIndex(ca)
    return m, g, P, ix, sh


def aggregate_particles_to_T_s_grid(part_x, part_y, part_z, part_alive,
                                    part_m_solid, part_m_water, part_m_char,
                                    part_T_s, dx, dy, z_face, T_s_grid, T_amb):
    """Mass-weighted particle T_s per cell; cells without mass unchanged."""
    m, g, P, ix, sh = _agg(part_x, part_y, part_z, part_alive, part_m_solid,
                           part_m_water, part_m_char, part_T_s, dx, dy, z_face,
                           T_s_grid, True, False)
    Tg = m.out(T_s_grid, shape=sh)
    _lib.call("fw_bed_aggregate", g.ref, C.byref(P.c), C.byref(ix.c), ptr(Tg), None,
              stream_ptr())
    m.writeback()


def aggregate_particles_to_M_local_grid(part_x, part_y, part_z, part_alive,
                                        part_m_solid, part_m_water, dx, dy, z_face,
                                        M_local_grid):
    """sum(m_water)/sum(m_solid) per cell (0 where no solid)."""
    m, g, P, ix, sh = _agg(part_x, part_y, part_z, part_alive, part_m_solid,
                           part_m_water, None, None, dx, dy, z_face, M_local_grid,
                           False, True)
    M = m.out(M_local_grid, shape=sh)
    M.zero_()
    _lib.call("fw_bed_aggregate", g.ref, C.byref(P.c), C.byref(ix.c), None, ptr(M),
              stream_ptr())
    m.writeback()


def step_horizontal_solid_conduction_scatter(part_x, part_y, part_z, part_alive,
                                             part_m_solid, part_m_water,
                                             part_m_char, part_T_s, T_s_grid,
                                             alpha_s_grid, dx, dy, z_face, k_solid,
                                             rho_solid_true, cp_solid, n_z_bed, dt):
    """5-point x/y conduction of the grid T_s on bed cells, scattered to the
    particles (lagrangian_bed_3d.py:427-497)."""
    m, g, P, ix, sh = _agg(part_x, part_y, part_z, part_alive, part_m_solid,
                           part_m_water, part_m_char, part_T_s, dx, dy, z_face,
                           T_s_grid, False, False)
    Tg = m.out(T_s_grid, shape=sh)
    T_old = Tg.clone()
    dT = torch.zeros(int(np.prod(sh)), dtype=torch.float64, device=device())
    _lib.call("fw_bed_horizontal_conduction", g.ref, C.byref(P.c), C.byref(ix.c),
              ptr(T_old), ptr(Tg), ptr(m.dev(alpha_s_grid, shape=sh)), float(k_solid),
              float(rho_solid_true), float(cp_solid), int(n_z_bed), float(dt), ptr(dT),
              stream_ptr())
    buf = {"T_s": part_T_s, "x": part_x, "y": part_y, "z": part_z, "alive": part_alive,
           "m_solid": part_m_solid, "m_water": part_m_water, "m_char": part_m_char,
           "m_water_0": part_x, "sav": part_x, "m_char_max": part_x}
    torch.cuda.current_stream().synchronize()
    if isinstance(part_T_s, np.ndarray):
        part_T_s[:P.n] = P.t["T_s"].cpu().numpy()
    del buf
    m.writeback()
