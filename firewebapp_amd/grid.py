"""Structured grid of the Tier-3 step (host side, one-shot, float64).

Grid3D.build: spread_3d.py:160-424 of the reference (both the segmented
path through build_z_axis_bed_atm, mesh.py:152-316, and the legacy
uniform/expanding path).  Wind helpers: boundary.py:40-111 and
boundary_conditions/outdoor_wind.py:80-106.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


# ---------------------------------------------------------------------------
# z-axis segments (mesh.py:43-147)
# ---------------------------------------------------------------------------

def _uniform(L, N):
    if N <= 0 or L <= 0.0:
        return np.empty(0)
    return np.full(N, L / N)


def _inflation(N, first_dz, growth, reverse):
    if N <= 0 or first_dz <= 0.0:
        return np.empty(0)
    if abs(growth - 1.0) < 1e-12:
        dzs = np.full(N, first_dz)
    else:
        dzs = first_dz * (growth ** np.arange(N, dtype=np.float64))
    return dzs[::-1].copy() if reverse else dzs


def _bulk(L, interface_dz, max_dz, growth):
    if L <= 0.0:
        return np.empty(0)
    cells = []
    cum = 0.0
    next_dz = min(interface_dz * growth, max_dz)
    while cum < L:
        dz = min(next_dz, max_dz)
        if cum + dz > L + 0.5 * dz:
            break
        cells.append(dz)
        cum += dz
        next_dz = dz * growth
    return np.array(cells, dtype=np.float64)


def build_z_axis_bed_atm(*, h_bed, Lz, n_z_bed, wall_bl_N=0, wall_bl_first_dz=0.0,
                         wall_bl_growth=1.3, bed_top_inner_bl_N=0,
                         bed_top_inner_bl_first_dz=0.0, bed_top_inner_bl_growth=1.3,
                         bed_top_outer_bl_N=0, bed_top_outer_bl_first_dz=0.0,
                         bed_top_outer_bl_growth=1.3, atm_max_dz=None, atm_growth=1.3,
                         atm_uniform_dz=None):
    """Segmented z axis (mesh.py:152-316); returns (dz_arr, n_z_bed)."""
    if h_bed < 0 or Lz <= h_bed or n_z_bed <= 0:
        raise ValueError(f"invalid bed/domain: h_bed={h_bed} Lz={Lz} n_z_bed={n_z_bed}")
    bed = []
    bed_used = 0.0
    if wall_bl_N > 0 and wall_bl_first_dz > 0.0:
        wall = _inflation(wall_bl_N, wall_bl_first_dz, wall_bl_growth, False)
        bed.append(wall)
        bed_used = float(wall.sum())
    top_inner = None
    top_inner_th = 0.0
    if bed_top_inner_bl_N > 0 and bed_top_inner_bl_first_dz > 0.0:
        top_inner = _inflation(bed_top_inner_bl_N, bed_top_inner_bl_first_dz,
                               bed_top_inner_bl_growth, True)
        top_inner_th = float(top_inner.sum())
    n_bulk = n_z_bed - wall_bl_N - bed_top_inner_bl_N
    if n_bulk < 0:
        raise ValueError("n_z_bed smaller than the boundary-layer cell counts")
    L_bulk = h_bed - bed_used - top_inner_th
    if L_bulk < 0:
        raise ValueError("boundary layers thicker than the bed")
    if n_bulk > 0:
        bed.append(_uniform(L_bulk, n_bulk))
    elif L_bulk > 1e-3 * h_bed:
        raise ValueError("bed thickness not covered by the boundary layers")
    if top_inner is not None:
        bed.append(top_inner)
    bed = [a for a in bed if a.size]
    bed_dz = np.concatenate(bed) if bed else np.empty(0)
    tot = bed_dz.sum()
    resid = h_bed - tot
    if abs(resid) > 1e-6 * h_bed:
        if abs(resid) > 0.05 * h_bed:
            raise ValueError("bed z segments do not add up to h_bed")
        bed_dz = bed_dz * (h_bed / tot)
    air = []
    interface_dz = bed_dz[-1]
    outer_th = 0.0
    if bed_top_outer_bl_N > 0 and bed_top_outer_bl_first_dz > 0.0:
        outer = _inflation(bed_top_outer_bl_N, bed_top_outer_bl_first_dz,
                           bed_top_outer_bl_growth, False)
        air.append(outer)
        interface_dz = outer[-1]
        outer_th = float(outer.sum())
    L_atm = Lz - h_bed - outer_th
    if atm_uniform_dz is not None and atm_uniform_dz > 0 and L_atm > 0:
        n_u = max(1, int(round(L_atm / atm_uniform_dz)))
        air.append(_uniform(n_u * atm_uniform_dz, n_u))
    elif L_atm > 0:
        if atm_max_dz is None:
            atm_max_dz = max(0.05, 4.0 * interface_dz)
        air.append(_bulk(L_atm, interface_dz, atm_max_dz, atm_growth))
    air = [a for a in air if a.size]
    air_dz = np.concatenate(air) if air else np.empty(0)
    dz_arr = np.concatenate([bed_dz, air_dz]) if air_dz.size else bed_dz.copy()
    return dz_arr, int(bed_dz.size)


@dataclass
class Grid3D:
    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    dz: float
    Lx: float
    Ly: float
    Lz: float
    n_z_bed: int
    x_mid: np.ndarray
    y_mid: np.ndarray
    z_mid: np.ndarray
    dz_arr: np.ndarray
    inv_dz_arr: np.ndarray
    z_face: np.ndarray
    d_face_above: np.ndarray
    d_face_below: np.ndarray

    @staticmethod
    def _faces(dz_arr):
        Nz = len(dz_arr)
        da = np.empty(Nz)
        db = np.empty(Nz)
        for k in range(Nz):
            da[k] = 0.5 * (dz_arr[k] + dz_arr[k + 1]) if k + 1 < Nz else 0.5 * dz_arr[k]
            db[k] = 0.5 * (dz_arr[k] + dz_arr[k - 1]) if k - 1 >= 0 else 0.5 * dz_arr[k]
        return da, db

    @classmethod
    def build(cls, Lx, Ly, Lz, dx, h_bed, n_z_bed, dy=None, dz_expansion=1.0,
              dz_first=None, bl_growth=1.0, dz_first_above=None, bl_growth_above=1.3,
              bed_refine_top=False, wall_bl_N=0, wall_bl_first_dz=0.0,
              wall_bl_growth=1.3, bed_top_inner_bl_N=0, bed_top_inner_bl_first_dz=0.0,
              bed_top_inner_bl_growth=1.3, bed_top_outer_bl_N=0,
              bed_top_outer_bl_first_dz=0.0, bed_top_outer_bl_growth=1.3,
              atm_max_dz=None, atm_growth=1.3, atm_uniform_dz=None):
        if Lz < h_bed:
            raise ValueError(f"Lz={Lz} must be >= h_bed={h_bed}")
        if dz_expansion < 1.0:
            raise ValueError("dz_expansion must be >= 1")
        if bl_growth < 1.0:
            raise ValueError("bl_growth must be >= 1")
        Nx = max(2, int(round(Lx / dx)))
        dyt = dy if dy is not None else dx
        Ny = max(1, int(round(Ly / dyt)))
        dy = Ly / Ny if Ny > 0 else 1.0
        if wall_bl_N > 0 or bed_top_inner_bl_N > 0 or bed_top_outer_bl_N > 0:
            dz_arr, n_z_bed = build_z_axis_bed_atm(
                h_bed=h_bed, Lz=Lz, n_z_bed=n_z_bed, wall_bl_N=wall_bl_N,
                wall_bl_first_dz=wall_bl_first_dz, wall_bl_growth=wall_bl_growth,
                bed_top_inner_bl_N=bed_top_inner_bl_N,
                bed_top_inner_bl_first_dz=bed_top_inner_bl_first_dz,
                bed_top_inner_bl_growth=bed_top_inner_bl_growth,
                bed_top_outer_bl_N=bed_top_outer_bl_N,
                bed_top_outer_bl_first_dz=bed_top_outer_bl_first_dz,
                bed_top_outer_bl_growth=bed_top_outer_bl_growth,
                atm_max_dz=atm_max_dz, atm_growth=atm_growth,
                atm_uniform_dz=atm_uniform_dz)
            dz_bed = float(h_bed) / max(n_z_bed, 1)
        else:
            if (n_z_bed > 0 and dz_first is not None and dz_first > 0.0
                    and bl_growth > 1.0 + 1e-12):
                lst = [float(dz_first)]
                cum = float(dz_first)
                for _ in range(n_z_bed - 1):
                    nd = lst[-1] * bl_growth
                    if cum + nd >= h_bed:
                        break
                    lst.append(nd)
                    cum += nd
                tail = h_bed - cum
                if tail > 0.0:
                    lst.append(tail)
                bed_dz = np.array(lst)
                if bed_refine_top:
                    bed_dz = bed_dz[::-1].copy()
                n_z_bed = len(bed_dz)
                dz_bed = bed_dz[-1]
            else:
                dz_bed = h_bed / max(n_z_bed, 1)
                bed_dz = np.full(n_z_bed, dz_bed)
            bl_dz = np.empty(0)
            if n_z_bed == 0 and dz_first is not None and dz_first > 0.0:
                if bl_growth <= 1.0 + 1e-12:
                    bl_dz = np.array([dz_first])
                else:
                    bl = []
                    d = float(dz_first)
                    while d < dx and len(bl) < 200:
                        bl.append(d)
                        d *= bl_growth
                    bl_dz = np.array(bl)
                dz_buf0 = dx
            else:
                dz_buf0 = dz_bed
            ab = np.empty(0)
            if (n_z_bed > 0 and dz_first_above is not None and dz_first_above > 0.0
                    and dz_first_above < dz_buf0 and bl_growth_above > 1.0 + 1e-12):
                al = [float(dz_first_above)]
                while al[-1] * bl_growth_above < dz_buf0:
                    al.append(al[-1] * bl_growth_above)
                    if len(al) >= 200:
                        break
                ab = np.array(al)
            target = Lz - h_bed - bl_dz.sum() - ab.sum()
            if dz_expansion <= 1.0 + 1e-12:
                n_buf = max(1, int(round(target / dz_buf0)))
                buf_dz = np.full(n_buf, dz_buf0)
            else:
                ratio = target * (dz_expansion - 1.0) / dz_buf0 + 1.0
                n_buf = max(1, int(np.ceil(np.log(ratio) / np.log(dz_expansion))))
                buf_dz = dz_buf0 * np.power(dz_expansion, np.arange(n_buf, dtype=np.float64))
            dz_arr = np.concatenate([bed_dz, bl_dz, ab, buf_dz])
        Nz = len(dz_arr)
        z_face = np.concatenate([[0.0], np.cumsum(dz_arr)])
        z_mid = z_face[:-1] + 0.5 * dz_arr
        da, db = cls._faces(dz_arr)
        x_mid = (np.arange(Nx) + 0.5) * (Lx / Nx)
        y_mid = (np.arange(Ny) + 0.5) * dy
        return cls(Nx=Nx, Ny=Ny, Nz=Nz, dx=Lx / Nx, dy=dy, dz=float(dz_bed), Lx=Lx,
                   Ly=Ly, Lz=float(z_face[-1]), n_z_bed=n_z_bed, x_mid=x_mid,
                   y_mid=y_mid, z_mid=z_mid, dz_arr=dz_arr, inv_dz_arr=1.0 / dz_arr,
                   z_face=z_face, d_face_above=da, d_face_below=db)


_WAF = {"open": 0.9, "shrub": 0.6}
Z_0_INLET = 0.01


def midflame_wind_speed(U_10m, terrain="open"):
    """boundary.py:49-65."""
    return U_10m * _WAF.get(terrain.lower(), 0.9)


def wind_profile_log_law(z, U_ref, z_ref=10.0, z_0=0.01):
    """boundary.py:68-111."""
    if z <= 0.0 or U_ref <= 0.0 or z_ref <= 0.0:
        return 0.0
    u_tk = U_ref / math.log((z_ref + z_0) / z_0)
    return max(0.0, u_tk * math.log((z + z_0) / z_0))


def build_u_inlet(grid: Grid3D, wind_speed_m_s: float) -> np.ndarray:
    """OutdoorWindBC.build_u_inlet, log-law profile (outdoor_wind.py:80-106)."""
    u = np.zeros((grid.Nz, grid.Ny))
    for k in range(grid.Nz):
        u[k, :] = wind_profile_log_law(grid.z_mid[k], wind_speed_m_s, z_ref=10.0,
                                       z_0=Z_0_INLET)
    return u
