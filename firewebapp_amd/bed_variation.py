"""Seeded variation of the fuel bed and the inflow (no reference counterpart:
the reference has a uniform bed of depth h_bed, a cosine alpha_s
perturbation at most, and a static log-law inlet).

* ``fuel_depth_field``  per-column fuel depth h(j, i): a lognormal field with
  the given mean and coefficient of variation whose logarithm is a Gaussian
  random field with correlation exp(-r^2 / l^2) (correlation length l).
* ``bed_variation``     what the bed set-up needs from the depth field and
  the per-particle jitter: which bed cells hold fuel (multiplies alpha_s and
  the Eulerian pools), the per-particle load factor and the per-particle
  fuel top h_top (the column's filled height, or h_bed, times the depth
  jitter) used by the canopy attenuation of the radiative view factor.
* ``GustSchedule``      inlet gust factor g(t): node values 1 + a_m at
  t = m T, a_0 = 0 and a_m = A (2 U_m - 1) for m >= 1 with U_m drawn in order
  from PCG64(seed), joined by half-cosine ramps.  The inlet profile of a step
  starting at time t is the static log law times g(t).
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np


def fuel_depth_field(Nx: int, Ny: int, dx: float, dy: float, mean: float, cv: float,
                     corr_m: float, seed: int) -> np.ndarray:
    """(Ny, Nx) lognormal depth with sample mean of log-depth exactly
    ln(mean) - s^2 / 2 and sample std s, s^2 = ln(1 + cv^2).

    White noise from PCG64(seed) on a grid padded by 3 l in x (so x is not
    periodic) is filtered in Fourier space by the Gaussian whose
    autocorrelation is exp(-r^2 / l^2); y is periodic like the domain."""
    if cv <= 0.0:
        return np.full((Ny, Nx), float(mean))
    pad = int(math.ceil(3.0 * corr_m / dx))
    nx = Nx + 2 * pad
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((Ny, nx))
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(Ny, d=dy)
    # spectrum of exp(-r^2 / l^2) is proportional to exp(-k^2 l^2 / 4); the
    # filter is its square root
    filt = np.exp(-(ky[:, None] ** 2 + kx[None, :] ** 2) * corr_m ** 2 / 8.0)
    G = np.fft.ifft2(np.fft.fft2(w) * filt).real[:, pad:pad + Nx]
    G = (G - G.mean()) / G.std()
    s2 = math.log1p(cv * cv)
    return np.exp(math.log(mean) - 0.5 * s2 + math.sqrt(s2) * G)


def bed_variation(alpha_mask: np.ndarray, z_face: np.ndarray, n_z_bed: int, n_per_cell: int,
                  h_bed: float, depth: Optional[np.ndarray], jitter: float,
                  jitter_seed: int) -> Tuple[Optional[np.ndarray], Optional[np.ndarray],
                                             Optional[np.ndarray]]:
    """(fill, m_scale, h_top).

    alpha_mask: (Nz, Ny, Nx) bool, the bed cells of the uniform bed.
    depth: (Ny, Nx) fuel depth per column (None: uniform h_bed).  Bed cell k
    of a column is filled (fill 1) when its mid-height lies below the
    column's depth, the lowest layer always: whole cells, so every particle
    keeps the uniform bed's share of its cell's volume and volumetric heat.
    The column's fuel top (h_top of its particles) is the top face of its
    last filled cell.
    jitter: per-particle factors 1 + jitter U(-1, 1) on the load, then on
    the fuel top, drawn from PCG64(jitter_seed) in slot order.
    Returns None for each part that is the reference's uniform bed."""
    Nz, Ny, Nx = alpha_mask.shape
    fill, m_scale, h_top = None, None, None
    if depth is not None:
        zf = np.asarray(z_face, np.float64)
        zmid = 0.5 * (zf[:n_z_bed] + zf[1:n_z_bed + 1])
        filled = zmid[:, None, None] < depth[None, :, :]
        filled[0] = True
        fill = np.zeros((Nz, Ny, Nx))
        fill[:n_z_bed] = np.where(alpha_mask[:n_z_bed] & filled, 1.0, 0.0)
        top = zf[filled.sum(axis=0)]                       # (Ny, Nx)
        sel = fill[:n_z_bed] > 0.0
        h_top = np.repeat(np.broadcast_to(top, (n_z_bed, Ny, Nx))[sel], n_per_cell)
        n = int(sel.sum()) * n_per_cell
    else:
        n = int(alpha_mask[:n_z_bed].sum()) * n_per_cell
    if jitter > 0.0:
        rng = np.random.default_rng(jitter_seed)
        m_scale = 1.0 + jitter * rng.uniform(-1.0, 1.0, n)
        dj = 1.0 + jitter * rng.uniform(-1.0, 1.0, n)
        h_top = h_bed * dj if h_top is None else h_top * dj
    return fill, m_scale, h_top


class GustSchedule:
    """g(t) = 1 + a_m + (a_{m+1} - a_m) (1 - cos(pi f)) / 2, m = floor(t / T),
    f = t / T - m; nodes drawn lazily, in order, so g is reproducible."""

    def __init__(self, amplitude: float, period_s: float, seed: int):
        self.A, self.T = float(amplitude), float(period_s)
        self._rng = np.random.default_rng(seed)
        self._a = [0.0]

    @property
    def enabled(self) -> bool:
        return self.A > 0.0 and self.T > 0.0

    def _node(self, m: int) -> float:
        while len(self._a) <= m:
            self._a.append(self.A * (2.0 * float(self._rng.random()) - 1.0))
        return self._a[m]

    def factor(self, t: float) -> float:
        if not self.enabled:
            return 1.0
        q = t / self.T
        m = int(math.floor(q))
        f = q - m
        a0, a1 = self._node(m), self._node(m + 1)
        return 1.0 + a0 + (a1 - a0) * 0.5 * (1.0 - math.cos(math.pi * f))
