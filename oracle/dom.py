"""FP64 oracle of the discrete-ordinates radiation solve (test infrastructure).

dom_3d.py:98-473 of the reference: S4 ordinates (:98-142), upwind
step-differencing sweep (:146-211), G accumulation (:214-224), the solve
with source iteration, ground reflection and under-relaxation (:286-473).

The sweep is evaluated hyperplane by hyperplane (i'+j'+k' = s in sweep
coordinates).  Every cell reads the same neighbour values as in the
reference's (k, j, i) loop: upwind x/z neighbours lie on hyperplane s-1; the
periodic-y wrap neighbour (row Ny-1 for eta >= 0) lies on a later hyperplane
and therefore still holds the previous iteration's value, exactly as the
reference's not-yet-updated read (dom_3d.py:189-202).  The arithmetic per
cell is the reference's, so the result is bit-identical to the scalar loop
(checked against a literal loop on small grids in the tests).

S8 is not in the reference (dom_3d.py:114-127 raises); ``sn_ordinates(8)``
is the level-symmetric LQ8 set (Lathrop & Carlson 1965; Lewis & Miller 1984
Table 4-1): mu_i^2 = (1 + 6 (i-1))/21, index triplets summing to 6, three
weight classes.
"""
from __future__ import annotations

import math

import numpy as np

SIGMA_SB = 5.67e-8
KAPPA_SOOT_HOT = 0.5
OMEGA_COMB_THRESH = 1.0e-3
T_GAS_RAD_REF = 600.0
T_GAS_RAD_DT = 400.0
EPS_W_GROUND = 0.85

_OCTANTS = [(1, 1, 1), (1, 1, -1), (1, -1, 1), (-1, 1, 1),
            (1, -1, -1), (-1, 1, -1), (-1, -1, 1), (-1, -1, -1)]


def sn_ordinates(N):
    if N == 4:
        mu1 = 0.295876
        mu2 = 0.908248
        per = [(mu1, mu1, mu2), (mu1, mu2, mu1), (mu2, mu1, mu1)]
        wper = [1.0 / 3.0] * 3
    elif N == 8:
        mu = [math.sqrt((1.0 + 6.0 * i) / 21.0) for i in range(4)]
        cls_w = {(0, 0, 3): 0.1209877, (0, 1, 2): 0.0907407, (1, 1, 1): 0.0925926}
        per, wper = [], []
        for a in range(4):
            for b in range(4):
                c = 3 - a - b          # level indices (a+1)+(b+1)+(c+1) = 6
                if 0 <= c <= 3:
                    per.append((mu[a], mu[b], mu[c]))
                    wper.append(cls_w[tuple(sorted((a, b, c)))])
        tot = sum(wper)
        wper = [x / tot for x in wper]
    else:
        raise NotImplementedError(f"S{N} not implemented")
    Om, w = [], []
    for sx, sy, sz in _OCTANTS:
        for o, wo in zip(per, wper):
            Om.append((sx * o[0], sy * o[1], sz * o[2]))
            w.append(wo)
    Om = np.array(Om, dtype=np.float64)
    M = Om.shape[0]
    if N == 4:
        w = np.full(M, 4.0 * math.pi / M)
    else:
        w = np.array(w) * (4.0 * math.pi / 8.0)
    return Om, w


def sweep_one_ordinate(I_n, kappa, B, xi, eta, mu, dx, dy, dz_arr, I_left,
                       I_right, I_top, I_ground_in):
    """Hyperplane-ordered restatement of dom_3d.py:146-211 (periodic y)."""
    Nz, Ny, Nx = I_n.shape
    aix, aiy, aiz = abs(xi), abs(eta), abs(mu)
    inv_dx = aix / dx
    inv_dy = aiy / dy
    inv_dz = aiz / np.asarray(dz_arr)
    # sweep coordinates
    ii = np.arange(Nx) if xi >= 0 else np.arange(Nx - 1, -1, -1)
    jj = np.arange(Ny) if eta >= 0 else np.arange(Ny - 1, -1, -1)
    kk = np.arange(Nz) if mu >= 0 else np.arange(Nz - 1, -1, -1)
    KP, JP, IP = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    S = (KP + JP + IP).ravel()
    order = np.argsort(S, kind="stable")
    Ssorted = S[order]
    bounds = np.searchsorted(Ssorted, np.arange(Nx + Ny + Nz - 1), side="left")
    bounds = np.append(bounds, Ssorted.size)
    kpf, jpf, ipf = KP.ravel(), JP.ravel(), IP.ravel()
    for s in range(Nx + Ny + Nz - 2):
        sel = order[bounds[s]:bounds[s + 1]]
        kp, jp, ip = kpf[sel], jpf[sel], ipf[sel]
        k, j, i = kk[kp], jj[jp], ii[ip]
        # x inflow
        if xi >= 0:
            Ix = np.where(i > 0, I_n[k, j, np.maximum(i - 1, 0)], I_left)
        else:
            Ix = np.where(i < Nx - 1, I_n[k, j, np.minimum(i + 1, Nx - 1)], I_right)
        # y inflow (periodic: wrap neighbour holds its pre-sweep value)
        if eta >= 0:
            Iy = I_n[k, (j - 1) % Ny, i]
        else:
            Iy = I_n[k, (j + 1) % Ny, i]
        if mu >= 0:
            Iz = np.where(k > 0, I_n[np.maximum(k - 1, 0), j, i], I_ground_in[j, i])
        else:
            Iz = np.where(k < Nz - 1, I_n[np.minimum(k + 1, Nz - 1), j, i], I_top)
        kap = kappa[k, j, i]
        num = kap * B[k, j, i] + inv_dx * Ix + inv_dy * Iy + inv_dz[k] * Iz
        den = kap + inv_dx + inv_dy + inv_dz[k]
        I_n[k, j, i] = num / den


def sweep_one_ordinate_loop(I_n, kappa, B, xi, eta, mu, dx, dy, dz_arr,
                            I_left, I_right, I_top, I_ground_in):
    """Literal scalar loop of dom_3d.py:146-211 (small grids only)."""
    Nz, Ny, Nx = I_n.shape
    inv_dx = abs(xi) / dx
    inv_dy = abs(eta) / dy
    ir = range(Nx) if xi >= 0 else range(Nx - 1, -1, -1)
    jr = range(Ny) if eta >= 0 else range(Ny - 1, -1, -1)
    kr = range(Nz) if mu >= 0 else range(Nz - 1, -1, -1)
    for k in kr:
        inv_dz_k = abs(mu) / dz_arr[k]
        for j in jr:
            for i in ir:
                if xi >= 0:
                    Ix = I_n[k, j, i - 1] if i > 0 else I_left
                else:
                    Ix = I_n[k, j, i + 1] if i < Nx - 1 else I_right
                if eta >= 0:
                    Iy = I_n[k, j - 1, i] if j > 0 else I_n[k, Ny - 1, i]
                else:
                    Iy = I_n[k, j + 1, i] if j < Ny - 1 else I_n[k, 0, i]
                if mu >= 0:
                    Iz = I_n[k - 1, j, i] if k > 0 else I_ground_in[j, i]
                else:
                    Iz = I_n[k + 1, j, i] if k < Nz - 1 else I_top
                num = kappa[k, j, i] * B[k, j, i] + inv_dx * Ix + inv_dy * Iy + inv_dz_k * Iz
                den = kappa[k, j, i] + inv_dx + inv_dy + inv_dz_k
                I_n[k, j, i] = num / den


def radiative_properties(T_s, T_g, alpha_s, omega_comb, sigma_sav, T_amb,
                         Y_H2O=None, rho=None, bed_moisture=None,
                         kappa_gas_max=KAPPA_SOOT_HOT):
    """kappa, kappa_solid, kappa_gas, kappa_safe, B  (dom_3d.py:315-369)."""
    kappa_solid = sigma_sav * alpha_s
    if bed_moisture is not None:
        kappa_solid = kappa_solid * (1.0 + 5.0 * bed_moisture)
    omega_factor = np.clip(omega_comb / OMEGA_COMB_THRESH, 0.0, 1.0)
    T_factor = np.clip((T_g - T_GAS_RAD_REF) / T_GAS_RAD_DT, 0.0, 1.0)
    kappa_gas = kappa_gas_max * np.maximum(omega_factor, T_factor)
    if Y_H2O is not None and rho is not None:
        kappa_gas = kappa_gas + 30.0 * rho * Y_H2O
    kappa = kappa_solid + kappa_gas
    kappa_safe = np.maximum(kappa, 1.0e-3)
    ks_T = np.where(kappa > 1.0e-9, kappa_safe, 1.0)
    T_rad4 = np.where(kappa > 1.0e-9,
                      (kappa_solid * T_s ** 4 + kappa_gas * T_g ** 4) / ks_T,
                      T_amb ** 4)
    B = SIGMA_SB * T_rad4 / math.pi
    return kappa, kappa_solid, kappa_gas, kappa_safe, B


class DOMRadiationSolver:
    """dom_3d.py:227-473 restated (periodic y)."""

    def __init__(self, Nz, Ny, Nx, dy, dx, dz_arr, N_quadrature=4,
                 eps_w_ground=EPS_W_GROUND, max_source_iter=30,
                 tol_source_iter=1.0e-3, omega_relax=0.7,
                 kappa_gas_max=KAPPA_SOOT_HOT, loop_sweep=False):
        self.Nz, self.Ny, self.Nx = Nz, Ny, Nx
        self.dx, self.dy = dx, dy
        self.dz_arr = np.asarray(dz_arr, np.float64).copy()
        self.eps_w_ground = eps_w_ground
        self.max_iter = max_source_iter
        self.tol = tol_source_iter
        self.omega = omega_relax
        self.kappa_gas_max = kappa_gas_max
        self.Omega, self.weights = sn_ordinates(N_quadrature)
        self.M = self.Omega.shape[0]
        self.I_set = np.zeros((self.M, Nz, Ny, Nx))
        self.G = np.zeros((Nz, Ny, Nx))
        self.G_prev = np.zeros((Nz, Ny, Nx))
        self.sweep = sweep_one_ordinate_loop if loop_sweep else sweep_one_ordinate
        self.last_iters = 0

    def solve(self, T_s, T_g, alpha_s, omega_comb, sigma_sav, T_amb,
              q_rad_solid_out, q_rad_gas_out, T_soil_surface=None,
              q_in_soil_out=None, Y_H2O=None, rho=None,
              bed_moisture_per_cell=None):
        Nz, Ny, Nx = self.Nz, self.Ny, self.Nx
        kappa, kappa_solid, kappa_gas, kappa_safe, B = radiative_properties(
            T_s, T_g, alpha_s, omega_comb, sigma_sav, T_amb, Y_H2O, rho,
            bed_moisture_per_cell, self.kappa_gas_max)
        I_amb = SIGMA_SB * T_amb ** 4 / math.pi
        if T_soil_surface is not None:
            I_soil_emit = self.eps_w_ground * SIGMA_SB * T_soil_surface ** 4 / math.pi
        else:
            I_soil_emit = np.full((Ny, Nx), self.eps_w_ground * I_amb)
        I_ground_out = I_soil_emit.copy()
        q_in_ground = np.zeros((Ny, Nx))
        it_done = 0
        for it in range(self.max_iter):
            np.copyto(self.G_prev, self.G)
            self.G.fill(0.0)
            for n in range(self.M):
                xi, eta, mu = (float(v) for v in self.Omega[n])
                I_g_in = I_ground_out if mu > 0 else np.zeros_like(I_ground_out)
                self.sweep(self.I_set[n], kappa_safe, B, xi, eta, mu, self.dx,
                           self.dy, self.dz_arr, I_amb, I_amb, I_amb, I_g_in)
            G = np.zeros((Nz, Ny, Nx))
            for n in range(self.M):
                G += self.weights[n] * self.I_set[n]
            self.G[...] = G
            q_in_ground = np.zeros((Ny, Nx))
            for n in range(self.M):
                mu_n = float(self.Omega[n, 2])
                if mu_n < 0:
                    q_in_ground += self.weights[n] * abs(mu_n) * self.I_set[n, 0]
            I_ground_new = I_soil_emit + (1.0 - self.eps_w_ground) / math.pi * q_in_ground
            I_ground_out = self.omega * I_ground_new + (1.0 - self.omega) * I_ground_out
            it_done = it + 1
            if it > 0:
                G_max = max(self.G.max(), 1e-12)
                err = np.abs(self.G - self.G_prev).max() / G_max
                if err < self.tol:
                    break
        self.last_iters = it_done
        net = kappa * (self.G - 4.0 * math.pi * B)
        net_h = net * self.dz_arr.reshape(-1, 1, 1)
        f_solid = np.where(kappa > 1.0e-9, kappa_solid / kappa_safe, 0.0)
        f_gas = np.where(kappa > 1.0e-9, kappa_gas / kappa_safe, 0.0)
        np.copyto(q_rad_solid_out, net_h * f_solid)
        np.copyto(q_rad_gas_out, net_h * f_gas)
        if q_in_soil_out is not None:
            np.copyto(q_in_soil_out, self.eps_w_ground * q_in_ground)
