"""FP64 oracle of the pressure projection (test infrastructure only).

* variable-density 7-point operator: projection_3d.py:65-175
  (_fill_var_density_data, alpha_g = 1, y periodic), applied matrix-free;
* compatible divergence: projection_3d.py:847-892;
* separable constant-coefficient preconditioner: fft_poisson_3d.py:45-228
  (eigh of L_x and of the symmetrised L_z, FFT in y);
* project(): projection_3d.py:677-844, fft_pcg branch :756-791 -- BiCGSTAB
  preconditioned by the separable solve, x0 = previous p, rtol on ||b||_2,
  then the forward-difference velocity correction :818-842.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import eigh
from scipy.sparse.linalg import LinearOperator, bicgstab

EPS_REG = 1.0e-6


def coefficients(rho, dx, dy, dz_arr, d_above, d_below, eps_reg=EPS_REG):
    """(cxm, cxp, cym, cyp, czm, czp, diag) of projection_3d.py:65-175.

    Coefficients of neighbours that do not exist are zero; the Dirichlet
    faces at the outlet (i = Nx-1) and the top (k = Nz-1) enter only the
    diagonal.
    """
    Nz, Ny, Nx = rho.shape
    ir = 1.0 / np.maximum(rho, 0.01)
    dx2 = dx * dx
    dy2 = dy * dy
    z = np.zeros_like(ir)
    cxm, cxp, cym, cyp, czm, czp = (z.copy() for _ in range(6))
    diag = np.full_like(ir, eps_reg)
    if Nx > 1:
        fx = 0.5 * (ir[:, :, :-1] + ir[:, :, 1:])           # face i+1/2
        cxp[:, :, :-1] = 1.0 * fx / dx2
        cxm[:, :, 1:] = 1.0 * fx / dx2
        diag[:, :, :-1] += -cxp[:, :, :-1]
        diag[:, :, 1:] += -cxm[:, :, 1:]
        diag[:, :, -1] += -(2.0 * 1.0 * ir[:, :, -1] / dx2)
    if Ny > 1:
        fy = 0.5 * (ir + np.roll(ir, -1, axis=1))            # face j+1/2
        cyp[...] = 1.0 * fy / dy2
        cym[...] = 1.0 * np.roll(fy, 1, axis=1) / dy2
        diag += -cyp - cym
    if Nz > 1:
        inv_dz = (1.0 / dz_arr).reshape(-1, 1, 1)
        da = d_above.reshape(-1, 1, 1)
        db = d_below.reshape(-1, 1, 1)
        fz = 0.5 * (ir[:-1] + ir[1:])                        # face k+1/2
        czp[:-1] = 1.0 * fz * inv_dz[:-1] / da[:-1]
        czm[1:] = 1.0 * fz * inv_dz[1:] / db[1:]
        diag[:-1] += -czp[:-1]
        diag[1:] += -czm[1:]
        diag[-1] += -(2.0 * 1.0 * ir[-1] * inv_dz[-1] * inv_dz[-1])
    return cxm, cxp, cym, cyp, czm, czp, diag


def matvec(coef, x):
    cxm, cxp, cym, cyp, czm, czp, diag = coef
    y = diag * x
    y[:, :, 1:] += cxm[:, :, 1:] * x[:, :, :-1]
    y[:, :, :-1] += cxp[:, :, :-1] * x[:, :, 1:]
    if x.shape[1] > 1:
        y += cym * np.roll(x, 1, axis=1) + cyp * np.roll(x, -1, axis=1)
    y[1:] += czm[1:] * x[:-1]
    y[:-1] += czp[:-1] * x[1:]
    return y


def divergence_compatible(u, v, w, dx, dy, dz_arr, u_inlet):
    """projection_3d.py:847-892 (alpha_g None, periodic y, w_inlet_zmin 0)."""
    Nz, Ny, Nx = u.shape
    div = np.zeros((Nz, Ny, Nx))
    dzk = dz_arr.reshape(-1, 1, 1)
    div[:, :, 1:] += (u[:, :, 1:] - u[:, :, :-1]) / dx
    div[:, :, 0] += (u[:, :, 0] - u_inlet) / dx
    if Ny > 1:
        div += (v - np.roll(v, 1, axis=1)) / dy
    div[1:] += (w[1:] - w[:-1]) / dzk[1:]
    div[0] += (w[0] - 0.0) / dzk[0]
    return div


def correct_velocity(u, v, w, p, rho, dt, dx, dy, dz_arr, d_above):
    """projection_3d.py:818-842, in place."""
    Nz, Ny, Nx = u.shape
    ir = 1.0 / np.maximum(rho, 0.01)
    if Nx > 2:
        irx = 0.5 * (ir[:, :, :-1] + ir[:, :, 1:])
        u[:, :, :-1] -= dt * (p[:, :, 1:] - p[:, :, :-1]) / dx * irx
        u[:, :, -1] += 2.0 * dt * p[:, :, -1] / dx * ir[:, :, -1]
    if Ny > 1:
        iry = 0.5 * (ir + np.roll(ir, -1, axis=1))
        v -= dt * (np.roll(p, -1, axis=1) - p) / dy * iry
    if Nz > 2:
        da = d_above[:-1].reshape(-1, 1, 1)
        irz = 0.5 * (ir[:-1] + ir[1:])
        w[:-1] -= dt * (p[1:] - p[:-1]) / da * irz
        w[-1] += 2.0 * dt * p[-1] / dz_arr[-1] * ir[-1]


# ---------------------------------------------------------------------------
# separable preconditioner  (fft_poisson_3d.py)
# ---------------------------------------------------------------------------


def build_L_x(Nx, dx):
    L = np.zeros((Nx, Nx))
    inv_dx2 = 1.0 / (dx * dx)
    for i in range(Nx):
        if 1 <= i <= Nx - 2:
            L[i, i] = -2.0 * inv_dx2
            L[i, i + 1] = inv_dx2
            L[i, i - 1] = inv_dx2
        elif i == 0:
            L[0, 0] = -1.0 * inv_dx2
            L[0, 1] = 1.0 * inv_dx2
        else:
            L[Nx - 1, Nx - 1] = -3.0 * inv_dx2
            L[Nx - 1, Nx - 2] = 1.0 * inv_dx2
    return L


def build_L_z(Nz, dz_arr, d_above, d_below):
    L = np.zeros((Nz, Nz))
    for k in range(Nz):
        inv_dz_k = 1.0 / dz_arr[k]
        if 1 <= k <= Nz - 2:
            ca = inv_dz_k / d_above[k]
            cb = inv_dz_k / d_below[k]
            L[k, k] = -ca - cb
            L[k, k + 1] = ca
            L[k, k - 1] = cb
        elif k == 0:
            ca = inv_dz_k / d_above[0]
            L[0, 0] = -ca
            L[0, 1] = ca
        else:
            cb = inv_dz_k / d_below[Nz - 1]
            face = 2.0 * inv_dz_k * inv_dz_k
            L[Nz - 1, Nz - 1] = -face - cb
            L[Nz - 1, Nz - 2] = cb
    return L


class SeparableLaplacian3D:
    """fft_poisson_3d.py:45-228 restated."""

    def __init__(self, Nz, Ny, Nx, dx, dy, dz_arr, d_above, d_below,
                 eps_reg=EPS_REG):
        self.Nz, self.Ny, self.Nx = Nz, Ny, Nx
        dz_arr = np.asarray(dz_arr, dtype=np.float64)
        L_x = build_L_x(Nx, dx)
        L_z = build_L_z(Nz, dz_arr, d_above, d_below)
        D = np.sqrt(dz_arr)
        L_zs = (D[:, None] * L_z) * (1.0 / D)[None, :]
        self.lambda_x, self.U_x = eigh(L_x)
        self.lambda_z, self.U_z = eigh(L_zs)
        j = np.arange(Ny)
        self.lambda_y = (2.0 / (dy * dy)) * (np.cos(2.0 * np.pi * j / Ny) - 1.0)
        lt = (self.lambda_z[:, None, None] + self.lambda_y[None, :, None]
              + self.lambda_x[None, None, :]) - eps_reg
        self.lambda_total_inv = 1.0 / lt
        self.D = D
        self.D_inv = 1.0 / D

    def solve(self, rhs):
        Nz, Ny, Nx = self.Nz, self.Ny, self.Nx
        q = rhs * self.D[:, None, None]
        q_hat = (q.reshape(Nz * Ny, Nx) @ self.U_x).reshape(Nz, Ny, Nx)
        q_hat = np.fft.fft(q_hat, axis=1)
        q_hat = (self.U_z.T @ q_hat.reshape(Nz, Ny * Nx)).reshape(Nz, Ny, Nx)
        q_hat = q_hat * self.lambda_total_inv
        q = (self.U_z @ q_hat.reshape(Nz, Ny * Nx)).reshape(Nz, Ny, Nx)
        q = np.fft.ifft(q, axis=1).real
        q = (q.reshape(Nz * Ny, Nx) @ self.U_x.T).reshape(Nz, Ny, Nx)
        return q * self.D_inv[:, None, None]


def dct4_basis(Nx, dx):
    """Closed-form eigenpairs of build_L_x: DCT-IV vectors
    cos(pi (m+1/2)(i+1/2)/Nx) sqrt(2/Nx), lambda_m = -(4/dx^2) sin^2(pi(2m+1)/(4Nx))."""
    i = np.arange(Nx)
    m = np.arange(Nx)
    U = np.sqrt(2.0 / Nx) * np.cos(np.pi * (m[None, :] + 0.5) * (i[:, None] + 0.5) / Nx)
    lam = -(4.0 / (dx * dx)) * np.sin(np.pi * (2 * m + 1) / (4.0 * Nx)) ** 2
    return lam, U


def real_periodic_basis(Ny, dy):
    """Real orthonormal eigenbasis of the periodic 1-D Laplacian: columns
    1, cos(2 pi m j/Ny), sin(2 pi m j/Ny) (and (-1)^j for even Ny), with
    lambda = (2/dy^2)(cos(2 pi m/Ny) - 1).  Equal to the FFT diagonalisation
    of fft_poisson_3d.py:210-223 because lambda(m) = lambda(Ny - m)."""
    j = np.arange(Ny)
    cols, lams = [], []
    lam = lambda m: (2.0 / (dy * dy)) * (np.cos(2.0 * np.pi * m / Ny) - 1.0)
    cols.append(np.full(Ny, 1.0 / np.sqrt(Ny)))
    lams.append(lam(0))
    for m in range(1, (Ny - 1) // 2 + 1):
        cols.append(np.sqrt(2.0 / Ny) * np.cos(2.0 * np.pi * m * j / Ny))
        lams.append(lam(m))
        cols.append(np.sqrt(2.0 / Ny) * np.sin(2.0 * np.pi * m * j / Ny))
        lams.append(lam(m))
    if Ny % 2 == 0 and Ny > 1:
        cols.append(((-1.0) ** j) / np.sqrt(Ny))
        lams.append(lam(Ny // 2))
    return np.array(lams), np.stack(cols, axis=1)


# ---------------------------------------------------------------------------
# project()
# ---------------------------------------------------------------------------


class ProjectionSolver3D:
    """fft_pcg ProjectionSolver3D restated (projection_3d.py:214-844)."""

    def __init__(self, Nz, Ny, Nx, dy, dx, dz_arr, d_above, d_below,
                 cg_rtol=1.0e-6):
        self.shape = (Nz, Ny, Nx)
        self.dx, self.dy = dx, dy
        self.dz_arr = np.asarray(dz_arr, np.float64)
        self.d_above = np.asarray(d_above, np.float64)
        self.d_below = np.asarray(d_below, np.float64)
        self.cg_rtol = cg_rtol
        self.sep = SeparableLaplacian3D(Nz, Ny, Nx, dx, dy, self.dz_arr,
                                        self.d_above, self.d_below)
        self.u_inlet = np.zeros((Nz, Ny))
        self.p_prev = None
        self.coef = None
        self.last_iters = 0

    def set_inlet_BC(self, u_inlet):
        self.u_inlet = np.asarray(u_inlet, np.float64)

    def rebuild_for_rho(self, rho):
        self.coef = coefficients(rho, self.dx, self.dy, self.dz_arr,
                                 self.d_above, self.d_below)

    def divergence(self, u, v, w):
        return divergence_compatible(u, v, w, self.dx, self.dy, self.dz_arr,
                                     self.u_inlet)

    def project(self, u, v, w, rho, dt, div_target=None):
        div = self.divergence(u, v, w)
        if div_target is not None:
            div = div - div_target
        rhs = (div / dt).reshape(-1)
        shp = self.shape
        A = LinearOperator((rhs.size, rhs.size),
                           matvec=lambda x: matvec(self.coef, x.reshape(shp)).reshape(-1))
        M = LinearOperator((rhs.size, rhs.size),
                           matvec=lambda r: self.sep.solve(r.reshape(shp)).reshape(-1))
        x0 = self.p_prev if self.p_prev is not None else np.zeros_like(rhs)
        n = [0]

        def cb(_):
            n[0] += 1
        p, info = bicgstab(A, rhs, x0=x0, M=M, rtol=self.cg_rtol, maxiter=200,
                           callback=cb)
        if info != 0:
            raise RuntimeError(f"FFT-PCG BiCGSTAB failed to converge: info={info}")
        self.last_iters = n[0]
        self.p_prev = p.copy()
        p = p.reshape(shp)
        correct_velocity(u, v, w, p, rho, dt, self.dx, self.dy, self.dz_arr,
                         self.d_above)
        return p
