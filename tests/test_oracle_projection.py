"""Oracle projection vs the golden vectors and closed forms.

* operator (matvec_out) and compatible divergence (div_out): 1e-13 relative
  (the reference sums CSR rows in column order; the stencil sum here is in a
  different order, so equality is to rounding);
* separable solve (poisson p_out): 1e-9 relative (eigh is LAPACK-dependent);
* project(): max|div(u_new) - target| within 10x the reference's residual
  (docs/kerneltest.mjs:612-627);
* closed forms: the x eigenbasis is DCT-IV; the real periodic basis
  reproduces the FFT diagonalisation; solve(L p) = p for the constant
  operator (tests/outdoor/test_fft_poisson_3d.py:125-196 of the reference);
  the ring-of-faces identity div(correction) = dt * A p.
"""
import numpy as np
import pytest

from oracle import projection as P
from tests.golden_io import arr, case_ids, cases, field, relerr


def _geo(c):
    return arr(c, "dz_arr"), arr(c, "d_face_above"), arr(c, "d_face_below")


@pytest.mark.parametrize("c", cases("poisson"), ids=case_ids("poisson"))
def test_poisson_golden(c):
    dz, da, db = _geo(c)
    s = P.SeparableLaplacian3D(c["nz"], c["ny"], c["nx"], c["dx"], c["dy"], dz, da, db)
    p = s.solve(field(c, "rhs"))
    assert relerr(p, field(c, "p_out")) <= 1e-9


@pytest.mark.parametrize("c", cases("projection"), ids=case_ids("projection"))
def test_projection_golden(c):
    dz, da, db = _geo(c)
    sh = (c["nz"], c["ny"], c["nx"])
    rho = field(c, "rho")
    coef = P.coefficients(rho, c["dx"], c["dy"], dz, da, db)
    y = P.matvec(coef, arr(c, "probe").reshape(sh))
    assert relerr(y, field(c, "matvec_out")) <= 1e-13
    uin = field(c, "u_inlet", (c["nz"], c["ny"]))
    u, v, w = field(c, "u_in"), field(c, "v_in"), field(c, "w_in")
    div = P.divergence_compatible(u, v, w, c["dx"], c["dy"], dz, uin)
    assert relerr(div, field(c, "div_out")) <= 1e-13
    sol = P.ProjectionSolver3D(*sh, c["dy"], c["dx"], dz, da, db, cg_rtol=c["cg_rtol"])
    sol.set_inlet_BC(uin)
    sol.rebuild_for_rho(rho)
    tgt = field(c, "div_target")
    sol.project(u, v, w, rho, c["dt"], div_target=tgt)
    resid = np.max(np.abs(sol.divergence(u, v, w) - tgt))
    assert resid <= max(10.0 * c["proj_resid"], 1e-9)


def test_dct4_is_the_x_eigenbasis():
    for Nx, dx in ((16, 0.3), (33, 0.1), (128, 0.1)):
        L = P.build_L_x(Nx, dx)
        lam, U = P.dct4_basis(Nx, dx)
        assert np.allclose(U.T @ U, np.eye(Nx), atol=1e-12)
        assert relerr(L @ U, U * lam[None, :]) <= 1e-12
        ev = np.linalg.eigvalsh(L)
        assert relerr(np.sort(lam), ev) <= 1e-12


def test_real_periodic_basis_matches_fft():
    for Ny in (1, 2, 5, 8):
        lam, B = P.real_periodic_basis(Ny, 0.1)
        assert np.allclose(B.T @ B, np.eye(Ny), atol=1e-12)
        L = np.zeros((Ny, Ny))
        for j in range(Ny):
            L[j, j] += -2.0 / 0.01
            L[j, (j + 1) % Ny] += 1.0 / 0.01
            L[j, (j - 1) % Ny] += 1.0 / 0.01
        if Ny == 1:
            L[:] = 0.0
        assert np.allclose(L @ B, B * lam[None, :], atol=1e-9)


def _constant_operator(Nz, Ny, Nx, dx, dy, dz, da, db):
    """Constant-density (rho = 1) operator of projection_3d.py:65-175, with
    the diagonal shift the separable inverse inverts: it divides by
    lambda - eps (fft_poisson_3d.py:137-146) while the matrix adds +eps."""
    return P.coefficients(np.ones((Nz, Ny, Nx)), dx, dy, dz, da, db,
                          eps_reg=-P.EPS_REG)


@pytest.mark.parametrize("Ny", [1, 8])
def test_separable_solve_inverts_constant_operator(Ny):
    Nz, Nx, dx, dy = 12, 20, 0.1, 0.1
    dz = np.concatenate([np.full(3, 0.04), 0.04 * 1.2 ** np.arange(1, Nz - 2)])
    zc = np.cumsum(dz) - 0.5 * dz
    da = np.empty(Nz)
    db = np.empty(Nz)
    da[:-1] = zc[1:] - zc[:-1]
    da[-1] = dz[-1]
    db[1:] = zc[1:] - zc[:-1]
    db[0] = dz[0]
    s = P.SeparableLaplacian3D(Nz, Ny, Nx, dx, dy, dz, da, db)
    coef = _constant_operator(Nz, Ny, Nx, dx, dy, dz, da, db)
    p = np.random.default_rng(3).standard_normal((Nz, Ny, Nx))
    assert relerr(s.solve(P.matvec(coef, p)), p) <= 1e-9


def test_correction_is_dt_times_operator():
    rng = np.random.default_rng(9)
    Nz, Ny, Nx = 7, 4, 9
    dz = 0.05 + 0.05 * rng.random(Nz)
    zc = np.cumsum(dz) - 0.5 * dz
    da = np.empty(Nz)
    db = np.empty(Nz)
    da[:-1] = zc[1:] - zc[:-1]
    da[-1] = dz[-1]
    db[1:] = zc[1:] - zc[:-1]
    db[0] = dz[0]
    rho = 0.3 + rng.random((Nz, Ny, Nx))
    p = rng.standard_normal((Nz, Ny, Nx))
    u, v, w = (rng.standard_normal((Nz, Ny, Nx)) for _ in range(3))
    uin = rng.standard_normal((Nz, Ny))
    d0 = P.divergence_compatible(u, v, w, 0.1, 0.1, dz, uin)
    dt = 0.01
    P.correct_velocity(u, v, w, p, rho, dt, 0.1, 0.1, dz, da)
    d1 = P.divergence_compatible(u, v, w, 0.1, 0.1, dz, uin)
    coef = P.coefficients(rho, 0.1, 0.1, dz, da, db, eps_reg=0.0)
    assert relerr(d0 - d1, dt * P.matvec(coef, p)) <= 1e-12
