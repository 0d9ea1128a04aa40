// Transport stencils: MUSCL advection, species, gas energy, turbulent
// diffusion, solid vertical conduction, drag, tentative velocity, sponge,
// equation of state, soil column.
//
// One thread per cell; a workgroup is a run of <=256 x-cells of one (k, j)
// row, so every neighbour read in x is a coalesced load of the same or the
// adjacent cache line and the y/z neighbours of a row are reused from L2 by
// the workgroups of the neighbouring rows.  All kernels are out-of-place
// (the reference double-buffers through dY/du temporaries, so the results
// are identical) and every expression keeps the reference's operation order
// so that the FP64 instantiation reproduces the reference bit-for-bit.
#include <map>
#include <mutex>

#include "fw_common.h"
#include "gas_solid.h"
#include "edc_cell.h"

namespace fw {

// ---------------------------------------------------------------------------
// MUSCL scalar advection: rhs -= u.grad(phi)   (muscl_3d.py:67-171)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_advect_muscl(GridArgs g, const T* __restrict__ phi, const T* __restrict__ u,
               const T* __restrict__ v, const T* __restrict__ w,
               T* __restrict__ rhs, T phi_inlet) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t c = cidx(g, k, j, i);
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  const T dx = T(g.dx), dy = T(g.dy);
  const T da = T(g.d_above[k]), db = T(g.d_below[k]);
  const T inv_db = T(g.inv_db[k]);
  const T inv_da = T(g.inv_da[k]);
  const T uc = u[c], vc = v[c], wc = w[c], pc = phi[c];
  const T xL = (i == 0) ? phi_inlet : phi[c - 1];
  const T xR = (i == Nx - 1) ? pc : phi[c + 1];
  const T zL = (k == 0) ? pc : phi[c - sz];
  const T zR = (k == Nz - 1) ? pc : phi[c + sz];
  T flux_x;
  if (i >= 2 && i <= Nx - 3) {
    T fp = muscl_face<T>(phi[c - 1], pc, phi[c + 1], phi[c + 2], uc);
    T fm = muscl_face<T>(phi[c - 2], phi[c - 1], pc, phi[c + 1], uc);
    flux_x = uc * (fp - fm) / dx;
  } else {
    flux_x = (uc >= T(0)) ? uc * (pc - xL) / dx : uc * (xR - pc) / dx;
  }
  const size_t rm2 = c + (size_t)(wrap(j - 2, Ny) - j) * sy;
  const size_t rm1 = c + (size_t)(wrap(j - 1, Ny) - j) * sy;
  const size_t rp1 = c + (size_t)(wrap(j + 1, Ny) - j) * sy;
  const size_t rp2 = c + (size_t)(wrap(j + 2, Ny) - j) * sy;
  T fyp = muscl_face<T>(phi[rm1], pc, phi[rp1], phi[rp2], vc);
  T fym = muscl_face<T>(phi[rm2], phi[rm1], pc, phi[rp1], vc);
  T flux_y = vc * (fyp - fym) / dy;
  T flux_z;
  if (k >= 2 && k <= Nz - 3) {
    T fp = muscl_face<T>(phi[c - sz], pc, phi[c + sz], phi[c + 2 * sz], wc);
    T fm = muscl_face<T>(phi[c - 2 * sz], phi[c - sz], pc, phi[c + sz], wc);
    flux_z = wc * (fp - fm) / (T(0.5) * (da + db));
  } else {
    flux_z = (wc >= T(0)) ? wc * (pc - zL) * inv_db : wc * (zR - pc) * inv_da;
  }
  rhs[c] -= flux_x + flux_y + flux_z;
}

// ---------------------------------------------------------------------------
// species transport (species_3d.py:40-175)
// ---------------------------------------------------------------------------
// New value of species Y at cell (i, j, k) (species_3d.py:40-175): MUSCL
// advection (x/z MUSCL on 2..N-3, upwind with ghosts elsewhere, y MUSCL
// periodic), D central diffusion, source / rho, clip [0, 1], then y_max.
// kind 0: the source is S[c]; kind 1/2/3: the dilution source of the time
// step (spread_3d.py:2990-3004) from S_pyro part sp and S_dry part sd.
// irc = 1/rho and rhz = 1/(0.5 (d_above + d_below)) are used by the float32
// build only (quot); float64 divides as the reference does.
template <typename T>
__device__ __forceinline__ T species_cell(const GridArgs& g, int i, int k, const Nbr& n,
                                          const T* __restrict__ Y, T rc, T irc, T rhz,
                                          T ui, T vi, T wi, T Sgiven, T sp, T sd, int kind,
                                          T dt, T D, T Y_inlet, T y_max) {
  const int Nx = g.nx, Nz = g.nz;
  const T inv_dx = T(g.inv_dx), inv_dy = T(g.inv_dy);
  const T inv_dx2 = inv_dx * inv_dx, inv_dy2 = inv_dy * inv_dy;
  const T Yi = at(Y, n.c);
  const T YL_x = (i == 0) ? Y_inlet : at(Y, n.xm1);
  const T YR_x = (i == Nx - 1) ? Yi : at(Y, n.xp1);
  const T YL_z = (k == 0) ? Yi : at(Y, n.zm1);
  const T YR_z = (k == Nz - 1) ? Yi : at(Y, n.zp1);
  T flux_x;
  if (i >= 2 && i <= Nx - 3) {
    T fp = muscl_face<T>(YL_x, Yi, YR_x, at(Y, n.xp2), ui);
    T fm = muscl_face<T>(at(Y, n.xm2), YL_x, Yi, YR_x, ui);
    flux_x = ui * (fp - fm) * inv_dx;
  } else {
    flux_x = (ui >= T(0)) ? ui * (Yi - YL_x) * inv_dx : ui * (YR_x - Yi) * inv_dx;
  }
  const T Yjm1 = at(Y, n.ym1), Yjp1 = at(Y, n.yp1);
  T fyp = muscl_face<T>(Yjm1, Yi, Yjp1, at(Y, n.yp2), vi);
  T fym = muscl_face<T>(at(Y, n.ym2), Yjm1, Yi, Yjp1, vi);
  T flux_y = vi * (fyp - fym) * inv_dy;
  const T inv_da = T(g.inv_da[k]);
  const T inv_db = T(g.inv_db[k]);
  const T inv_dz = T(g.inv_dz[k]);
  T flux_z;
  if (k >= 2 && k <= Nz - 3) {
    T fp = muscl_face<T>(YL_z, Yi, YR_z, at(Y, n.zp2), wi);
    T fm = muscl_face<T>(at(Y, n.zm2), YL_z, Yi, YR_z, wi);
    flux_z = quot(wi * (fp - fm), T(0.5) * (T(g.d_above[k]) + T(g.d_below[k])), rhz);
  } else {
    flux_z = (wi >= T(0)) ? wi * (Yi - YL_z) * inv_db : wi * (YR_z - Yi) * inv_da;
  }
  T adv = -(flux_x + flux_y + flux_z);
  T d2x = (YR_x - T(2) * Yi + YL_x) * inv_dx2;
  T d2y = (Yjp1 - T(2) * Yi + Yjm1) * inv_dy2;
  T d2z = ((YR_z - Yi) * inv_da - (Yi - YL_z) * inv_db) * inv_dz;
  T diff = D * (d2x + d2y + d2z);
  T Se;
  if (kind == 0) {
    Se = Sgiven;
  } else {
    const T st = sp + sd;
    Se = kind == 1 ? sp - Yi * st : (kind == 2 ? -Yi * st : sd - Yi * st);
  }
  T src = quot(Se, rc, irc);
  T dY = (adv + diff + src) * dt;
  T nv = Yi + dY;
  if (nv < T(0)) nv = T(0);
  else if (nv > T(1)) nv = T(1);
  if (nv > y_max) nv = y_max;
  return nv;
}

// 1 / (0.5 (d_above + d_below)) of level k, for quot() in float32
template <typename T>
__device__ __forceinline__ T rhz_of(const GridArgs& g, int k) {
  return recip(T(0.5) * (T(g.d_above[k]) + T(g.d_below[k])));
}

template <typename T>
__global__ void __launch_bounds__(256)
k_species(GridArgs g, const T* __restrict__ Y, T* __restrict__ Yo,
          const T* __restrict__ rho, const T* __restrict__ u,
          const T* __restrict__ v, const T* __restrict__ w,
          const T* __restrict__ S, const T* __restrict__ Sp,
          const T* __restrict__ Sd, int kind, T dt, T D, T Y_inlet, T y_max) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const Nbr n = nbr_offsets<T>(g, i, j, k);
  const T rc = at(rho, n.c);
  at(Yo, n.c) = species_cell<T>(g, i, k, n, Y, rc, recip(rc), rhz_of<T>(g, k), at(u, n.c),
                                at(v, n.c), at(w, n.c), kind == 0 ? at(S, n.c) : T(0),
                                kind == 0 ? T(0) : at(Sp, n.c), kind == 0 ? T(0) : at(Sd, n.c),
                                kind, dt, D, Y_inlet, y_max);
}

// The three species of a sub-step (fuel, O2, H2O) in one pass: rho, u, v, w,
// the two particle sources and the stencil offsets serve all three.  A
// thread updates CZ vertically adjacent cells (the grid's z dimension is
// ceil(Nz / CZ)): their z stencils overlap, and the compiler shares the
// common loads, differences and limited slopes between them (the same
// expressions on the same read-only data), and the launch has fewer
// wavefronts than cells, so they are all resident at once.
template <typename T>
__device__ __forceinline__ void species3_at(const GridArgs& g, int i, int j, int k,
                                            const T* __restrict__ Y1,
                                            const T* __restrict__ Y2,
                                            const T* __restrict__ Y3, T* __restrict__ O1,
                                            T* __restrict__ O2, T* __restrict__ O3,
                                            const T* __restrict__ rho,
                                            const T* __restrict__ u,
                                            const T* __restrict__ v,
                                            const T* __restrict__ w,
                                            const T* __restrict__ Sp,
                                            const T* __restrict__ Sd, T dt, T D, T in1,
                                            T in2, T in3, T mx1, T mx2, T mx3) {
  const Nbr n = nbr_offsets<T>(g, i, j, k);
  const T rc = at(rho, n.c), ui = at(u, n.c), vi = at(v, n.c), wi = at(w, n.c);
  const T sp = at(Sp, n.c), sd = at(Sd, n.c);
  const T irc = recip(rc), rhz = rhz_of<T>(g, k);
  at(O1, n.c) = species_cell<T>(g, i, k, n, Y1, rc, irc, rhz, ui, vi, wi, T(0), sp, sd, 1,
                                dt, D, in1, mx1);
  at(O2, n.c) = species_cell<T>(g, i, k, n, Y2, rc, irc, rhz, ui, vi, wi, T(0), sp, sd, 2,
                                dt, D, in2, mx2);
  at(O3, n.c) = species_cell<T>(g, i, k, n, Y3, rc, irc, rhz, ui, vi, wi, T(0), sp, sd, 3,
                                dt, D, in3, mx3);
}

template <typename T, int CZ>
__global__ void __launch_bounds__(256)
k_species3(GridArgs g, const T* __restrict__ Y1, const T* __restrict__ Y2,
           const T* __restrict__ Y3, T* __restrict__ O1, T* __restrict__ O2,
           T* __restrict__ O3, const T* __restrict__ rho, const T* __restrict__ u,
           const T* __restrict__ v, const T* __restrict__ w, const T* __restrict__ Sp,
           const T* __restrict__ Sd, T dt, T D, T in1, T in2, T in3, T mx1, T mx2,
           T mx3, const double* __restrict__ sb) {
  if (sb) dt = T(sb[SB_DT_SUB]);
  int i, j, kb;
  cell_ijk(i, j, kb);
  if (i >= g.nx) return;
#pragma unroll
  for (int q = 0; q < CZ; ++q) {
    const int k = kb * CZ + q;
    if (k < g.nz)
      species3_at<T>(g, i, j, k, Y1, Y2, Y3, O1, O2, O3, rho, u, v, w, Sp, Sd, dt, D, in1,
                     in2, in3, mx1, mx2, mx3);
  }
}

// ---------------------------------------------------------------------------
// gas energy: MUSCL advection + alpha_th FV diffusion + clamps
// (spread_3d.py:3827-3897)
// ---------------------------------------------------------------------------
// The new temperature of cell (i, j, k) = c.
template <typename T>
__device__ __forceinline__ T gas_energy_cell(const GridArgs& g, int i, int j, int k, size_t c,
                                             const T* __restrict__ Tg,
                                             const T* __restrict__ u,
                                             const T* __restrict__ v,
                                             const T* __restrict__ w, T dt, T alpha_th,
                                             T T_amb) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  const T dx = T(g.dx), dy = T(g.dy);
  const T da = T(g.d_above[k]), db = T(g.d_below[k]);
  const T inv_db = T(g.inv_db[k]);
  const T inv_da = T(g.inv_da[k]);
  const T uc = u[c], vc = v[c], wc = w[c], pc = Tg[c];
  const T xL = (i == 0) ? T_amb : Tg[c - 1];
  const T xR = (i == Nx - 1) ? pc : Tg[c + 1];
  const T zL = (k == 0) ? pc : Tg[c - sz];
  const T zR = (k == Nz - 1) ? pc : Tg[c + sz];
  T flux_x;
  if (i >= 2 && i <= Nx - 3) {
    T fp = muscl_face<T>(Tg[c - 1], pc, Tg[c + 1], Tg[c + 2], uc);
    T fm = muscl_face<T>(Tg[c - 2], Tg[c - 1], pc, Tg[c + 1], uc);
    flux_x = uc * (fp - fm) / dx;
  } else {
    flux_x = (uc >= T(0)) ? uc * (pc - xL) / dx : uc * (xR - pc) / dx;
  }
  const size_t rm2 = c + (size_t)(wrap(j - 2, Ny) - j) * sy;
  const size_t rm1 = c + (size_t)(wrap(j - 1, Ny) - j) * sy;
  const size_t rp1 = c + (size_t)(wrap(j + 1, Ny) - j) * sy;
  const size_t rp2 = c + (size_t)(wrap(j + 2, Ny) - j) * sy;
  const T Tjm1 = Tg[rm1], Tjp1 = Tg[rp1];
  T fyp = muscl_face<T>(Tjm1, pc, Tjp1, Tg[rp2], vc);
  T fym = muscl_face<T>(Tg[rm2], Tjm1, pc, Tjp1, vc);
  T flux_y = vc * (fyp - fym) / dy;
  T flux_z;
  if (k >= 2 && k <= Nz - 3) {
    T fp = muscl_face<T>(Tg[c - sz], pc, Tg[c + sz], Tg[c + 2 * sz], wc);
    T fm = muscl_face<T>(Tg[c - 2 * sz], Tg[c - sz], pc, Tg[c + sz], wc);
    flux_z = wc * (fp - fm) / (T(0.5) * (da + db));
  } else {
    flux_z = (wc >= T(0)) ? wc * (pc - zL) * inv_db : wc * (zR - pc) * inv_da;
  }
  T dT = T(0) - (flux_x + flux_y + flux_z);
  // x diffusion: interior / inlet ghost T_amb / outlet zero-flux
  const T inv_dx2 = T(g.inv_dx2);
  if (i >= 1 && i <= Nx - 2)
    dT += alpha_th * (Tg[c + 1] - T(2) * pc + Tg[c - 1]) * inv_dx2;
  else if (i == 0)
    dT += alpha_th * (Tg[c + 1] - T(2) * pc + T_amb) * inv_dx2;
  if (i == Nx - 1) dT += alpha_th * (Tg[c - 1] - pc) * inv_dx2;
  // y diffusion (periodic)
  dT += alpha_th * (Tjp1 - T(2) * pc + Tjm1) / (dy * dy);
  // z diffusion
  if (k >= 1 && k <= Nz - 2) {
    dT += alpha_th * ((Tg[c + sz] - pc) * inv_da - (pc - Tg[c - sz]) * inv_db) *
          T(g.inv_dz[k]);
  }
  if (k == 0) dT += alpha_th * ((Tg[c + sz] - pc) / da) / T(g.dz[0]);
  if (k == Nz - 1) dT += alpha_th * (-(pc - Tg[c - sz]) / db) / T(g.dz[Nz - 1]);
  T nv = pc + dT * dt;
  nv = nv < T_amb ? T_amb : nv;  // np.maximum then np.minimum
  nv = nv > T(10000) ? T(10000) : nv;
  return nv;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_gas_energy(GridArgs g, const T* __restrict__ Tg, T* __restrict__ To,
             const T* __restrict__ u, const T* __restrict__ v,
             const T* __restrict__ w, T dt, T alpha_th, T T_amb) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const size_t c = cidx(g, k, j, i);
  To[c] = gas_energy_cell<T>(g, i, j, k, c, Tg, u, v, w, dt, alpha_th, T_amb);
}

// The same update with the first sub-step's EDC of the cell fused in, as
// k_turb_diff_edc does for the later sub-steps (whole 256-cell segments per
// workgroup, checked by the host; the species' segment summaries from
// k_edc_summary).
template <typename T>
__global__ void __launch_bounds__(256)
k_gas_energy_edc(GridArgs g, const T* __restrict__ Tg, T* __restrict__ To,
                 const T* __restrict__ u, const T* __restrict__ v,
                 const T* __restrict__ w, T dt, T alpha_th, T T_amb, size_t n,
                 const T* __restrict__ rho, T* __restrict__ Yf, T* __restrict__ YO2,
                 const T* __restrict__ kt, const T* __restrict__ et,
                 const T* __restrict__ Yw, T* __restrict__ omega,
                 const int* __restrict__ summary, EdcParams P) {
  int i, j, k;
  cell_ijk(i, j, k);
  const size_t c = cidx(g, k, j, i);
  const EdcCellIn<T> in = edc_load<T>(true, c, rho, Yf, YO2, kt, et, Yw);
  const T tn = gas_energy_cell<T>(g, i, j, k, c, Tg, u, v, w, dt, alpha_th, T_amb);
  edc_segment<T>(n, (int)(c / EDC_SEG), tn, in, To, Yf, YO2, omega, summary, P);
}

// ---------------------------------------------------------------------------
// turbulent diffusion pass (turbulence_3d.py:920-949, one sub-step)
// ---------------------------------------------------------------------------
// One explicit turbulent-diffusion sub-step of f at cell c with zero-flux
// ghosts (turbulence_3d.py:920-949); optional clip folded from the caller
// (spread_3d.py:3079-3084), lo > hi disables it.
template <typename T>
__device__ __forceinline__ T turb_cell(const GridArgs& g, int i, int j, int k, size_t c,
                                       const T* __restrict__ f, T D_t, T dt_sub, T lo, T hi) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  const T inv_dx2 = T(g.inv_dx2), inv_dy2 = T(g.inv_dy2);
  const T inv_dz = T(g.inv_dz[k]);
  const T inv_da = T(g.inv_da[k]), inv_db = T(g.inv_db[k]);
  const T fc = f[c];
  const T fxL = (i == 0) ? fc : f[c - 1];
  const T fxR = (i == Nx - 1) ? fc : f[c + 1];
  const T fzL = (k == 0) ? fc : f[c - sz];
  const T fzR = (k == Nz - 1) ? fc : f[c + sz];
  const T fyp = f[c + (size_t)(wrap(j + 1, Ny) - j) * sy];
  const T fym = f[c + (size_t)(wrap(j - 1, Ny) - j) * sy];
  T d2x = (fxR - T(2) * fc + fxL) * inv_dx2;
  T d2y = (fyp - T(2) * fc + fym) * inv_dy2;
  T d2z = ((fzR - fc) * inv_da - (fc - fzL) * inv_db) * inv_dz;
  T df = D_t * (d2x + d2y + d2z) * dt_sub;
  T nv = fc + df;
  if (lo <= hi) {
    nv = nv < lo ? lo : nv;
    nv = nv > hi ? hi : nv;
  }
  return nv;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_turb_diff(GridArgs g, const T* __restrict__ f, T* __restrict__ fo,
            const T* __restrict__ nu_t, T sc_t, T dt_sub, T lo, T hi,
            const double* __restrict__ sb, int sb_slot) {
  if (sb) dt_sub = T(sb[sb_slot]);
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const size_t c = cidx(g, k, j, i);
  fo[c] = turb_cell<T>(g, i, j, k, c, f, nu_t[c] / sc_t, dt_sub, lo, hi);
}

// The temperature pass with the next EDC update of the same cell fused in:
// the workgroup is one whole 256-cell segment (256 threads along x, Nx a
// multiple of 256, checked by the host, so no thread returns early), the
// diffused temperature goes straight into the EDC update and its result to
// fo.  The species' segment summaries come from the species pass.
template <typename T>
__global__ void __launch_bounds__(256)
k_turb_diff_edc(GridArgs g, const T* __restrict__ f, T* __restrict__ fo,
                const T* __restrict__ nu_t, T sc_t, T dt_sub, T lo, T hi,
                const double* __restrict__ sb, int sb_slot, size_t n,
                const T* __restrict__ rho, T* __restrict__ Yf, T* __restrict__ YO2,
                const T* __restrict__ kt, const T* __restrict__ et,
                const T* __restrict__ Yw, T* __restrict__ omega,
                const int* __restrict__ summary, EdcParams P) {
  if (sb) dt_sub = T(sb[sb_slot]);
  P.dt = sb_get(sb, SB_DT_SUB, P.dt);
  P.relax = sb_get(sb, SB_EDC_RELAX, P.relax);
  int i, j, k;
  cell_ijk(i, j, k);
  const size_t c = cidx(g, k, j, i);
  // the EDC inputs issued with the stencil's loads
  const EdcCellIn<T> in = edc_load<T>(true, c, rho, Yf, YO2, kt, et, Yw);
  const T tn = turb_cell<T>(g, i, j, k, c, f, nu_t[c] / sc_t, dt_sub, lo, hi);
  edc_segment<T>(n, (int)(c / EDC_SEG), tn, in, fo, Yf, YO2, omega, summary, P);
}

// The same pass for three fields sharing nu_t (the species of a sub-step).
// GS: the sub-step's gas energy update and vertical solid conduction fused
// in (pointwise on the cell; they read its new Y_H2O, and nothing else this
// pass writes).
template <typename T, bool GS>
__global__ void __launch_bounds__(256)
k_turb_diff3(GridArgs g, const T* __restrict__ f1, const T* __restrict__ f2,
             const T* __restrict__ f3, T* __restrict__ o1, T* __restrict__ o2,
             T* __restrict__ o3, const T* __restrict__ nu_t, T sc_t, T dt_sub, T lo1,
             T hi1, T lo2, T hi2, T lo3, T hi3, int* __restrict__ edc_summary,
             const double* __restrict__ sb, int sb_slot, GasSolidArgs<T> gsa) {
  const T dt_gs = sb ? T(sb[SB_DT_SUB]) : gsa.dt_sub;
  if (sb) dt_sub = T(sb[sb_slot]);
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const size_t c = cidx(g, k, j, i);
  const T D_t = nu_t[c] / sc_t;
  const T y1 = turb_cell<T>(g, i, j, k, c, f1, D_t, dt_sub, lo1, hi1);
  const T y2 = turb_cell<T>(g, i, j, k, c, f2, D_t, dt_sub, lo2, hi2);
  const T y3 = turb_cell<T>(g, i, j, k, c, f3, D_t, dt_sub, lo3, hi3);
  o1[c] = y1;
  o2[c] = y2;
  o3[c] = y3;
  if (edc_summary) {
    // the EDC segment summary of the new (fuel, O2, H2O), as k_edc_summary
    // would form it: the launch is 256 threads along x and Nx % 256 == 0
    // (checked by the host), so the workgroup is one whole 256-cell segment
    // and no thread returned above
    __shared__ int wmax[4];
    const int last = block_last_true(edc_resets<T>(y1, y2), wmax);
    if (last < 0) {
      if (threadIdx.x == 0) edc_summary[c / EDC_SEG] = -1;
    } else if ((int)threadIdx.x == last) {
      edc_summary[c / EDC_SEG] = edc_own_flag<T>(y3) ? 1 : 0;
    }
  }
  if constexpr (GS) gas_solid_cell<T>(g, i, k, c, gsa, dt_gs, y3);
}

// Max of nu_t over interior cells 1<=k<=Nz-2, 1<=i<=Nx-2, all j, starting
// from 0 and skipping NaN exactly like the reference's `if v > max` scan
// (turbulence_3d.py:890-897).
template <typename T>
__global__ void __launch_bounds__(256)
k_nu_t_max_interior(GridArgs g, const T* __restrict__ nu_t, double* out) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double m = 0.0;
  const size_t nrow = (size_t)g.ny * g.nz;
  for (size_t r = blockIdx.x; r < nrow; r += gridDim.x) {
    const int k = (int)(r / g.ny);
    if (k < 1 || k > g.nz - 2) continue;
    const T* row = nu_t + r * g.nx;
    for (int i = tid + 1; i <= g.nx - 2; i += blockDim.x) {
      double x = (double)row[i];
      if (x > m) m = x;
    }
  }
  red[tid] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s && red[tid + s] > red[tid]) red[tid] = red[tid + s];
    __syncthreads();
  }
  if (tid == 0) out[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// solid vertical conduction (solid_conduction_3d.py:56-127)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_solid_cond(GridArgs g, const T* __restrict__ Ts, T* __restrict__ To,
             const T* __restrict__ alpha, T k_solid, T inv_rho_cp, T dt,
             T lo, T hi) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const int Nz = g.nz;
  const size_t c = cidx(g, k, j, i);
  const size_t sz = (size_t)g.nx * g.ny;
  const T a = alpha[c];
  T nv = Ts[c];
  if (a > T(0)) {
    T flux_up = T(0), flux_dn = T(0);
    if (k < Nz - 1) {
      T aa = alpha[c + sz];
      if (aa > T(0)) {
        T a_eff = T(2) * a * aa / (a + aa);
        flux_up = k_solid * a_eff * (Ts[c + sz] - Ts[c]) / T(g.d_above[k]);
      }
    }
    if (k > 0) {
      T ab = alpha[c - sz];
      if (ab > T(0)) {
        T a_eff = T(2) * a * ab / (a + ab);
        flux_dn = k_solid * a_eff * (Ts[c - sz] - Ts[c]) / T(g.d_below[k]);
      }
    }
    T src = (flux_up + flux_dn) / T(g.dz[k]);
    T dT = dt * src * inv_rho_cp / a;
    nv = Ts[c] + dT;
  }
  // folded T_s clamps of the sub-step (spread_3d.py:3195-3196); lo > hi off
  if (lo <= hi) {
    nv = nv > hi ? hi : nv;
    nv = nv < lo ? lo : nv;
  }
  To[c] = nv;
}

// ---------------------------------------------------------------------------
// drag (drag_3d.py:99-144)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_drag(GridArgs g, const T* __restrict__ u, const T* __restrict__ v,
       const T* __restrict__ w, const T* __restrict__ rho,
       const T* __restrict__ alpha, T sigma_sav, T C_D, T* __restrict__ Fx,
       T* __restrict__ Fy, T* __restrict__ Fz) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const size_t c = cidx(g, k, j, i);
  const T a = alpha[c];
  if (a <= T(0)) {
    Fx[c] = T(0); Fy[c] = T(0); Fz[c] = T(0);
    return;
  }
  const T ui = u[c], vi = v[c], wi = w[c];
  const T speed = sqrt(ui * ui + vi * vi + wi * wi);
  const T sigma2 = sigma_sav * sigma_sav;
  const T oma = T(1) - a;
  const T K_visc = T(50.0) * T(3.0e-5) * sigma2 * a * a / (oma * oma * oma);
  const T K_quad = C_D * sigma_sav * a * T(0.5) * rho[c] * speed;
  const T coeff = -(K_visc + K_quad);
  Fx[c] = coeff * ui; Fy[c] = coeff * vi; Fz[c] = coeff * wi;
}

// ---------------------------------------------------------------------------
// tentative velocity (momentum_3d.py:41-276)
// ---------------------------------------------------------------------------
template <typename T>
struct Vel3 { T u, v, w; };

template <typename T>
__global__ void __launch_bounds__(256)
k_momentum(GridArgs g, const T* __restrict__ u, const T* __restrict__ v,
           const T* __restrict__ w, T* __restrict__ uo, T* __restrict__ vo,
           T* __restrict__ wo, const T* __restrict__ rho,
           const T* __restrict__ Tg, const T* __restrict__ Fx,
           const T* __restrict__ Fy, const T* __restrict__ Fz,
           const T* __restrict__ u_in, const T* __restrict__ v_in,
           const T* __restrict__ w_in, T dt, T T_amb) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t c = cidx(g, k, j, i);
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  const T inv_dx = T(g.inv_dx), inv_dy = T(g.inv_dy);
  const T inv_dx2 = inv_dx * inv_dx, inv_dy2 = inv_dy * inv_dy;
  const T ui = u[c], vi = v[c], wi = w[c];
  const T rho_i = rho[c];
  const T nu_i = T(1.8e-5) / rho_i;
  const size_t rm2 = c + (size_t)(wrap(j - 2, Ny) - j) * sy;
  const size_t rm1 = c + (size_t)(wrap(j - 1, Ny) - j) * sy;
  const size_t rp1 = c + (size_t)(wrap(j + 1, Ny) - j) * sy;
  const size_t rp2 = c + (size_t)(wrap(j + 2, Ny) - j) * sy;
  const bool mx = (i >= 2 && i <= Nx - 3), mz = (k >= 2 && k <= Nz - 3);
  const T inv_da = T(g.inv_da[k]), inv_db = T(g.inv_db[k]);
  const T inv_dz = T(g.inv_dz[k]);
  const T inv_dz_eff = T(1.0 / (0.5 * (g.d_above[k] + g.d_below[k])));
  const T* F[3] = {u, v, w};
  const T* Gin[3] = {u_in, v_in, w_in};
  T res[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T* f = F[q];
    const T fi = f[c];
    const T xL = (i == 0) ? Gin[q][(size_t)k * Ny + j] : f[c - 1];
    const T xR = (i == Nx - 1) ? fi : f[c + 1];
    const T zL = (k == 0) ? T(0) : f[c - sz];
    const T zR = (k == Nz - 1) ? fi : f[c + sz];
    T dfdx, dfdz;
    if (mx) {
      T fp = muscl_face<T>(f[c - 1], fi, f[c + 1], f[c + 2], ui);
      T fm = muscl_face<T>(f[c - 2], f[c - 1], fi, f[c + 1], ui);
      dfdx = (fp - fm) * inv_dx;
    } else {
      dfdx = (ui >= T(0)) ? (fi - xL) * inv_dx : (xR - fi) * inv_dx;
    }
    const T fjm1 = f[rm1], fjp1 = f[rp1];
    T fyp = muscl_face<T>(fjm1, fi, fjp1, f[rp2], vi);
    T fym = muscl_face<T>(f[rm2], fjm1, fi, fjp1, vi);
    T dfdy = (fyp - fym) * inv_dy;
    if (mz) {
      T fp = muscl_face<T>(f[c - sz], fi, f[c + sz], f[c + 2 * sz], wi);
      T fm = muscl_face<T>(f[c - 2 * sz], f[c - sz], fi, f[c + sz], wi);
      dfdz = (fp - fm) * inv_dz_eff;
    } else {
      dfdz = (wi >= T(0)) ? (fi - zL) * inv_db : (zR - fi) * inv_da;
    }
    T adv = -(ui * dfdx + vi * dfdy + wi * dfdz);
    T d2x = (xR - T(2) * fi + xL) * inv_dx2;
    T d2y = (fjp1 - T(2) * fi + fjm1) * inv_dy2;
    T d2z = ((zR - fi) * inv_da - (fi - zL) * inv_db) * inv_dz;
    T visc = nu_i * (d2x + d2y + d2z);
    res[q] = adv + visc;
  }
  const T buoy_w = T(9.81) * (Tg[c] - T_amb) / T_amb;
  uo[c] = ui + (res[0] + Fx[c] / rho_i) * dt;
  vo[c] = vi + (res[1] + Fy[c] / rho_i) * dt;
  wo[c] = wi + (res[2] + buoy_w + Fz[c] / rho_i) * dt;
}

// outflow sponge (momentum_3d.py:292-317), in place
template <typename T>
__global__ void __launch_bounds__(256)
k_sponge(GridArgs g, T* __restrict__ u, const T* __restrict__ u_target,
         const T* __restrict__ sigma_x, const T* __restrict__ Y_F, T Y_F_skip,
         T dt) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const T sig = sigma_x[i];
  if (!(sig > T(0))) return;
  const size_t c = cidx(g, k, j, i);
  if (Y_F[c] < Y_F_skip) {
    const T ut = u_target[(size_t)k * g.ny + j];
    u[c] += sig * dt * (ut - u[c]);
  }
}

// rho = P0 / (R_air * max(T_g, T_amb))   (spread_3d.py:2042)
template <typename T>
__global__ void __launch_bounds__(256)
k_eos(size_t n, const T* __restrict__ Tg, T* __restrict__ rho, T T_amb) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const T t = Tg[c];
  const T tm = (t > T_amb || t != t) ? t : T_amb;
  rho[c] = T(101325.0) / (T(8.314 / 0.029) * tm);
}

// ---------------------------------------------------------------------------
// soil column, Gauss-Seidel in k (soil_3d.py:105-153); one thread per (j,i)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_soil(int nsoil, int ncol, T* __restrict__ Ts, const T* __restrict__ q_in,
       const double* __restrict__ sdz, const double* __restrict__ sda,
       const double* __restrict__ sdb, T dt, T k_s, T rho_cp_inv, T eps_s,
       T T_amb) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  for (int k = 0; k < nsoil; ++k) {
    const T Tk = Ts[(size_t)k * ncol + c];
    T q_top, q_bot;
    if (k == 0) {
      const T T4 = (Tk * Tk) * (Tk * Tk);
      q_top = q_in[c] - eps_s * T(5.67e-8) * T4;
    } else {
      q_top = k_s * (Ts[(size_t)(k - 1) * ncol + c] - Tk) * T(1.0 / sda[k]);
    }
    if (k == nsoil - 1)
      q_bot = k_s * (Tk - T_amb) * T(1.0 / sdb[k]);
    else
      q_bot = k_s * (Tk - Ts[(size_t)(k + 1) * ncol + c]) * T(1.0 / sdb[k]);
    Ts[(size_t)k * ncol + c] = Tk + dt * (q_top - q_bot) * rho_cp_inv / T(sdz[k]);
  }
}

// ---------------------------------------------------------------------------
// host implementations
// ---------------------------------------------------------------------------
template <typename T>
static int advect_impl(const fw_grid* g, const void* phi, const void* u,
                       const void* v, const void* w, void* rhs,
                       double phi_inlet, hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_advect_muscl<T>, L.grid, L.block, 0, s, grid_args(g),
                     (const T*)phi, (const T*)u, (const T*)v, (const T*)w,
                     (T*)rhs, T(phi_inlet));
  FW_CHECK_LAUNCH("advect_scalar_muscl");
  return FW_OK;
}

template <typename T>
static int species_impl(const fw_grid* g, const void* Y, void* Yo,
                        const void* rho, const void* u, const void* v,
                        const void* w, const void* S, const void* Sp,
                        const void* Sd, int kind, double dt, double D,
                        double Y_inlet, double y_max, hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_species<T>, L.grid, L.block, 0, s, grid_args(g),
                     (const T*)Y, (T*)Yo, (const T*)rho, (const T*)u,
                     (const T*)v, (const T*)w, (const T*)S, (const T*)Sp,
                     (const T*)Sd, kind, T(dt), T(D), T(Y_inlet), T(y_max));
  FW_CHECK_LAUNCH("species_transport");
  return FW_OK;
}


// Cells per thread of k_species3 per grid shape (fw_tune_species_cz; 2 when
// the shape has no entry).  Every CZ computes each cell with the same
// arithmetic, so the choice changes speed only.
namespace {
struct ShapeKey {
  int nx, ny, nz, dtype;
  bool operator<(const ShapeKey& o) const {
    if (nx != o.nx) return nx < o.nx;
    if (ny != o.ny) return ny < o.ny;
    if (nz != o.nz) return nz < o.nz;
    return dtype < o.dtype;
  }
};
std::mutex g_cz_mu;
std::map<ShapeKey, int> g_cz;
}  // namespace

static int species_cz(const fw_grid* g) {
  std::lock_guard<std::mutex> lock(g_cz_mu);
  const auto it = g_cz.find(ShapeKey{g->nx, g->ny, g->nz, g->dtype});
  return it == g_cz.end() ? 2 : it->second;
}

template <typename T, int CZ>
static int species3_cz(const fw_grid* g, const void* const* Y, void* const* Yo,
                       const void* rho, const void* u, const void* v, const void* w,
                       const void* Sp, const void* Sd, double dt, double D,
                       const double* y_in, const double* y_max, hipStream_t s,
                       const double* sb);

template <typename T>
static int species3_impl(const fw_grid* g, const void* const* Y, void* const* Yo,
                         const void* rho, const void* u, const void* v, const void* w,
                         const void* Sp, const void* Sd, double dt, double D,
                         const double* y_in, const double* y_max, hipStream_t s,
                         const double* sb) {
  switch (species_cz(g)) {
    case 1:
      return species3_cz<T, 1>(g, Y, Yo, rho, u, v, w, Sp, Sd, dt, D, y_in, y_max, s, sb);
    case 4:
      return species3_cz<T, 4>(g, Y, Yo, rho, u, v, w, Sp, Sd, dt, D, y_in, y_max, s, sb);
    default:
      return species3_cz<T, 2>(g, Y, Yo, rho, u, v, w, Sp, Sd, dt, D, y_in, y_max, s, sb);
  }
}

template <typename T, int CZ>
static int species3_cz(const fw_grid* g, const void* const* Y, void* const* Yo,
                       const void* rho, const void* u, const void* v, const void* w,
                       const void* Sp, const void* Sd, double dt, double D,
                       const double* y_in, const double* y_max, hipStream_t s,
                       const double* sb) {
  Launch L = cell_launch(g);
  L.grid.z = (g->nz + CZ - 1) / CZ;
  // the event-recording launch costs about twice the host time of a plain
  // one: use it only while a timer asked for the events (never for a launch
  // with a step scalar block, which a graph capture records)
  const KernelEvents ev = sb ? KernelEvents{} : timer_next(FW_TIMER_SPECIES3);
  if (ev.start)
    hipExtLaunchKernelGGL((k_species3<T, CZ>), L.grid, L.block, 0, s, ev.start, ev.stop, 0,
                          grid_args(g), (const T*)Y[0],
                          (const T*)Y[1], (const T*)Y[2], (T*)Yo[0], (T*)Yo[1], (T*)Yo[2],
                          (const T*)rho, (const T*)u, (const T*)v, (const T*)w,
                          (const T*)Sp, (const T*)Sd, T(dt), T(D), T(y_in[0]), T(y_in[1]),
                          T(y_in[2]), T(y_max[0]), T(y_max[1]), T(y_max[2]), sb);
  else
    hipLaunchKernelGGL((k_species3<T, CZ>), L.grid, L.block, 0, s, grid_args(g),
                       (const T*)Y[0], (const T*)Y[1], (const T*)Y[2], (T*)Yo[0], (T*)Yo[1],
                       (T*)Yo[2], (const T*)rho, (const T*)u, (const T*)v, (const T*)w,
                       (const T*)Sp, (const T*)Sd, T(dt), T(D), T(y_in[0]), T(y_in[1]),
                       T(y_in[2]), T(y_max[0]), T(y_max[1]), T(y_max[2]), sb);
  FW_CHECK_LAUNCH("species_transport3");
  return FW_OK;
}

int species3_launch(const fw_grid* g, const void* const* Y, void* const* Yo, const void* rho,
                    const void* u, const void* v, const void* w, const void* Sp,
                    const void* Sd, double dt, double D, const double* y_in,
                    const double* y_max, hipStream_t s, const double* sb) {
  return FW_DISPATCH(g, species3_impl, g, Y, Yo, rho, u, v, w, Sp, Sd, dt, D, y_in, y_max, s,
                     sb);
}

template <typename T>
static int turb3_impl(const fw_grid* g, const void* const* f, void* const* fo,
                      const void* nu_t, double sc_t, double dt_sub, const double* lo,
                      const double* hi, int* edc_summary, hipStream_t s, const double* sb,
                      const GasSolidParams* gs) {
  Launch L = cell_launch(g);
  if (edc_summary && (g->nx % EDC_SEG != 0 || L.block.x != EDC_SEG))
    return fw_set_error(FW_E_CONFIG, "turbulent_diffusion3: the EDC summary needs Nx %% 256 == 0");
  if (gs) {
    const void* reads[] = {f[0], f[1], f[2], nu_t, gs->T_s, gs->alpha_s, gs->rho,
                           gs->S_pyro_part, gs->S_dry_part, gs->Q_g_conv, gs->q_rad_gas};
    const void* writes[] = {fo[0], fo[1], fo[2], gs->T_g, gs->Q_comb_out};
    FW_REQUIRE(gs->T_g && gs->T_s && gs->T_s_out, "turbulent_diffusion3: bad gas-solid fields");
    for (const void* q : reads)
      FW_REQUIRE(q != gs->T_s_out, "turbulent_diffusion3: T_s_out aliases an input");
    for (const void* q : writes)
      FW_REQUIRE(q != gs->T_s_out, "turbulent_diffusion3: T_s_out aliases an output");
    hipLaunchKernelGGL((k_turb_diff3<T, true>), L.grid, L.block, 0, s, grid_args(g),
                       (const T*)f[0], (const T*)f[1], (const T*)f[2], (T*)fo[0], (T*)fo[1],
                       (T*)fo[2], (const T*)nu_t, T(sc_t), T(dt_sub), T(lo[0]), T(hi[0]),
                       T(lo[1]), T(hi[1]), T(lo[2]), T(hi[2]), edc_summary, sb, (int)SB_DSC,
                       gas_solid_args<T>(*gs));
  } else {
    hipLaunchKernelGGL((k_turb_diff3<T, false>), L.grid, L.block, 0, s, grid_args(g),
                       (const T*)f[0], (const T*)f[1], (const T*)f[2], (T*)fo[0], (T*)fo[1],
                       (T*)fo[2], (const T*)nu_t, T(sc_t), T(dt_sub), T(lo[0]), T(hi[0]),
                       T(lo[1]), T(hi[1]), T(lo[2]), T(hi[2]), edc_summary, sb, (int)SB_DSC,
                       GasSolidArgs<T>{});
  }
  FW_CHECK_LAUNCH("turbulent_diffusion3");
  return FW_OK;
}

int turb3_launch(const fw_grid* g, const void* const* f, void* const* fo, const void* nu_t,
                 double sc_t, double dt_sub, const double* lo, const double* hi,
                 int* edc_summary, hipStream_t s, const double* sb, const GasSolidParams* gs) {
  return FW_DISPATCH(g, turb3_impl, g, f, fo, nu_t, sc_t, dt_sub, lo, hi, edc_summary, s, sb,
                     gs);
}

template <typename T>
static int gas_energy_impl(const fw_grid* g, const void* Tg, void* To,
                           const void* u, const void* v, const void* w,
                           double dt, double alpha_th, double T_amb,
                           hipStream_t s, const EdcFuseArgs* e) {
  Launch L = cell_launch(g);
  if (e) {
    FW_REQUIRE(g->nx % EDC_SEG == 0 && L.block.x == EDC_SEG && e->summary,
               "gas_energy: the fused EDC needs whole 256-cell rows and summaries");
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    hipLaunchKernelGGL(k_gas_energy_edc<T>, L.grid, L.block, 0, s, grid_args(g),
                       (const T*)Tg, (T*)To, (const T*)u, (const T*)v, (const T*)w, T(dt),
                       T(alpha_th), T(T_amb), n, (const T*)e->rho, (T*)e->Y_fuel,
                       (T*)e->Y_O2, (const T*)e->k_turb, (const T*)e->eps_turb,
                       (const T*)e->Y_H2O, (T*)e->omega_out, e->summary, e->P);
  } else {
    hipLaunchKernelGGL(k_gas_energy<T>, L.grid, L.block, 0, s, grid_args(g),
                       (const T*)Tg, (T*)To, (const T*)u, (const T*)v,
                       (const T*)w, T(dt), T(alpha_th), T(T_amb));
  }
  FW_CHECK_LAUNCH("gas_energy");
  return FW_OK;
}

int gas_energy_launch(const fw_grid* g, const void* Tg, void* To, const void* u,
                      const void* v, const void* w, double dt, double alpha_th,
                      double T_amb, hipStream_t s, const EdcFuseArgs* edc) {
  return FW_DISPATCH(g, gas_energy_impl, g, Tg, To, u, v, w, dt, alpha_th, T_amb, s, edc);
}

template <typename T>
static int turb_diff_pass_impl(const fw_grid* g, const void* f, void* fo,
                               const void* nu_t, double sc_t, double dt_sub,
                               double lo, double hi, hipStream_t s, const double* sb,
                               const EdcFuseArgs* e) {
  Launch L = cell_launch(g);
  if (e) {
    FW_REQUIRE(g->nx % EDC_SEG == 0 && L.block.x == EDC_SEG && e->summary,
               "turbulent_diffusion: the fused EDC needs whole 256-cell rows and summaries");
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    hipLaunchKernelGGL(k_turb_diff_edc<T>, L.grid, L.block, 0, s, grid_args(g),
                       (const T*)f, (T*)fo, (const T*)nu_t, T(sc_t), T(dt_sub), T(lo), T(hi),
                       sb, (int)SB_DPR, n, (const T*)e->rho, (T*)e->Y_fuel, (T*)e->Y_O2,
                       (const T*)e->k_turb, (const T*)e->eps_turb, (const T*)e->Y_H2O,
                       (T*)e->omega_out, e->summary, e->P);
  } else {
    hipLaunchKernelGGL(k_turb_diff<T>, L.grid, L.block, 0, s, grid_args(g),
                       (const T*)f, (T*)fo, (const T*)nu_t, T(sc_t), T(dt_sub),
                       T(lo), T(hi), sb, (int)SB_DPR);
  }
  FW_CHECK_LAUNCH("turbulent_diffusion");
  return FW_OK;
}

int turb_pass_launch(const fw_grid* g, const void* f, void* fo, const void* nu_t, double sc_t,
                     double dt_sub, double lo, double hi, hipStream_t s, const double* sb,
                     const EdcFuseArgs* edc) {
  return FW_DISPATCH(g, turb_diff_pass_impl, g, f, fo, nu_t, sc_t, dt_sub, lo, hi, s, sb,
                     edc);
}

template <typename T>
static int nu_t_max_impl(const fw_grid* g, const void* nu_t, double* work,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_nu_t_max_interior<T>, dim3(FW_REDUCE_BLOCKS),
                     dim3(256), 0, s, grid_args(g), (const T*)nu_t, work);
  FW_CHECK_LAUNCH("nu_t_max_interior");
  return FW_OK;
}

int nu_t_max_partials(const fw_grid* g, const void* nu_t, double* work, hipStream_t s) {
  return FW_DISPATCH(g, nu_t_max_impl, g, nu_t, work, s);
}

template <typename T>
static int solid_cond_impl(const fw_grid* g, const void* Ts, void* To,
                           const void* alpha, double k_solid, double rho_solid,
                           double cp_solid, double dt, double lo, double hi,
                           hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_solid_cond<T>, L.grid, L.block, 0, s, grid_args(g),
                     (const T*)Ts, (T*)To, (const T*)alpha, T(k_solid),
                     T(1.0 / (rho_solid * cp_solid)), T(dt), T(lo), T(hi));
  FW_CHECK_LAUNCH("solid_conduction_vertical");
  return FW_OK;
}

template <typename T>
static int drag_impl(const fw_grid* g, const void* u, const void* v,
                     const void* w, const void* rho, const void* alpha,
                     double sigma_sav, double C_D, void* Fx, void* Fy, void* Fz,
                     hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_drag<T>, L.grid, L.block, 0, s, grid_args(g),
                     (const T*)u, (const T*)v, (const T*)w, (const T*)rho,
                     (const T*)alpha, T(sigma_sav), T(C_D), (T*)Fx, (T*)Fy,
                     (T*)Fz);
  FW_CHECK_LAUNCH("drag_force");
  return FW_OK;
}

template <typename T>
static int momentum_impl(const fw_grid* g, const void* u, const void* v,
                         const void* w, void* uo, void* vo, void* wo,
                         const void* rho, const void* Tg, const void* Fx,
                         const void* Fy, const void* Fz, const void* u_in,
                         const void* v_in, const void* w_in, double dt,
                         double T_amb, hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_momentum<T>, L.grid, L.block, 0, s, grid_args(g),
                     (const T*)u, (const T*)v, (const T*)w, (T*)uo, (T*)vo,
                     (T*)wo, (const T*)rho, (const T*)Tg, (const T*)Fx,
                     (const T*)Fy, (const T*)Fz, (const T*)u_in,
                     (const T*)v_in, (const T*)w_in, T(dt), T(T_amb));
  FW_CHECK_LAUNCH("tentative_velocity");
  return FW_OK;
}

template <typename T>
static int sponge_impl(const fw_grid* g, void* u, const void* u_target,
                       const void* sigma_x, const void* Y_F, double Y_F_skip,
                       double dt, hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_sponge<T>, L.grid, L.block, 0, s, grid_args(g), (T*)u,
                     (const T*)u_target, (const T*)sigma_x, (const T*)Y_F,
                     T(Y_F_skip), T(dt));
  FW_CHECK_LAUNCH("outflow_sponge");
  return FW_OK;
}

template <typename T>
static int eos_impl(const fw_grid* g, const void* Tg, void* rho, double T_amb,
                    hipStream_t s) {
  size_t n = (size_t)g->nx * g->ny * g->nz;
  hipLaunchKernelGGL(k_eos<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     s, n, (const T*)Tg, (T*)rho, T(T_amb));
  FW_CHECK_LAUNCH("eos");
  return FW_OK;
}

template <typename T>
static int soil_impl(const fw_grid* g, int nsoil, void* Ts, const void* q_in,
                     const double* sdz, const double* sda, const double* sdb,
                     double dt, double k_s, double rho_s, double cp_s,
                     double eps_s, double T_amb, hipStream_t s) {
  int ncol = g->nx * g->ny;
  hipLaunchKernelGGL(k_soil<T>, dim3((ncol + 255) / 256), dim3(256), 0, s,
                     nsoil, ncol, (T*)Ts, (const T*)q_in, sdz, sda, sdb, T(dt),
                     T(k_s), T(1.0 / (rho_s * cp_s)), T(eps_s), T(T_amb));
  FW_CHECK_LAUNCH("soil_conduction");
  return FW_OK;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_advect_scalar_muscl(const fw_grid* g, const void* phi, const void* u,
                           const void* v, const void* w, void* rhs,
                           double phi_inlet, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && g->nx >= 1 && g->ny >= 1 && g->nz >= 1, "bad grid");
  return FW_DISPATCH(g, advect_impl, g, phi, u, v, w, rhs, phi_inlet,
                     (hipStream_t)stream);
}

int fw_species_transport(const fw_grid* g, const void* Y, void* Y_out,
                         const void* rho, const void* u, const void* v,
                         const void* w, const void* S, double dt, double D,
                         double Y_inlet, double y_max, void* stream) {
  FW_GRID(g);
  FW_GRID_OFFSETS32(g);
  FW_REQUIRE(g && Y != Y_out && S, "species_transport: Y_out must not alias Y");
  return FW_DISPATCH(g, species_impl, g, Y, Y_out, rho, u, v, w, S, nullptr,
                     nullptr, 0, dt, D, Y_inlet, y_max, (hipStream_t)stream);
}

int fw_species_transport_dilution(const fw_grid* g, const void* Y, void* Y_out,
                                  const void* rho, const void* u, const void* v,
                                  const void* w, const void* S_pyro_part,
                                  const void* S_dry_part, int kind, double dt,
                                  double D, double Y_inlet, double y_max,
                                  void* stream) {
  FW_GRID(g);
  FW_GRID_OFFSETS32(g);
  FW_REQUIRE(g && Y != Y_out && S_pyro_part && S_dry_part && kind >= 1 && kind <= 3,
             "species_transport_dilution: bad argument");
  return FW_DISPATCH(g, species_impl, g, Y, Y_out, rho, u, v, w, nullptr,
                     S_pyro_part, S_dry_part, kind, dt, D, Y_inlet, y_max,
                     (hipStream_t)stream);
}

int fw_tune_species_cz(const fw_grid* g, int cz) {
  FW_REQUIRE(g, "tune_species_cz: null grid");
  FW_REQUIRE(cz == 0 || cz == 1 || cz == 2 || cz == 4,
             "tune_species_cz: cells per thread must be 0 (default), 1, 2 or 4");
  std::lock_guard<std::mutex> lock(g_cz_mu);
  const ShapeKey key{g->nx, g->ny, g->nz, g->dtype};
  if (cz == 0) g_cz.erase(key);
  else g_cz[key] = cz;
  return FW_OK;
}

int fw_species_transport3(const fw_grid* g, const void* const* Y, void* const* Y_out,
                          const void* rho, const void* u, const void* v, const void* w,
                          const void* S_pyro_part, const void* S_dry_part, double dt,
                          double D, const double* Y_inlet, const double* y_max,
                          void* stream) {
  FW_GRID(g);
  FW_GRID_OFFSETS32(g);
  FW_REQUIRE(g && Y && Y_out && Y_inlet && y_max && S_pyro_part && S_dry_part,
             "species_transport3: null argument");
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      FW_REQUIRE(Y[a] != Y_out[b], "species_transport3: outputs must not alias inputs");
  return FW_DISPATCH(g, species3_impl, g, Y, Y_out, rho, u, v, w, S_pyro_part,
                     S_dry_part, dt, D, Y_inlet, y_max, (hipStream_t)stream,
                     (const double*)nullptr);
}

int fw_turbulent_diffusion_pass3(const fw_grid* g, const void* const* f, void* const* f_out,
                                 const void* nu_t, double sc_t, double dt_sub,
                                 const double* clip_lo, const double* clip_hi,
                                 void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && f && f_out && clip_lo && clip_hi, "turbulent_diffusion3: null argument");
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      FW_REQUIRE(f[a] != f_out[b], "turbulent_diffusion3: outputs must not alias inputs");
  return FW_DISPATCH(g, turb3_impl, g, f, f_out, nu_t, sc_t, dt_sub, clip_lo, clip_hi,
                     (int*)nullptr, (hipStream_t)stream, (const double*)nullptr,
                     (const GasSolidParams*)nullptr);
}

int fw_turbulent_diffusion_pass3_summary(const fw_grid* g, const void* const* f,
                                         void* const* f_out, const void* nu_t, double sc_t,
                                         double dt_sub, const double* clip_lo,
                                         const double* clip_hi, int* edc_summary,
                                         void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && f && f_out && clip_lo && clip_hi && edc_summary,
             "turbulent_diffusion3_summary: null argument");
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      FW_REQUIRE(f[a] != f_out[b], "turbulent_diffusion3: outputs must not alias inputs");
  return FW_DISPATCH(g, turb3_impl, g, f, f_out, nu_t, sc_t, dt_sub, clip_lo, clip_hi,
                     edc_summary, (hipStream_t)stream, (const double*)nullptr,
                     (const GasSolidParams*)nullptr);
}

int fw_gas_energy(const fw_grid* g, const void* T_g, void* T_out,
                  const void* u, const void* v, const void* w, double dt,
                  double alpha_th, double T_amb, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && T_g != T_out, "gas_energy: T_out must not alias T_g");
  FW_REQUIRE(g->nx >= 2 && g->nz >= 2, "gas_energy: needs Nx, Nz >= 2");
  return FW_DISPATCH(g, gas_energy_impl, g, T_g, T_out, u, v, w, dt, alpha_th,
                     T_amb, (hipStream_t)stream, (const EdcFuseArgs*)nullptr);
}

int fw_turbulent_diffusion_pass(const fw_grid* g, const void* f, void* f_out,
                                const void* nu_t, double sc_t, double dt_sub,
                                double clip_lo, double clip_hi, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && f != f_out, "turbulent_diffusion: f_out must not alias f");
  return FW_DISPATCH(g, turb_diff_pass_impl, g, f, f_out, nu_t, sc_t, dt_sub,
                     clip_lo, clip_hi, (hipStream_t)stream, (const double*)nullptr,
                     (const EdcFuseArgs*)nullptr);
}

int fw_nu_t_max_interior(const fw_grid* g, const void* nu_t, double* work,
                         double* out_host, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && work && out_host, "nu_t_max_interior: null pointer");
  hipStream_t s = (hipStream_t)stream;
  int rc = nu_t_max_partials(g, nu_t, work, s);
  if (rc) return rc;
  double h[FW_REDUCE_BLOCKS];
  FW_CHECK_HIP(hipMemcpyAsync(h, work, sizeof(h), hipMemcpyDeviceToHost, s));
  FW_CHECK_HIP(hipStreamSynchronize(s));
  double m = 0.0;
  for (int b = 0; b < FW_REDUCE_BLOCKS; ++b)
    if (h[b] > m) m = h[b];
  *out_host = m;
  return FW_OK;
}

int fw_solid_conduction_vertical(const fw_grid* g, const void* T_s,
                                 void* T_out, const void* alpha_s,
                                 double k_solid, double rho_solid,
                                 double cp_solid, double dt, double clip_lo,
                                 double clip_hi, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && T_s != T_out, "solid_conduction: T_out must not alias T_s");
  return FW_DISPATCH(g, solid_cond_impl, g, T_s, T_out, alpha_s, k_solid,
                     rho_solid, cp_solid, dt, clip_lo, clip_hi,
                     (hipStream_t)stream);
}

int fw_drag_force(const fw_grid* g, const void* u, const void* v,
                  const void* w, const void* rho, const void* alpha_s,
                  double sigma_sav, double C_D, void* Fx, void* Fy, void* Fz,
                  void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g, "bad grid");
  return FW_DISPATCH(g, drag_impl, g, u, v, w, rho, alpha_s, sigma_sav, C_D,
                     Fx, Fy, Fz, (hipStream_t)stream);
}

int fw_tentative_velocity(const fw_grid* g, const void* u, const void* v,
                          const void* w, void* u_out, void* v_out,
                          void* w_out, const void* rho, const void* T_g,
                          const void* Fx, const void* Fy, const void* Fz,
                          const void* u_inlet, const void* v_inlet,
                          const void* w_inlet, double dt, double T_amb,
                          void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && u != u_out && v != v_out && w != w_out,
             "tentative_velocity: outputs must not alias inputs");
  return FW_DISPATCH(g, momentum_impl, g, u, v, w, u_out, v_out, w_out, rho,
                     T_g, Fx, Fy, Fz, u_inlet, v_inlet, w_inlet, dt, T_amb,
                     (hipStream_t)stream);
}

int fw_outflow_sponge(const fw_grid* g, void* u, const void* u_target,
                      const void* sigma_x, const void* Y_F, double Y_F_skip,
                      double dt, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g, "bad grid");
  return FW_DISPATCH(g, sponge_impl, g, u, u_target, sigma_x, Y_F, Y_F_skip,
                     dt, (hipStream_t)stream);
}

int fw_eos(const fw_grid* g, const void* T_g, void* rho, double T_amb,
           void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g, "bad grid");
  return FW_DISPATCH(g, eos_impl, g, T_g, rho, T_amb, (hipStream_t)stream);
}

int fw_soil_conduction(const fw_grid* g, int n_soil, void* T_soil,
                       const void* q_in_surface, const double* soil_dz,
                       const double* soil_d_above, const double* soil_d_below,
                       double dt, double k_s, double rho_s, double cp_s,
                       double eps_s, double T_amb, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && n_soil >= 1, "soil_conduction: n_soil >= 1");
  return FW_DISPATCH(g, soil_impl, g, n_soil, T_soil, q_in_surface, soil_dz,
                     soil_d_above, soil_d_below, dt, k_s, rho_s, cp_s, eps_s,
                     T_amb, (hipStream_t)stream);
}

}  // extern "C"
