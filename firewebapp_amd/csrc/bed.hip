// Lagrangian bed particles (lagrangian_bed_3d.py:427-1070 of the reference).
//
// Layout: particle state is float64 SoA (x, y, z, m_solid, m_water, m_char,
// T_s, m_water_0, sav, m_char_max) + int8 alive, exactly the reference's
// buffers.  Particles never move, so every particle's cell is computed once
// (fw_bed_locate, float64, the reference's int() truncation rules) and the
// particles of each cell are listed in slot order (CSR: seg_ptr / seg_idx).
//
// Step: one thread per particle evaluates drying, pyrolysis, char oxidation,
// smouldering and the 5-iteration implicit Newton solve for T_s, and writes
// its per-cell source contributions to a per-particle scratch array; one
// thread per cell then adds its particles' contributions in slot order.
// This reproduces the reference's sequential `cell += contribution` sums
// bit-for-bit without atomics.  Counters use integer atomics (exact).
#include <string.h>
#include "fw_common.h"

namespace fw {

#define NCONTRIB 10  // S_pyro, S_dry_a, S_dry_b, Q_pyro, Q_dry_a, Q_dry_b, Y_F, Q_char, Q_smold, Q_gconv
#define NDIAG 13

struct BedArgs {
  double* x; double* y; double* z;
  double* m_solid; double* m_water; double* m_char; double* T_s;
  double* m_water_0; double* sav; double* m_char_max;
  signed char* alive;
  int n;
  const double* h_top;   // per-particle fuel top, or null: the scalar h_bed
};

// kinetic constants (pyrolysis_3d.py:45, :117, :217-267, :446-562)
#define R_GAS 8.314
#define HEAT_OF_PYROLYSIS 4.0e5
#define A_MD2004 36280.0
#define E_MD2004 58600.0
#define CHAR_YIELD 0.15
#define ETA_MD (1.0 - CHAR_YIELD)
#define A_OP 7.26e4
#define E_OP 55000.0
#define HOR_OP (-1.0e6)
#define Y_O2_MIN_OP 1.0e-3
#define A_DRY 4.29e3
#define E_DRY 43800.0
#define L_VAP 2.26e6
#define A_CHAR 7.09e5
#define E_CHAR 94000.0
#define HOC_CHAR 32000000.0
#define T_CHAR_ONSET 600.0
#define Y_O2_MIN_CHAR 1.0e-3
#define A_SMOLD 1.0e6
#define E_SMOLD 80000.0
#define HOC_SMOLD 28000000.0
#define T_SMOLD_ONSET 473.0
#define Y_O2_MIN_SMOLD 1.0e-3
#define A_DRY_GRASS 4.29e3
#define E_DRY_GRASS 30000.0
#define T_BOIL 373.15
#define M_BURNOUT 1.0e-8
#define SIGMA_SB 5.67e-8

__device__ __forceinline__ int locate_k(double z, const double* zf, int Nz) {
  if (z < zf[0] || z >= zf[Nz]) return -1;
  for (int k = 0; k < Nz; ++k)
    if (z < zf[k + 1]) return k;
  return Nz - 1;
}

// cell_step: locate_cell (lagrangian_particles_3d.py:111-138): x < 0 -> -1
// before truncation.  cell_agg: int() first, then the range test
// (lagrangian_bed_3d.py:490-497, :528-534).
__global__ void __launch_bounds__(256)
k_bed_locate(GridArgs g, int n, const double* __restrict__ x,
             const double* __restrict__ y, const double* __restrict__ z,
             int* __restrict__ cell_step, int* __restrict__ cell_agg) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double xp = x[p], yp = y[p];
  const int k = locate_k(z[p], g.z_face, g.nz);
  int i = xp < 0.0 ? -1 : (int)(xp / g.dx);
  if (i >= g.nx) i = -1;
  int j = yp < 0.0 ? -1 : (int)(yp / g.dy);
  if (j >= g.ny) j = -1;
  cell_step[p] = (i < 0 || j < 0 || k < 0) ? -1 : (int)(((size_t)k * g.ny + j) * g.nx + i);
  // int() truncates toward zero; values beyond int range are out anyway
  const double xi = xp / g.dx, yj = yp / g.dy;
  int ia = (xi > -2147483648.0 && xi < 2147483647.0) ? (int)xi : -1;
  int ja = (yj > -2147483648.0 && yj < 2147483647.0) ? (int)yj : -1;
  const bool ok = ia >= 0 && ia < g.nx && ja >= 0 && ja < g.ny && k >= 0;
  cell_agg[p] = ok ? (int)(((size_t)k * g.ny + ja) * g.nx + ia) : -1;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_bed_particles(GridArgs g, BedArgs B, fw_bed_params P, const int* __restrict__ cell_step,
                const T* __restrict__ Tg_grid, const T* __restrict__ YO_grid,
                const T* __restrict__ Qext_grid, double* __restrict__ contrib,
                int* __restrict__ counters, double* __restrict__ diag_rec) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  __shared__ int s_alive, s_burned;
  if (threadIdx.x == 0) { s_alive = 0; s_burned = 0; }
  __syncthreads();
  const size_t n = (size_t)B.n;
  bool valid = p < B.n;
  double cv[NCONTRIB];
#pragma unroll
  for (int q = 0; q < NCONTRIB; ++q) cv[q] = 0.0;
  double dg[NDIAG + 3];
#pragma unroll
  for (int q = 0; q < NDIAG + 3; ++q) dg[q] = 0.0;
  bool has_diag = false;
  if (valid && B.alive[p] != 0) {
    const int c = cell_step[p];
    if (c < 0) {
      B.alive[p] = 0;
    } else {
      const int k = (int)(c / ((size_t)g.nx * g.ny));
      const double dt = P.dt;
      const double inv_dt = dt > 0.0 ? 1.0 / dt : 0.0;
      const double V_cell = g.dx * g.dy * g.dz[k];
      double T_s = B.T_s[p];
      const double T_g = (double)Tg_grid[c];
      const double Y_O2 = (double)YO_grid[c];
      double ms = B.m_solid[p], mw = B.m_water[p], mc = B.m_char[p];
      const double mw0 = B.m_water_0[p], sav = B.sav[p];
      double mcmax = B.m_char_max[p];
      // drying (lagrangian_bed_3d.py:756-767)
      double dm_evap = 0.0;
      if (P.do_drying && mw > 0.0 && T_s > 0.0) {
        if (P.drying_mode == 0) {
          const double kd = A_DRY * exp(-E_DRY / (R_GAS * T_s));
          const double mwn = mw * exp(-kd * dt);
          dm_evap = mw - mwn;
          mw = mwn;
        } else if (P.drying_mode == 2) {
          const double kd = A_DRY_GRASS * exp(-E_DRY_GRASS / (R_GAS * T_s));
          const double mwn = mw * exp(-kd * dt);
          dm_evap = mw - mwn;
          mw = mwn;
        }
      }
      // pyrolysis (:770-813)
      double rate = 0.0, f_th = 1.0, f_op = 0.0;
      if (P.do_pyrolysis && ms > 0.0 && T_s > 0.0) {
        double gate;
        if (mw0 > 0.0) {
          const double wet = mw / mw0;
          gate = 1.0 - wet;
          if (gate < 0.0) gate = 0.0;
        } else {
          gate = 1.0;
        }
        if (gate > 0.0) {
          const double kth = A_MD2004 * exp(-E_MD2004 / (R_GAS * T_s));
          const double kop = Y_O2 > Y_O2_MIN_OP ? A_OP * exp(-E_OP / (R_GAS * T_s)) * Y_O2 : 0.0;
          const double kt = kth + kop;
          if (kt > 0.0) {
            const double msn = ms * exp(-kt * dt);
            const double dm_full = ms - msn;
            const double dm_pyro = dm_full * gate;
            rate = dm_pyro * inv_dt;
            f_th = kth / kt;
            f_op = kop / kt;
            mc += CHAR_YIELD * dm_pyro;
            ms -= dm_pyro;
            if (mc > mcmax) mcmax = mc;
          }
        }
      }
      const double A_now = sav * (ms + mc) / P.rho_solid_true;
      // char oxidation (:836-857)
      double Q_ch = 0.0;
      if (P.do_char_ox && mc > 1.0e-12 && T_s >= T_CHAR_ONSET && Y_O2 >= Y_O2_MIN_CHAR) {
        const double kch = A_CHAR * exp(-E_CHAR / (R_GAS * T_s));
        const double mdot = kch * mc * Y_O2;
        double mcons = mdot * dt;
        if (mcons > 0.5 * mc) mcons = 0.5 * mc;
        double Ar = A_now;
        if (P.char_ox_ash_exp > 0.0 && mcmax > 1.0e-15) {
          double bf = mc / mcmax;
          if (bf < 0.0) bf = 0.0;
          Ar *= pow(bf, P.char_ox_ash_exp);
        }
        const double Qcap = P.char_ox_flux_cap * Ar;
        const double Qarr = (mcons * HOC_CHAR) * inv_dt;
        if (Qarr > Qcap) mcons = Qcap / (HOC_CHAR * inv_dt);
        Q_ch = (mcons * HOC_CHAR) * inv_dt;
        mc -= mcons;
      }
      // smouldering (:867-893)
      double Q_sm = 0.0;
      if (P.do_smolder && T_s >= T_SMOLD_ONSET && Y_O2 >= Y_O2_MIN_SMOLD) {
        const double avail = ms + mc;
        if (avail > 1.0e-12) {
          const double ksm = A_SMOLD * exp(-E_SMOLD / (R_GAS * T_s));
          const double mdot = ksm * avail * Y_O2;
          double mcons = mdot * dt;
          if (mcons > 0.5 * avail) mcons = 0.5 * avail;
          const double Qcap = 4.0e4 * A_now;
          const double Qarr = (mcons * HOC_SMOLD) * inv_dt;
          if (Qarr > Qcap) mcons = Qcap / (HOC_SMOLD * inv_dt);
          if (avail > 0.0) {
            const double fs = ms / avail;
            const double ds = fs * mcons;
            const double dc = (1.0 - fs) * mcons;
            ms -= ds;
            mc -= dc;
          }
          Q_sm = (mcons * HOC_SMOLD) * inv_dt;
        }
      }
      const double inv_V = 1.0 / V_cell;
      cv[0] = ETA_MD * rate * inv_V;
      cv[1] = dm_evap * inv_dt * inv_V;
      cv[3] = (rate * (f_th * HEAT_OF_PYROLYSIS + f_op * HOR_OP)) * inv_V;
      cv[4] = dm_evap * L_VAP * inv_dt * inv_V;
      cv[6] = ETA_MD * rate * inv_V;
      cv[7] = Q_ch * inv_V;
      cv[8] = Q_sm * inv_V;
      // energy balance (:927-1047)
      const double mt = ms + mw + mc;
      double Q_conv = 0.0;
      if (mt > 0.0) {
        const double A_p = sav * (ms + mc) / P.rho_solid_true;
        Q_conv = P.h_conv * A_p * (T_g - T_s);
        const double Q_rxn = -rate * (f_th * HEAT_OF_PYROLYSIS + f_op * HOR_OP);
        const double Q_dry = -dm_evap * L_VAP * inv_dt;
        double Q_ext_p = 0.0;
        if (P.n_per_cell_for_split > 0)
          Q_ext_p = (double)Qext_grid[c] * V_cell / P.n_per_cell_for_split;
        double fgeom = 1.0;
        if (P.view_factor_geometric) {
          double ha = (B.h_top ? B.h_top[p] : P.h_bed) - B.z[p];
          if (ha < 0.0) ha = 0.0;
          fgeom = exp(-P.kappa_bed_eff * ha);
        }
        const double C_rad = P.eps_solid * SIGMA_SB * A_p * P.view_factor * fgeom;
        const double mcp = mt * P.cp_solid;
        const double Ta = P.T_amb;
        const double Ta4 = Ta * Ta * Ta * Ta;
        const double Q_other = Q_conv + Q_rxn + Q_dry + Q_ch + Q_sm + Q_ext_p;
        const double mc_dt = mcp / dt;
        const double T_old = T_s;
        double Ti = T_s, F = 0.0;
        for (int it = 0; it < 5; ++it) {
          const double T2 = Ti * Ti;
          const double T3 = T2 * Ti;
          const double T4 = T3 * Ti;
          F = mc_dt * (Ti - T_old) - Q_other + C_rad * (T4 - Ta4);
          const double Fp = mc_dt + 4.0 * C_rad * T3;
          Ti = Ti - F / Fp;
          if (Ti < Ta) Ti = Ta;
        }
        T_s = Ti;
        if (diag_rec) {
          has_diag = true;
          const double Tq2 = T_s * T_s;
          dg[0] = T_s; dg[1] = Q_conv; dg[2] = Q_rxn; dg[3] = Q_dry; dg[4] = Q_ch;
          dg[5] = Q_sm; dg[6] = Q_ext_p; dg[7] = C_rad * (Tq2 * Tq2 - Ta4);
          dg[8] = Q_other; dg[9] = T_g; dg[10] = mc_dt; dg[11] = C_rad;
          dg[12] = F >= 0.0 ? F : -F; dg[13] = A_p; dg[14] = mt;
        }
        if (T_s > 1.0e4) T_s = 1.0e4;
        if (P.do_drying && (P.drying_mode == 1 || P.drying_mode == 2) && mw > 0.0 &&
            T_s > T_BOIL) {
          const double excess = mcp * (T_s - T_BOIL);
          double dm_eq = excess / L_VAP;
          if (dm_eq >= mw) {
            const double resid = excess - mw * L_VAP;
            dm_eq = mw;
            mw = 0.0;
            T_s = T_BOIL + (mcp > 0.0 ? resid / mcp : 0.0);
          } else {
            mw -= dm_eq;
            T_s = T_BOIL;
          }
          cv[2] = dm_eq * inv_dt * inv_V;
          cv[5] = dm_eq * L_VAP * inv_dt * inv_V;
        }
      }
      cv[9] = Q_conv * inv_V;
      B.T_s[p] = T_s;
      B.m_solid[p] = ms;
      B.m_water[p] = mw;
      B.m_char[p] = mc;
      B.m_char_max[p] = mcmax;
      if (mt < M_BURNOUT) {
        B.alive[p] = 0;
        atomicAdd(&s_burned, 1);
      } else {
        atomicAdd(&s_alive, 1);
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int q = 0; q < NCONTRIB; ++q) contrib[(size_t)q * n + p] = cv[q];
    if (diag_rec) {
      // record 0: T_s used for the argmax (-1 when the particle has no mass)
      diag_rec[p] = has_diag ? dg[0] : -1.0;
#pragma unroll
      for (int q = 1; q < NDIAG + 2; ++q) diag_rec[(size_t)q * n + p] = dg[q];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_alive) atomicAdd(&counters[0], s_alive);
    if (s_burned) atomicAdd(&counters[1], s_burned);
  }
}

// per-cell slot-ordered sums of the particle contributions
template <typename T>
__global__ void __launch_bounds__(256)
k_bed_segsum(int ncell_seg, size_t n, const int* __restrict__ seg_ptr,
             const int* __restrict__ seg_idx, const double* __restrict__ contrib,
             T* Sp, T* Sd, T* Qp, T* Qd, T* YF, T* Qch, T* Qsm, T* Qgc) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell_seg) return;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = seg_ptr[c]; q < seg_ptr[c + 1]; ++q) {
    const size_t p = (size_t)seg_idx[q];
    s[0] += contrib[0 * n + p];
    s[1] += contrib[1 * n + p];
    s[1] += contrib[2 * n + p];
    s[2] += contrib[3 * n + p];
    s[3] += contrib[4 * n + p];
    s[3] += contrib[5 * n + p];
    s[4] += contrib[6 * n + p];
    s[5] += contrib[7 * n + p];
    s[6] += contrib[8 * n + p];
    s[7] += contrib[9 * n + p];
  }
  Sp[c] = (T)s[0]; Sd[c] = (T)s[1]; Qp[c] = (T)s[2]; Qd[c] = (T)s[3];
  YF[c] = (T)s[4]; Qch[c] = (T)s[5]; Qsm[c] = (T)s[6]; Qgc[c] = (T)s[7];
}

// aggregate_particles_to_T_s_grid / _M_local_grid over the agg segments
template <typename T>
__global__ void __launch_bounds__(256)
k_bed_aggregate(int ncell_seg, const int* __restrict__ seg_ptr,
                const int* __restrict__ seg_idx, BedArgs B, T* Ts_grid, T* M_grid) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell_seg) return;
  double num = 0.0, den = 0.0, wn = 0.0, wd = 0.0;
  for (int q = seg_ptr[c]; q < seg_ptr[c + 1]; ++q) {
    const int p = seg_idx[q];
    if (B.alive[p] == 0) continue;
    if (Ts_grid) {
      const double mt = B.m_solid[p] + B.m_water[p] + B.m_char[p];
      if (mt > 0.0) {
        num += B.T_s[p] * mt;
        den += mt;
      }
    }
    if (M_grid) {
      wn += B.m_water[p];
      wd += B.m_solid[p];
    }
  }
  if (Ts_grid && den > 0.0) Ts_grid[c] = (T)(num / den);
  if (M_grid) M_grid[c] = (T)(wd > 0.0 ? wn / wd : 0.0);
}

// step_horizontal_solid_conduction_scatter, grid part (:452-483)
template <typename T>
__global__ void __launch_bounds__(256)
k_bed_hcond(GridArgs g, int nzb, const T* __restrict__ To, T* __restrict__ Tn,
            const T* __restrict__ alpha, double diff, double dt,
            double* __restrict__ dTc) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx || k >= nzb) return;
  const int Nx = g.nx, Ny = g.ny;
  const size_t c = cidx(g, k, j, i);
  if (!((double)alpha[c] > 0.0)) {
    dTc[c] = 0.0;
    return;
  }
  const double inv_dx2 = g.inv_dx2, inv_dy2 = g.inv_dy2;
  const double Tc = (double)To[c];
  const double Txp = (i + 1 < Nx && (double)alpha[c + 1] > 0.0) ? (double)To[c + 1] : Tc;
  const double Txm = (i - 1 >= 0 && (double)alpha[c - 1] > 0.0) ? (double)To[c - 1] : Tc;
  const int jp = j + 1 < Ny ? j + 1 : 0;
  const int jm = j - 1 >= 0 ? j - 1 : Ny - 1;
  const size_t cp = cidx(g, k, jp, i), cm = cidx(g, k, jm, i);
  const double Typ = (double)alpha[cp] > 0.0 ? (double)To[cp] : Tc;
  const double Tym = (double)alpha[cm] > 0.0 ? (double)To[cm] : Tc;
  const double lap = ((Txp - 2.0 * Tc + Txm) * inv_dx2 + (Typ - 2.0 * Tc + Tym) * inv_dy2);
  const double dT = diff * lap * dt;
  dTc[c] = dT;
  Tn[c] = (T)(Tc + dT);
}

__global__ void __launch_bounds__(256)
k_bed_hcond_particles(int n, int nzb_cells, const int* __restrict__ cell_agg,
                      const signed char* __restrict__ alive, double* __restrict__ Ts,
                      const double* __restrict__ dTc) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n || alive[p] == 0) return;
  const int c = cell_agg[p];
  if (c < 0 || c >= nzb_cells) return;
  Ts[p] += dTc[c];
}

// argmax of the recorded T_s (first slot among equals) and max |F|
__global__ void __launch_bounds__(256)
k_bed_diag(int n, const double* __restrict__ rec, double* __restrict__ diag) {
  __shared__ double sv[256];
  __shared__ int si[256];
  __shared__ double sf[256];
  const int t = threadIdx.x;
  double bv = 0.0, bf = 0.0;
  int bi = -1;
  for (int p = t; p < n; p += 256) {
    const double v = rec[p];
    if (v > bv || (v == bv && bi >= 0 && p < bi && v > 0.0)) { bv = v; bi = p; }
    if (v >= 0.0) {
      const double f = rec[(size_t)12 * n + p];
      if (f > bf) bf = f;
    }
  }
  sv[t] = bv; si[t] = bi; sf[t] = bf;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      const double v2 = sv[t + s];
      const int i2 = si[t + s];
      if (i2 >= 0 && (si[t] < 0 || v2 > sv[t] || (v2 == sv[t] && i2 < si[t]))) {
        sv[t] = v2;
        si[t] = i2;
      }
      if (sf[t + s] > sf[t]) sf[t] = sf[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int q = 0; q < 16; ++q) diag[q] = 0.0;
    const int q0 = si[0];
    if (q0 >= 0) {
      for (int q = 0; q < NDIAG; ++q) diag[q] = rec[(size_t)q * n + q0];
      diag[14] = rec[(size_t)13 * n + q0];
      diag[15] = rec[(size_t)14 * n + q0];
    }
    diag[13] = sf[0];
  }
}

template <typename T>
static int bed_step_impl(const fw_grid* g, const fw_bed_particles* bp,
                         const fw_bed_index* ix, const fw_bed_params* P,
                         const void* Tg, const void* YO, const void* Qext,
                         void* const* src, double* contrib, int* counters,
                         double* diag_rec, double* diag_out, hipStream_t s) {
  BedArgs B{bp->x, bp->y, bp->z, bp->m_solid, bp->m_water, bp->m_char, bp->T_s,
            bp->m_water_0, bp->sav, bp->m_char_max, bp->alive, bp->n, bp->h_top};
  GridArgs ga = grid_args(g);
  FW_CHECK_HIP(hipMemsetAsync(counters, 0, 2 * sizeof(int), s));
  if (bp->n > 0) {
    hipLaunchKernelGGL(k_bed_particles<T>, dim3((bp->n + 255) / 256), dim3(256), 0, s,
                       ga, B, *P, ix->cell, (const T*)Tg, (const T*)YO,
                       (const T*)Qext, contrib, counters, diag_rec);
    FW_CHECK_LAUNCH("bed_particles");
  }
  if (ix->n_seg_cells > 0) {
    hipLaunchKernelGGL(k_bed_segsum<T>, dim3((ix->n_seg_cells + 255) / 256), dim3(256), 0,
                       s, ix->n_seg_cells, (size_t)bp->n, ix->seg_ptr, ix->seg_idx,
                       (const double*)contrib, (T*)src[0], (T*)src[1], (T*)src[2],
                       (T*)src[3], (T*)src[4], (T*)src[5], (T*)src[6], (T*)src[7]);
    FW_CHECK_LAUNCH("bed_segsum");
  }
  if (diag_rec && diag_out && bp->n > 0) {
    hipLaunchKernelGGL(k_bed_diag, dim3(1), dim3(256), 0, s, bp->n,
                       (const double*)diag_rec, diag_out);
    FW_CHECK_LAUNCH("bed_diag");
  }
  return FW_OK;
}

// Dynamic bed seeding behind the empirical-ROS front (spread_3d.py:3271-3300)
__global__ void __launch_bounds__(256)
k_bed_seed(int n, const double* __restrict__ x, const signed char* __restrict__ alive,
           const double* __restrict__ m_solid, double* __restrict__ T_s, double x_lo,
           double x_front, double T_ign, double T_seed) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double xp = x[p];
  if (alive[p] == 1 && xp >= x_lo && xp <= x_front && T_s[p] < T_ign && m_solid[p] > 1.0e-9)
    T_s[p] = T_seed;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_bed_locate(const fw_grid* g, int n, const double* x, const double* y,
                  const double* z, int* cell_step, int* cell_agg, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && n >= 0, "bed_locate: bad argument");
  if (n == 0) return FW_OK;
  hipLaunchKernelGGL(k_bed_locate, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, grid_args(g), n, x, y, z, cell_step, cell_agg);
  FW_CHECK_LAUNCH("bed_locate");
  return FW_OK;
}

int fw_bed_step(const fw_grid* g, const fw_bed_particles* particles,
                const fw_bed_index* index, const fw_bed_params* params,
                const void* T_g, const void* Y_O2, const void* Q_solid_ext,
                void* const* sources, double* contrib_work, int* counters,
                double* diag_work, double* diag_out, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && particles && index && params && T_g && Y_O2 && Q_solid_ext && sources &&
                 contrib_work && counters,
             "bed_step: null argument");
  FW_REQUIRE(params->dt > 0.0, "bed_step: dt must be > 0");
  return FW_DISPATCH(g, bed_step_impl, g, particles, index, params, T_g, Y_O2,
                     Q_solid_ext, sources, contrib_work, counters, diag_work, diag_out,
                     (hipStream_t)stream);
}

int fw_bed_aggregate(const fw_grid* g, const fw_bed_particles* particles,
                     const fw_bed_index* agg_index, void* T_s_grid,
                     void* M_local_grid, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && particles && agg_index, "bed_aggregate: null argument");
  if (agg_index->n_seg_cells <= 0) return FW_OK;
  BedArgs B{particles->x, particles->y, particles->z, particles->m_solid,
            particles->m_water, particles->m_char, particles->T_s, particles->m_water_0,
            particles->sav, particles->m_char_max, particles->alive, particles->n};
  hipStream_t s = (hipStream_t)stream;
  dim3 gr((agg_index->n_seg_cells + 255) / 256);
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_bed_aggregate<double>, gr, dim3(256), 0, s,
                       agg_index->n_seg_cells, agg_index->seg_ptr, agg_index->seg_idx, B,
                       (double*)T_s_grid, (double*)M_local_grid);
  else
    hipLaunchKernelGGL(k_bed_aggregate<float>, gr, dim3(256), 0, s,
                       agg_index->n_seg_cells, agg_index->seg_ptr, agg_index->seg_idx, B,
                       (float*)T_s_grid, (float*)M_local_grid);
  FW_CHECK_LAUNCH("bed_aggregate");
  return FW_OK;
}

int fw_bed_horizontal_conduction(const fw_grid* g, const fw_bed_particles* particles,
                                 const fw_bed_index* agg_index, const void* T_s_old,
                                 void* T_s_grid, const void* alpha_s, double k_solid,
                                 double rho_solid_true, double cp_solid, int n_z_bed,
                                 double dt, double* dT_work, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && particles && agg_index && T_s_old && T_s_grid && alpha_s && dT_work,
             "bed_horizontal_conduction: null argument");
  FW_REQUIRE(T_s_old != T_s_grid, "bed_horizontal_conduction: T_s_old must not alias");
  hipStream_t s = (hipStream_t)stream;
  const int nzb = n_z_bed < g->nz ? n_z_bed : g->nz;
  if (nzb <= 0) return FW_OK;
  Launch L = cell_launch(g);
  L.grid.z = nzb;
  const double diff = k_solid / (rho_solid_true * cp_solid);
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_bed_hcond<double>, L.grid, L.block, 0, s, grid_args(g), nzb,
                       (const double*)T_s_old, (double*)T_s_grid, (const double*)alpha_s,
                       diff, dt, dT_work);
  else
    hipLaunchKernelGGL(k_bed_hcond<float>, L.grid, L.block, 0, s, grid_args(g), nzb,
                       (const float*)T_s_old, (float*)T_s_grid, (const float*)alpha_s,
                       diff, dt, dT_work);
  FW_CHECK_LAUNCH("bed_hcond");
  if (particles->n > 0) {
    hipLaunchKernelGGL(k_bed_hcond_particles, dim3((particles->n + 255) / 256), dim3(256),
                       0, s, particles->n, nzb * g->nx * g->ny, agg_index->cell,
                       particles->alive, particles->T_s, (const double*)dT_work);
    FW_CHECK_LAUNCH("bed_hcond_particles");
  }
  return FW_OK;
}

int fw_bed_seed(const fw_bed_particles* particles, double x_lo, double x_front,
                double T_ign, double T_seed, void* stream) {
  FW_REQUIRE(particles && particles->n >= 0, "bed_seed: bad argument");
  if (particles->n == 0) return FW_OK;
  FW_REQUIRE(particles->x && particles->alive && particles->m_solid && particles->T_s,
             "bed_seed: particle arrays missing");
  hipLaunchKernelGGL(k_bed_seed, dim3((particles->n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, particles->n, particles->x,
                     (const signed char*)particles->alive, particles->m_solid,
                     particles->T_s, x_lo, x_front, T_ign, T_seed);
  FW_CHECK_LAUNCH("bed_seed");
  return FW_OK;
}

}  // extern "C"
