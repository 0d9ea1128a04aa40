// Point-wise glue of the outer time step (spread_3d.py:1897-3683 of the
// reference) and the fused per-step reductions.  Each kernel replaces a
// chain of whole-array NumPy expressions of the reference with one pass,
// keeping the reference's per-element operation order.
#include "fw_common.h"
#include "gas_solid.h"

namespace fw {

// rho_prev <- rho ; rho <- EoS(T_g) ; Y_O2[:, :, 0] <- 0.232
// (spread_3d.py:1984-1987, :2042, :2049-2050)
template <typename T>
__global__ void __launch_bounds__(256)
k_prologue(GridArgs g, const T* __restrict__ Tg, T* __restrict__ rho,
           T* __restrict__ rho_prev, T* __restrict__ YO2, T T_amb) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const size_t c = cidx(g, k, j, i);
  rho_prev[c] = rho[c];
  const T t = Tg[c];
  const T tm = (t > T_amb || t != t) ? t : T_amb;
  rho[c] = T(101325.0) / (T(8.314 / 0.029) * tm);
  if (i == 0) YO2[c] = T(0.232);
}

// Q_solid_ext on the bed layers (spread_3d.py:2078-2103): drip on the
// ignition patch while t < t_ign, plus clip(q_rad/dz, +-Q_RAD_MAX).
template <typename T>
__global__ void __launch_bounds__(256)
k_bed_qext(GridArgs g, int nzb, const T* __restrict__ q_rad, T* __restrict__ Qext,
           int i0, int i1, double drip, double qmax) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx || k >= nzb) return;
  const size_t c = cidx(g, k, j, i);
  double v = 0.0;
  if (i >= i0 && i < i1) v += drip;
  double r = (double)q_rad[c] / g.dz[k];
  r = r < -qmax ? -qmax : (r > qmax ? qmax : r);
  v += r;
  Qext[c] = (T)v;
}

// S_pyro = Sp + Sd ; div_target = (S_pyro - (rho - rho_prev)/dt) / max(rho, 0.1)
// (spread_3d.py:2147-2154, :2300-2317)
template <typename T>
__global__ void __launch_bounds__(256)
k_div_target(size_t n, const T* __restrict__ Sp, const T* __restrict__ Sd,
             const T* __restrict__ rho, const T* __restrict__ rho_prev, T dt,
             T* __restrict__ S_pyro, T* __restrict__ tgt) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const T s = Sp[c] + Sd[c];
  S_pyro[c] = s;
  const T r = rho[c];
  const T drho = (r - rho_prev[c]) / dt;
  tgt[c] = (s - drho) / (r > T(0.1) ? r : T(0.1));
}

// Sub-step gas-energy update of the particle path (spread_3d.py:2916-2957,
// :3104-3105, :3129-3146):
//   Q_comb = [drip on the patch] + q_rad_gas/dz
//   T_g += (Q_comb - Q_gconv - debit) dt_sub / (max(rho,1e-3) cp_mix), T_g >= T_amb
template <typename T>
__global__ void __launch_bounds__(256)
k_gas_update(GridArgs g, T* __restrict__ Tg, const T* __restrict__ Ts,
             const T* __restrict__ rho, const T* __restrict__ Yw,
             const T* __restrict__ Sp, const T* __restrict__ Sd,
             const T* __restrict__ Qgc, const T* __restrict__ qrg,
             T* __restrict__ Qcomb_out, int nzb, int i0, int i1, T drip, T dt_sub,
             T T_amb) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const size_t c = cidx(g, k, j, i);
  T Q = T(0);
  if (k < nzb && i >= i0 && i < i1) Q += drip;
  Q += qrg[c] / T(g.dz[k]);
  if (Qcomb_out) Qcomb_out[c] = Q;
  const T y = Yw[c];
  const T cp = (T(1) - y) * T(1100.0) + y * T(2000.0);
  const T tg = Tg[c];
  T a = tg - Ts[c];
  a = a > T(0) ? a : T(0);
  T b = tg - T(373.15);
  b = b > T(0) ? b : T(0);
  const T debit = (Sp[c] * a + Sd[c] * b) * T(2000.0);
  const T r = rho[c];
  const T inv = dt_sub / ((r > T(1.0e-3) ? r : T(1.0e-3)) * cp);
  T nv = tg + (Q - Qgc[c] - debit) * inv;
  nv = (nv > T_amb || nv != nv) ? nv : T_amb;
  Tg[c] = nv;
}

// The sub-step's gas-energy update (as k_gas_update, in place on T_g) fused
// with the vertical solid conduction that follows it
// (solid_conduction_3d.py:56-127, out of place Ts -> Ts_out, with the T_s
// clamps of spread_3d.py:3195-3196).  Both read only the T_s of before the
// conduction, as in the reference order.
template <typename T>
__global__ void __launch_bounds__(256)
k_gas_solid(GridArgs g, GasSolidArgs<T> a, const T* __restrict__ Yw,
            const double* __restrict__ sb) {
  const T dt_sub = sb ? T(sb[SB_DT_SUB]) : a.dt_sub;
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const size_t c = cidx(g, k, j, i);
  gas_solid_cell<T>(g, i, k, c, a, dt_sub, Yw[c]);
}

// q_rad[k, :, i0:i1] += add   (spread_3d.py:2528-2532)
template <typename T>
__global__ void k_add_patch(GridArgs g, T* __restrict__ q, int k, int i0, int i1,
                            T add) {
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= i1) return;
  q[cidx(g, k, j, i)] += add;
}

// flame mask = active (omega > 1e-3 | (T_g > 1000 & Y_F > 1e-3)), identical
// to phi_flame <= 0 of compute_phi_flame_from_state + band 0
// (flame_front_3d.py:756-813); update_cell_age (:816-835).
template <typename T>
__global__ void __launch_bounds__(256)
k_flame_mask_age(size_t n, const T* __restrict__ om, const T* __restrict__ Tg,
                 const T* __restrict__ YF, T* __restrict__ age,
                 unsigned char* __restrict__ mask, T dt) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const bool act = (om[c] > T(1.0e-3)) || ((Tg[c] > T(1000.0)) && (YF[c] > T(1.0e-3)));
  if (mask) mask[c] = act ? 1 : 0;
  if (age) {
    const T a = age[c];
    if (act) age[c] = isinf(a) ? T(0) : a + dt;
    else age[c] = T(INFINITY);
  }
}

// Per-step reductions in one pass.  Output slots:
//  0 max|u|  1 max|v|  2 max|w|  3 max nu_t  4 max T_g  5 T_g non-finite (0/1)
//  6 max T_s  7 max S_pyro  8 max Q_comb  9 max omega  10 max Y_fuel
// 11 count(T_s >= T_ign)  12 count of flame-active cells  13 last i with a
// burning bed column (T_s >= T_front on alpha_s > 0, k < nzb), -1 if none
#define NRED 14
template <typename T>
__global__ void __launch_bounds__(256)
k_step_reduce(GridArgs g, size_t n, const T* __restrict__ u, const T* __restrict__ v,
              const T* __restrict__ w, const T* __restrict__ nut,
              const T* __restrict__ Tg, const T* __restrict__ Ts,
              const T* __restrict__ Sp, const T* __restrict__ Qc,
              const T* __restrict__ om, const T* __restrict__ YF,
              const T* __restrict__ alpha, double T_ign, double T_front, int nzb,
              double* __restrict__ part) {
  __shared__ double sh[NRED][4];
  double r[NRED];
  for (int q = 0; q < NRED; ++q) r[q] = -1e300;
  r[5] = 0.0; r[11] = 0.0; r[12] = 0.0; r[13] = -1.0;
  const size_t plane = (size_t)g.nx * g.ny;
  // unrolled so the loads of several cells are in flight together (one
  // wavefront per SIMD at this launch size)
#pragma unroll 4
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n;
       c += (size_t)gridDim.x * 256) {
    const double au = fabs((double)u[c]), av = fabs((double)v[c]), aw = fabs((double)w[c]);
    r[0] = au > r[0] ? au : r[0];
    r[1] = av > r[1] ? av : r[1];
    r[2] = aw > r[2] ? aw : r[2];
    const double nu = (double)nut[c];
    r[3] = nu > r[3] ? nu : r[3];
    const double tg = (double)Tg[c];
    if (!(tg == tg) || isinf(tg)) r[5] = 1.0;
    r[4] = tg > r[4] ? tg : r[4];
    const double ts = (double)Ts[c];
    r[6] = ts > r[6] ? ts : r[6];
    if (Sp) { const double x = (double)Sp[c]; r[7] = x > r[7] ? x : r[7]; }
    if (Qc) { const double x = (double)Qc[c]; r[8] = x > r[8] ? x : r[8]; }
    const double o = (double)om[c];
    r[9] = o > r[9] ? o : r[9];
    const double y = (double)YF[c];
    r[10] = y > r[10] ? y : r[10];
    if (ts >= T_ign) r[11] += 1.0;
    if (o > 1.0e-3 || (tg > 1000.0 && y > 1.0e-3)) r[12] += 1.0;
    const int k = (int)(c / plane);
    if (k < nzb && ts >= T_front && (double)alpha[c] > 0.0) {
      const double i = (double)(c % g.nx);
      r[13] = i > r[13] ? i : r[13];
    }
  }
  // maxima and counts (exact in any order) within each wavefront by
  // shuffles, then over the four wavefronts
  const int t = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
      const double o = __shfl_down(r[q], off, 64);
      if (q == 11 || q == 12) r[q] += o;
      else if (o > r[q]) r[q] = o;
    }
  }
  if ((t & 63) == 0)
    for (int q = 0; q < NRED; ++q) sh[q][t >> 6] = r[q];
  __syncthreads();
  if (t < NRED) {
    double a = sh[t][0];
    for (int wq = 1; wq < (int)(blockDim.x >> 6); ++wq) {
      const double o = sh[t][wq];
      if (t == 11 || t == 12) a += o;
      else if (o > a) a = o;
    }
    part[(size_t)t * gridDim.x + blockIdx.x] = a;
  }
}

__global__ void __launch_bounds__(256)
k_step_reduce_fin(const double* __restrict__ part, int nb, double* __restrict__ out) {
  const int q = blockIdx.x;  // one workgroup per slot
  __shared__ double sh[256];
  const int t = threadIdx.x;
  const bool sum = (q == 11 || q == 12);
  double a = sum ? 0.0 : -1e300;
  if (q == 13) a = -1.0;
  for (int b = t; b < nb; b += 256) {
    const double x = part[(size_t)q * nb + b];
    if (sum) a += x;
    else if (x > a) a = x;
  }
  sh[t] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      if (sum) sh[t] += sh[t + s];
      else if (sh[t + s] > sh[t]) sh[t] = sh[t + s];
    }
    __syncthreads();
  }
  if (t == 0) out[q] = sh[0];
}

// ---- level set (flame_front_3d.py:113-402) --------------------------------
template <typename T>
__device__ __forceinline__ void dz_inv_pm(const GridArgs& g, int k, T& im, T& ip) {
  const int Nz = g.nz;
  if (k == 0) {
    im = T(g.inv_dz[0]);
    ip = Nz > 1 ? T(2.0 / (g.dz[0] + g.dz[1])) : T(g.inv_dz[0]);
  } else if (k == Nz - 1) {
    ip = T(g.inv_dz[Nz - 1]);
    im = T(2.0 / (g.dz[Nz - 1] + g.dz[Nz - 2]));
  } else {
    im = T(2.0 / (g.dz[k] + g.dz[k - 1]));
    ip = T(2.0 / (g.dz[k] + g.dz[k + 1]));
  }
}

// mode 0: godunov_grad_norm + evolve: phi_out = phi - dt vn |grad|
// mode 1: reinit substep (sign from phi0): phi_out = phi - dtau s(|grad|-1)
// mode 2: godunov_grad_norm alone: phi_out = |grad|
// mode 3: reinit_godunov_grad alone (sign from phi0): phi_out = |grad|
template <typename T>
__global__ void __launch_bounds__(256)
k_lset(GridArgs g, const T* __restrict__ phi, const T* __restrict__ phi0,
       const T* __restrict__ vn, T* __restrict__ out, T step, int mode) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t c = cidx(g, k, j, i);
  const size_t sz = (size_t)Nx * Ny;
  const T inv_dx = T(g.inv_dx), inv_dy = T(g.inv_dy);
  T im, ip;
  dz_inv_pm<T>(g, k, im, ip);
  const T pc = phi[c];
  const T dmx = i > 0 ? (pc - phi[c - 1]) * inv_dx : T(0);
  const T dpx = i < Nx - 1 ? (phi[c + 1] - pc) * inv_dx : T(0);
  const int jm = j > 0 ? j - 1 : Ny - 1;
  const int jp = j < Ny - 1 ? j + 1 : 0;
  const T dmy = (pc - phi[cidx(g, k, jm, i)]) * inv_dy;
  const T dpy = (phi[cidx(g, k, jp, i)] - pc) * inv_dy;
  const T dmz = k > 0 ? (pc - phi[c - sz]) * im : T(0);
  const T dpz = k < Nz - 1 ? (phi[c + sz] - pc) * ip : T(0);
  T s = T(1);
  if (mode == 1 || mode == 3) {
    const T p0 = phi0[c];
    s = p0 > T(0) ? T(1) : (p0 < T(0) ? T(-1) : T(0));
  }
  T gxp, gxm, gyp, gym, gzp, gzm;
  if (mode == 0 || mode == 2 || s > T(0)) {
    gxp = fmax(dmx, T(0)); gxm = fmin(dpx, T(0));
    gyp = fmax(dmy, T(0)); gym = fmin(dpy, T(0));
    gzp = fmax(dmz, T(0)); gzm = fmin(dpz, T(0));
  } else {
    gxp = fmin(dmx, T(0)); gxm = fmax(dpx, T(0));
    gyp = fmin(dmy, T(0)); gym = fmax(dpy, T(0));
    gzp = fmin(dmz, T(0)); gzm = fmax(dpz, T(0));
  }
  const T grad = sqrt(gxp * gxp + gxm * gxm + gyp * gyp + gym * gym + gzp * gzp + gzm * gzm);
  if (mode == 0) out[c] = pc - step * vn[c] * grad;
  else if (mode == 1) out[c] = pc - step * s * (grad - T(1));
  else out[c] = grad;
}

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace fw

using namespace fw;

extern "C" {

int fw_step_prologue(const fw_grid* g, const void* T_g, void* rho, void* rho_prev,
                     void* Y_O2, double T_amb, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && T_g && rho && rho_prev && Y_O2, "step_prologue: null argument");
  Launch L = cell_launch(g);
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_prologue<double>, L.grid, L.block, 0, s, grid_args(g),
                       (const double*)T_g, (double*)rho, (double*)rho_prev,
                       (double*)Y_O2, T_amb);
  else
    hipLaunchKernelGGL(k_prologue<float>, L.grid, L.block, 0, s, grid_args(g),
                       (const float*)T_g, (float*)rho, (float*)rho_prev, (float*)Y_O2,
                       (float)T_amb);
  FW_CHECK_LAUNCH("step_prologue");
  return FW_OK;
}

int fw_bed_external_heat(const fw_grid* g, const void* q_rad, void* Q_ext, int n_z_bed,
                         int i_src0, int i_src1, double drip_W_m3, double q_rad_max,
                         void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && q_rad && Q_ext, "bed_external_heat: null argument");
  const int nzb = n_z_bed < g->nz ? n_z_bed : g->nz;
  if (nzb <= 0) return FW_OK;
  Launch L = cell_launch(g);
  L.grid.z = nzb;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_bed_qext<double>, L.grid, L.block, 0, s, grid_args(g), nzb,
                       (const double*)q_rad, (double*)Q_ext, i_src0, i_src1, drip_W_m3,
                       q_rad_max);
  else
    hipLaunchKernelGGL(k_bed_qext<float>, L.grid, L.block, 0, s, grid_args(g), nzb,
                       (const float*)q_rad, (float*)Q_ext, i_src0, i_src1, drip_W_m3,
                       q_rad_max);
  FW_CHECK_LAUNCH("bed_external_heat");
  return FW_OK;
}

int fw_div_target(const fw_grid* g, const void* S_pyro_part, const void* S_dry_part,
                  const void* rho, const void* rho_prev, double dt, void* S_pyro,
                  void* div_target, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && S_pyro_part && S_dry_part && rho && rho_prev && S_pyro && div_target,
             "div_target: null argument");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_div_target<double>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const double*)S_pyro_part, (const double*)S_dry_part,
                       (const double*)rho, (const double*)rho_prev, dt,
                       (double*)S_pyro, (double*)div_target);
  else
    hipLaunchKernelGGL(k_div_target<float>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const float*)S_pyro_part, (const float*)S_dry_part,
                       (const float*)rho, (const float*)rho_prev, (float)dt,
                       (float*)S_pyro, (float*)div_target);
  FW_CHECK_LAUNCH("div_target");
  return FW_OK;
}

int fw_gas_energy_update(const fw_grid* g, void* T_g, const void* T_s, const void* rho,
                         const void* Y_H2O, const void* S_pyro_part,
                         const void* S_dry_part, const void* Q_g_conv,
                         const void* q_rad_gas, void* Q_comb_out, int n_z_bed,
                         int i_src0, int i_src1, double drip_W_m3, double dt_sub,
                         double T_amb, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && T_g && T_s && rho && Y_H2O && S_pyro_part && S_dry_part && Q_g_conv &&
                 q_rad_gas,
             "gas_energy_update: null argument");
  Launch L = cell_launch(g);
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_gas_update<double>, L.grid, L.block, 0, s, grid_args(g),
                       (double*)T_g, (const double*)T_s, (const double*)rho,
                       (const double*)Y_H2O, (const double*)S_pyro_part,
                       (const double*)S_dry_part, (const double*)Q_g_conv,
                       (const double*)q_rad_gas, (double*)Q_comb_out, n_z_bed, i_src0,
                       i_src1, drip_W_m3, dt_sub, T_amb);
  else
    hipLaunchKernelGGL(k_gas_update<float>, L.grid, L.block, 0, s, grid_args(g),
                       (float*)T_g, (const float*)T_s, (const float*)rho,
                       (const float*)Y_H2O, (const float*)S_pyro_part,
                       (const float*)S_dry_part, (const float*)Q_g_conv,
                       (const float*)q_rad_gas, (float*)Q_comb_out, n_z_bed, i_src0,
                       i_src1, (float)drip_W_m3, (float)dt_sub, (float)T_amb);
  FW_CHECK_LAUNCH("gas_energy_update");
  return FW_OK;
}

int fw_gas_update_solid_conduction(const fw_grid* g, void* T_g, const void* T_s,
                                   void* T_s_out, const void* alpha_s, const void* rho,
                                   const void* Y_H2O, const void* S_pyro_part,
                                   const void* S_dry_part, const void* Q_g_conv,
                                   const void* q_rad_gas, void* Q_comb_out, int n_z_bed,
                                   int i_src0, int i_src1, double drip_W_m3, double dt_sub,
                                   double T_amb, double k_solid, double rho_solid,
                                   double cp_solid, double T_s_lo, double T_s_hi,
                                   void* stream) {
  FW_GRID(g);
  return fw::gas_solid_launch(g, T_g, T_s, T_s_out, alpha_s, rho, Y_H2O, S_pyro_part,
                              S_dry_part, Q_g_conv, q_rad_gas, Q_comb_out, n_z_bed, i_src0,
                              i_src1, drip_W_m3, dt_sub, T_amb, k_solid, rho_solid, cp_solid,
                              T_s_lo, T_s_hi, (hipStream_t)stream, nullptr);
}

}  // extern "C"

namespace fw {

int gas_solid_launch(const fw_grid* g, void* T_g, const void* T_s, void* T_s_out,
                     const void* alpha_s, const void* rho, const void* Y_H2O,
                     const void* S_pyro_part, const void* S_dry_part, const void* Q_g_conv,
                     const void* q_rad_gas, void* Q_comb_out, int n_z_bed, int i_src0,
                     int i_src1, double drip_W_m3, double dt_sub, double T_amb,
                     double k_solid, double rho_solid, double cp_solid, double T_s_lo,
                     double T_s_hi, hipStream_t s, const double* sb) {
  FW_REQUIRE(g && T_g && T_s && T_s_out && T_s != T_s_out && alpha_s && rho && Y_H2O &&
                 S_pyro_part && S_dry_part && Q_g_conv && q_rad_gas,
             "gas_update_solid_conduction: bad argument");
  Launch L = cell_launch(g);
  if (g->dtype == FW_F64) {
    const GasSolidArgs<double> a = gas_solid_args<double>(
        T_g, T_s, T_s_out, alpha_s, rho, S_pyro_part, S_dry_part, Q_g_conv, q_rad_gas,
        Q_comb_out, n_z_bed, i_src0, i_src1, drip_W_m3, dt_sub, T_amb, k_solid, rho_solid,
        cp_solid, T_s_lo, T_s_hi);
    hipLaunchKernelGGL(k_gas_solid<double>, L.grid, L.block, 0, s, grid_args(g), a,
                       (const double*)Y_H2O, sb);
  } else {
    const GasSolidArgs<float> a = gas_solid_args<float>(
        T_g, T_s, T_s_out, alpha_s, rho, S_pyro_part, S_dry_part, Q_g_conv, q_rad_gas,
        Q_comb_out, n_z_bed, i_src0, i_src1, drip_W_m3, dt_sub, T_amb, k_solid, rho_solid,
        cp_solid, T_s_lo, T_s_hi);
    hipLaunchKernelGGL(k_gas_solid<float>, L.grid, L.block, 0, s, grid_args(g), a,
                       (const float*)Y_H2O, sb);
  }
  FW_CHECK_LAUNCH("gas_update_solid_conduction");
  return FW_OK;
}

}  // namespace fw

extern "C" {

int fw_add_patch(const fw_grid* g, void* field, int k, int i0, int i1, double add,
                 void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && field && k >= 0 && k < g->nz && i0 >= 0 && i1 <= g->nx,
             "add_patch: bad argument");
  if (i1 <= i0) return FW_OK;
  dim3 gr((i1 - i0 + 255) / 256, g->ny);
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_add_patch<double>, gr, dim3(256), 0, s, grid_args(g),
                       (double*)field, k, i0, i1, add);
  else
    hipLaunchKernelGGL(k_add_patch<float>, gr, dim3(256), 0, s, grid_args(g),
                       (float*)field, k, i0, i1, (float)add);
  FW_CHECK_LAUNCH("add_patch");
  return FW_OK;
}

int fw_flame_mask_age(const fw_grid* g, const void* omega, const void* T_g,
                      const void* Y_fuel, void* cell_age, uint8_t* mask, double dt,
                      void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && omega && T_g && Y_fuel, "flame_mask_age: null argument");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_flame_mask_age<double>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const double*)omega, (const double*)T_g, (const double*)Y_fuel,
                       (double*)cell_age, mask, dt);
  else
    hipLaunchKernelGGL(k_flame_mask_age<float>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const float*)omega, (const float*)T_g, (const float*)Y_fuel,
                       (float*)cell_age, mask, (float)dt);
  FW_CHECK_LAUNCH("flame_mask_age");
  return FW_OK;
}

int fw_step_reduce(const fw_grid* g, const void* u, const void* v, const void* w,
                   const void* nu_t, const void* T_g, const void* T_s,
                   const void* S_pyro, const void* Q_comb, const void* omega,
                   const void* Y_fuel, const void* alpha_s, double T_ign,
                   double T_front, int n_z_bed, double* work,
                   double* out, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && u && v && w && nu_t && T_g && T_s && omega && Y_fuel && alpha_s &&
                 work && out,
             "step_reduce: null argument");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const int nb = FW_REDUCE_BLOCKS;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_step_reduce<double>, dim3(nb), dim3(256), 0, s, grid_args(g), n,
                       (const double*)u, (const double*)v, (const double*)w,
                       (const double*)nu_t, (const double*)T_g, (const double*)T_s,
                       (const double*)S_pyro, (const double*)Q_comb,
                       (const double*)omega, (const double*)Y_fuel,
                       (const double*)alpha_s, T_ign, T_front, n_z_bed, work);
  else
    hipLaunchKernelGGL(k_step_reduce<float>, dim3(nb), dim3(256), 0, s, grid_args(g), n,
                       (const float*)u, (const float*)v, (const float*)w,
                       (const float*)nu_t, (const float*)T_g, (const float*)T_s,
                       (const float*)S_pyro, (const float*)Q_comb, (const float*)omega,
                       (const float*)Y_fuel, (const float*)alpha_s, T_ign, T_front,
                       n_z_bed, work);
  hipLaunchKernelGGL(k_step_reduce_fin, dim3(NRED), dim3(256), 0, s, (const double*)work,
                     nb, out);
  FW_CHECK_LAUNCH("step_reduce");
  return FW_OK;
}

int fw_levelset_update(const fw_grid* g, const void* phi, const void* phi0,
                       const void* v_n, void* phi_out, double step, int mode,
                       void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && phi && phi_out && phi != phi_out && mode >= 0 && mode <= 3,
             "levelset_update: bad argument");
  FW_REQUIRE(mode == 1 || mode == 3 ? phi0 != nullptr : (mode == 2 || v_n != nullptr),
             "levelset_update: mode needs phi0 (reinit) or v_n (evolve)");
  Launch L = cell_launch(g);
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_lset<double>, L.grid, L.block, 0, s, grid_args(g),
                       (const double*)phi, (const double*)phi0, (const double*)v_n,
                       (double*)phi_out, step, mode);
  else
    hipLaunchKernelGGL(k_lset<float>, L.grid, L.block, 0, s, grid_args(g),
                       (const float*)phi, (const float*)phi0, (const float*)v_n,
                       (float*)phi_out, (float)step, mode);
  FW_CHECK_LAUNCH("levelset_update");
  return FW_OK;
}

}  // extern "C"
