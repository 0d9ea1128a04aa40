// The particle path's per-sub-step gas energy update and vertical solid
// conduction at one cell (spread_3d.py:2916-2957, :3104-3146 and
// solid_conduction_3d.py:56-127), shared by k_gas_solid and the last
// species turbulent-diffusion pass that it is fused into.
#pragma once
#include "fw_common.h"

namespace fw {

template <typename T>
struct GasSolidArgs {
  T* Tg;                    // in place
  const T* Ts;
  T* Ts_out;                // must not alias Ts or any field the launch reads
  const T *alpha, *rho, *Sp, *Sd, *Qgc, *qrg;
  T* Qcomb_out;             // may be null
  int nzb, i0, i1;
  T drip, dt_sub, T_amb, k_solid, inv_rho_cp, ts_lo, ts_hi;
};

// The same arguments on the host side, type-free (turb3_launch's fused pass).
struct GasSolidParams {
  void* T_g;
  const void* T_s;
  void* T_s_out;
  const void *alpha_s, *rho, *S_pyro_part, *S_dry_part, *Q_g_conv, *q_rad_gas;
  void* Q_comb_out;
  int n_z_bed, i_src0, i_src1;
  double drip_W_m3, dt_sub, T_amb, k_solid, rho_solid, cp_solid, T_s_lo, T_s_hi;
};

// The arguments in the kernel's storage type (the scalars rounded once, on
// the host, as the launch passed them before).
template <typename T>
inline GasSolidArgs<T> gas_solid_args(void* T_g, const void* T_s, void* T_s_out,
                                      const void* alpha_s, const void* rho,
                                      const void* S_pyro_part, const void* S_dry_part,
                                      const void* Q_g_conv, const void* q_rad_gas,
                                      void* Q_comb_out, int n_z_bed, int i_src0, int i_src1,
                                      double drip_W_m3, double dt_sub, double T_amb,
                                      double k_solid, double rho_solid, double cp_solid,
                                      double T_s_lo, double T_s_hi) {
  GasSolidArgs<T> a;
  a.Tg = (T*)T_g;
  a.Ts = (const T*)T_s;
  a.Ts_out = (T*)T_s_out;
  a.alpha = (const T*)alpha_s;
  a.rho = (const T*)rho;
  a.Sp = (const T*)S_pyro_part;
  a.Sd = (const T*)S_dry_part;
  a.Qgc = (const T*)Q_g_conv;
  a.qrg = (const T*)q_rad_gas;
  a.Qcomb_out = (T*)Q_comb_out;
  a.nzb = n_z_bed;
  a.i0 = i_src0;
  a.i1 = i_src1;
  a.drip = (T)drip_W_m3;
  a.dt_sub = (T)dt_sub;
  a.T_amb = (T)T_amb;
  a.k_solid = (T)k_solid;
  a.inv_rho_cp = (T)(1.0 / (rho_solid * cp_solid));
  a.ts_lo = (T)T_s_lo;
  a.ts_hi = (T)T_s_hi;
  return a;
}

template <typename T>
inline GasSolidArgs<T> gas_solid_args(const GasSolidParams& p) {
  return gas_solid_args<T>(p.T_g, p.T_s, p.T_s_out, p.alpha_s, p.rho, p.S_pyro_part,
                           p.S_dry_part, p.Q_g_conv, p.q_rad_gas, p.Q_comb_out, p.n_z_bed,
                           p.i_src0, p.i_src1, p.drip_W_m3, p.dt_sub, p.T_amb, p.k_solid,
                           p.rho_solid, p.cp_solid, p.T_s_lo, p.T_s_hi);
}

// y: the cell's Y_H2O after the sub-step's species transport and diffusion
template <typename T>
__device__ __forceinline__ void gas_solid_cell(const GridArgs& g, int i, int k, size_t c,
                                               const GasSolidArgs<T>& a, T dt_sub, T y) {
  const size_t sz = (size_t)g.nx * g.ny;
  const T tsc = a.Ts[c];
  // gas energy
  T Q = T(0);
  if (k < a.nzb && i >= a.i0 && i < a.i1) Q += a.drip;
  Q += a.qrg[c] / T(g.dz[k]);
  if (a.Qcomb_out) a.Qcomb_out[c] = Q;
  const T cp = (T(1) - y) * T(1100.0) + y * T(2000.0);
  const T tg = a.Tg[c];
  T da = tg - tsc;
  da = da > T(0) ? da : T(0);
  T db = tg - T(373.15);
  db = db > T(0) ? db : T(0);
  const T debit = (a.Sp[c] * da + a.Sd[c] * db) * T(2000.0);
  const T r = a.rho[c];
  const T inv = dt_sub / ((r > T(1.0e-3) ? r : T(1.0e-3)) * cp);
  T nv = tg + (Q - a.Qgc[c] - debit) * inv;
  nv = (nv > a.T_amb || nv != nv) ? nv : a.T_amb;
  a.Tg[c] = nv;
  // solid conduction
  const T al = a.alpha[c];
  T ns = tsc;
  if (al > T(0)) {
    T flux_up = T(0), flux_dn = T(0);
    if (k < g.nz - 1) {
      const T aa = a.alpha[c + sz];
      if (aa > T(0)) {
        const T a_eff = T(2) * al * aa / (al + aa);
        flux_up = a.k_solid * a_eff * (a.Ts[c + sz] - tsc) / T(g.d_above[k]);
      }
    }
    if (k > 0) {
      const T ab = a.alpha[c - sz];
      if (ab > T(0)) {
        const T a_eff = T(2) * al * ab / (al + ab);
        flux_dn = a.k_solid * a_eff * (a.Ts[c - sz] - tsc) / T(g.d_below[k]);
      }
    }
    const T src = (flux_up + flux_dn) / T(g.dz[k]);
    const T dT = dt_sub * src * a.inv_rho_cp / al;
    ns = tsc + dT;
  }
  ns = ns > a.ts_hi ? a.ts_hi : ns;
  ns = ns < a.ts_lo ? a.ts_lo : ns;
  a.Ts_out[c] = ns;
}

}  // namespace fw
