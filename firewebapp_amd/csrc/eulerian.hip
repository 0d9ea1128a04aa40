// Eulerian fuel bed (lagrangian_bed_enable = false): per-cell drying,
// single-pool MD2004 pyrolysis with char yield, char and smouldering
// oxidation (pyrolysis_3d.py:270-614), and the gas-solid coupling of the
// time step (coupling_3d.py:76-192).  Every kernel is one thread per cell
// with no neighbour reads, so the bed pools update in place.  Arithmetic is
// float64 for both grid types (the pools are float64 fields; the grid's
// fields are read and written in their own type), in the reference's
// operation order.
#include "fw_common.h"

namespace fw {

namespace eu {
// pyrolysis_3d.py:45 (R), :217-267 (MD2004 and the oxidative path), :255-267
// (drying), :446-562 (char and smouldering); coupling_3d.py:39-60
constexpr double R_GAS = 8.314;
constexpr double A_DRY = 4.29e3, E_DRY = 43800.0, L_VAP_WATER = 2.26e6;
constexpr double HEAT_OF_PYROLYSIS = 4.0e5;
constexpr double CHAR_YIELD = 0.15, ETA = 1.0 - CHAR_YIELD;
constexpr double A_OP = 7.26e4, E_OP = 55000.0, HOR_OP = -1.0e6, Y_O2_MIN_OP = 1.0e-3;
constexpr double A_CHAR = 7.09e5, E_CHAR = 94000.0, HOC_CHAR = 32000000.0;
constexpr double T_CHAR_ONSET = 600.0, Y_O2_MIN_CHAR = 1.0e-3, Q_CHAR_MAX = 5.0e5;
constexpr double A_SMOLD = 1.0e6, E_SMOLD = 80000.0, HOC_SMOLD = 28000000.0;
constexpr double T_SMOLD_ONSET = 473.0, Y_O2_MIN_SMOLD = 1.0e-3, Q_SMOLD_MAX = 2.0e5;
constexpr double MU_GAS = 1.8e-5, K_GAS = 0.026, RHO_SOLID = 500.0, CP_SOLID = 1300.0;
constexpr double EPS_SOLID = 0.9, SIGMA_SB = 5.67e-8, H_GROUND = 5.0, CP_GAS = 1100.0;
}  // namespace eu

// step_drying (pyrolysis_3d.py:270-309): returns Q_dry, updates m_water
__device__ __forceinline__ double eu_drying(double T, double a, double& mw, double dt) {
  if (a <= 0.0 || mw <= 0.0 || T <= 0.0) return 0.0;
  const double k = eu::A_DRY * exp(-eu::E_DRY / (eu::R_GAS * T));
  const double mw_new = mw * exp(-k * dt);
  const double dm = mw - mw_new;
  mw = mw_new;
  return dt > 0.0 ? dm * eu::L_VAP_WATER / dt : 0.0;
}

// step_pyrolysis_md2004 (pyrolysis_3d.py:311-462): S, Q out; m, m_char in/out
__device__ __forceinline__ void eu_md2004(double T, double a, double yO2, double mw,
                                          double inv_mw_init, double A_p, double E_p,
                                          double dt, double& m, double& mc, double& S,
                                          double& Q) {
  S = 0.0;
  Q = 0.0;
  if (a <= 0.0 || T <= 0.0) return;
  double gate = 1.0;
  if (inv_mw_init > 0.0) {
    gate = 1.0 - 100.0 * (mw * inv_mw_init);
    if (gate <= 0.0) return;
  }
  const double m0 = m;
  if (m0 <= 0.0) return;
  const double k_th = A_p * exp(-E_p / (eu::R_GAS * T));
  // yO2 ** N_O2_OP with N_O2_OP = 1 is yO2
  const double k_ox = yO2 > eu::Y_O2_MIN_OP ? eu::A_OP * exp(-eu::E_OP / (eu::R_GAS * T)) * yO2
                                            : 0.0;
  const double k_tot = k_th + k_ox;
  const double m_new = m0 * exp(-k_tot * dt);
  const double dm = (m0 - m_new) * gate;
  const double rate = dt > 0.0 ? dm / dt : 0.0;
  const double f_th = k_tot > 0.0 ? k_th / k_tot : 1.0;
  const double f_ox = k_tot > 0.0 ? k_ox / k_tot : 0.0;
  S = eu::ETA * rate;
  Q = rate * (f_th * eu::HEAT_OF_PYROLYSIS + f_ox * eu::HOR_OP);
  mc += eu::CHAR_YIELD * dm;
  m = m0 - dm;
}

// step_char_oxidation / step_smoldering_oxidation (pyrolysis_3d.py:465-614):
// the same Arrhenius surface oxidation with their own constants
__device__ __forceinline__ double eu_oxidation(double T, double a, double yO2, double dt,
                                               double& m, double A, double E, double HOC,
                                               double T_on, double y_min, double Q_max) {
  if (a <= 0.0) return 0.0;
  const double m0 = m;
  if (m0 <= 1.0e-6 || T < T_on || yO2 < y_min) return 0.0;
  const double k = A * exp(-E / (eu::R_GAS * T));
  double m_dot = k * m0 * yO2;
  if (m_dot * HOC > Q_max) m_dot = Q_max / HOC;
  double mc = m_dot * dt;
  if (mc > 0.5 * m0) mc = 0.5 * m0;
  m = m0 - mc;
  return dt > 0.0 ? mc * HOC / dt : 0.0;
}

// The four processes of one outer step in the reference's order
// (spread_3d.py:2183-2215): Q_pyro = Q_md + Q_dry - Q_char - Q_smold.
// Single processes: mask bit 0 drying, 1 MD2004, 2 char, 3 smouldering; the
// outputs of a process not run are left alone.
template <typename T>
__global__ void __launch_bounds__(256)
k_eu_bed(size_t n, const T* __restrict__ Ts, const T* __restrict__ alpha,
         const T* __restrict__ YO2, double* __restrict__ m_solid, double* __restrict__ m_char,
         double* __restrict__ m_water, int has_moisture, double inv_mw_init, double A_p,
         double E_p, double dt, int mask, T* __restrict__ S_out, T* __restrict__ Q_out,
         T* __restrict__ Q_dry, T* __restrict__ Q_char, T* __restrict__ Q_smold) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const double t = (double)Ts[c], a = (double)alpha[c];
  const double y = YO2 ? (double)YO2[c] : 0.0;
  double qd = 0.0, q = 0.0, s = 0.0;
  if (mask & 1) {
    double mw = m_water[c];
    qd = has_moisture ? eu_drying(t, a, mw, dt) : 0.0;
    m_water[c] = mw;
    if (Q_dry) Q_dry[c] = (T)qd;
  }
  if (mask & 2) {
    double m = m_solid[c], mc = m_char[c];
    eu_md2004(t, a, y, m_water[c], inv_mw_init, A_p, E_p, dt, m, mc, s, q);
    m_solid[c] = m;
    m_char[c] = mc;
  }
  double qc = 0.0, qs = 0.0;
  if (mask & 4) {
    double mc = m_char[c];
    qc = eu_oxidation(t, a, y, dt, mc, eu::A_CHAR, eu::E_CHAR, eu::HOC_CHAR, eu::T_CHAR_ONSET,
                      eu::Y_O2_MIN_CHAR, eu::Q_CHAR_MAX);
    m_char[c] = mc;
    if (Q_char) Q_char[c] = (T)qc;
  }
  if (mask & 8) {
    double mc = m_char[c];
    qs = eu_oxidation(t, a, y, dt, mc, eu::A_SMOLD, eu::E_SMOLD, eu::HOC_SMOLD,
                      eu::T_SMOLD_ONSET, eu::Y_O2_MIN_SMOLD, eu::Q_SMOLD_MAX);
    m_char[c] = mc;
    if (Q_smold) Q_smold[c] = (T)qs;
  }
  if (mask & 2) {
    if (S_out) S_out[c] = (T)s;
    if (Q_out) {
      double qp = q;
      if (mask & 1) qp += qd;
      if (mask & 4) qp -= qc;
      if (mask & 8) qp -= qs;
      Q_out[c] = (T)qp;
    }
  }
}

// step_gas_solid_coupling (coupling_3d.py:76-192), in place on T_g, T_s and
// m_water.  Q_comb either from the array, or (Q_comb == nullptr) formed as the
// time step forms it before the coupling (spread_3d.py:2790, :2916-2956,
// :3094-3099): the drip-torch gas share in the source bed cells plus
// q_rad_gas / dz; the solid share is then subtracted from Q_pyro in those
// cells first, in place, once per call.
struct CouplingArgs {
  double sigma_sav, d_p, pr13, L_v, dt, T_amb, h_conv_mult;
  int q_loss;
  int nzb, i0, i1;
  double drip_gas, drip_solid;
};

template <typename T>
__global__ void __launch_bounds__(256)
k_coupling(GridArgs g, CouplingArgs A, T* __restrict__ Tg_f, T* __restrict__ Ts_f,
           const T* __restrict__ rho_f, const T* __restrict__ u_f, const T* __restrict__ v_f,
           const T* __restrict__ w_f, const T* __restrict__ alpha_f,
           const T* __restrict__ q_rad_in, T* __restrict__ Q_pyro,
           const T* __restrict__ Q_comb, const T* __restrict__ q_rad_gas,
           T* __restrict__ Q_comb_out, double* __restrict__ m_water) {
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int i = (int)(c % g.nx);
  const int k = (int)(c / ((size_t)g.nx * g.ny));
  const bool src = k < A.nzb && i >= A.i0 && i < A.i1;
  double qc;
  if (Q_comb) {
    qc = (double)Q_comb[c];
  } else {
    qc = (src && A.drip_gas != 0.0) ? 0.0 + A.drip_gas : 0.0;
    qc += (double)q_rad_gas[c] / g.dz[k];
    if (Q_comb_out) Q_comb_out[c] = (T)qc;
  }
  double qp = (double)Q_pyro[c];
  if (!Q_comb && src && A.drip_solid != 0.0) {
    qp -= A.drip_solid;
    Q_pyro[c] = (T)qp;
  }
  const double a_s = (double)alpha_f[c], Tg = (double)Tg_f[c], Ts = (double)Ts_f[c];
  const double rho = (double)rho_f[c];
  double Tg_new = Tg + qc / (rho * eu::CP_GAS) * A.dt;
  if (a_s > 0.0) {
    const double u = (double)u_f[c], v = (double)v_f[c], w = (double)w_f[c];
    const double speed = sqrt(u * u + v * v + w * w);
    double Re = rho * speed * A.d_p / eu::MU_GAS;
    if (Re < 0.1) Re = 0.1;
    const double Nu = 2.0 + 0.6 * sqrt(Re) * A.pr13;
    const double h_p = Nu * eu::K_GAS / A.d_p * A.h_conv_mult;
    const double a_v = A.sigma_sav * a_s;
    const double q_conv = h_p * a_v * (Tg_new - Ts);
    double q_loss = 0.0;
    if (A.q_loss) {
      const double t2 = Ts * Ts, a2 = A.T_amb * A.T_amb;
      q_loss = eu::EPS_SOLID * eu::SIGMA_SB * (t2 * t2 - a2 * a2) * a_v;
    }
    const double q_rad_vol = (double)q_rad_in[c] / g.dz[k];
    const double q_ground = k == 0 ? eu::H_GROUND * (Ts - A.T_amb) / g.dz[0] : 0.0;
    const double q_in = q_rad_vol + q_conv - q_loss - q_ground;
    const double mw = m_water[c];
    double q_res;
    if (mw > 0.0 && A.L_v > 0.0 && A.dt > 0.0 && q_in > 0.0) {
      const double q_max = mw * A.L_v / A.dt;
      const double q_use = q_in < q_max ? q_in : q_max;
      double mw_new = mw - q_use * A.dt / A.L_v;
      if (mw_new < 0.0) mw_new = 0.0;
      m_water[c] = mw_new;
      q_res = q_in - q_use;
    } else {
      q_res = q_in;
    }
    const double C_s = eu::RHO_SOLID * eu::CP_SOLID * a_s;
    const double Ts_new = C_s > 0.0 ? Ts + (q_res - qp) / C_s * A.dt : Ts;
    Tg_new = Tg_new - q_conv / (rho * eu::CP_GAS) * A.dt;
    Ts_f[c] = (T)Ts_new;
  }
  Tg_f[c] = (T)Tg_new;
}

static inline unsigned eu_blocks(size_t n) { return (unsigned)((n + 255) / 256); }

int eu_bed_launch(const fw_grid* g, const void* T_s, const void* alpha_s, const void* Y_O2,
                  double* m_solid, double* m_char, double* m_water, int has_moisture,
                  double m_water_init, double A_p, double E_p, double dt, int mask,
                  void* S_out, void* Q_out, void* Q_dry, void* Q_char, void* Q_smold,
                  hipStream_t s) {
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const double inv = m_water_init > 0.0 ? 1.0 / m_water_init : 0.0;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_eu_bed<double>, dim3(eu_blocks(n)), dim3(256), 0, s, n,
                       (const double*)T_s, (const double*)alpha_s, (const double*)Y_O2,
                       m_solid, m_char, m_water, has_moisture, inv, A_p, E_p, dt, mask,
                       (double*)S_out, (double*)Q_out, (double*)Q_dry, (double*)Q_char,
                       (double*)Q_smold);
  else
    hipLaunchKernelGGL(k_eu_bed<float>, dim3(eu_blocks(n)), dim3(256), 0, s, n,
                       (const float*)T_s, (const float*)alpha_s, (const float*)Y_O2, m_solid,
                       m_char, m_water, has_moisture, inv, A_p, E_p, dt, mask, (float*)S_out,
                       (float*)Q_out, (float*)Q_dry, (float*)Q_char, (float*)Q_smold);
  FW_CHECK_LAUNCH("eulerian_bed");
  return FW_OK;
}

int coupling_launch(const fw_grid* g, void* T_g, void* T_s, const void* rho, const void* u,
                    const void* v, const void* w, const void* alpha_s, double sigma_sav,
                    const void* q_rad_in, void* Q_pyro, const void* Q_comb,
                    const void* q_rad_gas, void* Q_comb_out, double* m_water, double L_v,
                    double dt, double T_amb, int q_loss, double h_conv_mult, int nzb, int i0,
                    int i1, double drip_gas, double drip_solid, hipStream_t s) {
  if (sigma_sav <= 0.0) return FW_OK;  // the reference returns without a change
  CouplingArgs A;
  A.sigma_sav = sigma_sav;
  A.d_p = 4.0 / sigma_sav;
  A.pr13 = pow(0.7, 1.0 / 3.0);
  A.L_v = L_v;
  A.dt = dt;
  A.T_amb = T_amb;
  A.h_conv_mult = h_conv_mult;
  A.q_loss = q_loss;
  A.nzb = nzb;
  A.i0 = i0;
  A.i1 = i1;
  A.drip_gas = drip_gas;
  A.drip_solid = drip_solid;
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_coupling<double>, dim3(eu_blocks(n)), dim3(256), 0, s, grid_args(g),
                       A, (double*)T_g, (double*)T_s, (const double*)rho, (const double*)u,
                       (const double*)v, (const double*)w, (const double*)alpha_s,
                       (const double*)q_rad_in, (double*)Q_pyro, (const double*)Q_comb,
                       (const double*)q_rad_gas, (double*)Q_comb_out, m_water);
  else
    hipLaunchKernelGGL(k_coupling<float>, dim3(eu_blocks(n)), dim3(256), 0, s, grid_args(g),
                       A, (float*)T_g, (float*)T_s, (const float*)rho, (const float*)u,
                       (const float*)v, (const float*)w, (const float*)alpha_s,
                       (const float*)q_rad_in, (float*)Q_pyro, (const float*)Q_comb,
                       (const float*)q_rad_gas, (float*)Q_comb_out, m_water);
  FW_CHECK_LAUNCH("gas_solid_coupling");
  return FW_OK;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_pyro_step(const fw_grid* g, const void* T_s, const void* alpha_s, const void* Y_O2,
                 double* m_solid, double* m_char, double* m_water, int has_moisture,
                 double m_water_init, double A_p, double E_p, double dt, int processes,
                 void* S_pyro_out, void* Q_pyro_out, void* Q_dry_out, void* Q_char_out,
                 void* Q_smold_out, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(T_s && alpha_s, "pyro_step: null argument");
  FW_REQUIRE(processes > 0 && processes < 16, "pyro_step: processes must be a 1..15 mask");
  FW_REQUIRE(!(processes & 1) || m_water, "pyro_step: drying needs m_water");
  FW_REQUIRE(!(processes & 2) || (m_solid && m_char && m_water && S_pyro_out && Q_pyro_out),
             "pyro_step: MD2004 needs m_solid, m_char, m_water, S_pyro_out, Q_pyro_out");
  FW_REQUIRE(!(processes & 12) || (m_char && Y_O2), "pyro_step: oxidation needs m_char, Y_O2");
  FW_REQUIRE(!(processes & 2) || Y_O2, "pyro_step: MD2004 needs Y_O2");
  return eu_bed_launch(g, T_s, alpha_s, Y_O2, m_solid, m_char, m_water, has_moisture,
                       m_water_init, A_p, E_p, dt, processes, S_pyro_out, Q_pyro_out,
                       Q_dry_out, Q_char_out, Q_smold_out, (hipStream_t)stream);
}

int fw_gas_solid_coupling(const fw_grid* g, void* T_g, void* T_s, const void* rho,
                          const void* u, const void* v, const void* w, const void* alpha_s,
                          double sigma_sav, const void* q_rad_in, const void* Q_pyro,
                          const void* Q_comb, double* m_water, double L_v, double dt,
                          double T_amb, int q_loss_enable, double h_conv_mult, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(T_g && T_s && rho && u && v && w && alpha_s && q_rad_in && Q_pyro && Q_comb &&
                 m_water,
             "gas_solid_coupling: null argument");
  return coupling_launch(g, T_g, T_s, rho, u, v, w, alpha_s, sigma_sav, q_rad_in,
                         (void*)Q_pyro, Q_comb, nullptr, nullptr, m_water, L_v, dt, T_amb,
                         q_loss_enable, h_conv_mult, 0, 0, 0, 0.0, 0.0, (hipStream_t)stream);
}

}  // extern "C"
