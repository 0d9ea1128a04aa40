// Realizable k-epsilon step (turbulence_3d.py:481-858) with the strain and
// vorticity magnitudes (turbulence_3d.py:391-478) computed in the same
// kernel: they are consumed only at their own cell, so fusing them removes
// two full-field writes and reads per step without changing any value.
// dT/dz uses one-sided differences at k=0 and k=Nz-1, the fixed upstream form
// encoded by the golden vectors (docs/js/physics/kepsilon.js:24-31,233-236).
// Production call: bvg_factor = 0, L-min and Durbin caps off
// (spread_3d.py:2409-2426).
#include "fw_common.h"

namespace fw {

template <typename T>
__device__ __forceinline__ T inv_dz_central(const GridArgs& g, int k) {
  if (k == 0) return T(g.inv_da[0]);
  if (k == g.nz - 1) return T(g.inv_db[g.nz - 1]);
  return T(1.0 / (g.d_below[k] + g.d_above[k]));
}

template <typename T>
__global__ void __launch_bounds__(256)
k_kepsilon(GridArgs g, const T* __restrict__ kt, const T* __restrict__ et,
           T* __restrict__ kt_o, T* __restrict__ et_o, T* __restrict__ nut_o,
           const T* __restrict__ u, const T* __restrict__ v,
           const T* __restrict__ w, const T* __restrict__ Tg,
           const T* __restrict__ alpha, const T* __restrict__ u_inlet,
           const T* __restrict__ k_wall, const T* __restrict__ e_wall,
           T sigma_sav, T dt, T T_amb, T beta_p, T beta_d) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t c = cidx(g, k, j, i);
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  const size_t rm2 = c + (size_t)(wrap(j - 2, Ny) - j) * sy;
  const size_t rm1 = c + (size_t)(wrap(j - 1, Ny) - j) * sy;
  const size_t rp1 = c + (size_t)(wrap(j + 1, Ny) - j) * sy;
  const size_t rp2 = c + (size_t)(wrap(j + 2, Ny) - j) * sy;
  const T inv_dx = T(g.inv_dx), inv_dy = T(g.inv_dy);
  const T idzc = inv_dz_central<T>(g, k);

  // ---- strain / vorticity (Way-B ghosts) ----
  const T ui = u[c], vi = v[c], wi = w[c];
  T S2, O2;
  {
    T uxL, vxL, wxL, uxR, vxR, wxR, uzL, vzL, wzL, uzR, vzR, wzR;
    if (i == 0) { uxL = u_inlet[(size_t)k * Ny + j]; vxL = T(0); wxL = T(0); }
    else { uxL = u[c - 1]; vxL = v[c - 1]; wxL = w[c - 1]; }
    if (i == Nx - 1) { uxR = ui; vxR = vi; wxR = wi; }
    else { uxR = u[c + 1]; vxR = v[c + 1]; wxR = w[c + 1]; }
    if (k == 0) { uzL = T(0); vzL = T(0); wzL = T(0); }
    else { uzL = u[c - sz]; vzL = v[c - sz]; wzL = w[c - sz]; }
    if (k == Nz - 1) { uzR = ui; vzR = vi; wzR = wi; }
    else { uzR = u[c + sz]; vzR = v[c + sz]; wzR = w[c + sz]; }
    T dudx = (uxR - uxL) * T(0.5) * inv_dx;
    T dudy = (u[rp1] - u[rm1]) * T(0.5) * inv_dy;
    T dudz = (uzR - uzL) * idzc;
    T dvdx = (vxR - vxL) * T(0.5) * inv_dx;
    T dvdy = (v[rp1] - v[rm1]) * T(0.5) * inv_dy;
    T dvdz = (vzR - vzL) * idzc;
    T dwdx = (wxR - wxL) * T(0.5) * inv_dx;
    T dwdy = (w[rp1] - w[rm1]) * T(0.5) * inv_dy;
    T dwdz = (wzR - wzL) * idzc;
    T S12 = T(0.5) * (dudy + dvdx);
    T S13 = T(0.5) * (dudz + dwdx);
    T S23 = T(0.5) * (dvdz + dwdy);
    S2 = T(2) * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) +
         T(4) * (S12 * S12 + S13 * S13 + S23 * S23);
    T O12 = T(0.5) * (dudy - dvdx);
    T O13 = T(0.5) * (dudz - dwdx);
    T O23 = T(0.5) * (dvdz - dwdy);
    O2 = T(4) * (O12 * O12 + O13 * O13 + O23 * O23);
  }

  const T K_MIN = T(1.0e-8), EPS_MIN = T(1.0e-8);
  const T kraw = kt[c], eraw = et[c];
  const T k_loc = kraw < K_MIN ? K_MIN : kraw;
  const T e_loc = eraw < EPS_MIN ? EPS_MIN : eraw;
  const T kxL = (i == 0) ? K_MIN : kt[c - 1];
  const T exL = (i == 0) ? EPS_MIN : et[c - 1];
  const T kxR = (i == Nx - 1) ? k_loc : kt[c + 1];
  const T exR = (i == Nx - 1) ? e_loc : et[c + 1];
  const T kzL = (k == 0) ? k_wall[(size_t)j * Nx + i] : kt[c - sz];
  const T ezL = (k == 0) ? e_wall[(size_t)j * Nx + i] : et[c - sz];
  const T kzR = (k == Nz - 1) ? k_loc : kt[c + sz];
  const T ezR = (k == Nz - 1) ? e_loc : et[c + sz];

  const T U_star = sqrt(T(0.5) * (S2 + O2));
  const T A0 = T(4.04), AS = T(4.5);
  T C_mu = T(1) / (A0 + AS * U_star * k_loc / e_loc);
  T nu_t = C_mu * k_loc * k_loc / e_loc;
  if (nu_t < T(0)) nu_t = T(0);
  const T P_k = nu_t * S2;
  T P_k_lim = P_k;
  if (P_k_lim > T(10.0) * e_loc) P_k_lim = T(10.0) * e_loc;

  T dTdz;
  if (k == 0) dTdz = (Tg[c + sz] - Tg[c]) * idzc;
  else if (k == Nz - 1) dTdz = (Tg[c] - Tg[c - sz]) * idzc;
  else dTdz = (Tg[c + sz] - Tg[c - sz]) * idzc;
  T Tb = Tg[c];
  if (Tb < T_amb) Tb = T_amb;
  T G_k = (nu_t / T(0.85)) * (T(9.81) / Tb) * dTdz;
  if (G_k < T(0)) G_k = T(0);
  T G_kk = G_k;
  if (G_kk > P_k_lim) G_kk = P_k_lim;

  T u_h = sqrt(ui * ui + vi * vi);
  if (u_h < T(0.01)) u_h = T(0.01);
  const T C_3eps = tanh(fabs(wi) / u_h);

  const T a_s = alpha[c];
  T P_kc = T(0), D_kc = T(0);
  if (a_s > T(0)) {
    const T a_v = sigma_sav * a_s;
    const T speed = sqrt(ui * ui + vi * vi + wi * wi);
    const T cd = T(1.0) * a_v * speed;
    P_kc = beta_p * cd * speed * speed;
    D_kc = beta_d * cd * k_loc;
  }

  const bool mx = (i >= 2 && i <= Nx - 3), mz = (k >= 2 && k <= Nz - 3);
  const T inv_dx2 = T(g.inv_dx2), inv_dy2 = T(g.inv_dy2);
  const T inv_dz = T(g.inv_dz[k]);
  const T inv_da = T(g.inv_da[k]), inv_db = T(g.inv_db[k]);
  const T dzm = T(0.5) * (T(g.d_above[k]) + T(g.d_below[k]));

  // advection + diffusion of one of (k, eps)
  auto transport = [&](const T* __restrict__ f, T loc, T fxL, T fxR, T fzL,
                       T fzR, T& adv, T& lap) {
    T ax, az;
    if (mx) {
      T fp = muscl_face<T>(f[c - 1], loc, f[c + 1], f[c + 2], ui);
      T fm = muscl_face<T>(f[c - 2], f[c - 1], loc, f[c + 1], ui);
      ax = ui * (fp - fm) * inv_dx;
    } else {
      ax = (ui >= T(0)) ? ui * (loc - fxL) * inv_dx : ui * (fxR - loc) * inv_dx;
    }
    const T fjm1 = f[rm1], fjp1 = f[rp1];
    T fyp = muscl_face<T>(fjm1, loc, fjp1, f[rp2], vi);
    T fym = muscl_face<T>(f[rm2], fjm1, loc, fjp1, vi);
    T ay = vi * (fyp - fym) * inv_dy;
    if (mz) {
      T fp = muscl_face<T>(f[c - sz], loc, f[c + sz], f[c + 2 * sz], wi);
      T fm = muscl_face<T>(f[c - 2 * sz], f[c - sz], loc, f[c + sz], wi);
      az = wi * (fp - fm) / dzm;
    } else {
      az = (wi >= T(0)) ? wi * (loc - fzL) * inv_db : wi * (fzR - loc) * inv_da;
    }
    adv = ax + ay + az;
    T d2x = (fxR - T(2) * loc + fxL) * inv_dx2;
    T d2y = (fjp1 - T(2) * loc + fjm1) * inv_dy2;
    T d2z = ((fzR - loc) * inv_da - (loc - fzL) * inv_db) * inv_dz;
    lap = d2x + d2y + d2z;
  };

  T adv_k, lap_k, adv_e, lap_e;
  transport(kt, k_loc, kxL, kxR, kzL, kzR, adv_k, lap_k);
  const T alpha_k = T(1.5e-5) + nu_t / T(1.0);
  const T diff_k = alpha_k * lap_k;
  const T S_pos_k = P_k_lim + G_kk + P_kc + T(0);
  const T S_neg_k = e_loc + D_kc;
  T k_new = (k_loc + (-adv_k + diff_k + S_pos_k) * dt) / (T(1) + S_neg_k * dt / k_loc);
  if (k_new < K_MIN) k_new = K_MIN;

  transport(et, e_loc, exL, exR, ezL, ezR, adv_e, lap_e);
  const T alpha_eps = T(1.5e-5) + nu_t / T(1.3);
  const T diff_e = alpha_eps * lap_e;
  const T S_mag = sqrt(S2);
  const T eta = S_mag * k_loc / e_loc;
  const T eta3 = eta * (eta * eta);
  T rng = T(0.09) * eta3 * (T(1) - eta / T(4.38)) / (T(1) + T(0.012) * eta3);
  if (rng < T(0)) rng = T(0);
  const T C2 = T(1.92) + rng;
  const T ek = e_loc / k_loc;
  const T S_pos_e = T(1.44) * (P_k_lim + C_3eps * G_k) * ek + T(0.9) * P_kc * ek;
  const T S_neg_e = C2 * e_loc * ek + T(0.9) * D_kc * ek;
  T e_new = (e_loc + (-adv_e + diff_e + S_pos_e) * dt) / (T(1) + S_neg_e * dt / e_loc);
  if (e_new < EPS_MIN) e_new = EPS_MIN;

  T C_mu_n = T(1) / (A0 + AS * U_star * k_new / e_new);
  T nu_n = C_mu_n * k_new * k_new / e_new;
  if (nu_n < T(0)) nu_n = T(0);
  kt_o[c] = k_new;
  et_o[c] = e_new;
  nut_o[c] = nu_n;
}

template <typename T>
static int kepsilon_impl(const fw_grid* g, const void* kt, const void* et,
                         void* kt_o, void* et_o, void* nut_o, const void* u,
                         const void* v, const void* w, const void* Tg,
                         const void* alpha, const void* u_inlet,
                         const void* k_wall, const void* e_wall,
                         double sigma_sav, double dt, double T_amb,
                         double beta_p, double beta_d, hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_kepsilon<T>, L.grid, L.block, 0, s, grid_args(g),
                     (const T*)kt, (const T*)et, (T*)kt_o, (T*)et_o, (T*)nut_o,
                     (const T*)u, (const T*)v, (const T*)w, (const T*)Tg,
                     (const T*)alpha, (const T*)u_inlet, (const T*)k_wall,
                     (const T*)e_wall, T(sigma_sav), T(dt), T(T_amb),
                     T(beta_p), T(beta_d));
  FW_CHECK_LAUNCH("k_epsilon");
  return FW_OK;
}

}  // namespace fw

extern "C" int fw_k_epsilon(const fw_grid* g, const void* k_turb,
                            const void* eps_turb, void* k_out, void* eps_out,
                            void* nu_t_out, const void* u, const void* v,
                            const void* w, const void* T_g,
                            const void* alpha_s, const void* u_inlet,
                            const void* k_wall_ghost,
                            const void* eps_wall_ghost, double sigma_sav,
                            double dt, double T_amb, double beta_p,
                            double beta_d, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(k_turb != k_out && eps_turb != eps_out,
             "k_epsilon: outputs must not alias inputs");
  FW_REQUIRE(g->nz >= 2, "k_epsilon: needs Nz >= 2");
  return FW_DISPATCH(g, fw::kepsilon_impl, g, k_turb, eps_turb, k_out, eps_out,
                     nu_t_out, u, v, w, T_g, alpha_s, u_inlet, k_wall_ghost,
                     eps_wall_ghost, sigma_sav, dt, T_amb, beta_p, beta_d,
                     (hipStream_t)stream);
}
