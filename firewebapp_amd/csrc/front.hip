// Level-set front and flame-front helpers of flame_front_3d.py, and the O2
// supply limiter of combustion_3d.py: the parts of those modules' call
// surface that the outer step does not fuse into its own kernels.
//
// The level set phi is float64 in every build (LevelSetFront3D keeps it so);
// the other fields are in the grid's storage type.  Each kernel is one
// elementwise or per-column pass: memory-bound, one read and one write per
// field touched.
#include <math.h>
#include <stdlib.h>

#include "fw_common.h"

namespace fw {

static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// LevelSetFront3D.initialize_source_patch (flame_front_3d.py:304-330)
__global__ void __launch_bounds__(256)
k_lset_init_patch(int nx, int ny, int nz, double* __restrict__ phi,
                  const double* __restrict__ x_mid, int i0, int i1, int k_top,
                  double x_front, double half_burnout) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t n = (size_t)nx * ny * nz;
  if (c >= n) return;
  const int i = (int)(c % nx);
  const int k = (int)(c / ((size_t)nx * ny));
  double v;
  if (i0 <= i && i < i1 && k <= k_top) {
    v = -half_burnout;
  } else {
    const double d = x_mid[i] - x_front;
    v = d > 0.0 ? d : fmax(d, -half_burnout);
  }
  phi[c] = v;
}

// mask = (lo_strict ? phi > lo : phi >= lo) && phi <= hi on the first
// n_top cells (the layers below n_z_top), 0 above
__global__ void __launch_bounds__(256)
k_phi_window(size_t n, size_t n_top, const double* __restrict__ phi, double lo,
             int lo_strict, double hi, uint8_t* __restrict__ mask) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  bool m = false;
  if (c < n_top) {
    const double p = phi[c];
    m = (lo_strict ? p > lo : p >= lo) && p <= hi;
  }
  mask[c] = m ? 1 : 0;
}

// update_cell_age (flame_front_3d.py:816-835)
template <typename T>
__global__ void __launch_bounds__(256)
k_cell_age(size_t n, const uint8_t* __restrict__ mask, T* __restrict__ age, T dt) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const T a = age[c];
  if (mask[c]) age[c] = isinf(a) ? T(0) : a + dt;
  else age[c] = T(INFINITY);
}

// compute_q_in_at_front_3d (flame_front_3d.py:440-471): one thread per
// (j, i) column, k ascending; the burst flux goes to the highest band cell
template <typename T>
__global__ void __launch_bounds__(256)
k_q_in_3d(size_t nxy, int nz, const T* __restrict__ q_fr, const T* __restrict__ q_dom,
          const uint8_t* __restrict__ mask, const T* __restrict__ q_burst,
          T* __restrict__ out) {
  const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= nxy) return;
  int k_top = -1;
  for (int k = 0; k < nz; ++k) {
    const size_t c = (size_t)k * nxy + col;
    const bool in = mask[c] != 0;
    out[c] = in ? (q_fr ? q_fr[c] : T(0)) + q_dom[c] : T(0);
    if (in) k_top = k;
  }
  if (q_burst && k_top >= 0) {
    const size_t c = (size_t)k_top * nxy + col;
    out[c] = out[c] + q_burst[col];
  }
}

// LevelSetFront3D.enforce_z_uniformity (flame_front_3d.py:362-377): the z
// mean of each column (summed in k order, as numpy reduces the outer axis),
// written back to every layer
__global__ void __launch_bounds__(256)
k_phi_z_mean(size_t nxy, int nz, double* __restrict__ phi) {
  const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= nxy) return;
  double s = phi[col];
  for (int k = 1; k < nz; ++k) s += phi[(size_t)k * nxy + col];
  const double m = s / (double)nz;
  for (int k = 0; k < nz; ++k) phi[(size_t)k * nxy + col] = m;
}

// combustion_3d.step_o2_supply_rate (combustion_3d.py:184-260): first-order
// upwind O2 inflow over the six faces of each interior cell (1 <= k <= Nz-2,
// 1 <= j <= Ny-2, 1 <= i <= Nx-2; y is not wrapped here), divided by the
// stoichiometric ratio.  Other cells are left as the caller filled them.
template <typename T>
__global__ void __launch_bounds__(256)
k_o2_supply(GridArgs g, const T* __restrict__ rho, const T* __restrict__ u,
            const T* __restrict__ v, const T* __restrict__ w, const T* __restrict__ Y,
            T* __restrict__ out, T s_stoich) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i < 1 || i > g.nx - 2 || j < 1 || j > g.ny - 2 || k < 1 || k > g.nz - 2) return;
  const size_t c = cidx(g, k, j, i);
  const size_t sy = (size_t)g.nx, sz = (size_t)g.nx * g.ny;
  const T inv_dx = T(g.inv_dx), inv_dy = T(g.inv_dy), inv_dz = T(g.inv_dz[k]);
  T m = T(0);
  T f = T(0.5) * (u[c - 1] + u[c]);
  if (f > T(0)) m += rho[c - 1] * f * Y[c - 1] * inv_dx;
  f = T(0.5) * (u[c] + u[c + 1]);
  if (f < T(0)) m += rho[c + 1] * (-f) * Y[c + 1] * inv_dx;
  f = T(0.5) * (v[c - sy] + v[c]);
  if (f > T(0)) m += rho[c - sy] * f * Y[c - sy] * inv_dy;
  f = T(0.5) * (v[c] + v[c + sy]);
  if (f < T(0)) m += rho[c + sy] * (-f) * Y[c + sy] * inv_dy;
  f = T(0.5) * (w[c - sz] + w[c]);
  if (f > T(0)) m += rho[c - sz] * f * Y[c - sz] * inv_dz;
  f = T(0.5) * (w[c] + w[c + sz]);
  if (f < T(0)) m += rho[c + sz] * (-f) * Y[c + sz] * inv_dz;
  out[c] = m / s_stoich;
}

template <typename T>
static int cell_age_impl(const fw_grid* g, const uint8_t* mask, void* age, double dt,
                         hipStream_t s) {
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  hipLaunchKernelGGL(k_cell_age<T>, dim3(nblk(n)), dim3(256), 0, s, n, mask, (T*)age, T(dt));
  FW_CHECK_LAUNCH("update_cell_age");
  return FW_OK;
}

template <typename T>
static int q_in_3d_impl(const fw_grid* g, const void* q_fr, const void* q_dom,
                        const uint8_t* mask, const void* q_burst, void* out, hipStream_t s) {
  const size_t nxy = (size_t)g->nx * g->ny;
  hipLaunchKernelGGL(k_q_in_3d<T>, dim3(nblk(nxy)), dim3(256), 0, s, nxy, g->nz,
                     (const T*)q_fr, (const T*)q_dom, mask, (const T*)q_burst, (T*)out);
  FW_CHECK_LAUNCH("q_in_at_front_3d");
  return FW_OK;
}

template <typename T>
static int o2_supply_impl(const fw_grid* g, const void* rho, const void* u, const void* v,
                          const void* w, const void* Y, void* out, double s_stoich,
                          hipStream_t s) {
  Launch L = cell_launch(g);
  hipLaunchKernelGGL(k_o2_supply<T>, L.grid, L.block, 0, s, grid_args(g), (const T*)rho,
                     (const T*)u, (const T*)v, (const T*)w, (const T*)Y, (T*)out,
                     T(s_stoich));
  FW_CHECK_LAUNCH("o2_supply_rate");
  return FW_OK;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_lset_init_patch(const fw_grid* g, double* phi, const double* x_mid, int i_start,
                       int i_end, int k_top_bed, double L_burnout, void* stream) {
  FW_REQUIRE(g && phi && x_mid && g->nx > 0 && g->ny > 0 && g->nz > 0,
             "lset_init_patch: null argument or empty grid");
  FW_REQUIRE(i_end >= 1 && i_end <= g->nx, "lset_init_patch: i_end %d outside 1..%d", i_end,
             g->nx);
  // x_front = x_mid[i_end - 1] + 0.5 dx: read the one value from the device
  double xm = 0.0;
  hipStream_t s = (hipStream_t)stream;
  FW_CHECK_HIP(hipMemcpyAsync(&xm, x_mid + (i_end - 1), sizeof(double), hipMemcpyDeviceToHost,
                              s));
  FW_CHECK_HIP(hipStreamSynchronize(s));
  const double x_front = xm + 0.5 * g->dx;
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  hipLaunchKernelGGL(k_lset_init_patch, dim3(nblk(n)), dim3(256), 0, s, g->nx, g->ny, g->nz,
                     phi, x_mid, i_start, i_end, k_top_bed, x_front, L_burnout / 2.0);
  FW_CHECK_LAUNCH("lset_init_patch");
  return FW_OK;
}

int fw_phi_window_mask(const fw_grid* g, const double* phi, double lo, int lo_strict,
                       double hi, int n_z_top, uint8_t* mask, void* stream) {
  FW_REQUIRE(g && phi && mask, "phi_window_mask: null argument");
  FW_REQUIRE(n_z_top >= 0, "phi_window_mask: negative layer count");
  const size_t nxy = (size_t)g->nx * g->ny, n = nxy * g->nz;
  const size_t n_top = nxy * (size_t)(n_z_top < g->nz ? n_z_top : g->nz);
  hipLaunchKernelGGL(k_phi_window, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, n_top,
                     phi, lo, lo_strict, hi, mask);
  FW_CHECK_LAUNCH("phi_window_mask");
  return FW_OK;
}

int fw_lset_front_x(const fw_grid* g, const double* phi, int k, int j, double* out_host,
                    void* stream) {
  FW_REQUIRE(g && phi && out_host, "lset_front_x: null argument");
  FW_REQUIRE(k >= 0 && k < g->nz && j >= 0 && j < g->ny, "lset_front_x: row (%d, %d) outside "
             "the grid", k, j);
  const int Nx = g->nx;
  double* h = (double*)malloc((size_t)Nx * sizeof(double));
  FW_REQUIRE(h, "lset_front_x: out of host memory");
  hipStream_t s = (hipStream_t)stream;
  const size_t row = ((size_t)k * g->ny + j) * Nx;
  hipError_t e = hipMemcpyAsync(h, phi + row, Nx * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    free(h);
    return fw_set_error(FW_E_HIP, "lset_front_x: %s", hipGetErrorString(e));
  }
  *out_host = front_x_row(h, Nx, g->dx);
  free(h);
  return FW_OK;
}

int fw_phi_z_mean(const fw_grid* g, double* phi, void* stream) {
  FW_REQUIRE(g && phi, "phi_z_mean: null argument");
  const size_t nxy = (size_t)g->nx * g->ny;
  hipLaunchKernelGGL(k_phi_z_mean, dim3(nblk(nxy)), dim3(256), 0, (hipStream_t)stream, nxy,
                     g->nz, phi);
  FW_CHECK_LAUNCH("phi_z_mean");
  return FW_OK;
}

int fw_update_cell_age(const fw_grid* g, const uint8_t* mask, void* cell_age, double dt,
                       void* stream) {
  FW_REQUIRE(g && mask && cell_age, "update_cell_age: null argument");
  return FW_DISPATCH(g, cell_age_impl, g, mask, cell_age, dt, (hipStream_t)stream);
}

int fw_q_in_at_front_3d(const fw_grid* g, const void* q_frankman, const void* q_dom_fwd,
                        const uint8_t* mask, const void* q_burst_2d, void* out, void* stream) {
  FW_REQUIRE(g && q_dom_fwd && mask && out, "q_in_at_front_3d: null argument");
  return FW_DISPATCH(g, q_in_3d_impl, g, q_frankman, q_dom_fwd, mask, q_burst_2d, out,
                     (hipStream_t)stream);
}

int fw_o2_supply_rate(const fw_grid* g, const void* rho, const void* u, const void* v,
                      const void* w, const void* Y_O2, void* omega_O2_out, double s_stoich,
                      void* stream) {
  FW_GRID(g);
  FW_REQUIRE(rho && u && v && w && Y_O2 && omega_O2_out, "o2_supply_rate: null argument");
  return FW_DISPATCH(g, o2_supply_impl, g, rho, u, v, w, Y_O2, omega_O2_out, s_stoich,
                     (hipStream_t)stream);
}

}  // extern "C"
