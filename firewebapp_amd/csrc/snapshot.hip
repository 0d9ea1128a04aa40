// Level-set forcing and the fields the periodic snapshot writes besides the
// state itself (spread_3d.py:3227-3256, :3537-3670): the flame signed
// distance, the ahead-band mask and the forward heat flux into it, the front
// speed of the active level set, the mixing time scale and the per-cell
// aggregates of the bed particles.  With a passive level set nothing in the
// step reads them and they are computed only when a snapshot is due.
#include "fw_common.h"

namespace fw {

// ---- flame signed distance (flame_front_3d.py:756-800) ---------------------
//
// phi = +dist(cell, nearest active cell) outside the flame and
//       -dist(cell, nearest inactive cell) inside, distances between cell
// centres with per-axis sampling (dz_eff, dy, dx); +-1e6 when no cell / every
// cell is active.  The exact Euclidean transform is separable in the squared
// distance: D(k,j,i) = min over (k',j',i') of (dz dk)^2 + (dy dj)^2 + (dx di)^2
// is three 1-D passes, each D_out(e) = min_e' D_in(e') + (h (e - e'))^2 along
// one axis.  Both fields (distance to the active set, A, and to the inactive
// set, B) go through the same passes.  Each 1-D pass takes the minimum by
// brute force over the line, staged in LDS: O(L) per cell, exact, and run
// only at snapshot time.

#define EDT_SENTINEL 1.0e6

template <typename T>
__global__ void __launch_bounds__(256)
k_edt_init(size_t n, const T* __restrict__ om, const T* __restrict__ Tg,
           const T* __restrict__ YF, double* __restrict__ A, double* __restrict__ B) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  // the reference's float64 comparisons, on the stored values
  const bool act = ((double)om[c] > 1.0e-3) ||
                   (((double)Tg[c] > 1000.0) && ((double)YF[c] > 1.0e-3));
  const double inf = __builtin_huge_val();
  A[c] = act ? 0.0 : inf;
  B[c] = act ? inf : 0.0;
}

// Lines of one pass.  Line l starts at (l / q) * P + (l % q) * r and its
// elements are es apart.  A workgroup takes TX consecutive lines whole, so the
// pass can write in place: every element it overwrites belongs to a line it
// has already read into LDS.  With es == 1 the elements of a line are
// contiguous and the staging walks line by line; otherwise consecutive lines
// are contiguous (r == 1) and the staging walks across lines.
struct EdtPass {
  size_t q, P, r, es;
  int L, nlines, TX;
  double h;
};

__device__ __forceinline__ size_t edt_addr(const EdtPass& p, int l, int e) {
  return ((size_t)l / p.q) * p.P + ((size_t)l % p.q) * p.r + (size_t)e * p.es;
}

// FINAL: the last pass also forms phi from A and B into A (A == 0 exactly on
// active cells: their own distance term is zero and the others are positive).
template <bool FINAL>
__global__ void __launch_bounds__(256)
k_edt_line(double* __restrict__ A, double* __restrict__ B, EdtPass p) {
  extern __shared__ double lds[];
  // LDS tile [TX][L] when walking along lines (lanes of one line read the
  // same candidate: a broadcast), [L][TX] when walking across them (lanes of
  // consecutive lines read consecutive words)
  double* la = lds;
  double* lb = lds + (size_t)p.TX * p.L;
  const int l0 = blockIdx.x * p.TX;
  const int nt = min(p.TX, p.nlines - l0);
  const int tot = nt * p.L;
  const bool along = p.es == 1;
  const int st = along ? 1 : p.TX;      // LDS stride between elements
  const int sl = along ? p.L : 1;       // LDS stride between lines
  for (int x = threadIdx.x; x < tot; x += 256) {
    const int t = along ? x / p.L : x % nt;
    const int e = along ? x % p.L : x / nt;
    const size_t a = edt_addr(p, l0 + t, e);
    la[t * sl + e * st] = A[a];
    lb[t * sl + e * st] = B[a];
  }
  __syncthreads();
  for (int x = threadIdx.x; x < tot; x += 256) {
    const int t = along ? x / p.L : x % nt;
    const int e = along ? x % p.L : x / nt;
    const double* ra = la + t * sl;
    const double* rb = lb + t * sl;
    double da = ra[e * st], db = rb[e * st];
    for (int f = 0; f < p.L; ++f) {
      const double d = p.h * (double)(e - f);
      const double d2 = d * d;
      da = fmin(da, ra[f * st] + d2);
      db = fmin(db, rb[f * st] + d2);
    }
    const size_t a = edt_addr(p, l0 + t, e);
    if (FINAL) {
      const double inf = __builtin_huge_val();
      const bool act = ra[e * st] == 0.0;
      A[a] = act ? (db < inf ? -sqrt(db) : -EDT_SENTINEL)
                 : (da < inf ? sqrt(da) : EDT_SENTINEL);
    } else {
      A[a] = da;
      B[a] = db;
    }
  }
}

static int edt_tile(int L) {
  int tx = 64;
  while (tx > 1 && (size_t)tx * L * 16 > 32768) tx /= 2;
  return tx;
}

static int edt_pass(double* A, double* B, EdtPass p, bool final_pass, hipStream_t s) {
  p.TX = edt_tile(p.L);
  const size_t lds = (size_t)p.TX * p.L * 2 * sizeof(double);
  const dim3 grid((p.nlines + p.TX - 1) / p.TX);
  if (final_pass)
    hipLaunchKernelGGL(k_edt_line<true>, grid, dim3(256), lds, s, A, B, p);
  else
    hipLaunchKernelGGL(k_edt_line<false>, grid, dim3(256), lds, s, A, B, p);
  FW_CHECK_LAUNCH("phi_flame");
  return FW_OK;
}

// ---- LevelSet3D.ahead_band_mask + bed-top clip -----------------------------
// (flame_front_3d.py:387-389, spread_3d.py:2555-2565)
__global__ void __launch_bounds__(256)
k_ahead_band(size_t n, size_t n_bed, const double* __restrict__ phi, double band,
             unsigned char* __restrict__ mask) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const double v = phi[c];
  mask[c] = (c < n_bed && v > 0.0 && v <= band) ? 1 : 0;
}

// ---- q_in_at_front_2d (spread_3d.py:3227-3233) -----------------------------
// compute_q_in_at_front_3d with no Frankman flux and no burst term is
// where(mask, 0 + q_dom_fwd, 0); its z sum, accumulated k = 0, 1, ...
template <typename T>
__global__ void __launch_bounds__(256)
k_q_in_2d(size_t nxy, int nz, const T* __restrict__ q, const unsigned char* __restrict__ mask,
          double* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= nxy) return;
  double s = 0.0;
  for (int k = 0; k < nz; ++k) {
    const size_t a = (size_t)k * nxy + c;
    s += (mask == nullptr || mask[a]) ? 0.0 + (double)q[a] : 0.0;
  }
  out[c] = s;
}

// ---- tau_mix = max(k / max(eps, 1e-12), tau_buoy) (spread_3d.py:2454-2455)
template <typename T>
__global__ void __launch_bounds__(256)
k_tau_mix(size_t n, const T* __restrict__ k, const T* __restrict__ eps, double tau_buoy,
          T* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  out[c] = (T)fmax((double)k[c] / fmax((double)eps[c], 1.0e-12), tau_buoy);
}

// ---- per-cell bed-particle aggregates of the snapshot (spread_3d.py:3569-3610)
// Live particles of each cell in slot order: float32 count and masses
// accumulated as float32 += float64 (each sum rounded back to float32), the
// mass-weighted T_s in float64.
__global__ void __launch_bounds__(256)
k_bed_snapshot(size_t n, int n_seg, const int* __restrict__ seg_ptr,
               const int* __restrict__ seg_idx, const double* __restrict__ m_solid,
               const double* __restrict__ m_water, const double* __restrict__ m_char,
               const double* __restrict__ T_s, const signed char* __restrict__ alive,
               float* __restrict__ count, float* __restrict__ ms, float* __restrict__ mw,
               float* __restrict__ mc, float* __restrict__ tavg) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  float cnt = 0.0f, s = 0.0f, w = 0.0f, ch = 0.0f;
  double tw = 0.0, mt_sum = 0.0;
  if (c < (size_t)n_seg) {
    for (int q = seg_ptr[c]; q < seg_ptr[c + 1]; ++q) {
      const int p = seg_idx[q];
      if (alive[p] == 0) continue;
      const double a = m_solid[p], b = m_water[p], d = m_char[p];
      const double mt = a + b + d;
      cnt = (float)((double)cnt + 1.0);
      s = (float)((double)s + a);
      w = (float)((double)w + b);
      ch = (float)((double)ch + d);
      tw += T_s[p] * mt;
      mt_sum += mt;
    }
  }
  count[c] = cnt;
  ms[c] = s;
  mw[c] = w;
  mc[c] = ch;
  tavg[c] = mt_sum > 0.0 ? (float)(tw / fmax(mt_sum, 1.0e-30)) : 0.0f;
}

// ---- front speed of the active level set ------------------------------------
// compute_v_n_3d (flame_front_3d.py:474-514) with q_in = q_dom_fwd, then the
// empirical-ROS blend (spread_3d.py:3252-3255), in the reference's order:
// E = max(E_sens + E_lat_coef M, 1); v = max(q / E, 0); v *= 1 - w; v += w ros.
template <typename T>
__global__ void __launch_bounds__(256)
k_lset_speed(size_t n, const T* __restrict__ q, const T* __restrict__ M, double E_sens,
             double E_lat_coef, double one_minus_w, double w_ros, int blend,
             double* __restrict__ v_n) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const double E = fmax(E_sens + E_lat_coef * (double)M[c], 1.0);
  double v = fmax((double)q[c] / E, 0.0);
  if (blend) v = v * one_minus_w + w_ros;
  v_n[c] = v;
}

// numpy's mean over axis 0: the z sum in k order, divided by Nz
__global__ void __launch_bounds__(256)
k_v_n_2d(size_t nxy, int nz, const double* __restrict__ v, double* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= nxy) return;
  double s = 0.0;
  for (int k = 0; k < nz; ++k) s += v[(size_t)k * nxy + c];
  out[c] = s / (double)nz;
}

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace fw

using namespace fw;

extern "C" {

int fw_phi_flame(const fw_grid* g, const void* omega, const void* T_g, const void* Y_fuel,
                 double dz_eff, double* phi, double* work, void* stream) {
  FW_REQUIRE(g && omega && T_g && Y_fuel && phi && work, "phi_flame: null argument");
  FW_REQUIRE(g->nx > 0 && g->ny > 0 && g->nz > 0, "phi_flame: empty grid");
  FW_REQUIRE(g->nx <= 2048 && g->ny <= 2048 && g->nz <= 2048,
             "phi_flame: a grid axis is longer than 2048 cells");
  FW_REQUIRE(dz_eff > 0.0 && g->dx > 0.0 && g->dy > 0.0, "phi_flame: bad sampling");
  const size_t nx = g->nx, ny = g->ny, nz = g->nz, n = nx * ny * nz;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_edt_init<double>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const double*)omega, (const double*)T_g, (const double*)Y_fuel,
                       phi, work);
  else
    hipLaunchKernelGGL(k_edt_init<float>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const float*)omega, (const float*)T_g, (const float*)Y_fuel, phi,
                       work);
  FW_CHECK_LAUNCH("phi_flame");
  const size_t big = (size_t)1 << 62;
  // x: line l = (k, j) row, contiguous
  EdtPass px{big, 0, nx, 1, (int)nx, (int)(ny * nz), 0, g->dx};
  // y: line l = (k, i), elements nx apart
  EdtPass py{nx, nx * ny, 1, nx, (int)ny, (int)(nz * nx), 0, g->dy};
  // z: line l = (j, i), elements nx ny apart
  EdtPass pz{big, 0, 1, nx * ny, (int)nz, (int)(nx * ny), 0, dz_eff};
  int rc = edt_pass(phi, work, px, false, s);
  if (rc == FW_OK) rc = edt_pass(phi, work, py, false, s);
  if (rc == FW_OK) rc = edt_pass(phi, work, pz, true, s);
  return rc;
}

int fw_ahead_band_mask(const fw_grid* g, const double* phi, double band_m, int n_z_top,
                       uint8_t* mask, void* stream) {
  FW_REQUIRE(g && phi && mask, "ahead_band_mask: null argument");
  FW_REQUIRE(n_z_top >= 0, "ahead_band_mask: negative layer count");
  const size_t nxy = (size_t)g->nx * g->ny, n = nxy * g->nz;
  const size_t n_bed = nxy * (size_t)(n_z_top < g->nz ? n_z_top : g->nz);
  hipLaunchKernelGGL(k_ahead_band, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, n,
                     n_bed, phi, band_m, mask);
  FW_CHECK_LAUNCH("ahead_band_mask");
  return FW_OK;
}

int fw_q_in_at_front_2d(const fw_grid* g, const void* q_dom_fwd, const uint8_t* mask,
                        double* out, void* stream) {
  FW_REQUIRE(g && q_dom_fwd && out, "q_in_at_front_2d: null argument");
  const size_t nxy = (size_t)g->nx * g->ny;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_q_in_2d<double>, dim3(nb256(nxy)), dim3(256), 0, s, nxy, g->nz,
                       (const double*)q_dom_fwd, mask, out);
  else
    hipLaunchKernelGGL(k_q_in_2d<float>, dim3(nb256(nxy)), dim3(256), 0, s, nxy, g->nz,
                       (const float*)q_dom_fwd, mask, out);
  FW_CHECK_LAUNCH("q_in_at_front_2d");
  return FW_OK;
}

int fw_tau_mix(const fw_grid* g, const void* k, const void* eps, double tau_buoy, void* out,
               void* stream) {
  FW_REQUIRE(g && k && eps && out, "tau_mix: null argument");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_tau_mix<double>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const double*)k, (const double*)eps, tau_buoy, (double*)out);
  else
    hipLaunchKernelGGL(k_tau_mix<float>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const float*)k, (const float*)eps, tau_buoy, (float*)out);
  FW_CHECK_LAUNCH("tau_mix");
  return FW_OK;
}

int fw_bed_snapshot(const fw_grid* g, const fw_bed_particles* particles,
                    const fw_bed_index* index, float* count, float* m_solid,
                    float* m_water, float* m_char, float* T_s_avg, void* stream) {
  FW_REQUIRE(g && particles && index && count && m_solid && m_water && m_char && T_s_avg,
             "bed_snapshot: null argument");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  FW_REQUIRE(index->n_seg_cells >= 0 && (size_t)index->n_seg_cells <= n,
             "bed_snapshot: index has more cells than the grid");
  FW_REQUIRE(index->n_seg_cells == 0 || (index->seg_ptr && index->seg_idx),
             "bed_snapshot: index without segments");
  FW_REQUIRE(index->n_seg_cells == 0 || (particles->m_solid && particles->m_water &&
                                         particles->m_char && particles->T_s &&
                                         particles->alive),
             "bed_snapshot: particle arrays missing");
  hipLaunchKernelGGL(k_bed_snapshot, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, n,
                     index->n_seg_cells, index->seg_ptr, index->seg_idx, particles->m_solid,
                     particles->m_water, particles->m_char, particles->T_s,
                     (const signed char*)particles->alive, count, m_solid, m_water, m_char,
                     T_s_avg);
  FW_CHECK_LAUNCH("bed_snapshot");
  return FW_OK;
}

int fw_lset_speed(const fw_grid* g, const void* q_dom_fwd, const void* M_local,
                  double E_sens, double E_lat_coef, double emp_w, double emp_ros,
                  double* v_n, void* stream) {
  FW_REQUIRE(g && q_dom_fwd && M_local && v_n, "lset_speed: null argument");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const int blend = emp_w > 0.0 ? 1 : 0;
  // the reference forms (1 - w) and w * ros as Python floats first
  const double omw = 1.0 - emp_w, wr = emp_w * emp_ros;
  hipStream_t s = (hipStream_t)stream;
  if (g->dtype == FW_F64)
    hipLaunchKernelGGL(k_lset_speed<double>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const double*)q_dom_fwd, (const double*)M_local, E_sens, E_lat_coef,
                       omw, wr, blend, v_n);
  else
    hipLaunchKernelGGL(k_lset_speed<float>, dim3(nb256(n)), dim3(256), 0, s, n,
                       (const float*)q_dom_fwd, (const float*)M_local, E_sens, E_lat_coef,
                       omw, wr, blend, v_n);
  FW_CHECK_LAUNCH("lset_speed");
  return FW_OK;
}

int fw_v_n_2d(const fw_grid* g, const double* v_n, double* out, void* stream) {
  FW_REQUIRE(g && v_n && out, "v_n_2d: null argument");
  const size_t nxy = (size_t)g->nx * g->ny;
  hipLaunchKernelGGL(k_v_n_2d, dim3(nb256(nxy)), dim3(256), 0, (hipStream_t)stream, nxy,
                     g->nz, v_n, out);
  FW_CHECK_LAUNCH("v_n_2d");
  return FW_OK;
}

}  // extern "C"
