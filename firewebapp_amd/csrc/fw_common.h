// Shared definitions for the firewave HIP kernels (gfx950 / CDNA4).
//
// Field layout: every 3-D field is a dense (Nz, Ny, Nx) C-order array, flat
// index (k*Ny + j)*Nx + i, x fastest, so a wavefront of 64 lanes reads 64
// consecutive x cells (one 256 B / 512 B coalesced line group for f32 / f64).
// Storage type is selected per grid (fw_grid.dtype); every stencil kernel is
// a template over that type and is instantiated for float and double.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
#include <stdio.h>
#include "../../include/firewave.h"

namespace fw {

// Number of workgroups (and partial results) of the grid-wide reductions.
#define FW_REDUCE_BLOCKS 256

// Launch shape for the cell kernels: one thread per cell, a workgroup spans a
// run of x cells of a single (k, j) row.  gridDim = (ceil(Nx/BX), Ny, Nz).
struct Launch {
  dim3 grid, block;
};

// The gas-solid coupling of the Eulerian bed (eulerian.hip), for the step
// driver: Q_comb from the array, or (Q_comb == nullptr) formed from the
// drip-torch gas share and q_rad_gas / dz with the solid share subtracted
// from Q_pyro in the source bed cells first.
// HIP refuses a synchronous copy or set on the legacy default stream while
// any stream of the process records a graph.  Graph recording (GraphCache)
// holds this lock exclusively; the entry points that issue such copies
// (handle creation and release) hold it shared, so on concurrent host
// threads the two never overlap.  Scopes do not nest.
void record_lock(bool exclusive);
void record_unlock(bool exclusive);
struct LegacyScope {
  LegacyScope() { record_lock(false); }
  ~LegacyScope() { record_unlock(false); }
};
struct RecordScope {
  RecordScope() { record_lock(true); }
  ~RecordScope() { record_unlock(true); }
};

// Device-resident step scalars of the sub-step loop (a graph-captured loop
// replays the same kernel arguments every step, so the values that change
// from step to step are read from this block instead).  Each slot holds the
// float64 value the host formed; a kernel given a null block uses its
// by-value argument.
enum { SB_DT_SUB = 0, SB_EDC_RELAX = 1, SB_DSC = 2, SB_DPR = 3, SB_SLOTS = 8 };
__device__ __forceinline__ double sb_get(const double* sb, int slot, double v) {
  return sb ? sb[slot] : v;
}

// The species turbulent-diffusion pass; with edc_summary (Nx % 256 == 0) it
// also leaves the EDC segment summaries of its output in edc_summary.
// sb (may be null): dt_sub from sb[SB_DSC].
struct GasSolidParams;
// gs (may be null): the gas energy update and vertical solid conduction of
// gas_solid_launch fused into this pass, with the new Y_H2O (fo[2]) of each
// cell; gs->T_s_out must not alias any field the pass reads or writes.
int turb3_launch(const fw_grid* g, const void* const* f, void* const* fo, const void* nu_t,
                 double sc_t, double dt_sub, const double* lo, const double* hi,
                 int* edc_summary, hipStream_t s, const double* sb = nullptr,
                 const GasSolidParams* gs = nullptr);
// fw_edc; summary_ready: the summaries of the current species are in work.
// sb (may be null): dt from sb[SB_DT_SUB], the relaxation from sb[SB_EDC_RELAX].
int edc_launch(const fw_grid* g, const void* rho, void* T_g, void* Y_fuel, void* Y_O2,
               const void* k_turb, const void* eps_turb, const void* Y_H2O, void* omega_out,
               double chi_rad, double cp_g, double dt, int n_substeps, int extinction_enable,
               double s_stoich, double hoc_J, int* work, bool summary_ready, hipStream_t s,
               const double* sb = nullptr);
double edc_relax(double dt, int n_substeps);
// fw_species_transport3; sb (may be null): dt from sb[SB_DT_SUB] (and no
// kernel timer events).
int species3_launch(const fw_grid* g, const void* const* Y, void* const* Yo, const void* rho,
                    const void* u, const void* v, const void* w, const void* Sp,
                    const void* Sd, double dt, double D, const double* y_in,
                    const double* y_max, hipStream_t s, const double* sb);
// fw_turbulent_diffusion_pass; sb (may be null): dt_sub from sb[SB_DPR].
// edc (may be null): the EDC update of the new field (the temperature) fused
// into the pass, one 256-cell segment per workgroup (needs Nx % 256 == 0 and
// the species' segment summaries in edc->summary).
struct EdcFuseArgs;
// fw_gas_energy; edc (may be null): the first sub-step's EDC fused in, as
// for turb_pass_launch.
int gas_energy_launch(const fw_grid* g, const void* Tg, void* To, const void* u,
                      const void* v, const void* w, double dt, double alpha_th,
                      double T_amb, hipStream_t s, const EdcFuseArgs* edc);
// The EDC segment summaries of the current species (k_edc_summary).
int edc_summary_launch(const fw_grid* g, const void* Y_fuel, const void* Y_O2,
                       const void* Y_H2O, int* work, hipStream_t s);
int turb_pass_launch(const fw_grid* g, const void* f, void* fo, const void* nu_t, double sc_t,
                     double dt_sub, double lo, double hi, hipStream_t s, const double* sb,
                     const EdcFuseArgs* edc = nullptr);
// fw_gas_update_solid_conduction; sb (may be null): dt_sub from sb[SB_DT_SUB].
int gas_solid_launch(const fw_grid* g, void* T_g, const void* T_s, void* T_s_out,
                     const void* alpha_s, const void* rho, const void* Y_H2O,
                     const void* S_pyro_part, const void* S_dry_part, const void* Q_g_conv,
                     const void* q_rad_gas, void* Q_comb_out, int n_z_bed, int i_src0,
                     int i_src1, double drip_gas, double dt_sub, double T_amb,
                     double k_solid, double rho_solid, double cp_solid, double clip_lo,
                     double clip_hi, hipStream_t s, const double* sb);
int coupling_launch(const fw_grid* g, void* T_g, void* T_s, const void* rho, const void* u,
                    const void* v, const void* w, const void* alpha_s, double sigma_sav,
                    const void* q_rad_in, void* Q_pyro, const void* Q_comb,
                    const void* q_rad_gas, void* Q_comb_out, double* m_water, double L_v,
                    double dt, double T_amb, int q_loss, double h_conv_mult, int nzb, int i0,
                    int i1, double drip_gas, double drip_solid, hipStream_t s);

// The learned preconditioner's network (unet.hip), for projection.hip.
int net_apply(void* net, const float* r, const float* q, const float** c, int* n,
              hipStream_t s);
int net_dims(void* net, int* nx, int* ny, int* nz);
int density_feature(const float* rho, float* q, size_t n, hipStream_t s);

// Events on which a timed launch records its own start and stop (the
// dispatch's timestamps, as a kernel trace sees them), or null.
struct KernelEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
// The pair for the next launch of kernel `which` (FW_TIMER_*) from this host
// thread while its timer is on and the launch is one of the timed ones;
// {null, null} otherwise (timer.hip).
KernelEvents timer_next(int which);
// Whether this host thread's timer of kernel `which` is on (launches that
// record events are not captured into graphs).
bool timer_on(int which);
// The calling thread's timers, for a worker thread that launches on its
// behalf to adopt (timer_adopt(null) returns the worker to its own); the two
// threads must not launch timed kernels of one kind at the same time.
void* timer_share();
void timer_adopt(void* timers_of_another_thread);

// The DOM solve on the handle's own stream, concurrent with the caller's
// stream (dom.hip): dom_fork orders the side stream after the work already
// on s and returns it; dom_join_record marks the end of the side stream's
// work; dom_join makes s wait for it.
int dom_fork(void* dom, hipStream_t s, hipStream_t* side);
int dom_join_record(void* dom);
int dom_join(void* dom, hipStream_t s);

// Per-workgroup maxima of the interior nu_t into work[FW_REDUCE_BLOCKS]
// (transport.hip), without the host read of fw_nu_t_max_interior.
int nu_t_max_partials(const fw_grid* g, const void* nu_t, double* work, hipStream_t s);

inline Launch cell_launch(const fw_grid* g) {
  int bx = 64;
  while (bx < g->nx && bx < 256) bx *= 2;
  Launch L;
  L.block = dim3(bx, 1, 1);
  L.grid = dim3((g->nx + bx - 1) / bx, g->ny, g->nz);
  return L;
}

// Plain-value view of fw_grid passed to kernels by value.
// The reciprocals are the host's IEEE float64 quotients (1.0 / dx, ...), the
// same values the kernels would divide out per thread, so using them keeps
// every float64 result unchanged while saving the per-thread divisions.
struct GridArgs {
  int nx, ny, nz, n_z_bed;
  double dx, dy;
  double inv_dx, inv_dy, inv_dx2, inv_dy2;  // 1/dx, 1/dy, 1/(dx dx), 1/(dy dy)
  const double* dz;
  const double* d_above;
  const double* d_below;
  const double* z_face;
  const double* inv_dz;
  const double* inv_da;
  const double* inv_db;
};

inline GridArgs grid_args(const fw_grid* g) {
  GridArgs a;
  a.nx = g->nx; a.ny = g->ny; a.nz = g->nz; a.n_z_bed = g->n_z_bed;
  a.dx = g->dx; a.dy = g->dy;
  a.inv_dx = 1.0 / g->dx; a.inv_dy = 1.0 / g->dy;
  a.inv_dx2 = 1.0 / (g->dx * g->dx); a.inv_dy2 = 1.0 / (g->dy * g->dy);
  a.dz = g->dz; a.d_above = g->d_above; a.d_below = g->d_below;
  a.z_face = g->z_face;
  a.inv_dz = g->inv_dz; a.inv_da = g->inv_d_above; a.inv_db = g->inv_d_below;
  return a;
}

// Cell coordinates of the calling thread for a cell_launch grid, with an
// XCD-aware block order.  Workgroups are handed to the 8 XCDs round-robin by
// linear id, so in launch order consecutive rows (j, k) would sit in
// different L2 caches and every y/z halo row would be fetched from HBM once
// per XCD.  The remap gives each XCD a contiguous run of rows instead, so
// neighbouring rows share one L2.  It is a bijection on the block ids, so
// correctness does not depend on the dispatch order.
__device__ __forceinline__ void cell_ijk(int& i, int& j, int& k) {
  const unsigned bx = gridDim.x, by = gridDim.y;
  const unsigned nb = bx * by * gridDim.z;
  const unsigned lin = blockIdx.x + bx * (blockIdx.y + by * blockIdx.z);
  const unsigned q = nb >> 3, r = nb & 7u;
  const unsigned xcd = lin & 7u, pos = lin >> 3;
  const unsigned m = xcd * q + (xcd < r ? xcd : r) + pos;
  const unsigned t = m / bx;
  i = (int)((m - t * bx) * blockDim.x + threadIdx.x);
  j = (int)(t % by);
  k = (int)(t / by);
}

__device__ __forceinline__ int wrap(int j, int n) {
  j %= n;
  return j < 0 ? j + n : j;
}

__device__ __forceinline__ size_t cidx(const GridArgs& g, int k, int j, int i) {
  return ((size_t)k * g.ny + j) * g.nx + i;
}

// 32-bit byte offsets of a cell's 13-point stencil (x, z: +-1, +-2; y: +-1,
// +-2 periodic), shared by every field read at the cell: a load is then the
// field's SGPR base plus one VGPR offset, with no 64-bit address per field
// and neighbour.  Valid while a field is smaller than 4 GiB (the entry points
// that use it check this).  Offsets past a non-periodic edge wrap and are
// only dereferenced under the same edge tests as the index form.
struct Nbr {
  uint32_t c, xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2, zm2, zm1, zp1, zp2;
};

template <typename T>
__device__ __forceinline__ Nbr nbr_offsets(const GridArgs& g, int i, int j, int k) {
  const uint32_t e = sizeof(T);
  const uint32_t sy = (uint32_t)g.nx * e, sz = (uint32_t)g.nx * (uint32_t)g.ny * e;
  Nbr n;
  n.c = (((uint32_t)k * (uint32_t)g.ny + (uint32_t)j) * (uint32_t)g.nx + (uint32_t)i) * e;
  n.xm2 = n.c - 2 * e;
  n.xm1 = n.c - e;
  n.xp1 = n.c + e;
  n.xp2 = n.c + 2 * e;
  n.ym2 = n.c + (uint32_t)(wrap(j - 2, g.ny) - j) * sy;
  n.ym1 = n.c + (uint32_t)(wrap(j - 1, g.ny) - j) * sy;
  n.yp1 = n.c + (uint32_t)(wrap(j + 1, g.ny) - j) * sy;
  n.yp2 = n.c + (uint32_t)(wrap(j + 2, g.ny) - j) * sy;
  n.zm2 = n.c - 2 * sz;
  n.zm1 = n.c - sz;
  n.zp1 = n.c + sz;
  n.zp2 = n.c + 2 * sz;
  return n;
}

// p[off / sizeof(T)] for a byte offset from nbr_offsets
template <typename T>
__device__ __forceinline__ T at(const T* __restrict__ p, uint32_t off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + off);
}
template <typename T>
__device__ __forceinline__ T& at(T* __restrict__ p, uint32_t off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + off);
}

// minmod limiter (muscl_3d.py:36-45): smaller-magnitude argument when both
// have the same sign, zero otherwise.
// Written as selects (no branches): the same values as the if/else form.
template <typename T>
__device__ __forceinline__ T minmod(T a, T b) {
  const T lo = a < b ? a : b;
  const T hi = a > b ? a : b;
  const T m = a > T(0) ? lo : hi;
  return a * b <= T(0) ? T(0) : m;
}

// float32: minmod(a, b) is the median of a, b and 0 (one v_med3_f32).  Same
// value as above for finite arguments, except that a product a*b which
// underflows float32 no longer zeroes the slope (the float64 reference's
// product does not underflow there either).
template <>
__device__ __forceinline__ float minmod<float>(float a, float b) {
  return __builtin_amdgcn_fmed3f(a, b, 0.0f);
}

// 1 / y for quot() below: v_rcp_f32 (1 ulp) in float32; float64 never uses
// it (quot divides), so the division is dead code there.
template <typename T>
__device__ __forceinline__ T recip(T y) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(y);
  else return T(1) / y;
}

// x / y: float32 as x * v_rcp_f32(y) (within 2 ulp), float64 the IEEE
// quotient (the reference's rounding)
template <typename T>
__device__ __forceinline__ T fdiv(T x, T y) {
  if constexpr (sizeof(T) == 4) return x * __builtin_amdgcn_rcpf(y);
  else return x / y;
}

// x / y.  The float32 build multiplies by ry = 1 / y, formed once by the
// caller and shared by several quotients (one rounding more, inside the
// float32 tolerance); float64 keeps the reference's division bit for bit.
template <typename T>
__device__ __forceinline__ T quot(T x, T y, T ry) {
  if constexpr (sizeof(T) == 4) return x * ry;
  else return x / y;
}

// MUSCL face value at face i+1/2 from the upwind side (muscl_3d.py:48-64).
// The upwind side is chosen from the sign of `uf` (the caller passes the
// cell-centre velocity, as the reference does).  Both one-sided candidates
// are formed and one is selected, which keeps the wavefront converged.
template <typename T>
__device__ __forceinline__ T muscl_face(T pm1, T p0, T pp1, T pp2, T uf) {
  const T fwd = p0 + T(0.5) * minmod<T>(p0 - pm1, pp1 - p0);
  const T bwd = pp1 - T(0.5) * minmod<T>(pp1 - p0, pp2 - pp1);
  return uf >= T(0) ? fwd : bwd;
}

// ---- EDC carried-flag scan helpers (edc.py:223) ----------------------------
// The flag scan runs over fixed 256-cell segments of the flat index; a
// segment summary is the flag of its last resetting cell, or -1.
#define EDC_SEG 256

template <typename T>
__device__ __forceinline__ bool edc_resets(T Yf, T YO2) {
  return Yf > T(1e-9) && YO2 > T(1e-9);
}

template <typename T>
__device__ __forceinline__ bool edc_own_flag(T Yw) {
  T f = T(1) - fdiv(Yw, T(0.18));   // as f_h2o in k_edc
  if (f < T(0)) f = T(0);
  return Yw > T(0) && f < T(0.5);
}

// Index (0..255) of the last thread of the 256-thread workgroup whose
// predicate holds, or -1; wmax[4] receives each wavefront's last index.
__device__ __forceinline__ int block_last_true(bool pred, int* wmax) {
  const int t = threadIdx.x, w = t >> 6;
  const unsigned long long m = __ballot(pred);
  if ((t & 63) == 0) wmax[w] = m ? w * 64 + 63 - __clzll(m) : -1;
  __syncthreads();
  int last = -1;
  for (int q = 0; q < 4; ++q) last = wmax[q] > last ? wmax[q] : last;
  return last;
}

// LevelSetFront3D.front_x (flame_front_3d.py:391-402) on one host row of
// phi: the linear zero crossing of the first sign change, +inf if none.
inline double front_x_row(const double* row, int Nx, double dx) {
  for (int i = 0; i < Nx - 1; ++i) {
    const double p0 = row[i], p1 = row[i + 1];
    if ((p0 < 0.0 && 0.0 < p1) || (p1 < 0.0 && 0.0 < p0))
      return (i + (-p0 / (p1 - p0)) + 0.5) * dx;
  }
  return INFINITY;
}

}  // namespace fw

// ---- error plumbing (fw_error.cpp) -----------------------------------------
int fw_set_error(int code, const char* fmt, ...);

#define FW_CHECK_LAUNCH(name)                                                 \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess)                                                     \
      return fw_set_error(FW_E_HIP, "%s: %s", name, hipGetErrorString(_e));   \
  } while (0)

#define FW_CHECK_HIP(call)                                                    \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess)                                                     \
      return fw_set_error(FW_E_HIP, "%s: %s", #call, hipGetErrorString(_e));  \
  } while (0)

#define FW_REQUIRE(cond, ...)                                                 \
  do {                                                                        \
    if (!(cond)) return fw_set_error(FW_E_CONFIG, __VA_ARGS__);               \
  } while (0)

// Every metric array of a grid must be present (the kernels read them
// unconditionally).
#define FW_GRID(g)                                                            \
  FW_REQUIRE((g) && (g)->dz && (g)->d_above && (g)->d_below && (g)->inv_dz &&  \
                 (g)->inv_d_above && (g)->inv_d_below && (g)->nx > 0 &&       \
                 (g)->ny > 0 && (g)->nz > 0,                                  \
             "fw_grid: missing metric arrays (dz, d_above, d_below and their "  \
             "reciprocals are required) or empty grid")

// Kernels that address fields through 32-bit byte offsets (nbr_offsets)
// need every field to be smaller than 4 GiB.
#define FW_GRID_OFFSETS32(g)                                                  \
  FW_REQUIRE((size_t)(g)->nx * (g)->ny * (g)->nz *                            \
                     ((g)->dtype == FW_F64 ? 8 : 4) <= 0xFFFFFFFFull,         \
             "fw_grid: a field of this grid exceeds 4 GiB")

// Dispatch a templated implementation on the grid's storage type.
#define FW_DISPATCH(g, fn, ...)                                               \
  ((g)->dtype == FW_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))
