// EDC closure (chemistry_closures/edc.py:133-334).
//
// The reference's wet-bulb flag `_h2o_quench_substantial` is assigned only in
// cells with Yf > 1e-9 and YO2 > 1e-9, so every other cell inherits the flag
// of the previous cell in (k, j, i) order (edc.py:223).  A resetting cell's
// flag depends only on its own Y_H2O, hence the inherited value of cell c is
// the flag of the nearest preceding resetting cell.  That is a "last valid"
// prefix scan over fixed 256-cell segments of the flat index, in two
// launches:
//   1. k_edc_summary: per segment, the flag of its last resetting cell
//      (or "none"), from one ballot per wavefront;
//   2. k_edc: the cell update.  Inside the segment a cell's nearest
//      preceding resetting cell comes from the wavefront ballots; a segment
//      whose leading cells have none looks back through the summaries of
//      the earlier segments, 256 at a time, for the last one with a
//      resetting cell (the flag is false before the first).
// No atomics and no inter-workgroup waits: the result is deterministic.
#include <mutex>

#include "fw_common.h"
#include "edc_cell.h"

namespace fw {

// summary[b] = -1 (no resetting cell), 0 or 1 (flag of the last one)
template <typename T>
__global__ void __launch_bounds__(256)
k_edc_summary(size_t n, const T* __restrict__ Yf, const T* __restrict__ YO2,
              const T* __restrict__ Yw, int* __restrict__ summary) {
  __shared__ int wmax[4];
  const size_t c = (size_t)blockIdx.x * EDC_SEG + threadIdx.x;
  const bool r = c < n && edc_resets<T>(Yf[c], YO2[c]);
  const int last = block_last_true(r, wmax);
  if (threadIdx.x == 0)
    summary[blockIdx.x] =
        last < 0 ? -1 : (edc_own_flag<T>(Yw[(size_t)blockIdx.x * EDC_SEG + last]) ? 1 : 0);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_edc(size_t n, const T* __restrict__ rho, T* __restrict__ Tg,
      T* __restrict__ Yf_io, T* __restrict__ YO2_io,
      const T* __restrict__ kturb, const T* __restrict__ eturb,
      const T* __restrict__ Yw_in, T* __restrict__ omega_out,
      const int* __restrict__ summary, EdcParams P, const double* __restrict__ sb) {
  P.dt = sb_get(sb, SB_DT_SUB, P.dt);
  P.relax = sb_get(sb, SB_EDC_RELAX, P.relax);
  const size_t c = (size_t)blockIdx.x * EDC_SEG + threadIdx.x;
  const bool valid = c < n;
  const EdcCellIn<T> in = edc_load<T>(valid, c, rho, Yf_io, YO2_io, kturb, eturb, Yw_in);
  edc_segment<T>(n, (int)blockIdx.x, valid ? Tg[c] : T(0), in, Tg, Yf_io, YO2_io, omega_out,
                 summary, P);
}

// summary_ready: the segment summaries of the current Y_F, Y_O2, Y_H2O are
// already in work (left by the species turbulent diffusion's last pass)
template <typename T>
static int edc_impl(const fw_grid* g, const void* rho, void* Tg, void* Yf,
                    void* YO2, const void* kt, const void* et, const void* Yw,
                    void* omega, int* work, const EdcParams& P, bool summary_ready,
                    hipStream_t s, const double* sb) {
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const int nseg = (int)((n + EDC_SEG - 1) / EDC_SEG);
  int* summary = work;
  if (!summary_ready) {
    hipLaunchKernelGGL(k_edc_summary<T>, dim3(nseg), dim3(256), 0, s, n,
                       (const T*)Yf, (const T*)YO2, (const T*)Yw, summary);
    FW_CHECK_LAUNCH("edc_summary");
  }
  hipLaunchKernelGGL(k_edc<T>, dim3(nseg), dim3(256), 0, s, n, (const T*)rho,
                     (T*)Tg, (T*)Yf, (T*)YO2, (const T*)kt, (const T*)et,
                     (const T*)Yw, (T*)omega, (const int*)summary, P, sb);
  FW_CHECK_LAUNCH("edc");
  return FW_OK;
}

template <typename T>
static int edc_summary_impl(const fw_grid* g, const void* Yf, const void* YO2,
                            const void* Yw, int* work, hipStream_t s) {
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const int nseg = (int)((n + EDC_SEG - 1) / EDC_SEG);
  hipLaunchKernelGGL(k_edc_summary<T>, dim3(nseg), dim3(256), 0, s, n, (const T*)Yf,
                     (const T*)YO2, (const T*)Yw, work);
  FW_CHECK_LAUNCH("edc_summary");
  return FW_OK;
}

int edc_summary_launch(const fw_grid* g, const void* Y_fuel, const void* Y_O2,
                       const void* Y_H2O, int* work, hipStream_t s) {
  return FW_DISPATCH(g, edc_summary_impl, g, Y_fuel, Y_O2, Y_H2O, work, s);
}

// The relaxation factor of the wet-bulb clamp for dt (formed on the host,
// also for the step scalar block).
double edc_relax(double dt, int n_substeps) {
  const int nsub = n_substeps > 1 ? n_substeps : 1;
  return 1.0 - exp(-(dt / nsub) / 0.5);
}

int edc_launch(const fw_grid* g, const void* rho, void* T_g, void* Y_fuel, void* Y_O2,
               const void* k_turb, const void* eps_turb, const void* Y_H2O, void* omega_out,
               double chi_rad, double cp_g, double dt, int n_substeps, int extinction_enable,
               double s_stoich, double hoc_J, int* work, bool summary_ready, hipStream_t s,
               const double* sb) {
  const EdcParams P = edc_params(chi_rad, cp_g, dt, n_substeps, extinction_enable, s_stoich,
                                 hoc_J);
  return FW_DISPATCH(g, edc_impl, g, rho, T_g, Y_fuel, Y_O2, k_turb, eps_turb, Y_H2O,
                     omega_out, work, P, summary_ready, s, sb);
}

}  // namespace fw

extern "C" int fw_edc_workspace_ints(const fw_grid* g) {
  if (!g) return fw_set_error(FW_E_CONFIG, "bad grid");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  return (int)((n + EDC_SEG - 1) / EDC_SEG);
}

extern "C" int fw_edc(const fw_grid* g, const void* rho, void* T_g,
                      void* Y_fuel, void* Y_O2, const void* k_turb,
                      const void* eps_turb, const void* Y_H2O, void* omega_out,
                      double chi_rad, double cp_g, double dt, int n_substeps,
                      int extinction_enable, double s_stoich, double hoc_J,
                      int* work, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(work, "edc: null workspace");
  return fw::edc_launch(g, rho, T_g, Y_fuel, Y_O2, k_turb, eps_turb, Y_H2O, omega_out,
                        chi_rad, cp_g, dt, n_substeps, extinction_enable, s_stoich, hoc_J,
                        work, false, (hipStream_t)stream);
}

// ---- fw_edc_static: the same closure with library-owned scratch ----------
// The segment summaries live in a file-scope device array instead of a
// caller workspace, so the entry point allocates nothing.  Each GPU holds its
// own copy; its address on a device comes from hipGetSymbolAddress (the
// array's name in host code is only a host-side stand-in, never a device
// address), looked up once per device with that device current and kept in
// g_edc_scratch_ptr.
static constexpr int EDC_STATIC_SEGS = FW_EDC_STATIC_MAX_CELLS / EDC_SEG;
static constexpr int EDC_STATIC_DEVICES = 64;
static __device__ __align__(256) int scratch[EDC_STATIC_SEGS];
static std::mutex g_edc_scratch_mu;
static int* g_edc_scratch_ptr[EDC_STATIC_DEVICES] = {};

static int edc_scratch(int** out) {
  int dev = -1;
  FW_CHECK_HIP(hipGetDevice(&dev));
  FW_REQUIRE(dev >= 0 && dev < EDC_STATIC_DEVICES, "edc_static: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_edc_scratch_mu);
  if (!g_edc_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fw_set_error(FW_E_HIP, "edc_static: hipGetSymbolAddress(scratch): %s",
                          hipGetErrorString(e));
    g_edc_scratch_ptr[dev] = (int*)p;
  }
  *out = g_edc_scratch_ptr[dev];
  return FW_OK;
}

extern "C" int fw_edc_static(const fw_grid* g, const void* rho, void* T_g, void* Y_fuel,
                             void* Y_O2, const void* k_turb, const void* eps_turb,
                             const void* Y_H2O, void* omega_out, double chi_rad, double cp_g,
                             double dt, int n_substeps, int extinction_enable,
                             double s_stoich, double hoc_J, void* stream) {
  FW_GRID(g);
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  FW_REQUIRE(n <= (size_t)FW_EDC_STATIC_MAX_CELLS,
             "edc_static: %zu cells exceed the static scratch's maximum of %d cells "
             "(use fw_edc with a workspace)", n, FW_EDC_STATIC_MAX_CELLS);
  int* summary = nullptr;
  const int rc = edc_scratch(&summary);
  if (rc) return rc;
  // k_edc_summary writes every segment's summary before k_edc reads it, so
  // the array needs no clearing between calls
  return fw::edc_launch(g, rho, T_g, Y_fuel, Y_O2, k_turb, eps_turb, Y_H2O, omega_out,
                        chi_rad, cp_g, dt, n_substeps, extinction_enable, s_stoich, hoc_J,
                        summary, false, (hipStream_t)stream);
}
