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

namespace fw {

struct EdcParams {
  double chi_rad, cp_g, dt, s_stoich, hoc_J;
  double relax;   // 1 - exp(-(dt / n_substeps) / 0.5), formed on the host
  int n_substeps, extinction;
};

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
  __shared__ int wmax[4];
  __shared__ unsigned char own_sh[256];
  __shared__ int s_carry;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const size_t c = (size_t)blockIdx.x * EDC_SEG + t;
  const bool valid = c < n;
  // every load of the cell is issued before the workgroup exchange below
  T Yf = valid ? Yf_io[c] : T(0);
  T YO2 = valid ? YO2_io[c] : T(0);
  const T Yw = valid ? Yw_in[c] : T(0);
  T Tgv = valid ? Tg[c] : T(0);
  const T rho_i = valid ? rho[c] : T(1);
  T k_t = valid ? kturb[c] : T(1);
  T e_t = valid ? eturb[c] : T(1);
  const bool resets = valid && edc_resets<T>(Yf, YO2);
  own_sh[t] = edc_own_flag<T>(Yw) ? 1 : 0;
  // nearest preceding resetting cell of the segment: in my wavefront (ballot
  // below my lane), else the last one of the wavefronts before mine
  const unsigned long long m = __ballot(resets);
  if (lane == 0) wmax[w] = m ? w * 64 + 63 - __clzll(m) : -1;
  const unsigned long long below = lane ? (m & ((1ull << lane) - 1)) : 0ull;
  __syncthreads();
  int prev = below ? w * 64 + 63 - __clzll(below) : -1;
  for (int q = 0; q < w && prev < 0; ++q) prev = wmax[w - 1 - q];
  // flag entering the segment, needed unless its first cell resets: the
  // last earlier segment with a resetting cell, looked up 256 at a time
  const bool need_carry = __syncthreads_or(prev < 0 && valid);
  if (need_carry) {
    // round 1: the 256 segments just before this one; round 2 (rare): all
    // earlier segments at once, every thread keeping the largest index it
    // saw, so the look-back costs at most two dependent load rounds
    const int b0 = (int)blockIdx.x - 256 + t;   // thread t, segment b0: ascending
    const bool has = b0 >= 0 && summary[b0] >= 0;
    int last = block_last_true(has, wmax);
    __syncthreads();
    int found = last >= 0 ? (int)blockIdx.x - 256 + last : -1;
    if (found < 0 && (int)blockIdx.x > 256) {
      int best = -1;
#pragma unroll 4
      for (int b = t; b < (int)blockIdx.x - 256; b += 256)
        if (summary[b] >= 0) best = b;
      // block max of best: per-wavefront shuffle, then the four wavefronts
      for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_down(best, off, 64);
        best = o > best ? o : best;
      }
      if ((t & 63) == 0) wmax[t >> 6] = best;
      __syncthreads();
      for (int q = 0; q < 4; ++q) found = wmax[q] > found ? wmax[q] : found;
    }
    if (t == 0) s_carry = found >= 0 ? summary[found] : 0;
    __syncthreads();
  }
  bool flag = prev >= 0 ? (own_sh[prev] != 0) : (s_carry != 0);
  if (!valid) return;

  const int nsub = P.n_substeps > 1 ? P.n_substeps : 1;
  const T h = T(P.dt / nsub);
  const T HoC = T(P.hoc_J * (1.0 - P.chi_rad));
  const T s = T(P.s_stoich);
  const T cp_g = T(P.cp_g);
  if (k_t < T(1.0e-4)) k_t = T(1.0e-4);
  if (e_t < T(1.0e-6)) e_t = T(1.0e-6);
  T ratio = fdiv(T(1.5e-5) * e_t, k_t * k_t);
  if (ratio < T(1e-30)) ratio = T(1e-30);
  const T gb = T(2.1377) * sqrt(sqrt(ratio));   // ratio^(1/4)
  T gamma = gb * (gb * gb);
  if (gamma > T(1)) gamma = T(1);
  const T tau = T(0.4083) * sqrt(fdiv(T(1.5e-5), e_t));
  const T relax = T(P.relax);
  T f_h2o = T(1) - fdiv(Yw, T(0.18));
  if (f_h2o < T(0)) f_h2o = T(0);
  T acc = T(0);
  for (int q = 0; q < nsub; ++q) {
    T omega;
    if (Yf <= T(1e-9) || YO2 <= T(1e-9)) {
      omega = T(0);
    } else {
      const T yo = fdiv(YO2, s);
      const T Ylim = Yf < yo ? Yf : yo;
      omega = gamma * fdiv(rho_i * Ylim, tau);
      flag = false;
      if (Yw > T(0)) {
        omega = omega * f_h2o;
        if (f_h2o < T(0.5)) flag = true;
      }
    }
    bool ext = false;
    if (P.extinction && omega > T(0)) {
      const T Yib = T(1) - Yf - YO2;
      if (Yib > T(0.88)) {
        T supp = (T(1) - Yib) / (T(1) - T(0.88));
        if (supp < T(0)) supp = T(0);
        omega = omega * supp;
        if (supp < T(0.5)) ext = true;
      }
      if (omega > T(0) && Tgv < T(1200.0)) { omega = T(0); ext = true; }
      if (omega > T(0) && omega * HoC < T(1.5) * T(5.0e4)) { omega = T(0); ext = true; }
    }
    if ((ext || flag) && Tgv > T(373.15) && Yw > T(0)) {
      T wb = fdiv(Yw, T(0.10));
      if (wb > T(1)) wb = T(1);
      Tgv = Tgv - (Tgv - T(373.15)) * relax * wb;
    }
    const T dY = fdiv(-omega * h, rho_i);
    Yf = Yf + dY;
    if (Yf < T(0)) Yf = T(0);
    YO2 = YO2 + s * dY;
    if (YO2 < T(0)) YO2 = T(0);
    const T cp_mix = (T(1) - Yw) * cp_g + Yw * T(2000.0);
    Tgv = Tgv + fdiv(omega * HoC * h, rho_i * cp_mix);
    if (Tgv > T(2400.0)) Tgv = T(2400.0);
    acc += omega * h;
  }
  Yf_io[c] = Yf;
  YO2_io[c] = YO2;
  Tg[c] = Tgv;
  omega_out[c] = P.dt > 0.0 ? fdiv(acc, T(P.dt)) : T(0);
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
  EdcParams P{chi_rad, cp_g, dt, s_stoich, hoc_J, edc_relax(dt, n_substeps), n_substeps,
              extinction_enable ? 1 : 0};
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
