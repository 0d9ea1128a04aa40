// EDC closure (chemistry_closures/edc.py:133-334).
//
// The reference's wet-bulb flag `_h2o_quench_substantial` is assigned only in
// cells with Yf > 1e-9 and YO2 > 1e-9, so every other cell inherits the flag
// of the previous cell in (k, j, i) order (edc.py:223).  A resetting cell's
// flag depends only on its own Y_H2O, hence the inherited value of cell c is
// the flag of the nearest preceding resetting cell.  That is a "last valid"
// prefix scan, computed here in three launches over fixed 256-cell segments
// of the flat index:
//   1. k_edc_summary: per segment, the flag of its last resetting cell
//      (or "none");
//   2. k_edc_carry:   one workgroup scans the segment summaries into the flag
//      entering each segment;
//   3. k_edc:         the cell update; a workgroup max-scan inside the segment
//      finds each cell's nearest preceding resetting cell.
// No atomics and no inter-workgroup waits: the result is deterministic.
#include "fw_common.h"

namespace fw {

#define EDC_SEG 256

struct EdcParams {
  double chi_rad, cp_g, dt, s_stoich, hoc_J;
  int n_substeps, extinction;
};

template <typename T>
__device__ __forceinline__ bool edc_resets(T Yf, T YO2) {
  return Yf > T(1e-9) && YO2 > T(1e-9);
}

template <typename T>
__device__ __forceinline__ bool edc_own_flag(T Yw) {
  T f = T(1) - Yw / T(0.18);
  if (f < T(0)) f = T(0);
  return Yw > T(0) && f < T(0.5);
}

// block-wide inclusive max-scan of an int over 256 threads (Hillis-Steele)
__device__ __forceinline__ int block_max_scan(int v, int* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int o = (t >= off) ? sh[t - off] : -1;
    __syncthreads();
    if (o > sh[t]) sh[t] = o;
    __syncthreads();
  }
  return sh[t];
}

// summary[b] = -1 (no resetting cell), 0 or 1 (flag of the last one)
template <typename T>
__global__ void __launch_bounds__(256)
k_edc_summary(size_t n, const T* __restrict__ Yf, const T* __restrict__ YO2,
              const T* __restrict__ Yw, int* __restrict__ summary) {
  __shared__ int sh[256];
  const size_t c = (size_t)blockIdx.x * EDC_SEG + threadIdx.x;
  int v = -1;
  if (c < n && edc_resets<T>(Yf[c], YO2[c])) v = threadIdx.x;
  // reduction of the max index (plain tree)
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s && sh[threadIdx.x + s] > sh[threadIdx.x])
      sh[threadIdx.x] = sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int last = sh[0];
    summary[blockIdx.x] =
        last < 0 ? -1 : (edc_own_flag<T>(Yw[(size_t)blockIdx.x * EDC_SEG + last]) ? 1 : 0);
  }
}

// carry[b] = flag entering segment b.  One workgroup of 256 threads; each
// thread owns a contiguous run of segments.
__global__ void __launch_bounds__(256)
k_edc_carry(int nseg, const int* __restrict__ summary, int* __restrict__ carry) {
  __shared__ int sh[256];
  const int t = threadIdx.x;
  const int per = (nseg + 255) / 256;
  const int b0 = t * per, b1 = min(nseg, b0 + per);
  int last = -1;  // last segment index with a summary in my run
  for (int b = b0; b < b1; ++b)
    if (summary[b] >= 0) last = b;
  int incl = block_max_scan(last, sh);
  int excl = (t > 0) ? sh[t - 1] : -1;
  (void)incl;
  int cur = excl;
  for (int b = b0; b < b1; ++b) {
    carry[b] = cur >= 0 ? summary[cur] : 0;
    if (summary[b] >= 0) cur = b;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_edc(size_t n, const T* __restrict__ rho, T* __restrict__ Tg,
      T* __restrict__ Yf_io, T* __restrict__ YO2_io,
      const T* __restrict__ kturb, const T* __restrict__ eturb,
      const T* __restrict__ Yw_in, T* __restrict__ omega_out,
      const int* __restrict__ carry, EdcParams P) {
  __shared__ int sh[256];
  __shared__ unsigned char own_sh[256];
  const int t = threadIdx.x;
  const size_t c = (size_t)blockIdx.x * EDC_SEG + t;
  const bool valid = c < n;
  T Yf = valid ? Yf_io[c] : T(0);
  T YO2 = valid ? YO2_io[c] : T(0);
  const T Yw = valid ? Yw_in[c] : T(0);
  const bool resets = valid && edc_resets<T>(Yf, YO2);
  const bool own = edc_own_flag<T>(Yw);
  own_sh[t] = own ? 1 : 0;
  // exclusive scan of the last resetting thread index before me
  int incl = block_max_scan(resets ? t : -1, sh);
  (void)incl;
  int prev = (t > 0) ? sh[t - 1] : -1;
  bool flag = prev >= 0 ? (own_sh[prev] != 0) : (carry[blockIdx.x] != 0);
  if (!valid) return;

  const int nsub = P.n_substeps > 1 ? P.n_substeps : 1;
  const T h = T(P.dt / nsub);
  const T HoC = T(P.hoc_J * (1.0 - P.chi_rad));
  const T s = T(P.s_stoich);
  const T cp_g = T(P.cp_g);
  T Tgv = Tg[c];
  const T rho_i = rho[c];
  T k_t = kturb[c];
  if (k_t < T(1.0e-4)) k_t = T(1.0e-4);
  T e_t = eturb[c];
  if (e_t < T(1.0e-6)) e_t = T(1.0e-6);
  T ratio = T(1.5e-5) * e_t / (k_t * k_t);
  if (ratio < T(1e-30)) ratio = T(1e-30);
  const T gb = T(2.1377) * pow(ratio, T(0.25));
  T gamma = gb * (gb * gb);
  if (gamma > T(1)) gamma = T(1);
  const T tau = T(0.4083) * sqrt(T(1.5e-5) / e_t);
  const T relax = T(1.0 - exp(-(P.dt / nsub) / 0.5));
  T f_h2o = T(1) - Yw / T(0.18);
  if (f_h2o < T(0)) f_h2o = T(0);
  T acc = T(0);
  for (int q = 0; q < nsub; ++q) {
    T omega;
    if (Yf <= T(1e-9) || YO2 <= T(1e-9)) {
      omega = T(0);
    } else {
      const T Ylim = Yf < (YO2 / s) ? Yf : (YO2 / s);
      omega = gamma * (rho_i * Ylim / tau);
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
      T wb = Yw / T(0.10);
      if (wb > T(1)) wb = T(1);
      Tgv = Tgv - (Tgv - T(373.15)) * relax * wb;
    }
    const T dY = -omega * h / rho_i;
    Yf = Yf + dY;
    if (Yf < T(0)) Yf = T(0);
    YO2 = YO2 + s * dY;
    if (YO2 < T(0)) YO2 = T(0);
    const T cp_mix = (T(1) - Yw) * cp_g + Yw * T(2000.0);
    Tgv = Tgv + omega * HoC * h / (rho_i * cp_mix);
    if (Tgv > T(2400.0)) Tgv = T(2400.0);
    acc += omega * h;
  }
  Yf_io[c] = Yf;
  YO2_io[c] = YO2;
  Tg[c] = Tgv;
  omega_out[c] = P.dt > 0.0 ? acc / T(P.dt) : T(0);
}

template <typename T>
static int edc_impl(const fw_grid* g, const void* rho, void* Tg, void* Yf,
                    void* YO2, const void* kt, const void* et, const void* Yw,
                    void* omega, int* work, const EdcParams& P, hipStream_t s) {
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const int nseg = (int)((n + EDC_SEG - 1) / EDC_SEG);
  int* summary = work;
  int* carry = work + nseg;
  hipLaunchKernelGGL(k_edc_summary<T>, dim3(nseg), dim3(256), 0, s, n,
                     (const T*)Yf, (const T*)YO2, (const T*)Yw, summary);
  FW_CHECK_LAUNCH("edc_summary");
  hipLaunchKernelGGL(k_edc_carry, dim3(1), dim3(256), 0, s, nseg,
                     (const int*)summary, carry);
  FW_CHECK_LAUNCH("edc_carry");
  hipLaunchKernelGGL(k_edc<T>, dim3(nseg), dim3(256), 0, s, n, (const T*)rho,
                     (T*)Tg, (T*)Yf, (T*)YO2, (const T*)kt, (const T*)et,
                     (const T*)Yw, (T*)omega, (const int*)carry, P);
  FW_CHECK_LAUNCH("edc");
  return FW_OK;
}

}  // namespace fw

extern "C" int fw_edc_workspace_ints(const fw_grid* g) {
  if (!g) return fw_set_error(FW_E_CONFIG, "bad grid");
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  return (int)(2 * ((n + EDC_SEG - 1) / EDC_SEG));
}

extern "C" int fw_edc(const fw_grid* g, const void* rho, void* T_g,
                      void* Y_fuel, void* Y_O2, const void* k_turb,
                      const void* eps_turb, const void* Y_H2O, void* omega_out,
                      double chi_rad, double cp_g, double dt, int n_substeps,
                      int extinction_enable, double s_stoich, double hoc_J,
                      int* work, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(work, "edc: null workspace");
  fw::EdcParams P{chi_rad, cp_g, dt, s_stoich, hoc_J, n_substeps,
                  extinction_enable ? 1 : 0};
  return FW_DISPATCH(g, fw::edc_impl, g, rho, T_g, Y_fuel, Y_O2, k_turb,
                     eps_turb, Y_H2O, omega_out, work, P, (hipStream_t)stream);
}
