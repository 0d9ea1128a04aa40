// EDC cell update of one 256-cell segment (chemistry_closures/edc.py:133-334),
// shared by k_edc and the temperature diffusion pass that carries the next
// sub-step's EDC.  The carried wet-bulb flag of the reference (edc.py:223)
// is a "last valid" scan over the flat cell order: inside the segment from
// the wavefronts' ballots, across segments from the per-segment summaries
// (-1: no resetting cell, else the flag of the last one).  Every thread of
// the calling workgroup must call it (it has workgroup barriers); thread t
// owns cell seg * EDC_SEG + t and passes that cell's temperature in Tg_in
// (anything when the cell is past n).
#pragma once
#include "fw_common.h"

namespace fw {

struct EdcParams {
  double chi_rad, cp_g, dt, s_stoich, hoc_J;
  double relax;   // 1 - exp(-(dt / n_substeps) / 0.5), formed on the host
  int n_substeps, extinction;
};

inline EdcParams edc_params(double chi_rad, double cp_g, double dt, int n_substeps,
                            int extinction_enable, double s_stoich, double hoc_J) {
  return EdcParams{chi_rad, cp_g,  dt, s_stoich, hoc_J, edc_relax(dt, n_substeps),
                   n_substeps, extinction_enable ? 1 : 0};
}

// The EDC fused into a temperature diffusion pass (turb_pass_launch): the
// fields of fw_edc, the pass's output being the temperature it updates.
struct EdcFuseArgs {
  const void *rho, *k_turb, *eps_turb, *Y_H2O;
  void *Y_fuel, *Y_O2, *omega_out;
  const int* summary;
  EdcParams P;
};

// A cell's EDC inputs besides its temperature, loaded by the caller (so a
// fused kernel can issue them with its own loads); past n: the neutral values.
template <typename T>
struct EdcCellIn {
  T Yf, YO2, Yw, rho, k, eps;
};

template <typename T>
__device__ __forceinline__ EdcCellIn<T> edc_load(bool valid, size_t c, const T* __restrict__ rho,
                                                 const T* __restrict__ Yf,
                                                 const T* __restrict__ YO2,
                                                 const T* __restrict__ kturb,
                                                 const T* __restrict__ eturb,
                                                 const T* __restrict__ Yw) {
  EdcCellIn<T> a;
  a.Yf = valid ? Yf[c] : T(0);
  a.YO2 = valid ? YO2[c] : T(0);
  a.Yw = valid ? Yw[c] : T(0);
  a.rho = valid ? rho[c] : T(1);
  a.k = valid ? kturb[c] : T(1);
  a.eps = valid ? eturb[c] : T(1);
  return a;
}

template <typename T>
__device__ __forceinline__ void edc_segment(size_t n, int seg, T Tg_in, const EdcCellIn<T>& in,
                                            T* __restrict__ Tg_out, T* __restrict__ Yf_io,
                                            T* __restrict__ YO2_io,
                                            T* __restrict__ omega_out,
                                            const int* __restrict__ summary,
                                            const EdcParams& P) {
  __shared__ int wmax[4];
  __shared__ unsigned char own_sh[256];
  __shared__ int s_carry;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const size_t c = (size_t)seg * EDC_SEG + t;
  const bool valid = c < n;
  // every load of the cell is issued before the workgroup exchange below
  T Yf = in.Yf;
  T YO2 = in.YO2;
  const T Yw = in.Yw;
  T Tgv = Tg_in;
  const T rho_i = in.rho;
  T k_t = in.k;
  T e_t = in.eps;
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
    const int b0 = (int)seg - 256 + t;   // thread t, segment b0: ascending
    const bool has = b0 >= 0 && summary[b0] >= 0;
    int last = block_last_true(has, wmax);
    __syncthreads();
    int found = last >= 0 ? (int)seg - 256 + last : -1;
    if (found < 0 && (int)seg > 256) {
      int best = -1;
#pragma unroll 4
      for (int b = t; b < (int)seg - 256; b += 256)
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
  Tg_out[c] = Tgv;
  omega_out[c] = P.dt > 0.0 ? fdiv(acc, T(P.dt)) : T(0);
}

}  // namespace fw
