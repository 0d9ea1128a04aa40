// Learned correction of the separable pressure preconditioner: a 4-level
// 3-D U-Net (16/32/64/64 channels, 3x3x3 kernels) with bfloat16 weights and
// activations and float32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16), its training kernels and the Adam step.
//
// The reference has no learned preconditioner (projection_3d.py:756-791
// preconditions BiCGSTAB with the separable inverse only); the network and
// its use in flexible PCG are this build's.  Levels L0 (the grid), L1 = L0/2,
// L2 = L0/4, L3 = L0/8 (grid dims divisible by 8); y periodic, zero padding
// in x and z.  Inputs: the scaled residual r and the density feature
// q = RHO_REF / rho - 1.
//   e0 = relu(conv([r, q]))        L0   2 -> 16   (VALU: K = 54)
//   e1 = relu(conv(pool(e0)))      L1  16 -> 32   (MFMA)
//   e2 = relu(conv(pool(e1)))      L2  32 -> 64
//   e3 = relu(conv(pool(e2)))      L3  64 -> 64
//   d2 = relu(conv([e2, up(e3)]))  L2 128 -> 64
//   d1 = relu(conv([e1, up(d2)]))  L1  96 -> 32
//   d0 = relu(conv([e0, up(d1)]))  L0  48 -> 16
//   c  = conv(d0)                  L0  16 -> 1    (VALU: N = 1)
// conv: y[p, co] = b[co] + sum_t sum_ci w[t, ci, co] x[p + off_t, ci], tap
// t = (dk+1) 9 + (dj+1) 3 + (di+1); pool = mean of 8 children; up = nearest.
// Parameters: one flat float32 vector, per layer w (27, ci, co) then b (co).
//
// Layout and precision.  Activations are channel-last bfloat16 [pixel][C],
// so the implicit GEMM's A operand (8 consecutive (tap, channel) values of a
// pixel) is one 16-byte load.  Weights are packed per layer into the B
// operand's lane order (K chunk, 16-column block, lane, 8 values), one
// 16-byte load per lane.  Rounding: weights, the inputs r and q, every stored
// activation (after bias and relu, and after pooling) and every stored
// activation gradient (after the relu mask) are rounded to bfloat16 (nearest,
// ties to even); everything else is float32, and every reduction (weight
// gradients, loss) is a fixed-order sum, so training is deterministic.
#include <math.h>
#include <vector>

#include "fw_common.h"

namespace fw {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;
union Frag {
  uint4 q;
  uint16_t h[8];
  bf16x8 v;
};

__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Lv {
  int nx, ny, nz;
};

// flat index of the neighbour (dk, dj, di) of (k, j, i), -1 outside in x, z
__device__ __forceinline__ int nbr_of(const Lv& L, int k, int j, int i, int dk, int dj, int di) {
  const int kk = k + dk, ii = i + di;
  if (kk < 0 || kk >= L.nz || ii < 0 || ii >= L.nx) return -1;
  int jj = j + dj;
  jj = jj < 0 ? jj + L.ny : (jj >= L.ny ? jj - L.ny : jj);
  return (kk * L.ny + jj) * L.nx + ii;
}
// the coarse (parent) pixel of fine pixel q on level L
__device__ __forceinline__ int parent_of(const Lv& L, int q) {
  const int i = q % L.nx, t = q / L.nx, j = t % L.ny, k = t / L.ny;
  return ((k >> 1) * (L.ny >> 1) + (j >> 1)) * (L.nx >> 1) + (i >> 1);
}
__device__ __forceinline__ void tap_off(int t, int& dk, int& dj, int& di) {
  dk = t / 9 - 1;
  dj = (t / 3) % 3 - 1;
  di = t % 3 - 1;
}

// ---- forward -----------------------------------------------------------

// layer e0 (2 -> 16) on the VALU: the inputs r, q (float32) rounded to
// bfloat16; weights (bfloat16-rounded, float) in LDS as [t][ci][co]
__global__ void __launch_bounds__(256)
k_e0_fwd(Lv L, const float* __restrict__ r, const float* __restrict__ q,
         const float* __restrict__ w, const float* __restrict__ b, uint16_t* __restrict__ out) {
  __shared__ float sw[27 * 2 * 16 + 16];
  for (int e = threadIdx.x; e < 27 * 32 + 16; e += 256) sw[e] = e < 27 * 32 ? w[e] : b[e - 27 * 32];
  __syncthreads();
  const int n = L.nx * L.ny * L.nz;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const int i = p % L.nx, tt = p / L.nx, j = tt % L.ny, k = tt / L.ny;
  float acc[16];
#pragma unroll
  for (int co = 0; co < 16; ++co) acc[co] = sw[27 * 32 + co];
  for (int t = 0; t < 27; ++t) {
    int dk, dj, di;
    tap_off(t, dk, dj, di);
    const int m = nbr_of(L, k, j, i, dk, dj, di);
    if (m < 0) continue;
    const float x0 = bf2f(f2bf(r[m])), x1 = bf2f(f2bf(q[m]));
    const float* wt = sw + t * 32;
#pragma unroll
    for (int co = 0; co < 16; ++co) acc[co] = fmaf(wt[co], x0, fmaf(wt[16 + co], x1, acc[co]));
  }
  Frag o0, o1;
#pragma unroll
  for (int co = 0; co < 8; ++co) {
    o0.h[co] = f2bf(acc[co] > 0.0f ? acc[co] : 0.0f);
    o1.h[co] = f2bf(acc[8 + co] > 0.0f ? acc[8 + co] : 0.0f);
  }
  uint4* dst = reinterpret_cast<uint4*>(out + (size_t)p * 16);
  dst[0] = o0.q;
  dst[1] = o1.q;
}

// 2x2x2 mean pooling of a C-channel bfloat16 field from level L to L/2;
// one thread per (coarse pixel, 8-channel group)
template <int C>
__global__ void __launch_bounds__(256)
k_pool(Lv L, const uint16_t* __restrict__ x, uint16_t* __restrict__ y) {
  const Lv H{L.nx / 2, L.ny / 2, L.nz / 2};
  const int nh = H.nx * H.ny * H.nz;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nh * (C / 8)) return;
  const int P = e / (C / 8), g = e % (C / 8);
  const int i = P % H.nx, t = P / H.nx, j = t % H.ny, k = t / H.ny;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = 0; c < 8; ++c) {
    const int q = ((2 * k + (c >> 2)) * L.ny + 2 * j + ((c >> 1) & 1)) * L.nx + 2 * i + (c & 1);
    Frag f;
    f.q = *reinterpret_cast<const uint4*>(x + (size_t)q * C + g * 8);
#pragma unroll
    for (int v = 0; v < 8; ++v) s[v] += bf2f(f.h[v]);
  }
  Frag o;
#pragma unroll
  for (int v = 0; v < 8; ++v) o.h[v] = f2bf(s[v] * 0.125f);
  *reinterpret_cast<uint4*>(y + (size_t)P * C + g * 8) = o.q;
}

// Implicit-GEMM 3x3x3 convolution on the matrix cores: out[p, n] =
// sum_k A[p, k] B[k, n], k = t CIN + ci (padded to 32-chunks with zeros).
// Input channels [0, C1) are x1 on level L, [C1, C1 + C2) are x2 on level
// L/2 read through nearest upsampling.  A wavefront computes MB 16-pixel
// blocks x CO channels; the workgroup 4 wavefronts.  FWD: + bias, relu,
// bfloat16 store; otherwise a float32 store of the plain product (the data
// gradient, with the packed transposed-and-flipped weights).
template <int C1, int C2, int CO, bool FWD, int MB>
__global__ void __launch_bounds__(256)
k_conv(Lv L, const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2,
       const uint4* __restrict__ wf, const float* __restrict__ bias, void* __restrict__ out) {
  constexpr int CIN = C1 + C2, KC = (27 * CIN + 31) / 32, NB = CO / 16;
  static_assert(CIN % 8 == 0 && C1 % 8 == 0 && CO % 16 == 0, "channel counts");
  const int npix = L.nx * L.ny * L.nz;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, kb = lane >> 4;
  const int m0 = (blockIdx.x * 4 + wave) * MB * 16;
  if (m0 >= npix) return;   // whole wavefront; no barriers in this kernel
  int pk[MB], pj[MB], pi[MB];
  bool pv[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    const int p = m0 + mb * 16 + (lane & 15);
    pv[mb] = p < npix;
    const int pp = pv[mb] ? p : 0;
    pi[mb] = pp % L.nx;
    const int t = pp / L.nx;
    pj[mb] = t % L.ny;
    pk[mb] = t / L.ny;
  }
  f4 acc[MB][NB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int kc = 0; kc < KC; ++kc) {
    const int k0 = kc * 32 + kb * 8;
    const int t = k0 / CIN, ci = k0 - t * CIN;
    Frag a[MB];
    if (t < 27) {
      int dk, dj, di;
      tap_off(t, dk, dj, di);
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const int m = pv[mb] ? nbr_of(L, pk[mb], pj[mb], pi[mb], dk, dj, di) : -1;
        if (m < 0) {
          a[mb].q = make_uint4(0, 0, 0, 0);
        } else if (ci < C1) {
          a[mb].q = *reinterpret_cast<const uint4*>(x1 + (size_t)m * C1 + ci);
        } else {
          a[mb].q = *reinterpret_cast<const uint4*>(x2 + (size_t)parent_of(L, m) * C2 + (ci - C1));
        }
      }
    } else {
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) a[mb].q = make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      Frag bf;
      bf.q = wf[((size_t)kc * NB + nb) * 64 + lane];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb].v, bf.v, acc[mb][nb], 0, 0, 0);
    }
  }
  // D[m = 4 kb + e][n = lane & 15] of each 16 x 16 tile
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int p = m0 + mb * 16 + 4 * kb + e;
        if (p >= npix) continue;
        const int co = nb * 16 + (lane & 15);
        float v = acc[mb][nb][e];
        if constexpr (FWD) {
          v += bias[co];
          reinterpret_cast<uint16_t*>(out)[(size_t)p * CO + co] = f2bf(v > 0.0f ? v : 0.0f);
        } else {
          reinterpret_cast<float*>(out)[(size_t)p * CO + co] = v;
        }
      }
}

// The same convolution for the coarse levels, where one wavefront per 16 x
// 16 output tile and a whole K loop would leave most of the chip idle: a
// workgroup computes one tile (16 pixels x 16 channels) and its 4
// wavefronts split the K chunks (kc = wave, wave + 4, ...); the partial
// tiles are summed in LDS in wavefront order (deterministic).
template <int C1, int C2, int CO, bool FWD>
__global__ void __launch_bounds__(256)
k_conv_sk(Lv L, const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2,
          const uint4* __restrict__ wf, const float* __restrict__ bias, void* __restrict__ out) {
  constexpr int CIN = C1 + C2, KC = (27 * CIN + 31) / 32, NB = CO / 16;
  static_assert(CIN % 8 == 0 && C1 % 8 == 0 && CO % 16 == 0, "channel counts");
  __shared__ float red[3][256];
  const int npix = L.nx * L.ny * L.nz;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, kb = lane >> 4;
  const int nb = blockIdx.y;
  const int p = blockIdx.x * 16 + (lane & 15);
  const bool pv = p < npix;
  const int pp = pv ? p : 0;
  const int pi = pp % L.nx, tq = pp / L.nx, pj = tq % L.ny, pk = tq / L.ny;
  f4 acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int kc = wave; kc < KC; kc += 4) {
    const int k0 = kc * 32 + kb * 8;
    const int t = k0 / CIN, ci = k0 - t * CIN;
    Frag a;
    a.q = make_uint4(0, 0, 0, 0);
    if (t < 27 && pv) {
      int dk, dj, di;
      tap_off(t, dk, dj, di);
      const int m = nbr_of(L, pk, pj, pi, dk, dj, di);
      if (m >= 0) {
        if (ci < C1)
          a.q = *reinterpret_cast<const uint4*>(x1 + (size_t)m * C1 + ci);
        else
          a.q = *reinterpret_cast<const uint4*>(x2 + (size_t)parent_of(L, m) * C2 + (ci - C1));
      }
    }
    Frag bf;
    bf.q = wf[((size_t)kc * NB + nb) * 64 + lane];
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, bf.v, acc, 0, 0, 0);
  }
  if (wave > 0)
#pragma unroll
    for (int e = 0; e < 4; ++e) red[wave - 1][lane * 4 + e] = acc[e];
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float v0 = ((acc[e] + red[0][lane * 4 + e]) + red[1][lane * 4 + e]) + red[2][lane * 4 + e];
    const int q = blockIdx.x * 16 + 4 * kb + e;
    if (q >= npix) continue;
    const int co = nb * 16 + (lane & 15);
    if constexpr (FWD) {
      const float v = v0 + bias[co];
      reinterpret_cast<uint16_t*>(out)[(size_t)q * CO + co] = f2bf(v > 0.0f ? v : 0.0f);
    } else {
      reinterpret_cast<float*>(out)[(size_t)q * CO + co] = v0;
    }
  }
}

// output layer c = b + sum_t sum_ci w[t, ci] d0[p + off_t, ci] (float32 out)
__global__ void __launch_bounds__(256)
k_c_fwd(Lv L, const uint16_t* __restrict__ d0, const float* __restrict__ w,
        const float* __restrict__ bias, float* __restrict__ c) {
  __shared__ float sw[27 * 16];
  for (int e = threadIdx.x; e < 27 * 16; e += 256) sw[e] = w[e];
  __syncthreads();
  const int n = L.nx * L.ny * L.nz;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const int i = p % L.nx, tt = p / L.nx, j = tt % L.ny, k = tt / L.ny;
  float acc = bias[0];
  for (int t = 0; t < 27; ++t) {
    int dk, dj, di;
    tap_off(t, dk, dj, di);
    const int m = nbr_of(L, k, j, i, dk, dj, di);
    if (m < 0) continue;
    Frag f0, f1;
    f0.q = *reinterpret_cast<const uint4*>(d0 + (size_t)m * 16);
    f1.q = *reinterpret_cast<const uint4*>(d0 + (size_t)m * 16 + 8);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      acc = fmaf(sw[t * 16 + v], bf2f(f0.h[v]), acc);
      acc = fmaf(sw[t * 16 + 8 + v], bf2f(f1.h[v]), acc);
    }
  }
  c[p] = acc;
}

// ---- weight packing ------------------------------------------------------

// B operand of the forward convolution: wf[kc][nb][lane][8] = w[k][n],
// k = kc 32 + 8 (lane / 16) + e, n = nb 16 + lane % 16, k < 27 CIN (else 0)
__global__ void __launch_bounds__(256)
k_pack_fwd(const float* __restrict__ w, int cin, int co, int kc_n, uint16_t* __restrict__ wf) {
  const int nb_n = co / 16;
  const int total = kc_n * nb_n * 64 * 8;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  const int v = e & 7, lane = (e >> 3) & 63, blk = e >> 9;
  const int nb = blk % nb_n, kc = blk / nb_n;
  const int k = kc * 32 + 8 * (lane >> 4) + v, n = nb * 16 + (lane & 15);
  wf[e] = k < 27 * cin ? f2bf(w[(size_t)k * co + n]) : (uint16_t)0;
}

// B operand of the data gradient: the forward layer's weights transposed
// and flipped, w'[t'][c'][n'] = w[26 - t'][n'][c'] (c' < co: the forward
// output channel, n' < cin: the forward input channel)
__global__ void __launch_bounds__(256)
k_pack_bwd(const float* __restrict__ w, int cin, int co, int kc_n, uint16_t* __restrict__ wb) {
  const int nb_n = cin / 16;
  const int total = kc_n * nb_n * 64 * 8;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  const int v = e & 7, lane = (e >> 3) & 63, blk = e >> 9;
  const int nb = blk % nb_n, kc = blk / nb_n;
  const int k = kc * 32 + 8 * (lane >> 4) + v, n = nb * 16 + (lane & 15);
  uint16_t out = 0;
  if (k < 27 * co) {
    const int t = k / co, c = k - t * co;
    out = f2bf(w[((size_t)(26 - t) * cin + n) * co + c]);
  }
  wb[e] = out;
}

// bfloat16-rounded float copy of n weights (the VALU layers)
__global__ void __launch_bounds__(256)
k_round_copy(int n, const float* __restrict__ w, float* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < n) out[e] = bf2f(f2bf(w[e]));
}

// ---- loss ----------------------------------------------------------------

// per-workgroup partials of sum dz (c - t)^2 and sum dz t^2
__global__ void __launch_bounds__(256)
k_loss_part(Lv L, const float* __restrict__ c, const float* __restrict__ t,
            const double* __restrict__ dz, double* __restrict__ part) {
  __shared__ double sh[2][256];
  const int n = L.nx * L.ny * L.nz;
  double a = 0.0, b = 0.0;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
    const double w = dz[p / (L.nx * L.ny)];
    const double d = (double)c[p] - (double)t[p];
    a += w * d * d;
    b += w * (double)t[p] * (double)t[p];
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// S[0] = sum dz (c - t)^2, S[1] = sum dz t^2 over nb partials, in order
__global__ void k_loss_fin(int nb, const double* __restrict__ part, double* __restrict__ S) {
  if (threadIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int q = 0; q < nb; ++q) {
    a += part[2 * q];
    b += part[2 * q + 1];
  }
  S[0] = a;
  S[1] = b;
}

// gc = dL/dc = 2 dz (c - t) / sum dz t^2
__global__ void __launch_bounds__(256)
k_loss_grad(Lv L, const float* __restrict__ c, const float* __restrict__ t,
            const double* __restrict__ dz, const double* __restrict__ S, float* __restrict__ gc) {
  const int n = L.nx * L.ny * L.nz;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double den = S[1] > 0.0 ? S[1] : 1.0;
  gc[p] = (float)(2.0 * dz[p / (L.nx * L.ny)] * ((double)c[p] - (double)t[p]) / den);
}

// ---- backward --------------------------------------------------------------

// data gradient of the output layer, masked by d0 > 0 and rounded:
// g[p, ci] = bf16((sum_t w[t, ci] gc[p - off_t]) (d0[p, ci] > 0))
__global__ void __launch_bounds__(256)
k_c_bwd(Lv L, const float* __restrict__ gc, const float* __restrict__ w,
        const uint16_t* __restrict__ d0, uint16_t* __restrict__ g) {
  __shared__ float sw[27 * 16];
  for (int e = threadIdx.x; e < 27 * 16; e += 256) sw[e] = w[e];
  __syncthreads();
  const int n = L.nx * L.ny * L.nz;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const int i = p % L.nx, tt = p / L.nx, j = tt % L.ny, k = tt / L.ny;
  float acc[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
  for (int t = 0; t < 27; ++t) {
    int dk, dj, di;
    tap_off(t, dk, dj, di);
    const int m = nbr_of(L, k, j, i, -dk, -dj, -di);
    if (m < 0) continue;
    const float gv = gc[m];
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = fmaf(sw[t * 16 + v], gv, acc[v]);
  }
  Frag a0, a1, o0, o1;
  a0.q = *reinterpret_cast<const uint4*>(d0 + (size_t)p * 16);
  a1.q = *reinterpret_cast<const uint4*>(d0 + (size_t)p * 16 + 8);
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    o0.h[v] = f2bf(bf2f(a0.h[v]) > 0.0f ? acc[v] : 0.0f);
    o1.h[v] = f2bf(bf2f(a1.h[v]) > 0.0f ? acc[8 + v] : 0.0f);
  }
  uint4* dst = reinterpret_cast<uint4*>(g + (size_t)p * 16);
  dst[0] = o0.q;
  dst[1] = o1.q;
}

// gradient of the coarse (upsampled) input of a concat layer, masked by the
// coarse activation and rounded: g[P, c] = bf16((sum over the 8 children q
// of dx[q, off + c]) (act[P, c] > 0)); dx has CT channels on level L
template <int C>
__global__ void __launch_bounds__(256)
k_sum8_mask(Lv L, const float* __restrict__ dx, int CT, int off,
            const uint16_t* __restrict__ act, uint16_t* __restrict__ g) {
  const Lv H{L.nx / 2, L.ny / 2, L.nz / 2};
  const int nh = H.nx * H.ny * H.nz;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nh * C) return;
  const int P = e / C, c = e % C;
  const int i = P % H.nx, t = P / H.nx, j = t % H.ny, k = t / H.ny;
  float s = 0.0f;
  for (int ch = 0; ch < 8; ++ch) {
    const int q = ((2 * k + (ch >> 2)) * L.ny + 2 * j + ((ch >> 1) & 1)) * L.nx + 2 * i + (ch & 1);
    s += dx[(size_t)q * CT + off + c];
  }
  g[e] = f2bf(bf2f(act[e]) > 0.0f ? s : 0.0f);
}

// gradient of a skip tensor (consumed by a concat layer and by pooling):
// g[p, c] = bf16((dxs[p, c] + dpool[parent(p), c] / 8) (act[p, c] > 0));
// dxs has CT channels (the skip part first), dpool C channels on L/2
template <int C>
__global__ void __launch_bounds__(256)
k_skip_mask(Lv L, const float* __restrict__ dxs, int CT, const float* __restrict__ dpool,
            const uint16_t* __restrict__ act, uint16_t* __restrict__ g) {
  const int n = L.nx * L.ny * L.nz;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * C) return;
  const int p = e / C, c = e % C;
  const float v = dxs[(size_t)p * CT + c] + dpool[(size_t)parent_of(L, p) * C + c] * 0.125f;
  g[e] = f2bf(bf2f(act[e]) > 0.0f ? v : 0.0f);
}

// Weight gradient of an MFMA layer on the matrix cores, partial over a chunk
// of pixels: part[chunk][t][ci][co] = sum_p x[p + off_t, ci] g[p, co] (the
// GEMM's K dimension is the pixel).  A wavefront owns one (tap, 16-channel
// input block) and all output channels.
template <int C1, int C2, int CO>
__global__ void __launch_bounds__(256)
k_dw(Lv L, const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2,
     const uint16_t* __restrict__ g, int chunk, float* __restrict__ part) {
  constexpr int CIN = C1 + C2, CIB = CIN / 16, NB = CO / 16;
  static_assert(C1 % 16 == 0 && CIN % 16 == 0 && CO % 16 == 0, "channel counts");
  const int npix = L.nx * L.ny * L.nz;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, kb = lane >> 4;
  const int grp = blockIdx.y * 4 + wave;
  if (grp >= 27 * CIB) return;   // whole wavefront; no barriers in this kernel
  const int t = grp / CIB, cb = grp - t * CIB;
  int dk, dj, di;
  tap_off(t, dk, dj, di);
  const int ci = cb * 16 + (lane & 15);
  const bool coarse = cb * 16 >= C1;   // uniform per wavefront (C1 % 16 == 0)
  const int p0 = blockIdx.x * chunk;
  const int p1 = p0 + chunk < npix ? p0 + chunk : npix;
  f4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int pb = p0; pb < p1; pb += 32) {
    Frag a;
    int pe[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int p = pb + 8 * kb + e;
      pe[e] = p;
      uint16_t v = 0;
      if (p < p1) {
        const int i = p % L.nx, tt = p / L.nx, j = tt % L.ny, k = tt / L.ny;
        const int m = nbr_of(L, k, j, i, dk, dj, di);
        if (m >= 0)
          v = coarse ? x2[(size_t)parent_of(L, m) * C2 + (ci - C1)] : x1[(size_t)m * C1 + ci];
      }
      a.h[e] = v;
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      Frag b;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        b.h[e] = pe[e] < p1 ? g[(size_t)pe[e] * CO + nb * 16 + (lane & 15)] : (uint16_t)0;
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[nb], 0, 0, 0);
    }
  }
  float* out = part + (size_t)blockIdx.x * 27 * CIN * CO;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int e = 0; e < 4; ++e)
      out[((size_t)t * CIN + cb * 16 + 4 * kb + e) * CO + nb * 16 + (lane & 15)] = acc[nb][e];
}

// Weight gradient of the VALU layers, partial over a chunk of pixels:
// part[chunk][t][ci][co] = sum_p x[p + off_t, ci] g[p, co].  E0: x = the
// bfloat16-rounded (r, q), g = the 16-channel bfloat16 gradient; else
// (the output layer) x = d0 (16 channels, bfloat16), g = gc (float32).
// grid (nchunk, 27); per-thread sums, then a fixed-order workgroup tree.
template <bool E0>
__global__ void __launch_bounds__(256)
k_dw_valu(Lv L, const float* __restrict__ r, const float* __restrict__ q,
          const uint16_t* __restrict__ xd0, const void* __restrict__ gv, int chunk,
          float* __restrict__ part) {
  constexpr int CI = E0 ? 2 : 16, CO = E0 ? 16 : 1, NO = CI * CO;
  __shared__ double sh[256];
  const int npix = L.nx * L.ny * L.nz;
  const int t = blockIdx.y;
  int dk, dj, di;
  tap_off(t, dk, dj, di);
  float acc[NO];
#pragma unroll
  for (int o = 0; o < NO; ++o) acc[o] = 0.0f;
  const int p0 = blockIdx.x * chunk;
  const int p1 = p0 + chunk < npix ? p0 + chunk : npix;
  for (int p = p0 + threadIdx.x; p < p1; p += 256) {
    const int i = p % L.nx, tt = p / L.nx, j = tt % L.ny, k = tt / L.ny;
    const int m = nbr_of(L, k, j, i, dk, dj, di);
    if (m < 0) continue;
    if constexpr (E0) {
      const float x0 = bf2f(f2bf(r[m])), x1 = bf2f(f2bf(q[m]));
      const uint16_t* g = reinterpret_cast<const uint16_t*>(gv) + (size_t)p * 16;
#pragma unroll
      for (int co = 0; co < 16; ++co) {
        const float gg = bf2f(g[co]);
        acc[co] = fmaf(x0, gg, acc[co]);
        acc[16 + co] = fmaf(x1, gg, acc[16 + co]);
      }
    } else {
      const float gg = reinterpret_cast<const float*>(gv)[p];
#pragma unroll
      for (int ci = 0; ci < 16; ++ci) acc[ci] = fmaf(bf2f(xd0[(size_t)m * 16 + ci]), gg, acc[ci]);
    }
  }
  float* out = part + ((size_t)blockIdx.x * 27 + t) * NO;
  for (int o = 0; o < NO; ++o) {
    sh[threadIdx.x] = acc[o];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[o] = (float)sh[0];
    __syncthreads();
  }
}

// bias gradient partials: part[block][co] = sum over the block's pixels of
// g[p, co] (bfloat16, or float32 when CO == 1)
template <int CO>
__global__ void __launch_bounds__(256)
k_db(int npix, const void* __restrict__ gv, int chunk, float* __restrict__ part) {
  constexpr int PER = 256 / CO;   // pixels handled side by side
  __shared__ double sh[256];
  const int co = threadIdx.x % CO, sub = threadIdx.x / CO;
  const int p0 = blockIdx.x * chunk;
  const int p1 = p0 + chunk < npix ? p0 + chunk : npix;
  double a = 0.0;
  for (int p = p0 + sub; p < p1; p += PER)
    a += CO == 1 ? (double)reinterpret_cast<const float*>(gv)[p]
                 : (double)bf2f(reinterpret_cast<const uint16_t*>(gv)[(size_t)p * CO + co]);
  sh[threadIdx.x] = a;
  __syncthreads();
  if ((int)threadIdx.x < CO) {
    double s = 0.0;
    for (int u = 0; u < PER; ++u) s += sh[u * CO + threadIdx.x];
    part[(size_t)blockIdx.x * CO + threadIdx.x] = (float)s;
  }
}

// out[e] = sum over nchunk partials of part[c][e], in chunk order
__global__ void __launch_bounds__(256)
k_reduce_parts(int n, int nchunk, const float* __restrict__ part, float* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * n + e];
  out[e] = (float)s;
}

// q = RHO_REF / max(rho, 0.01) - 1, the network's density feature
#define NET_RHO_REF (101325.0 / (8.314 / 0.029 * 300.0))
__global__ void __launch_bounds__(256)
k_density_feature(size_t n, const float* __restrict__ rho, float* __restrict__ q) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) {
    const float r = rho[e] > 0.01f ? rho[e] : 0.01f;
    q[e] = (float)NET_RHO_REF / r - 1.0f;
  }
}

int density_feature(const float* rho, float* q, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_density_feature, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     n, rho, q);
  FW_CHECK_LAUNCH("density feature");
  return FW_OK;
}

// Adam (Kingma & Ba 2015) on the flat parameter vector; c1 = 1 - b1^it and
// c2 = 1 - b2^it formed on the host
__global__ void __launch_bounds__(256)
k_adam(int n, float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.0f - b1) * gg;
  const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  th[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void __launch_bounds__(256) k_fill0(size_t n, float* __restrict__ a) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) a[e] = 0.0f;
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
enum { LE0, LE1, LE2, LE3, LD2, LD1, LD0, LC, NLAYER };
struct LayerDef {
  int level, c1, c2, co, woff, boff, kc_fwd, kc_bwd;
  size_t wf_off, wb_off;   // bfloat16 element offsets of the packed B operands
};
static const int LAYER_IO[NLAYER][4] = {   // level, c1, c2, co
    {0, 2, 0, 16}, {1, 16, 0, 32}, {2, 32, 0, 64}, {3, 64, 0, 64},
    {2, 64, 64, 64}, {1, 32, 64, 32}, {0, 16, 32, 16}, {0, 16, 0, 1}};

struct NetHandle {
  Lv lv[4];
  int npix[4];
  LayerDef L[NLAYER];
  int nparam;
  size_t nwf;
  float *theta, *grad, *m, *v, *wr;   // wr: bfloat16-rounded copy (the VALU layers)
  uint16_t* wpack;                    // packed B operands, forward and data gradient
  bool packed = false;
  // activations (bfloat16, channel-last)
  uint16_t *e0, *pe0, *e1, *pe1, *e2, *pe2, *e3, *d2, *d1, *d0;
  float* c;
  // backward buffers (allocated on the first training step)
  float *gc = nullptr, *dx6, *dx5, *dx4, *dx3, *dx2, *dx1, *part;
  uint16_t *gd0, *gd1, *gd2, *ge3, *ge2, *ge1, *ge0;
  double *lpart, *S;
  size_t part_n;
};

static void layer_defs(NetHandle* h) {
  int o = 0;
  size_t wo = 0;
  for (int l = 0; l < NLAYER; ++l) {
    LayerDef& d = h->L[l];
    d.level = LAYER_IO[l][0];
    d.c1 = LAYER_IO[l][1];
    d.c2 = LAYER_IO[l][2];
    d.co = LAYER_IO[l][3];
    const int cin = d.c1 + d.c2;
    d.woff = o;
    d.boff = o + 27 * cin * d.co;
    o = d.boff + d.co;
    d.kc_fwd = (27 * cin + 31) / 32;
    d.kc_bwd = (27 * d.co + 31) / 32;
    d.wf_off = d.wb_off = 0;
    if (l != LE0 && l != LC) {
      d.wf_off = wo;
      wo += (size_t)d.kc_fwd * (d.co / 16) * 512;
      d.wb_off = wo;
      wo += (size_t)d.kc_bwd * (cin / 16) * 512;
    }
  }
  h->nparam = o;
  h->nwf = wo;
}

static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

static int pack_weights(NetHandle* h, hipStream_t s) {
  for (int l = 0; l < NLAYER; ++l) {
    const LayerDef& d = h->L[l];
    const int cin = d.c1 + d.c2;
    if (l == LE0 || l == LC) {
      const int n = 27 * cin * d.co;   // weights only: biases stay float32
      hipLaunchKernelGGL(k_round_copy, dim3(nblk(n)), dim3(256), 0, s, n, h->theta + d.woff,
                         h->wr + d.woff);
      continue;
    }
    const int nf = d.kc_fwd * (d.co / 16) * 512, nbk = d.kc_bwd * (cin / 16) * 512;
    hipLaunchKernelGGL(k_pack_fwd, dim3(nblk(nf)), dim3(256), 0, s, h->theta + d.woff, cin, d.co,
                       d.kc_fwd, h->wpack + d.wf_off);
    hipLaunchKernelGGL(k_pack_bwd, dim3(nblk(nbk)), dim3(256), 0, s, h->theta + d.woff, cin,
                       d.co, d.kc_bwd, h->wpack + d.wb_off);
  }
  FW_CHECK_LAUNCH("net weight packing");
  h->packed = true;
  return FW_OK;
}

// Coarse levels (fewer than 512 workgroups of 128 pixels) take the
// split-K kernel, one 16 x 16 output tile per workgroup.
template <int C1, int C2, int CO, bool FWD>
static void conv_launch(const Lv& L, const uint16_t* x1, const uint16_t* x2, const uint16_t* wf,
                        const float* bias, void* out, hipStream_t s) {
  constexpr int MB = 2;
  const int npix = L.nx * L.ny * L.nz;
  const int per = 4 * MB * 16;
  const int nblk_mb = (npix + per - 1) / per;
  if (nblk_mb < 512)
    hipLaunchKernelGGL((k_conv_sk<C1, C2, CO, FWD>), dim3((npix + 15) / 16, CO / 16), dim3(256),
                       0, s, L, x1, x2, reinterpret_cast<const uint4*>(wf), bias, out);
  else
    hipLaunchKernelGGL((k_conv<C1, C2, CO, FWD, MB>), dim3(nblk_mb), dim3(256), 0, s, L, x1, x2,
                       reinterpret_cast<const uint4*>(wf), bias, out);
}

template <int C>
static void pool_launch(const Lv& L, const uint16_t* x, uint16_t* y, hipStream_t s) {
  const int nh = (L.nx / 2) * (L.ny / 2) * (L.nz / 2);
  hipLaunchKernelGGL(k_pool<C>, dim3(nblk((size_t)nh * (C / 8))), dim3(256), 0, s, L, x, y);
}

// forward of the whole network on (r, q); h->c = output
static int net_forward(NetHandle* h, const float* r, const float* q, hipStream_t s) {
  if (!h->packed) {
    int rc = pack_weights(h, s);
    if (rc) return rc;
  }
  const Lv* V = h->lv;
  const float* th = h->theta;
  const uint16_t* W = h->wpack;
  const LayerDef* D = h->L;
  hipLaunchKernelGGL(k_e0_fwd, dim3(nblk(h->npix[0])), dim3(256), 0, s, V[0], r, q,
                     h->wr + D[LE0].woff, th + D[LE0].boff, h->e0);
  pool_launch<16>(V[0], h->e0, h->pe0, s);
  conv_launch<16, 0, 32, true>(V[1], h->pe0, nullptr, W + D[LE1].wf_off, th + D[LE1].boff,
                               h->e1, s);
  pool_launch<32>(V[1], h->e1, h->pe1, s);
  conv_launch<32, 0, 64, true>(V[2], h->pe1, nullptr, W + D[LE2].wf_off, th + D[LE2].boff,
                               h->e2, s);
  pool_launch<64>(V[2], h->e2, h->pe2, s);
  conv_launch<64, 0, 64, true>(V[3], h->pe2, nullptr, W + D[LE3].wf_off, th + D[LE3].boff,
                               h->e3, s);
  conv_launch<64, 64, 64, true>(V[2], h->e2, h->e3, W + D[LD2].wf_off, th + D[LD2].boff, h->d2,
                                s);
  conv_launch<32, 64, 32, true>(V[1], h->e1, h->d2, W + D[LD1].wf_off, th + D[LD1].boff, h->d1,
                                s);
  conv_launch<16, 32, 16, true>(V[0], h->e0, h->d1, W + D[LD0].wf_off, th + D[LD0].boff, h->d0,
                                s);
  hipLaunchKernelGGL(k_c_fwd, dim3(nblk(h->npix[0])), dim3(256), 0, s, V[0], h->d0,
                     h->wr + D[LC].woff, th + D[LC].boff, h->c);
  FW_CHECK_LAUNCH("net forward");
  return FW_OK;
}

static int alloc_backward(NetHandle* h) {
  const size_t n0 = h->npix[0], n1 = h->npix[1], n2 = h->npix[2], n3 = h->npix[3];
  struct {
    float** p;
    size_t n;
  } f[] = {{&h->gc, n0}, {&h->dx6, 48 * n0}, {&h->dx5, 96 * n1}, {&h->dx4, 128 * n2},
           {&h->dx3, 64 * n3}, {&h->dx2, 32 * n2}, {&h->dx1, 16 * n1}};
  for (auto& a : f) FW_CHECK_HIP(hipMalloc((void**)a.p, a.n * sizeof(float)));
  struct {
    uint16_t** p;
    size_t n;
  } b[] = {{&h->gd0, 16 * n0}, {&h->gd1, 32 * n1}, {&h->gd2, 64 * n2}, {&h->ge3, 64 * n3},
           {&h->ge2, 64 * n2}, {&h->ge1, 32 * n1}, {&h->ge0, 16 * n0}};
  for (auto& a : b) FW_CHECK_HIP(hipMalloc((void**)a.p, a.n * sizeof(uint16_t)));
  // weight-gradient partials: at most 64 chunks of every layer's 27 CIN CO
  size_t mx = 0;
  for (int l = 0; l < NLAYER; ++l) {
    const size_t w = (size_t)27 * (h->L[l].c1 + h->L[l].c2) * h->L[l].co;
    mx = w > mx ? w : mx;
  }
  h->part_n = 64 * mx;
  FW_CHECK_HIP(hipMalloc((void**)&h->part, h->part_n * sizeof(float)));
  FW_CHECK_HIP(hipMalloc((void**)&h->lpart, 2 * 256 * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->S, 4 * sizeof(double)));
  return FW_OK;
}

// pixel chunk of the weight-gradient partials: at most 64 chunks, a
// multiple of 32 pixels
static int dw_chunk(int npix, int* nchunk) {
  int c = (npix + 63) / 64;
  c = (c + 31) / 32 * 32;
  *nchunk = (npix + c - 1) / c;
  return c;
}

template <int C1, int C2, int CO>
static void dw_launch(NetHandle* h, int l, const uint16_t* x1, const uint16_t* x2,
                      const uint16_t* g, hipStream_t s) {
  const LayerDef& d = h->L[l];
  const Lv& L = h->lv[d.level];
  const int npix = h->npix[d.level];
  int nchunk;
  const int chunk = dw_chunk(npix, &nchunk);
  constexpr int CIB = (C1 + C2) / 16;
  hipLaunchKernelGGL((k_dw<C1, C2, CO>), dim3(nchunk, (27 * CIB + 3) / 4), dim3(256), 0, s, L,
                     x1, x2, g, chunk, h->part);
  const int nw = 27 * (C1 + C2) * CO;
  hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(nw)), dim3(256), 0, s, nw, nchunk,
                     (const float*)h->part, h->grad + d.woff);
  hipLaunchKernelGGL(k_db<CO>, dim3(nchunk), dim3(256), 0, s, npix, (const void*)g, chunk,
                     h->part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, CO, nchunk,
                     (const float*)h->part, h->grad + d.boff);
}

// backward from the loss to h->grad, using the activations of the last
// forward; gc (dL/dc) is in h->gc
static int net_backward(NetHandle* h, const float* r, const float* q, hipStream_t s) {
  const Lv* V = h->lv;
  const int* P = h->npix;
  const uint16_t* W = h->wpack;
  const LayerDef* D = h->L;
  int nchunk;
  int chunk = dw_chunk(P[0], &nchunk);
  // output layer: weights and bias, then the gradient of d0
  hipLaunchKernelGGL(k_dw_valu<false>, dim3(nchunk, 27), dim3(256), 0, s, V[0], nullptr,
                     nullptr, h->d0, (const void*)h->gc, chunk, h->part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(27 * 16)), dim3(256), 0, s, 27 * 16, nchunk,
                     (const float*)h->part, h->grad + D[LC].woff);
  hipLaunchKernelGGL(k_db<1>, dim3(nchunk), dim3(256), 0, s, P[0], (const void*)h->gc, chunk,
                     h->part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, 1, nchunk, (const float*)h->part,
                     h->grad + D[LC].boff);
  hipLaunchKernelGGL(k_c_bwd, dim3(nblk(P[0])), dim3(256), 0, s, V[0], (const float*)h->gc,
                     h->wr + D[LC].woff, h->d0, h->gd0);
  // d0: [e0, up(d1)] -> 16
  dw_launch<16, 32, 16>(h, LD0, h->e0, h->d1, h->gd0, s);
  conv_launch<16, 0, 48, false>(V[0], h->gd0, nullptr, W + D[LD0].wb_off, nullptr, h->dx6, s);
  hipLaunchKernelGGL(k_sum8_mask<32>, dim3(nblk((size_t)P[1] * 32)), dim3(256), 0, s, V[0],
                     (const float*)h->dx6, 48, 16, h->d1, h->gd1);
  // d1: [e1, up(d2)] -> 32
  dw_launch<32, 64, 32>(h, LD1, h->e1, h->d2, h->gd1, s);
  conv_launch<32, 0, 96, false>(V[1], h->gd1, nullptr, W + D[LD1].wb_off, nullptr, h->dx5, s);
  hipLaunchKernelGGL(k_sum8_mask<64>, dim3(nblk((size_t)P[2] * 64)), dim3(256), 0, s, V[1],
                     (const float*)h->dx5, 96, 32, h->d2, h->gd2);
  // d2: [e2, up(e3)] -> 64
  dw_launch<64, 64, 64>(h, LD2, h->e2, h->e3, h->gd2, s);
  conv_launch<64, 0, 128, false>(V[2], h->gd2, nullptr, W + D[LD2].wb_off, nullptr, h->dx4, s);
  hipLaunchKernelGGL(k_sum8_mask<64>, dim3(nblk((size_t)P[3] * 64)), dim3(256), 0, s, V[2],
                     (const float*)h->dx4, 128, 64, h->e3, h->ge3);
  // e3: pool(e2) -> 64 (L3)
  dw_launch<64, 0, 64>(h, LE3, h->pe2, nullptr, h->ge3, s);
  conv_launch<64, 0, 64, false>(V[3], h->ge3, nullptr, W + D[LE3].wb_off, nullptr, h->dx3, s);
  hipLaunchKernelGGL(k_skip_mask<64>, dim3(nblk((size_t)P[2] * 64)), dim3(256), 0, s, V[2],
                     (const float*)h->dx4, 128, (const float*)h->dx3, h->e2, h->ge2);
  // e2: pool(e1) -> 64 (L2)
  dw_launch<32, 0, 64>(h, LE2, h->pe1, nullptr, h->ge2, s);
  conv_launch<64, 0, 32, false>(V[2], h->ge2, nullptr, W + D[LE2].wb_off, nullptr, h->dx2, s);
  hipLaunchKernelGGL(k_skip_mask<32>, dim3(nblk((size_t)P[1] * 32)), dim3(256), 0, s, V[1],
                     (const float*)h->dx5, 96, (const float*)h->dx2, h->e1, h->ge1);
  // e1: pool(e0) -> 32 (L1)
  dw_launch<16, 0, 32>(h, LE1, h->pe0, nullptr, h->ge1, s);
  conv_launch<32, 0, 16, false>(V[1], h->ge1, nullptr, W + D[LE1].wb_off, nullptr, h->dx1, s);
  hipLaunchKernelGGL(k_skip_mask<16>, dim3(nblk((size_t)P[0] * 16)), dim3(256), 0, s, V[0],
                     (const float*)h->dx6, 48, (const float*)h->dx1, h->e0, h->ge0);
  // e0: [r, q] -> 16, weights and bias only
  hipLaunchKernelGGL(k_dw_valu<true>, dim3(nchunk, 27), dim3(256), 0, s, V[0], r, q, nullptr,
                     (const void*)h->ge0, chunk, h->part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(27 * 32)), dim3(256), 0, s, 27 * 32, nchunk,
                     (const float*)h->part, h->grad + D[LE0].woff);
  hipLaunchKernelGGL(k_db<16>, dim3(nchunk), dim3(256), 0, s, P[0], (const void*)h->ge0, chunk,
                     h->part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, 16, nchunk, (const float*)h->part,
                     h->grad + D[LE0].boff);
  FW_CHECK_LAUNCH("net backward");
  return FW_OK;
}

// For the projection's learned preconditioner: N(r) of the attached net;
// *c points at the net's output buffer (valid until its next forward).
int net_apply(void* net, const float* r, const float* q, const float** c, int* n,
              hipStream_t s) {
  NetHandle* h = (NetHandle*)net;
  *c = h->c;
  *n = h->npix[0];
  return net_forward(h, r, q, s);
}

int net_dims(void* net, int* nx, int* ny, int* nz) {
  NetHandle* h = (NetHandle*)net;
  *nx = h->lv[0].nx;
  *ny = h->lv[0].ny;
  *nz = h->lv[0].nz;
  return FW_OK;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_net_density_feature(const void* rho, void* q, int64_t n, void* stream) {
  FW_REQUIRE(rho && q && n > 0, "net_density_feature: bad argument");
  return density_feature((const float*)rho, (float*)q, (size_t)n, (hipStream_t)stream);
}

int fw_net_nparams(void) {
  NetHandle h;
  layer_defs(&h);
  return h.nparam;
}

int fw_net_create(int nx, int ny, int nz, const float* theta_host, void** handle) {
  LegacyScope legacy_scope;
  FW_REQUIRE(handle && theta_host, "net_create: null argument");
  FW_REQUIRE(nx >= 8 && ny >= 8 && nz >= 8 && nx % 8 == 0 && ny % 8 == 0 && nz % 8 == 0,
             "net_create: grid dims must be multiples of 8 (got %d x %d x %d)", nx, ny, nz);
  FW_REQUIRE((size_t)nx * ny * nz * 128 <= 0x7FFFFFFFull,
             "net_create: the grid is too large for the network's 32-bit indexing");
  NetHandle* h = new NetHandle();
  for (int l = 0; l < 4; ++l) {
    h->lv[l] = Lv{nx >> l, ny >> l, nz >> l};
    h->npix[l] = (nx >> l) * (ny >> l) * (nz >> l);
  }
  layer_defs(h);
  float** vec[] = {&h->theta, &h->grad, &h->m, &h->v, &h->wr};
  for (float** p : vec) FW_CHECK_HIP(hipMalloc((void**)p, h->nparam * sizeof(float)));
  FW_CHECK_HIP(hipMalloc((void**)&h->wpack, h->nwf * sizeof(uint16_t)));
  FW_CHECK_HIP(hipMemcpy(h->theta, theta_host, h->nparam * sizeof(float),
                         hipMemcpyHostToDevice));
  FW_CHECK_HIP(hipMemset(h->m, 0, h->nparam * sizeof(float)));
  FW_CHECK_HIP(hipMemset(h->v, 0, h->nparam * sizeof(float)));
  FW_CHECK_HIP(hipMemset(h->grad, 0, h->nparam * sizeof(float)));
  const size_t n0 = h->npix[0], n1 = h->npix[1], n2 = h->npix[2], n3 = h->npix[3];
  struct {
    uint16_t** p;
    size_t n;
  } acts[] = {{&h->e0, 16 * n0}, {&h->pe0, 16 * n1}, {&h->e1, 32 * n1}, {&h->pe1, 32 * n2},
              {&h->e2, 64 * n2}, {&h->pe2, 64 * n3}, {&h->e3, 64 * n3}, {&h->d2, 64 * n2},
              {&h->d1, 32 * n1}, {&h->d0, 16 * n0}};
  for (auto& a : acts) FW_CHECK_HIP(hipMalloc((void**)a.p, a.n * sizeof(uint16_t)));
  FW_CHECK_HIP(hipMalloc((void**)&h->c, n0 * sizeof(float)));
  *handle = h;
  return FW_OK;
}

int fw_net_destroy(void* handle) {
  LegacyScope legacy_scope;
  NetHandle* h = (NetHandle*)handle;
  if (!h) return FW_OK;
  void* bufs[] = {h->theta, h->grad, h->m,   h->v,   h->wr,  h->wpack, h->e0,  h->pe0,
                  h->e1,    h->pe1,  h->e2,  h->pe2, h->e3,  h->d2,    h->d1,  h->d0,
                  h->c,     h->gc,   h->dx6, h->dx5, h->dx4, h->dx3,   h->dx2, h->dx1,
                  h->part,  h->gd0,  h->gd1, h->gd2, h->ge3, h->ge2,   h->ge1, h->ge0,
                  h->lpart, h->S};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete h;
  return FW_OK;
}

int fw_net_get_params(void* handle, float* theta_host) {
  LegacyScope legacy_scope;
  NetHandle* h = (NetHandle*)handle;
  FW_REQUIRE(h && theta_host, "net_get_params: null argument");
  FW_CHECK_HIP(hipMemcpy(theta_host, h->theta, h->nparam * sizeof(float),
                         hipMemcpyDeviceToHost));
  return FW_OK;
}

int fw_net_forward(void* handle, const void* r, const void* q, void* c_out, void* stream) {
  NetHandle* h = (NetHandle*)handle;
  FW_REQUIRE(h && r && q && c_out, "net_forward: null argument");
  hipStream_t s = (hipStream_t)stream;
  int rc = net_forward(h, (const float*)r, (const float*)q, s);
  if (rc) return rc;
  FW_CHECK_HIP(hipMemcpyAsync(c_out, h->c, (size_t)h->npix[0] * 4, hipMemcpyDeviceToDevice, s));
  return FW_OK;
}

int fw_net_set_params(void* handle, const float* theta_host) {
  LegacyScope legacy_scope;
  NetHandle* h = (NetHandle*)handle;
  FW_REQUIRE(h && theta_host, "net_set_params: null argument");
  FW_CHECK_HIP(hipMemcpy(h->theta, theta_host, h->nparam * sizeof(float),
                         hipMemcpyHostToDevice));
  FW_CHECK_HIP(hipMemset(h->m, 0, h->nparam * sizeof(float)));
  FW_CHECK_HIP(hipMemset(h->v, 0, h->nparam * sizeof(float)));
  h->packed = false;
  return FW_OK;
}

int fw_net_train_step(void* handle, const void* r, const void* q, const void* t,
                      const double* dz, double lr, int it, double* loss, void* stream) {
  NetHandle* h = (NetHandle*)handle;
  FW_REQUIRE(h && r && q && t && dz && loss, "net_train_step: null argument");
  FW_REQUIRE(it >= 1, "net_train_step: Adam step number must be >= 1");
  if (!h->gc) {
    int rc = alloc_backward(h);
    if (rc) return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = net_forward(h, (const float*)r, (const float*)q, s);
  if (rc) return rc;
  const Lv L0 = h->lv[0];
  hipLaunchKernelGGL(k_loss_part, dim3(256), dim3(256), 0, s, L0, (const float*)h->c,
                     (const float*)t, dz, h->lpart);
  hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(64), 0, s, 256, (const double*)h->lpart, h->S);
  hipLaunchKernelGGL(k_loss_grad, dim3(nblk(h->npix[0])), dim3(256), 0, s, L0,
                     (const float*)h->c, (const float*)t, dz, (const double*)h->S, h->gc);
  if ((rc = net_backward(h, (const float*)r, (const float*)q, s))) return rc;
  if (lr > 0.0) {
    const double c1 = 1.0 - pow(0.9, (double)it), c2 = 1.0 - pow(0.999, (double)it);
    hipLaunchKernelGGL(k_adam, dim3(nblk(h->nparam)), dim3(256), 0, s, h->nparam, h->theta,
                       (const float*)h->grad, h->m, h->v, (float)lr, 0.9f, 0.999f, 1.0e-8f,
                       (float)c1, (float)c2);
    h->packed = false;
  }
  FW_CHECK_LAUNCH("net train step");
  double sc[2];
  FW_CHECK_HIP(hipMemcpyAsync(sc, h->S, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  FW_CHECK_HIP(hipStreamSynchronize(s));
  *loss = sc[1] > 0.0 ? sc[0] / sc[1] : 0.0;
  return FW_OK;
}

int fw_net_grad_copy(void* handle, float* grad_host) {
  NetHandle* h = (NetHandle*)handle;
  FW_REQUIRE(h && grad_host, "net_grad_copy: null argument");
  FW_CHECK_HIP(hipMemcpy(grad_host, h->grad, h->nparam * sizeof(float),
                         hipMemcpyDeviceToHost));
  return FW_OK;
}

}  // extern "C"
