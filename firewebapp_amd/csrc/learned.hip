// Learned correction of the separable preconditioner: a two-level 3-D
// convolutional network (float32), its training kernels and the Adam step.
//
// The reference has no learned preconditioner (projection_3d.py:756-791
// preconditions BiCGSTAB with the separable inverse only); the network and
// its use in flexible PCG are this build's.  Network, with C0 = 8, C1 = 16,
// y periodic and zero padding in x and z (grid dims even), on the residual r
// and the density feature q = RHO_REF / rho - 1:
//   h0  = relu(conv3([r, q], w0, b0))        2 -> C0
//   d1  = avgpool2(h0)
//   h1  = relu(conv3(d1, w1, b1))            C0 -> C1    (coarse grid)
//   h1b = relu(conv3(h1, w2, b2))            C1 -> C1
//   h2  = relu(conv3([h0, up2(h1b)], w3, b3))   C0 + C1 -> C0
//   c   = conv3(h2, w4, b4)                  C0 -> 1
// conv3: y[co, p] = b[co] + sum_ci sum_t w[co, ci, t] x[ci, p + off_t], tap
// t = (dk+1) 9 + (dj+1) 3 + (di+1).  Activations are channel-major
// [C][Nz][Ny][Nx].  Parameters are one flat float32 vector, per layer w
// (co, ci, 27) then b (co).  Every reduction (weight gradients, loss) is a
// fixed-order two-level sum, so training is deterministic.
#include <vector>
#include "fw_common.h"

namespace fw {

struct Dims {
  int nx, ny, nz;
};

__device__ __forceinline__ int wrapj(int j, int n) { return j < 0 ? j + n : (j >= n ? j - n : j); }

// y = conv3(x) (+ b) with CO output channels, optional relu, optional mask
// (y *= mask > 0, for the backward pass).  Input channels [0, c1) come from
// x1 on the same grid, [c1, c1 + c2) from x2, which is on the same grid or,
// with x2_coarse, on the half-resolution grid read through up2 (nearest).
// flip: use w[ci][co][26 - t] (the adjoint of the forward convolution,
// with the weight tensor of the forward layer whose input has CO channels).
// One thread per cell; the weights are staged in LDS as [ci][t][CO].
template <int CO>
__global__ void __launch_bounds__(256)
k_conv3(Dims d, const float* __restrict__ x1, int c1, const float* __restrict__ x2, int c2,
        int x2_coarse, const float* __restrict__ w, const float* __restrict__ b,
        int flip, int relu, const float* __restrict__ mask, float* __restrict__ y) {
  extern __shared__ float sw[];
  const int cin = c1 + c2;
  const int nw = CO * cin * 27;
  for (int e = threadIdx.x; e < nw; e += blockDim.x) {
    // e = (ci * 27 + t) * CO + co
    const int co = e % CO, rest = e / CO, t = rest % 27, ci = rest / 27;
    // forward: w[co][ci][t]; flipped (adjoint): the forward layer is
    // (out = ci-space here, in = co-space here): w[ci][co][26 - t]
    sw[e] = flip ? w[((size_t)ci * CO + co) * 27 + (26 - t)]
                 : w[((size_t)co * cin + ci) * 27 + t];
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= d.nx) return;   // after the only barrier
  const int N = d.nx * d.ny * d.nz;
  const int hx = d.nx / 2, hy = d.ny / 2, hz = d.nz / 2, Nh = hx * hy * hz;
  float acc[CO];
#pragma unroll
  for (int co = 0; co < CO; ++co) acc[co] = b ? b[co] : 0.0f;
  for (int ci = 0; ci < cin; ++ci) {
    const bool second = ci >= c1;
    const bool coarse = second && x2_coarse;
    const float* src = second ? x2 + (size_t)(ci - c1) * (coarse ? Nh : N)
                              : x1 + (size_t)ci * N;
    const float* wr = sw + (size_t)ci * 27 * CO;
#pragma unroll 3
    for (int dk = -1; dk <= 1; ++dk) {
      const int kk = k + dk;
      if (kk < 0 || kk >= d.nz) continue;
      for (int dj = -1; dj <= 1; ++dj) {
        const int jj = wrapj(j + dj, d.ny);
        for (int di = -1; di <= 1; ++di) {
          const int ii = i + di;
          if (ii < 0 || ii >= d.nx) continue;
          const float v = coarse ? src[((size_t)(kk >> 1) * hy + (jj >> 1)) * hx + (ii >> 1)]
                                 : src[((size_t)kk * d.ny + jj) * d.nx + ii];
          const int t = (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
          const float* wt = wr + t * CO;
#pragma unroll
          for (int co = 0; co < CO; ++co) acc[co] = fmaf(wt[co], v, acc[co]);
        }
      }
    }
  }
  const size_t c = ((size_t)k * d.ny + j) * d.nx + i;
#pragma unroll
  for (int co = 0; co < CO; ++co) {
    float v = acc[co];
    if (relu) v = v > 0.0f ? v : 0.0f;
    if (mask && !(mask[(size_t)co * N + c] > 0.0f)) v = 0.0f;
    y[(size_t)co * N + c] = v;
  }
}

// Weight gradient of one layer, partial over a chunk of cells:
// part[((co * cin + ci) * nchunk + chunk) * 28 + t] = sum_q g[co, q] x[ci, q + off_t]
// (t < 27) and, for ci = 0, slot 27 = sum_q g[co, q] (the bias gradient).
// grid (nchunk, co, ci); x channels as in k_conv3 (x2 may be coarse).
__global__ void __launch_bounds__(256)
k_conv3_dw(Dims d, const float* __restrict__ g, const float* __restrict__ x1, int c1,
           const float* __restrict__ x2, int c2, int x2_coarse, int nchunk,
           double* __restrict__ part) {
  __shared__ double sh[4][28];
  const int chunk = blockIdx.x, co = blockIdx.y, ci = blockIdx.z;
  const int cin = c1 + c2;
  const int N = d.nx * d.ny * d.nz;
  const int hx = d.nx / 2, hy = d.ny / 2, hz = d.nz / 2, Nh = hx * hy * hz;
  const bool second = ci >= c1;
  const bool coarse = second && x2_coarse;
  const float* src = second ? x2 + (size_t)(ci - c1) * (coarse ? Nh : N) : x1 + (size_t)ci * N;
  const float* gc = g + (size_t)co * N;
  double acc[28];
#pragma unroll
  for (int t = 0; t < 28; ++t) acc[t] = 0.0;
  const int per = (N + nchunk - 1) / nchunk;
  const int q0 = chunk * per, q1 = q0 + per < N ? q0 + per : N;
  for (int q = q0 + threadIdx.x; q < q1; q += 256) {
    const float gv = gc[q];
    if (gv == 0.0f) continue;
    const int i = q % d.nx, j = (q / d.nx) % d.ny, k = q / (d.nx * d.ny);
    acc[27] += gv;
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int dk = t / 9 - 1, dj = (t / 3) % 3 - 1, di = t % 3 - 1;
      const int kk = k + dk, ii = i + di;
      if (kk < 0 || kk >= d.nz || ii < 0 || ii >= d.nx) continue;
      const int jj = wrapj(j + dj, d.ny);
      const float v = coarse ? src[((size_t)(kk >> 1) * hy + (jj >> 1)) * hx + (ii >> 1)]
                             : src[((size_t)kk * d.ny + jj) * d.nx + ii];
      acc[t] += (double)gv * (double)v;
    }
  }
  // workgroup sums of the 28 values in a fixed order
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < 28; ++t) {
    double v = acc[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) sh[wv][t] = v;
  }
  __syncthreads();
  if (threadIdx.x < 28) {
    const int t = threadIdx.x;
    const double s = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
    part[(((size_t)co * cin + ci) * nchunk + chunk) * 28 + t] = s;
  }
}

// grad (flat, per layer as the parameters) of the layer at offset woff:
// gw[co][ci][t] = sum over chunks; gb[co] = sum over chunks of slot 27 (ci 0)
__global__ void __launch_bounds__(256)
k_dw_finish(int co_n, int cin, int nchunk, const double* __restrict__ part,
            float* __restrict__ grad, int woff) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int nw = co_n * cin * 27;
  if (e < nw) {
    const int t = e % 27, ci = (e / 27) % cin, co = e / (27 * cin);
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(((size_t)co * cin + ci) * nchunk + c) * 28 + t];
    grad[woff + e] = (float)s;
  } else if (e < nw + co_n) {
    const int co = e - nw;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(((size_t)co * cin + 0) * nchunk + c) * 28 + 27];
    grad[woff + e] = (float)s;
  }
}

// y[c, K, J, I] = mean of x over the 2x2x2 children
__global__ void __launch_bounds__(256)
k_pool2(Dims d, int C, const float* __restrict__ x, float* __restrict__ y) {
  const int hx = d.nx / 2, hy = d.ny / 2, hz = d.nz / 2, Nh = hx * hy * hz;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= C * Nh) return;
  const int c = e / Nh, p = e % Nh;
  const int I = p % hx, J = (p / hx) % hy, K = p / (hx * hy);
  const float* xc = x + (size_t)c * d.nx * d.ny * d.nz;
  float s = 0.0f;
  for (int a = 0; a < 2; ++a)
    for (int bq = 0; bq < 2; ++bq)
      for (int q = 0; q < 2; ++q)
        s += xc[((size_t)(2 * K + a) * d.ny + (2 * J + bq)) * d.nx + (2 * I + q)];
  y[e] = s * 0.125f;
}

// dh1b[c, P] = (sum over the children of du[c]) * (h1b > 0)
__global__ void __launch_bounds__(256)
k_up2_T(Dims d, int C, const float* __restrict__ du, const float* __restrict__ h,
        float* __restrict__ out) {
  const int hx = d.nx / 2, hy = d.ny / 2, hz = d.nz / 2, Nh = hx * hy * hz;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= C * Nh) return;
  const int c = e / Nh, p = e % Nh;
  const int I = p % hx, J = (p / hx) % hy, K = p / (hx * hy);
  const float* dc = du + (size_t)c * d.nx * d.ny * d.nz;
  float s = 0.0f;
  for (int a = 0; a < 2; ++a)
    for (int bq = 0; bq < 2; ++bq)
      for (int q = 0; q < 2; ++q)
        s += dc[((size_t)(2 * K + a) * d.ny + (2 * J + bq)) * d.nx + (2 * I + q)];
  out[e] = h[e] > 0.0f ? s : 0.0f;
}

// dh0[c, p] = (dcat[c, p] + dd1[c, parent(p)] / 8) * (h0 > 0)
__global__ void __launch_bounds__(256)
k_mix_dh0(Dims d, int C, const float* __restrict__ dcat, const float* __restrict__ dd1,
          const float* __restrict__ h0, float* __restrict__ out) {
  const int N = d.nx * d.ny * d.nz, hx = d.nx / 2, hy = d.ny / 2, hz = d.nz / 2;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= C * N) return;
  const int c = e / N, p = e % N;
  const int i = p % d.nx, j = (p / d.nx) % d.ny, k = p / (d.nx * d.ny);
  const float v = dcat[e] + dd1[(size_t)c * hx * hy * hz +
                                ((size_t)(k >> 1) * hy + (j >> 1)) * hx + (i >> 1)] * 0.125f;
  out[e] = h0[e] > 0.0f ? v : 0.0f;
}

// partial sums of dz t^2 (den_only) or of dz (c - t)^2 with the output
// gradient gc = 2 dz (c - t) / den, den = *den_p (L = sum dz (c - t)^2 / den)
__global__ void __launch_bounds__(256)
k_loss_grad(Dims d, const float* __restrict__ c, const float* __restrict__ t,
            const double* __restrict__ dz, const double* __restrict__ den_p, int den_only,
            float* __restrict__ gc, double* __restrict__ part) {
  __shared__ double sh[4];
  const int N = d.nx * d.ny * d.nz;
  const int e = blockIdx.x * 256 + threadIdx.x;
  double a = 0.0;
  if (e < N) {
    const double w = dz[e / (d.nx * d.ny)];
    if (den_only) {
      a = w * (double)t[e] * (double)t[e];
    } else {
      const double r = (double)c[e] - (double)t[e];
      a = w * r * r;
      gc[e] = (float)(2.0 * w * r / *den_p);
    }
  }
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__global__ void __launch_bounds__(256)
k_sum_parts(int n, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[4];
  double a = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) a += part[b];
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) *out = ((sh[0] + sh[1]) + sh[2]) + sh[3];
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

// Adam (Kingma & Ba 2015) on the flat parameter vector, step number it >= 1
__global__ void __launch_bounds__(256)
k_adam(int n, float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
       float* __restrict__ v, float lr, float b1, float b2, float eps, int it) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.0f - b1) * gg;
  const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  const float mh = mm / (1.0f - powf(b1, (float)it));
  const float vh = vv / (1.0f - powf(b2, (float)it));
  th[e] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
static const int NC0 = 8, NC1 = 16;
struct LayerDef {
  int cin, co, woff;
};

struct NetHandle {
  Dims d;
  int N, Nh;
  LayerDef L[5];
  int nparam;
  float *theta, *grad, *m, *v;
  // activations and backward buffers (float32)
  float *h0, *d1, *h1, *h1b, *h2, *c, *gc, *dh2, *dcat, *dh1b, *dh1, *dd1, *dh0;
  double *part, *dwpart, *scal;
  int nchunk;
};

static void layer_defs(NetHandle* h) {
  const int io[5][2] = {{2, NC0}, {NC0, NC1}, {NC1, NC1}, {NC0 + NC1, NC0}, {NC0, 1}};
  int o = 0;
  for (int l = 0; l < 5; ++l) {
    h->L[l].cin = io[l][0];
    h->L[l].co = io[l][1];
    h->L[l].woff = o;
    o += io[l][0] * io[l][1] * 27 + io[l][1];
  }
  h->nparam = o;
}

static size_t conv_shm(int co, int cin) { return (size_t)co * cin * 27 * sizeof(float); }

// y = conv3 of layer l (forward or flipped/adjoint) on grid dd
static void conv(NetHandle* h, Dims dd, int co, const float* x1, int c1, const float* x2,
                 int c2, int x2c, const float* w, const float* b, int flip, int relu,
                 const float* mask, float* y, hipStream_t s) {
  const int bx = dd.nx <= 64 ? 64 : (dd.nx <= 128 ? 128 : 256);
  const dim3 grid((dd.nx + bx - 1) / bx, dd.ny, dd.nz);
  const size_t shm = conv_shm(co, c1 + c2);
  switch (co) {
    case 1: hipLaunchKernelGGL(k_conv3<1>, grid, dim3(bx), shm, s, dd, x1, c1, x2, c2, x2c, w, b, flip, relu, mask, y); break;
    case 8: hipLaunchKernelGGL(k_conv3<8>, grid, dim3(bx), shm, s, dd, x1, c1, x2, c2, x2c, w, b, flip, relu, mask, y); break;
    case 16: hipLaunchKernelGGL(k_conv3<16>, grid, dim3(bx), shm, s, dd, x1, c1, x2, c2, x2c, w, b, flip, relu, mask, y); break;
    default: hipLaunchKernelGGL(k_conv3<24>, grid, dim3(bx), shm, s, dd, x1, c1, x2, c2, x2c, w, b, flip, relu, mask, y); break;
  }
  (void)h;
}

static Dims half(Dims d) { return Dims{d.nx / 2, d.ny / 2, d.nz / 2}; }

// forward of the whole network on (r, q) (N values each); c = output
static int net_forward(NetHandle* h, const float* r, const float* q, hipStream_t s) {
  const Dims d = h->d, dh = half(d);
  const float* th = h->theta;
  auto W = [&](int l) { return th + h->L[l].woff; };
  auto B = [&](int l) { return th + h->L[l].woff + h->L[l].cin * h->L[l].co * 27; };
  conv(h, d, NC0, r, 1, q, 1, 0, W(0), B(0), 0, 1, nullptr, h->h0, s);
  hipLaunchKernelGGL(k_pool2, dim3((NC0 * h->Nh + 255) / 256), dim3(256), 0, s, d, NC0,
                     h->h0, h->d1);
  conv(h, dh, NC1, h->d1, NC0, nullptr, 0, 0, W(1), B(1), 0, 1, nullptr, h->h1, s);
  conv(h, dh, NC1, h->h1, NC1, nullptr, 0, 0, W(2), B(2), 0, 1, nullptr, h->h1b, s);
  conv(h, d, NC0, h->h0, NC0, h->h1b, NC1, 1, W(3), B(3), 0, 1, nullptr, h->h2, s);
  conv(h, d, 1, h->h2, NC0, nullptr, 0, 0, W(4), B(4), 0, 0, nullptr, h->c, s);
  FW_CHECK_LAUNCH("net forward");
  return FW_OK;
}

static void dw(NetHandle* h, Dims dd, int l, const float* g, const float* x1, int c1,
               const float* x2, int c2, int x2c, hipStream_t s) {
  const int co = h->L[l].co, cin = h->L[l].cin;
  hipLaunchKernelGGL(k_conv3_dw, dim3(h->nchunk, co, cin), dim3(256), 0, s, dd, g, x1, c1, x2,
                     c2, x2c, h->nchunk, h->dwpart);
  const int n = co * cin * 27 + co;
  hipLaunchKernelGGL(k_dw_finish, dim3((n + 255) / 256), dim3(256), 0, s, co, cin, h->nchunk,
                     (const double*)h->dwpart, h->grad, h->L[l].woff);
}

// backward from gc (dL/dc) to h->grad, using the activations of the last
// forward
static int net_backward(NetHandle* h, const float* r, const float* q, hipStream_t s) {
  const Dims d = h->d, dh = half(d);
  const float* th = h->theta;
  auto W = [&](int l) { return th + h->L[l].woff; };
  dw(h, d, 4, h->gc, h->h2, NC0, nullptr, 0, 0, s);
  conv(h, d, NC0, h->gc, 1, nullptr, 0, 0, W(4), nullptr, 1, 0, h->h2, h->dh2, s);
  dw(h, d, 3, h->dh2, h->h0, NC0, h->h1b, NC1, 1, s);
  conv(h, d, NC0 + NC1, h->dh2, NC0, nullptr, 0, 0, W(3), nullptr, 1, 0, nullptr, h->dcat, s);
  hipLaunchKernelGGL(k_up2_T, dim3((NC1 * h->Nh + 255) / 256), dim3(256), 0, s, d, NC1,
                     h->dcat + (size_t)NC0 * h->N, h->h1b, h->dh1b);
  dw(h, dh, 2, h->dh1b, h->h1, NC1, nullptr, 0, 0, s);
  conv(h, dh, NC1, h->dh1b, NC1, nullptr, 0, 0, W(2), nullptr, 1, 0, h->h1, h->dh1, s);
  dw(h, dh, 1, h->dh1, h->d1, NC0, nullptr, 0, 0, s);
  conv(h, dh, NC0, h->dh1, NC1, nullptr, 0, 0, W(1), nullptr, 1, 0, nullptr, h->dd1, s);
  hipLaunchKernelGGL(k_mix_dh0, dim3((NC0 * h->N + 255) / 256), dim3(256), 0, s, d, NC0,
                     h->dcat, h->dd1, h->h0, h->dh0);
  dw(h, d, 0, h->dh0, r, 1, q, 1, 0, s);
  FW_CHECK_LAUNCH("net backward");
  return FW_OK;
}

// For the projection's learned preconditioner: N(r) of the attached net;
// *c points at the net's output buffer (valid until its next forward).
int net_apply(void* net, const float* r, const float* q, const float** c, int* n,
              hipStream_t s) {
  NetHandle* h = (NetHandle*)net;
  *c = h->c;
  *n = h->N;
  return net_forward(h, r, q, s);
}

int net_dims(void* net, int* nx, int* ny, int* nz) {
  NetHandle* h = (NetHandle*)net;
  *nx = h->d.nx;
  *ny = h->d.ny;
  *nz = h->d.nz;
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
  FW_REQUIRE(handle && theta_host, "net_create: null argument");
  FW_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0,
             "net_create: grid dims must be even and >= 2 (got %d x %d x %d)", nx, ny, nz);
  FW_REQUIRE((size_t)nx * ny * nz * (NC0 + NC1) * 4 <= 0xFFFFFFFFull,
             "net_create: activations exceed 4 GiB");
  NetHandle* h = new NetHandle();
  h->d = Dims{nx, ny, nz};
  h->N = nx * ny * nz;
  h->Nh = h->N / 8;
  layer_defs(h);
  h->nchunk = 64;
  const size_t N = h->N, Nh = h->Nh;
  float** vec[] = {&h->theta, &h->grad, &h->m, &h->v};
  for (float** p : vec) FW_CHECK_HIP(hipMalloc((void**)p, h->nparam * sizeof(float)));
  FW_CHECK_HIP(hipMemcpy(h->theta, theta_host, h->nparam * sizeof(float),
                         hipMemcpyHostToDevice));
  FW_CHECK_HIP(hipMemset(h->m, 0, h->nparam * sizeof(float)));
  FW_CHECK_HIP(hipMemset(h->v, 0, h->nparam * sizeof(float)));
  struct {
    float** p;
    size_t n;
  } acts[] = {{&h->h0, NC0 * N},   {&h->d1, NC0 * Nh},  {&h->h1, NC1 * Nh},
              {&h->h1b, NC1 * Nh}, {&h->h2, NC0 * N},   {&h->c, N},
              {&h->gc, N},         {&h->dh2, NC0 * N},  {&h->dcat, (NC0 + NC1) * N},
              {&h->dh1b, NC1 * Nh}, {&h->dh1, NC1 * Nh}, {&h->dd1, NC0 * Nh},
              {&h->dh0, NC0 * N}};
  for (auto& a : acts) FW_CHECK_HIP(hipMalloc((void**)a.p, a.n * sizeof(float)));
  FW_CHECK_HIP(hipMalloc((void**)&h->part, ((N + 255) / 256) * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->dwpart,
                         (size_t)(NC0 + NC1) * NC0 * h->nchunk * 28 * sizeof(double) +
                             (size_t)NC1 * NC1 * h->nchunk * 28 * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->scal, 4 * sizeof(double)));
  *handle = h;
  return FW_OK;
}

int fw_net_destroy(void* handle) {
  NetHandle* h = (NetHandle*)handle;
  if (!h) return FW_OK;
  void* bufs[] = {h->theta, h->grad, h->m,   h->v,    h->h0,   h->d1,  h->h1,
                  h->h1b,   h->h2,   h->c,   h->gc,   h->dh2,  h->dcat, h->dh1b,
                  h->dh1,   h->dd1,  h->dh0, h->part, h->dwpart, h->scal};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete h;
  return FW_OK;
}

int fw_net_get_params(void* handle, float* theta_host) {
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
  FW_CHECK_HIP(hipMemcpyAsync(c_out, h->c, (size_t)h->N * 4, hipMemcpyDeviceToDevice, s));
  return FW_OK;
}

int fw_net_set_params(void* handle, const float* theta_host) {
  NetHandle* h = (NetHandle*)handle;
  FW_REQUIRE(h && theta_host, "net_set_params: null argument");
  FW_CHECK_HIP(hipMemcpy(h->theta, theta_host, h->nparam * sizeof(float),
                         hipMemcpyHostToDevice));
  FW_CHECK_HIP(hipMemset(h->m, 0, h->nparam * sizeof(float)));
  FW_CHECK_HIP(hipMemset(h->v, 0, h->nparam * sizeof(float)));
  return FW_OK;
}

int fw_net_train_step(void* handle, const void* r, const void* q, const void* t,
                      const double* dz, double lr, int it, double* loss, void* stream) {
  NetHandle* h = (NetHandle*)handle;
  FW_REQUIRE(h && r && q && t && dz && loss, "net_train_step: null argument");
  FW_REQUIRE(it >= 1, "net_train_step: Adam step number must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  int rc = net_forward(h, (const float*)r, (const float*)q, s);
  if (rc) return rc;
  const int nb = (h->N + 255) / 256;
  hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, s, h->d, (const float*)h->c,
                     (const float*)t, dz, (const double*)h->scal, 1, h->gc, h->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s, nb, (const double*)h->part,
                     h->scal);
  hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, s, h->d, (const float*)h->c,
                     (const float*)t, dz, (const double*)h->scal, 0, h->gc, h->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s, nb, (const double*)h->part,
                     h->scal + 1);
  if ((rc = net_backward(h, (const float*)r, (const float*)q, s))) return rc;
  if (lr > 0.0)
    hipLaunchKernelGGL(k_adam, dim3((h->nparam + 255) / 256), dim3(256), 0, s, h->nparam,
                       h->theta, (const float*)h->grad, h->m, h->v, (float)lr, 0.9f, 0.999f,
                       1.0e-8f, it);
  FW_CHECK_LAUNCH("net train step");
  double sc[2];
  FW_CHECK_HIP(hipMemcpyAsync(sc, h->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  FW_CHECK_HIP(hipStreamSynchronize(s));
  *loss = sc[0] > 0.0 ? sc[1] / sc[0] : 0.0;
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
