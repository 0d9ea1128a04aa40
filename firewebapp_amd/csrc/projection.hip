// Pressure projection (projection_3d.py:65-175, :677-892 of the reference),
// fft_pcg path, re-designed for the GPU:
//
// * the variable-density 7-point operator is applied matrix-free: the face
//   coefficients 0.5*(1/rho_a + 1/rho_b) are recomputed from rho in
//   registers, so a matvec streams rho and x once instead of 7 coefficient
//   arrays;
// * the separable constant-coefficient preconditioner (fft_poisson_3d.py)
//   is three dense 1-D transforms: x (DCT-IV eigenbasis, Nx x Nx), y (real
//   periodic eigenbasis, Ny x Ny; identical operator to the reference's
//   FFT), z (eigenbasis of the sqrt(dz)-symmetrised L_z with sqrt(dz) folded
//   into the matrices), each a batched LDS-tiled GEMM; the spectral divide
//   is fused into the epilogue of the forward z transform;
// * the Krylov method is PCG on the symmetrised system -W A p = -W b with
//   W = diag(dz) (W A is symmetric because d_above[k] = d_below[k+1]); with
//   the preconditioner above every W-inner product reduces to
//   sum dz_k a b over the original residual;
// * mixed precision: the solution, the right-hand side and the refinement
//   residual b - A x are float64; the inner PCG runs in the grid's storage
//   type; all dot products accumulate in float64 with a fixed-order
//   (deterministic) two-level reduction.
#include <string.h>
#include <functional>
#include <vector>
#include "fw_common.h"
#include "graph.h"

namespace fw {

#define PJ_RED 512  // partial-sum workgroups of the projection reductions
#define PJ_MV_MAX FW_PROJ_MV_MAX  // most workgroups of the fused PCG operator launch
#define PJ_MV_DEFAULT 1024        // its default (two rows or more per workgroup at 256x64x48)

// S_DONE: 0 while the inner PCG runs, 1 once ||r|| <= S_TOL, 2 on a NaN
// residual, 3 when k_round_init found the solve converged, 4 on a NaN
// ||b||; S_IT counts the iterations taken.  The iteration kernels leave
// the Krylov state untouched once S_DONE is set, so iterations can be
// enqueued in batches and the host reads S_DONE once per batch.
enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_TMP, S_DONE, S_IT, S_TOL, S_SIG, S_DMAX,
       S_ITL, S_SPEC, S_NS };
// what a final sum is stored as (k_finalize)
enum { F_STORE = 0, F_ALPHA = 1, F_BETA = 2, F_RR = 3 };

// ---------------------------------------------------------------------------
// operator
// ---------------------------------------------------------------------------
template <typename C, typename TR>
__device__ __forceinline__ C inv_rho_at(const TR* rho, size_t c) {
  C r = C(rho[c]);
  return C(1) / (r > C(0.01) ? r : C(0.01));
}

// y = A x at cell (k, j, i), computed in precision C from rho (type TR).
template <typename C, typename TX, typename TR>
__device__ __forceinline__ C apply_A(const GridArgs& g, const TR* __restrict__ rho,
                                     const TX* __restrict__ x, int k, int j,
                                     int i, size_t c, double eps_reg) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  const C irc = inv_rho_at<C>(rho, c);
  const C xc = C(x[c]);
  const C dx2 = C(g.dx) * C(g.dx), dy2 = C(g.dy) * C(g.dy);
  C diag = C(eps_reg);
  C acc = C(0);
  if (Nx > 1) {
    if (i < Nx - 1) {
      C cxp = (C(0.5) * (irc + inv_rho_at<C>(rho, c + 1))) / dx2;
      diag -= cxp;
      acc += cxp * C(x[c + 1]);
    } else {
      diag -= C(2) * irc / dx2;
    }
    if (i > 0) {
      C cxm = (C(0.5) * (inv_rho_at<C>(rho, c - 1) + irc)) / dx2;
      diag -= cxm;
      acc += cxm * C(x[c - 1]);
    }
  }
  if (Ny > 1) {
    const size_t cp = c + (size_t)(wrap(j + 1, Ny) - j) * sy;
    const size_t cm = c + (size_t)(wrap(j - 1, Ny) - j) * sy;
    C cyp = (C(0.5) * (irc + inv_rho_at<C>(rho, cp))) / dy2;
    C cym = (C(0.5) * (inv_rho_at<C>(rho, cm) + irc)) / dy2;
    diag -= cyp + cym;
    acc += cyp * C(x[cp]) + cym * C(x[cm]);
  }
  if (Nz > 1) {
    const C inv_dz = C(g.inv_dz[k]);
    if (k < Nz - 1) {
      C czp = (C(0.5) * (irc + inv_rho_at<C>(rho, c + sz))) * inv_dz / C(g.d_above[k]);
      diag -= czp;
      acc += czp * C(x[c + sz]);
    } else {
      diag -= C(2) * irc * inv_dz * inv_dz;
    }
    if (k > 0) {
      C czm = (C(0.5) * (inv_rho_at<C>(rho, c - sz) + irc)) * inv_dz / C(g.d_below[k]);
      diag -= czm;
      acc += czm * C(x[c - sz]);
    }
  }
  return diag * xc + acc;
}

// Block reduction of one double over 256 threads into partial[blockIdx.x].
__device__ __forceinline__ void block_sum_store(double v, double* sh, double* out) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = sh[0];
}

__device__ __forceinline__ void block_max_store(double v, double* sh, double* out) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s && sh[t + s] > sh[t]) sh[t] = sh[t + s];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = sh[0];
}

// Sum of v over the workgroup (64, 128 or 256 threads), in a fixed order: a
// shuffle tree inside each wavefront, then the wavefront sums in order.
// Valid on every thread; sh holds one double per wavefront.
__device__ __forceinline__ double block_sum_wg(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double a = sh[0];
  for (int q = 1; q < nw; ++q) a += sh[q];
  return a;
}

// Rows (k, j) of the grid for the PJ_RED workgroups of a row-blocked
// reduction kernel: contiguous runs of rows, with the same XCD-aware order as
// cell_ijk (each XCD gets a contiguous range of runs, so the y/z neighbour
// rows of a run mostly sit in the same L2).
__device__ __forceinline__ void row_range(int rows, int& r0, int& r1) {
  const unsigned nb = gridDim.x;
  const unsigned q = nb >> 3, rr = nb & 7u;
  const unsigned xcd = blockIdx.x & 7u, pos = blockIdx.x >> 3;
  const unsigned m = xcd * q + (xcd < rr ? xcd : rr) + pos;
  const int R = (rows + (int)nb - 1) / (int)nb;
  r0 = (int)m * R;
  r1 = r0 + R < rows ? r0 + R : rows;
}

// Fixed-order sum of np (<= a few thousand) partials, computed identically
// by every workgroup that calls it (so all of them agree bit for bit).
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int np,
                                               double* sh) {
  double a = 0.0;
  for (int b = threadIdx.x; b < np; b += blockDim.x) a += part[b];
  return block_sum_wg(a, sh);
}

// Two such sums at once (their loads in flight together, one barrier), each
// in the same order as sum_partials; sh holds two doubles per wavefront.
__device__ __forceinline__ void sum_partials2(const double* __restrict__ p1, int n1,
                                              const double* __restrict__ p2, int n2,
                                              double* sh, double& s1, double& s2) {
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n1; i += blockDim.x) a += p1[i];
  for (int i = threadIdx.x; i < n2; i += blockDim.x) b += p2[i];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = a;
    sh[nw + (threadIdx.x >> 6)] = b;
  }
  __syncthreads();
  double x = sh[0], y = sh[nw];
  for (int q = 1; q < nw; ++q) {
    x += sh[q];
    y += sh[nw + q];
  }
  s1 = x;
  s2 = y;
}

// Orders a wavefront's LDS accesses before and after it, without a
// workgroup barrier: for data that only this wavefront writes and reads (a
// wavefront's LDS operations complete in issue order).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The inner (Krylov) operator: A x at cell (i, j, k) in precision T from a
// precomputed 1 / max(rho, 0.01) field, with the metric reciprocals
// multiplied in.  It equals apply_A up to rounding; the refinement's float64
// residual uses apply_A itself, so the solution is unchanged.
// The operator on values that x(offset) supplies (a stored vector, or one
// formed on the fly).
template <typename T, typename X>
__device__ __forceinline__ T apply_A_irho_f(const GridArgs& g, const T* __restrict__ ir, X x,
                                            int i, int k, const Nbr& n, T eps_reg) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const T irc = at(ir, n.c), xc = x(n.c);
  const T idx2 = T(g.inv_dx2), idy2 = T(g.inv_dy2);
  T diag = eps_reg;
  T acc = T(0);
  if (Nx > 1) {
    if (i < Nx - 1) {
      const T cxp = T(0.5) * (irc + at(ir, n.xp1)) * idx2;
      diag -= cxp;
      acc += cxp * x(n.xp1);
    } else {
      diag -= T(2) * irc * idx2;
    }
    if (i > 0) {
      const T cxm = T(0.5) * (at(ir, n.xm1) + irc) * idx2;
      diag -= cxm;
      acc += cxm * x(n.xm1);
    }
  }
  if (Ny > 1) {
    const T cyp = T(0.5) * (irc + at(ir, n.yp1)) * idy2;
    const T cym = T(0.5) * (at(ir, n.ym1) + irc) * idy2;
    diag -= cyp + cym;
    acc += cyp * x(n.yp1) + cym * x(n.ym1);
  }
  if (Nz > 1) {
    const T idz = T(g.inv_dz[k]);
    if (k < Nz - 1) {
      const T czp = T(0.5) * (irc + at(ir, n.zp1)) * (idz * T(g.inv_da[k]));
      diag -= czp;
      acc += czp * x(n.zp1);
    } else {
      diag -= T(2) * irc * idz * idz;
    }
    if (k > 0) {
      const T czm = T(0.5) * (at(ir, n.zm1) + irc) * (idz * T(g.inv_db[k]));
      diag -= czm;
      acc += czm * x(n.zm1);
    }
  }
  return diag * xc + acc;
}

// The operator on a stored vector x.
template <typename T>
__device__ __forceinline__ T apply_A_irho(const GridArgs& g, const T* __restrict__ ir,
                                          const T* __restrict__ x, int i, int k,
                                          const Nbr& n, T eps_reg) {
  return apply_A_irho_f<T>(g, ir, [x](uint32_t o) { return at(x, o); }, i, k, n, eps_reg);
}

// The operator's seven points of one field, read ahead of their use.
template <typename T>
struct Pts7 {
  T c, xm, xp, ym, yp, zm, zp;
};

// Byte offsets of cell (i, j, k)'s stencil points with the missing x and z
// neighbours of the boundary cells replaced by the cell itself (always a
// valid address; apply_A_pts never uses those values).
template <typename T>
__device__ __forceinline__ Nbr nbr_offsets_in(const GridArgs& g, int i, int j, int k) {
  Nbr n = nbr_offsets<T>(g, i, j, k);
  if (i == 0) n.xm1 = n.c;
  if (i == g.nx - 1) n.xp1 = n.c;
  if (k == 0) n.zm1 = n.c;
  if (k == g.nz - 1) n.zp1 = n.c;
  return n;
}

template <typename T>
__device__ __forceinline__ Pts7<T> load_pts(const T* __restrict__ f, const Nbr& n) {
  Pts7<T> v;
  v.c = at(f, n.c);
  v.xm = at(f, n.xm1);
  v.xp = at(f, n.xp1);
  v.ym = at(f, n.ym1);
  v.yp = at(f, n.yp1);
  v.zm = at(f, n.zm1);
  v.zp = at(f, n.zp1);
  return v;
}

// apply_A_irho_f on preloaded point values: the same operations in the same
// order.
template <typename T>
__device__ __forceinline__ T apply_A_pts(const GridArgs& g, const Pts7<T>& ir,
                                         const Pts7<T>& x, int i, int k, T eps_reg) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const T irc = ir.c, xc = x.c;
  const T idx2 = T(g.inv_dx2), idy2 = T(g.inv_dy2);
  T diag = eps_reg;
  T acc = T(0);
  if (Nx > 1) {
    if (i < Nx - 1) {
      const T cxp = T(0.5) * (irc + ir.xp) * idx2;
      diag -= cxp;
      acc += cxp * x.xp;
    } else {
      diag -= T(2) * irc * idx2;
    }
    if (i > 0) {
      const T cxm = T(0.5) * (ir.xm + irc) * idx2;
      diag -= cxm;
      acc += cxm * x.xm;
    }
  }
  if (Ny > 1) {
    const T cyp = T(0.5) * (irc + ir.yp) * idy2;
    const T cym = T(0.5) * (ir.ym + irc) * idy2;
    diag -= cyp + cym;
    acc += cyp * x.yp + cym * x.ym;
  }
  if (Nz > 1) {
    const T idz = T(g.inv_dz[k]);
    if (k < Nz - 1) {
      const T czp = T(0.5) * (irc + ir.zp) * (idz * T(g.inv_da[k]));
      diag -= czp;
      acc += czp * x.zp;
    } else {
      diag -= T(2) * irc * idz * idz;
    }
    if (k > 0) {
      const T czm = T(0.5) * (ir.zm + irc) * (idz * T(g.inv_db[k]));
      diag -= czm;
      acc += czm * x.zm;
    }
  }
  return diag * xc + acc;
}

struct Cell {
  int i, j, k;
};
__device__ __forceinline__ Cell cell_of(const GridArgs& g, size_t c) {
  Cell r;
  r.i = (int)(c % g.nx);
  size_t t = c / g.nx;
  r.j = (int)(t % g.ny);
  r.k = (int)(t / g.ny);
  return r;
}

// q = A p (precision T), partial sum of dz_k * p * q
template <typename T, typename TR>
__global__ void __launch_bounds__(256)
k_matvec_pq(GridArgs g, size_t n, const TR* __restrict__ rho,
            const T* __restrict__ p, T* __restrict__ q, double eps_reg,
            double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n;
       c += (size_t)gridDim.x * 256) {
    Cell e = cell_of(g, c);
    T y = apply_A<T, T, TR>(g, rho, p, e.k, e.j, e.i, c, eps_reg);
    q[c] = y;
    acc += g.dz[e.k] * (double)p[c] * (double)y;
  }
  block_sum_store(acc, sh, part);
}

// ir = 1 / max(rho, 0.01) in the Krylov precision (once per solve)
template <typename T, typename TR>
__global__ void __launch_bounds__(256)
k_irho(size_t n, const TR* __restrict__ rho, T* __restrict__ ir) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c < n) ir[c] = inv_rho_at<T>(rho, c);
}

// z = r / diag(A) (Jacobi), diag from the same 1/rho field as the operator
template <typename T>
__global__ void __launch_bounds__(256)
k_jacobi(GridArgs g, const T* __restrict__ ir, const T* __restrict__ r, T* __restrict__ z,
         T eps_reg) {
  int i, j, k;
  cell_ijk(i, j, k);
  if (i >= g.nx) return;
  const Nbr n = nbr_offsets<T>(g, i, j, k);
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const T irc = at(ir, n.c);
  const T idx2 = T(g.inv_dx2), idy2 = T(g.inv_dy2);
  T diag = eps_reg;
  if (Nx > 1) {
    diag -= i < Nx - 1 ? T(0.5) * (irc + at(ir, n.xp1)) * idx2 : T(2) * irc * idx2;
    if (i > 0) diag -= T(0.5) * (at(ir, n.xm1) + irc) * idx2;
  }
  if (Ny > 1)
    diag -= T(0.5) * (irc + at(ir, n.yp1)) * idy2 + T(0.5) * (at(ir, n.ym1) + irc) * idy2;
  if (Nz > 1) {
    const T idz = T(g.inv_dz[k]);
    diag -= k < Nz - 1 ? T(0.5) * (irc + at(ir, n.zp1)) * (idz * T(g.inv_da[k]))
                       : T(2) * irc * idz * idz;
    if (k > 0) diag -= T(0.5) * (at(ir, n.zm1) + irc) * (idz * T(g.inv_db[k]));
  }
  at(z, n.c) = at(r, n.c) / diag;
}

// The inner PCG iteration.  Each reduction leaves PJ_RED partials; the next
// kernel that needs the total sums them itself (sum_partials, the same fixed
// order in every workgroup), so no final-sum launch and no cross-workgroup
// synchronisation is needed:
//   k_matvec_rows   q = A p                          -> Ppq
//   k_update_xr2    alpha = sum Prz / sum Ppq; d += alpha p; r -= alpha q
//                                                    -> Prr
//   (preconditioner z = M^-1 r)
//   k_dot_rows      convergence test on sum Prr (workgroup 0); <r, z>_W
//                                                    -> Prz
//   k_p_update2     beta = sum Prz / rz; p = z + beta p
// S_DONE is written only by workgroup 0 of k_dot_rows, which is also the
// only workgroup of that launch to read it.

// q = A p over the workgroup's rows; partial <p, q>_W
template <typename T>
__global__ void __launch_bounds__(256)
k_matvec_rows(GridArgs g, const T* __restrict__ ir, const T* __restrict__ p,
              T* __restrict__ q, T eps_reg, double* __restrict__ Ppq) {
  __shared__ double sh[4];
  int r0, r1;
  row_range(g.ny * g.nz, r0, r1);
  double acc = 0.0;
  for (int row = r0; row < r1; ++row) {
    const int k = row / g.ny, j = row - k * g.ny;
    const double dzk = g.dz[k];
    for (int i = threadIdx.x; i < g.nx; i += 256) {
      const Nbr nb = nbr_offsets<T>(g, i, j, k);
      const T y = apply_A_irho<T>(g, ir, p, i, k, nb, eps_reg);
      at(q, nb.c) = y;
      acc += dzk * (double)at(p, nb.c) * (double)y;
    }
  }
  const double t = block_sum_wg(acc, sh);
  if (threadIdx.x == 0) Ppq[blockIdx.x] = t;
}

// The direction update of the previous iteration fused into this
// iteration's operator: p+ = z + beta p, beta = sum Prz / rz (the new
// <r, z>_W partials over the rz the previous update stored, exactly as
// k_p_update2 forms it), is formed at every stencil point from z and p;
// q = A p+, p+ of the workgroup's rows -> pn (another buffer: neighbours
// still read p), partial <p+, q>_W.  Returns at once when S_DONE is set
// (pn is then stale; the next round's set-up rewrites the direction).
template <typename T>
__global__ void __launch_bounds__(256)
k_matvec_pz(GridArgs g, const T* __restrict__ ir, const T* __restrict__ p,
            const T* __restrict__ z, T* __restrict__ pn, T* __restrict__ q, T eps_reg,
            const double* __restrict__ Prz, int npz, double* __restrict__ S,
            double* __restrict__ Ppq) {
  __shared__ double sh[4];
  if (S[S_DONE] != 0.0) return;
  int r0, r1;
  row_range(g.ny * g.nz, r0, r1);
  // items: (row, cell chunk) in row order, one cell per thread each; the
  // first MVI items' stencil points of ir, z and p are loaded before the
  // sum of the <r, z> partials (beta), so one memory latency covers both
  const int cpr = (g.nx + 255) >> 8;
  const int nit = (r1 - r0) * cpr;
  constexpr int MVI = sizeof(T) == 4 ? 4 : 2;
  Pts7<T> vi[MVI], vz[MVI], vp[MVI];
#pragma unroll
  for (int it = 0; it < MVI; ++it) {
    if (it < nit) {
      const int row = r0 + it / cpr, i = threadIdx.x + 256 * (it % cpr);
      if (i < g.nx) {
        const int k = row / g.ny, j = row - k * g.ny;
        const Nbr nb = nbr_offsets_in<T>(g, i, j, k);
        vi[it] = load_pts(ir, nb);
        vz[it] = load_pts(z, nb);
        vp[it] = load_pts(p, nb);
      }
    }
  }
  const double bd = sum_partials(Prz, npz, sh) / S[S_RZ];
  __syncthreads();   // sh is reused by the block sum below
  if (blockIdx.x == 0 && threadIdx.x == 0) S[S_BETA] = bd;
  const T beta = (T)bd;
  // p+ = z + beta p at one point
  auto pz = [beta](T zv, T pv) { return zv + beta * pv; };
  double acc = 0.0;
  auto item = [&](int it, const Pts7<T>& ri, const Pts7<T>& zz, const Pts7<T>& pp) {
    const int row = r0 + it / cpr, i = threadIdx.x + 256 * (it % cpr);
    if (i >= g.nx) return;
    const int k = row / g.ny, j = row - k * g.ny;
    const uint32_t c = (((uint32_t)k * (uint32_t)g.ny + (uint32_t)j) * (uint32_t)g.nx +
                        (uint32_t)i) * (uint32_t)sizeof(T);
    Pts7<T> x;
    x.c = pz(zz.c, pp.c);
    x.xm = pz(zz.xm, pp.xm);
    x.xp = pz(zz.xp, pp.xp);
    x.ym = pz(zz.ym, pp.ym);
    x.yp = pz(zz.yp, pp.yp);
    x.zm = pz(zz.zm, pp.zm);
    x.zp = pz(zz.zp, pp.zp);
    const T y = apply_A_pts<T>(g, ri, x, i, k, eps_reg);
    at(q, c) = y;
    at(pn, c) = x.c;
    acc += g.dz[k] * (double)x.c * (double)y;
  };
#pragma unroll
  for (int it = 0; it < MVI; ++it)
    if (it < nit) item(it, vi[it], vz[it], vp[it]);
  for (int it = MVI; it < nit; ++it) {
    const int row = r0 + it / cpr, i = threadIdx.x + 256 * (it % cpr);
    if (i >= g.nx) continue;
    const int k = row / g.ny, j = row - k * g.ny;
    const Nbr nb = nbr_offsets_in<T>(g, i, j, k);
    item(it, load_pts(ir, nb), load_pts(z, nb), load_pts(p, nb));
  }
  const double t = block_sum_wg(acc, sh);
  if (threadIdx.x == 0) Ppq[blockIdx.x] = t;
}

// alpha = rz / <p, q>_W; d += alpha p ; r -= alpha q ; partial ||r||^2.
// Nothing moves once S_DONE is set.
template <typename T>
__global__ void __launch_bounds__(256)
k_update_xr2(size_t n, T* __restrict__ d, T* __restrict__ r, const T* __restrict__ p,
             const T* __restrict__ q, const double* __restrict__ Prz, int npz,
             const double* __restrict__ Ppq, double* __restrict__ S,
             double* __restrict__ Prr) {
  __shared__ double sh[4];
  if (S[S_DONE] != 0.0) return;
  const double rz = sum_partials(Prz, npz, sh);
  __syncthreads();
  const double pq = sum_partials(Ppq, PJ_RED, sh);
  const double alpha_d = rz / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S[S_RZ] = rz;
    S[S_PQ] = pq;
    S[S_ALPHA] = alpha_d;
  }
  const T alpha = (T)alpha_d;
  double acc = 0.0;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n;
       c += (size_t)gridDim.x * 256) {
    d[c] += alpha * p[c];
    const T rv = r[c] - alpha * q[c];
    r[c] = rv;
    acc += (double)rv * (double)rv;
  }
  __syncthreads();
  const double t = block_sum_wg(acc, sh);
  if (threadIdx.x == 0) Prr[blockIdx.x] = t;
}

// <a, b>_W over the workgroup's rows -> Prz.  check != 0: workgroup 0 first
// applies the convergence test of the iteration to sum Prr (stores ||r||^2,
// counts the iteration, raises S_DONE; 2 on a NaN).
template <typename T>
__global__ void __launch_bounds__(256)
k_dot_rows(GridArgs g, const T* __restrict__ a, const T* __restrict__ b,
           double* __restrict__ Prz, const double* __restrict__ Prr,
           double* __restrict__ S, int check) {
  __shared__ double sh[4];
  if (check && blockIdx.x == 0) {
    const double rr = sum_partials(Prr, PJ_RED, sh);
    if (threadIdx.x == 0 && S[S_DONE] == 0.0) {
      S[S_RR] = rr;
      S[S_IT] += 1.0;
      if (!(rr == rr)) S[S_DONE] = 2.0;
      else if (sqrt(rr) <= S[S_TOL]) S[S_DONE] = 1.0;
    }
    __syncthreads();
  }
  int r0, r1;
  row_range(g.ny * g.nz, r0, r1);
  double acc = 0.0;
  for (int row = r0; row < r1; ++row) {
    const int k = row / g.ny;
    const double dzk = g.dz[k];
    const size_t base = (size_t)row * g.nx;
    for (int i = threadIdx.x; i < g.nx; i += 256)
      acc += dzk * (double)a[base + i] * (double)b[base + i];
  }
  const double t = block_sum_wg(acc, sh);
  if (threadIdx.x == 0) Prz[blockIdx.x] = t;
}

// p = z + beta p with beta = sum Prz / rz (first: p = z).  Nothing moves
// once S_DONE is set.
template <typename T>
__global__ void __launch_bounds__(256)
k_p_update2(size_t n, T* __restrict__ p, const T* __restrict__ z,
            const double* __restrict__ Prz, int npz, double* __restrict__ S, int first) {
  __shared__ double sh[4];
  if (S[S_DONE] != 0.0) return;
  T beta = T(0);
  if (!first) {
    const double rz_new = sum_partials(Prz, npz, sh);
    const double bd = rz_new / S[S_RZ];
    if (blockIdx.x == 0 && threadIdx.x == 0) S[S_BETA] = bd;
    beta = (T)bd;
  }
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n;
       c += (size_t)gridDim.x * 256)
    p[c] = first ? z[c] : z[c] + beta * p[c];
}

// Learned preconditioner (flexible PCG).  The first direction of each inner
// round is z = M_sep r + s N(r / s, q), s the RMS of r (the network is
// trained on the rounds' initial residuals, the projection right-hand
// sides); later directions use M_sep.  The preconditioner therefore varies
// (and the network is nonlinear), hence the flexible (Polak-Ribiere) beta:
//   beta = <z+, r+ - r>_W / <z, r>_W = -alpha <z+, q>_W / rz,
// with q = A p of the iteration (r+ = r - alpha q).

// partial sums of r^2 (unweighted) over PJ_RED workgroups
__global__ void __launch_bounds__(256)
k_rr_part(size_t n, const float* __restrict__ r, double* __restrict__ P) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256)
    acc += (double)r[c] * (double)r[c];
  const double t = block_sum_wg(acc, sh);
  if (threadIdx.x == 0) P[blockIdx.x] = t;
}

// rs = r / s with s = sqrt(sum P / n) (workgroup 0 stores s in S[S_SIG])
__global__ void __launch_bounds__(256)
k_scale_rms(size_t n, const float* __restrict__ r, const double* __restrict__ P, int np,
            double* __restrict__ S, float* __restrict__ rs) {
  __shared__ double sh[4];
  const double rr = sum_partials(P, np, sh);
  const double sig = rr > 0.0 ? sqrt(rr / (double)n) : 1.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) S[S_SIG] = sig;
  const float is = (float)(1.0 / sig);
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256)
    rs[c] = r[c] * is;
}

// z += s c
__global__ void __launch_bounds__(256)
k_add_scaled(size_t n, const float* __restrict__ c, const double* __restrict__ S,
             float* __restrict__ z) {
  const float sg = (float)S[S_SIG];
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256)
    z[e] += sg * c[e];
}

// <r, z>_W -> Prz and <q, z>_W -> Pqz over the workgroup's rows; check as
// in k_dot_rows
template <typename T>
__global__ void __launch_bounds__(256)
k_dot2_rows(GridArgs g, const T* __restrict__ r, const T* __restrict__ q,
            const T* __restrict__ z, double* __restrict__ Prz, double* __restrict__ Pqz,
            const double* __restrict__ Prr, double* __restrict__ S, int check) {
  __shared__ double sh[4];
  if (check && blockIdx.x == 0) {
    const double rr = sum_partials(Prr, PJ_RED, sh);
    if (threadIdx.x == 0 && S[S_DONE] == 0.0) {
      S[S_RR] = rr;
      S[S_IT] += 1.0;
      if (!(rr == rr)) S[S_DONE] = 2.0;
      else if (sqrt(rr) <= S[S_TOL]) S[S_DONE] = 1.0;
    }
    __syncthreads();
  }
  int r0, r1;
  row_range(g.ny * g.nz, r0, r1);
  double a1 = 0.0, a2 = 0.0;
  for (int row = r0; row < r1; ++row) {
    const double dzk = g.dz[row / g.ny];
    const size_t base = (size_t)row * g.nx;
    for (int i = threadIdx.x; i < g.nx; i += 256) {
      const double zv = (double)z[base + i];
      a1 += dzk * (double)r[base + i] * zv;
      a2 += dzk * (double)q[base + i] * zv;
    }
  }
  const double t1 = block_sum_wg(a1, sh);
  __syncthreads();
  const double t2 = block_sum_wg(a2, sh);
  if (threadIdx.x == 0) {
    Prz[blockIdx.x] = t1;
    Pqz[blockIdx.x] = t2;
  }
}

// p = z + beta p with the flexible beta = -alpha sum Pqz / rz
template <typename T>
__global__ void __launch_bounds__(256)
k_p_update_flex(size_t n, T* __restrict__ p, const T* __restrict__ z,
                const double* __restrict__ Pqz, double* __restrict__ S) {
  __shared__ double sh[4];
  if (S[S_DONE] != 0.0) return;
  const double qz = sum_partials(Pqz, PJ_RED, sh);
  const double bd = -S[S_ALPHA] * qz / S[S_RZ];
  if (blockIdx.x == 0 && threadIdx.x == 0) S[S_BETA] = bd;
  const T beta = (T)bd;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256)
    p[c] = z[c] + beta * p[c];
}

// r = b64 - A64 x64 (float64 operator from rho), r_t = (T) r, partial sum
// of r^2 (the float64 residual itself is not kept: only its norm and its
// rounding to T are used)
template <typename T, typename TR>
__global__ void __launch_bounds__(256)
k_residual64(GridArgs g, size_t n, const TR* __restrict__ rho,
             const double* __restrict__ b, const double* __restrict__ x,
             T* __restrict__ r_t, double eps_reg, double* __restrict__ part,
             const double* __restrict__ gate) {
  __shared__ double sh[256];
  (void)n;
  if (gate && gate[S_DONE] != 1.0) return;   // k_finalize_round's round-end gate
  // contiguous runs of rows per workgroup in the XCD-aware order, so the
  // y and z neighbour rows of the stencil mostly come from the same L2
  int r0, r1;
  row_range(g.ny * g.nz, r0, r1);
  double acc = 0.0;
  for (int row = r0; row < r1; ++row) {
    const int k = row / g.ny, j = row - k * g.ny;
    for (int i = threadIdx.x; i < g.nx; i += 256) {
      const size_t c = ((size_t)k * g.ny + j) * g.nx + i;
      const double ax = apply_A<double, double, TR>(g, rho, x, k, j, i, c, eps_reg);
      const double r = b[c] - ax;
      r_t[c] = (T)r;
      acc += r * r;
    }
  }
  block_sum_store(acc, sh, part);
}

// partial sum of a^2 (unweighted)
__global__ void __launch_bounds__(256)
k_sumsq(size_t n, const double* __restrict__ a, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n;
       c += (size_t)gridDim.x * 256)
    acc += a[c] * a[c];
  block_sum_store(acc, sh, part);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_x64_add(size_t n, double* __restrict__ x, const T* __restrict__ d,
          const double* __restrict__ gate) {
  if (gate && gate[S_DONE] != 1.0) return;   // k_finalize_round's round-end gate
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c < n) x[c] += (double)d[c];
}

// The solve's scalars say it has converged with b != 0: the correction and
// the divergence check issued ahead of the host's read of that decision run
// only then (with b = 0 the host zeroes the pressure first).
__device__ __forceinline__ bool solve_done(const double* __restrict__ S) {
  return S[S_DONE] == 3.0 && S[S_BB] != 0.0;
}

// Set-up of a refinement round on the device, from ||b||^2 (S_BB) and
// ||b - A x||^2 (S_TMP) as the last finalize left them: S_DONE = 3 when the
// solve has converged (||r|| <= rtol ||b||, or b = 0), 2 on a NaN norm,
// else 0 with the inner tolerance S_TOL = max(rtol ||b|| / ||r||, floor) ||r||
// (the host's former arithmetic, in its order).  The iteration kernels do
// nothing while S_DONE != 0, so a round launched ahead of this decision is
// harmless when the solve has already converged.
__device__ __forceinline__ void round_init(double* __restrict__ S, double rtol, double floor) {
  const double bnorm = sqrt(S[S_BB]), rnorm = sqrt(S[S_TMP]);
  const double tol_abs = rtol * bnorm;
  S[S_ITL] = S[S_IT];   // the ended round's inner iterations
  S[S_IT] = 0.0;
  if (!(bnorm == bnorm)) {
    S[S_DONE] = 4.0;
    return;
  }
  if (!(rnorm == rnorm)) {
    S[S_DONE] = 2.0;
    return;
  }
  if (rnorm <= tol_abs || bnorm == 0.0) {
    S[S_DONE] = 3.0;
    return;
  }
  double inner = tol_abs / rnorm;
  if (inner < floor) inner = floor;
  S[S_TOL] = inner * rnorm;
  S[S_DONE] = 0.0;
}

__global__ void k_round_init(double* __restrict__ S, double rtol, double floor) {
  if (threadIdx.x == 0) round_init(S, rtol, floor);
}

// k_finalize's fixed-order sum of the residual partials into S_TMP, then the
// round set-up above, in one launch.
//
// Round-end gate: a round's end (x64 += d, the float64 residual, this
// finalize) may be enqueued behind an iteration batch before the host knows
// whether the inner solve has converged.  With a gate those kernels run only
// when it has (S_DONE = 1), and this one then leaves `tag` in S_SPEC so the
// host's next read tells whether the round ended.
__global__ void __launch_bounds__(256)
k_finalize_round(const double* __restrict__ part, int np, double* __restrict__ S, double rtol,
                 double floor, int gated, double tag, const double* __restrict__ part_b) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  if (part_b) {
    // ||b||^2 into S_BB first, k_finalize's fixed-order sum (F_STORE)
    double a = 0.0;
    for (int b = t; b < np; b += 256) a += part_b[b];
    sh[t] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) sh[t] += sh[t + w];
      __syncthreads();
    }
    if (t == 0) S[S_BB] = sh[0];
    __syncthreads();
  }
  // uniform; every thread reads it before the reduction's barriers, thread 0
  // writes S after them
  if (gated && S[S_DONE] != 1.0) return;
  double a = 0.0;
  for (int b = t; b < np; b += 256) a += part[b];
  sh[t] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  if (t == 0) {
    S[S_TMP] = sh[0];
    round_init(S, rtol, floor);
    if (gated) S[S_SPEC] = tag;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_zero(size_t n, T* __restrict__ a) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c < n) a[c] = T(0);
}

// Fixed-order final sum of PJ_RED partials; mode selects what is stored.
__global__ void __launch_bounds__(256)
k_finalize(const double* __restrict__ part, int np, double* __restrict__ S,
           int slot, int mode) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double a = 0.0;
  for (int b = t; b < np; b += 256) a += part[b];
  sh[t] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  if (t == 0) {
    const double v = sh[0];
    const bool done = S[S_DONE] != 0.0;
    if (mode == F_ALPHA) {          // v = <p, Ap>_W
      if (!done) {
        S[S_PQ] = v;
        S[S_ALPHA] = S[S_RZ] / v;
      }
    } else if (mode == F_BETA) {    // v = <r, z>_W (new)
      if (!done) {
        S[S_BETA] = v / S[S_RZ];
        S[S_RZ] = v;
      }
    } else if (mode == F_RR) {      // v = ||r||^2 after an iteration
      if (!done) {
        S[S_RR] = v;
        S[S_IT] += 1.0;
        if (!(v == v)) S[S_DONE] = 2.0;
        else if (sqrt(v) <= S[S_TOL]) S[S_DONE] = 1.0;
      }
    } else {
      S[slot] = v;
    }
  }
}

__global__ void __launch_bounds__(256)
k_finalize_max(const double* __restrict__ part, int np, double* __restrict__ S,
               int slot, int gated) {
  __shared__ double sh[256];
  if (gated && !solve_done(S)) return;
  const int t = threadIdx.x;
  double a = 0.0;
  for (int b = t; b < np; b += 256) a = part[b] > a ? part[b] : a;
  sh[t] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s && sh[t + s] > sh[t]) sh[t] = sh[t + s];
    __syncthreads();
  }
  if (t == 0) S[slot] = sh[0];
}

// ---------------------------------------------------------------------------
// compatible divergence (projection_3d.py:847-892), velocity correction
// (:818-842) and the post-correction divergence check
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ double div_at(const GridArgs& g, const T* u,
                                         const T* v, const T* w,
                                         const T* u_inlet, int k, int j, int i,
                                         size_t c) {
  const int Nx = g.nx, Ny = g.ny;
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  double d = 0.0;
  if (i > 0) d += ((double)u[c] - (double)u[c - 1]) / g.dx;
  else d += ((double)u[c] - (double)u_inlet[(size_t)k * Ny + j]) / g.dx;
  if (Ny > 1) {
    const size_t cm = c + (size_t)(wrap(j - 1, Ny) - j) * sy;
    d += ((double)v[c] - (double)v[cm]) / g.dy;
  }
  if (k > 0) d += ((double)w[c] - (double)w[c - sz]) / g.dz[k];
  else d += ((double)w[c] - 0.0) / g.dz[0];
  return d;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_div_rhs(GridArgs g, size_t n, const T* __restrict__ u, const T* __restrict__ v,
          const T* __restrict__ w, const T* __restrict__ u_inlet,
          const T* __restrict__ target, double inv_dt, double* __restrict__ b,
          double* __restrict__ part, double* __restrict__ x, double* __restrict__ xp,
          int xmode) {
  __shared__ double sh[256];
  (void)n;
  // contiguous runs of rows in the XCD-aware order (the stencil's y and z
  // neighbour rows mostly in the same L2), two rows' loads in flight
  int r0, r1;
  row_range(g.ny * g.nz, r0, r1);
  double acc = 0.0;
#pragma unroll 2
  for (int row = r0; row < r1; ++row) {
    const int k = row / g.ny, j = row - k * g.ny;
    for (int i = threadIdx.x; i < g.nx; i += 256) {
      const size_t c = ((size_t)k * g.ny + j) * g.nx + i;
      double d = div_at<T>(g, u, v, w, u_inlet, k, j, i, c);
      if (target) d = d - (double)target[c];
      const double r = d * inv_dt;
      if (b) b[c] = r;
      acc += r * r;
      // the warm start's pointwise update on the same cells: xmode 2
      // extrapolates x <- 2 x - xp, xp <- x; 1 keeps x in xp
      if (xmode == 2) {
        const double t = x[c];
        x[c] = 2.0 * t - xp[c];
        xp[c] = t;
      } else if (xmode == 1) {
        xp[c] = x[c];
      }
    }
  }
  block_sum_store(acc, sh, part);
}

// plain divergence into an output array (ProjectionSolver3D.divergence)
template <typename T>
__global__ void __launch_bounds__(256)
k_divergence(GridArgs g, size_t n, const T* __restrict__ u, const T* __restrict__ v,
             const T* __restrict__ w, const T* __restrict__ u_inlet,
             T* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  Cell e = cell_of(g, c);
  out[c] = (T)div_at<T>(g, u, v, w, u_inlet, e.k, e.j, e.i, c);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_div_check(GridArgs g, size_t n, const T* __restrict__ u, const T* __restrict__ v,
            const T* __restrict__ w, const T* __restrict__ u_inlet,
            const T* __restrict__ target, double* __restrict__ part,
            const double* __restrict__ gate) {
  __shared__ double sh[256];
  (void)n;
  if (gate && !solve_done(gate)) return;
  int r0, r1;
  row_range(g.ny * g.nz, r0, r1);
  double m = 0.0;
#pragma unroll 2
  for (int row = r0; row < r1; ++row) {
    const int k = row / g.ny, j = row - k * g.ny;
    for (int i = threadIdx.x; i < g.nx; i += 256) {
      const size_t c = ((size_t)k * g.ny + j) * g.nx + i;
      double d = div_at<T>(g, u, v, w, u_inlet, k, j, i, c);
      if (target) d -= (double)target[c];
      d = fabs(d);
      if (d > m || d != d) m = (d != d) ? __builtin_inf() : d;
    }
  }
  block_max_store(m, sh, part);
}

// Velocity correction: reads u, v, w (old) and writes the corrected fields
// to uo, vo, wo (may alias u, v, w: each cell writes only itself).
template <typename T, typename TR>
__global__ void __launch_bounds__(256)
k_correct(GridArgs g, size_t n, T* u, T* v, T* w, const double* __restrict__ p,
          const TR* __restrict__ rho, double dt, const double* __restrict__ gate) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  if (gate && !solve_done(gate)) return;
  Cell e = cell_of(g, c);
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const size_t sy = Nx, sz = (size_t)Nx * Ny;
  const double ir = inv_rho_at<double>(rho, c);
  if (Nx > 2) {
    double uu = (double)u[c];
    if (e.i < Nx - 1) {
      double irx = 0.5 * (ir + inv_rho_at<double>(rho, c + 1));
      uu -= dt * (p[c + 1] - p[c]) / g.dx * irx;
    } else {
      uu += 2.0 * dt * p[c] / g.dx * ir;
    }
    u[c] = (T)uu;
  }
  if (Ny > 1) {
    const size_t cp = c + (size_t)(wrap(e.j + 1, Ny) - e.j) * sy;
    double iry = 0.5 * (ir + inv_rho_at<double>(rho, cp));
    v[c] = (T)((double)v[c] - dt * (p[cp] - p[c]) / g.dy * iry);
  }
  if (Nz > 2) {
    double ww = (double)w[c];
    if (e.k < Nz - 1) {
      double irz = 0.5 * (ir + inv_rho_at<double>(rho, c + sz));
      ww -= dt * (p[c + sz] - p[c]) / g.d_above[e.k] * irz;
    } else {
      ww += 2.0 * dt * p[c] / g.dz[Nz - 1] * ir;
    }
    w[c] = (T)ww;
  }
}

// ---------------------------------------------------------------------------
// batched strided GEMM  C[b][m][n] = sum_k A[b][m][k] B[b][k][n]
// 32x32 output tile per workgroup of 256 lanes (2x2 per lane), K in steps of
// 16 staged through LDS.  Optional epilogue: multiply by the inverse
// spectral eigenvalue 1/(lz[m] + ly[j] + lx[i] - eps), n = j*Nx + i.
// ---------------------------------------------------------------------------
struct GemmArgs {
  int M, N, K;
  long sAb, sAm, sAk, sBb, sBk, sBn, sCb, sCm, sCn;
  const double* lz;
  const double* ly;
  const double* lx;
  int nx;
  double eps;
};

#define GT 32
#define GK 16
template <typename T>
__global__ void __launch_bounds__(256)
k_gemm(GemmArgs a, const T* __restrict__ A, const T* __restrict__ B,
       T* __restrict__ C) {
  __shared__ T As[GK][GT + 1];
  __shared__ T Bs[GK][GT + 1];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const long b = blockIdx.z;
  const T* Ab = A + b * a.sAb;
  const T* Bb = B + b * a.sBb;
  T acc[2][2] = {{T(0), T(0)}, {T(0), T(0)}};
  const bool a_kc = (a.sAk == 1);
  const bool b_nc = (a.sBn == 1);
  for (int k0 = 0; k0 < a.K; k0 += GK) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int e = tid + 256 * r;  // 512 elements per tile
      int mm, kk;
      if (a_kc) { kk = e % GK; mm = e / GK; } else { mm = e % GT; kk = e / GT; }
      const int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < a.M && gk < a.K) ? Ab[gm * a.sAm + gk * a.sAk] : T(0);
      int nn, kb;
      if (b_nc) { nn = e % GT; kb = e / GT; } else { kb = e % GK; nn = e / GK; }
      const int gn = n0 + nn, gkb = k0 + kb;
      Bs[kb][nn] = (gn < a.N && gkb < a.K) ? Bb[gkb * a.sBk + gn * a.sBn] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      const T a0 = As[kk][ty], a1 = As[kk][ty + 16];
      const T b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
      acc[0][0] += a0 * b0;
      acc[0][1] += a0 * b1;
      acc[1][0] += a1 * b0;
      acc[1][1] += a1 * b1;
    }
    __syncthreads();
  }
  T* Cb = C + b * a.sCb;
#pragma unroll
  for (int ii = 0; ii < 2; ++ii) {
    const int gm = m0 + ty + 16 * ii;
    if (gm >= a.M) continue;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int gn = n0 + tx + 16 * jj;
      if (gn >= a.N) continue;
      T v = acc[ii][jj];
      if (a.lz) {
        const int j = gn / a.nx, i = gn % a.nx;
        v = v * T(1.0 / (a.lz[gm] + a.ly[j] + a.lx[i] - a.eps));
      }
      Cb[gm * a.sCm + gn * a.sCn] = v;
    }
  }
}

// Column transform of the separable preconditioner along y or z on the
// matrix cores (float32): Out (K x ncol per slab) = Mat (K x KP, zero-padded
// past K) * X (K rows of ncol values at row stride sq).  One wavefront per
// 16 x 16 output tile with v_mfma_f32_16x16x4f32 over KP/4 steps (small
// tiles: many wavefronts in flight to cover the load latency):
//   A (16 x 4): lane l holds Mat[m0 + l%16][k0 + l/16] (stored on the host
//     in that fragment order, so a lane's KP/4 values are contiguous),
//   B (4 x 16): lane l holds X[k0 + l/16][n0 + l%16],
//   D (16 x 16): register r of lane l holds row 4(l/16) + r, column l%16.
// Every operand is loaded before the first MFMA.  spectral: out[m] /=
// lz[m] + ly[j] + lx[i] - eps for the column c = j Nx + i.
struct ColArgs {
  int K, ncol, nx;
  long sq, souter;
  const double *lz, *ly, *lx;
  double eps;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int KP>
__global__ void __launch_bounds__(256)
k_coltrans_mfma(ColArgs a, const float* __restrict__ Mat, const float* __restrict__ in,
                float* __restrict__ out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n0 = (blockIdx.x * 4 + wv) * 16;
  if (n0 >= a.ncol) return;   // uniform per wavefront; no barriers below
  const int m0 = blockIdx.y * 16;
  const float* inb = in + (size_t)blockIdx.z * a.souter;
  float* outb = out + (size_t)blockIdx.z * a.souter;
  const int col = n0 + (lane & 15);
  const int colc = col < a.ncol ? col : a.ncol - 1;
  const int kq = lane >> 4;
  const int nst = (a.K + 3) >> 2;   // steps holding rows < K (uniform)
  const float* af = Mat + ((size_t)blockIdx.y * 64 + lane) * (KP / 4);
  float av[KP / 4], bv[KP / 4];
#pragma unroll
  for (int st = 0; st < KP / 4; ++st) {
    if (st < nst) {
      // 32-bit element offsets (a field is smaller than 4 GiB); a row k >= K
      // of the last step re-reads row K-1, cancelled by Mat's zero padding
      const int k = 4 * st + kq;
      const uint32_t e = (uint32_t)(k < a.K ? k : a.K - 1) * (uint32_t)a.sq + (uint32_t)colc;
      bv[st] = at(inb, e * 4u);
      av[st] = af[st];
    }
  }
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int st = 0; st < KP / 4; ++st)
    if (st < nst) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[st], bv[st], acc, 0, 0, 0);
  double sd = 0.0;
  if (a.lz) {
    const int j = colc / a.nx, i = colc - j * a.nx;
    sd = a.ly[j] + a.lx[i] - a.eps;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = m0 + 4 * kq + r;
    if (m < a.K && col < a.ncol) {
      float v = acc[r];
      if (a.lz) v = v * recip((float)(a.lz[m] + sd));
      at(outb, ((uint32_t)m * (uint32_t)a.sq + (uint32_t)col) * 4u) = v;
    }
  }
}

static int col_pad(int K) { return K <= 16 ? 16 : K <= 32 ? 32 : K <= 64 ? 64 : 128; }

// The z pair of the separable solve in one pass (float32, matrix cores):
// out = Zi * D^-1 (Zf * X) for every column, D^-1 the spectral divide.
// One wavefront per 16 columns: the forward product (all ceil(K/16) row
// tiles) leaves the scaled intermediate in the wavefront's own LDS slab,
// re-read as the B operand of the inverse product, so the intermediate
// never goes to HBM and the pair is one launch.  Operand layouts as in
// k_coltrans_mfma; Zf, Zi in its A-fragment order.
template <int KP>
__global__ void __launch_bounds__(256)
k_zpair_mfma(ColArgs a, const float* __restrict__ Zf, const float* __restrict__ Zi,
             const float* __restrict__ in, float* __restrict__ out) {
  __shared__ float tl[4][KP * 16];   // per wavefront: rows 0..KP-1 x 16 columns
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n0 = (blockIdx.x * 4 + wv) * 16;
  const bool live = n0 < a.ncol;           // uniform per wavefront
  const int col = n0 + (lane & 15);
  const int colc = col < a.ncol ? col : a.ncol - 1;
  const int kq = lane >> 4;
  const int nst = (a.K + 3) >> 2;
  const int MT = (a.K + 15) >> 4;
  float* T = tl[wv];
  if (live) {
    float bv[KP / 4];
#pragma unroll
    for (int st = 0; st < KP / 4; ++st) {
      if (st < nst) {
        const int k = 4 * st + kq;
        const uint32_t e = (uint32_t)(k < a.K ? k : a.K - 1) * (uint32_t)a.sq + (uint32_t)colc;
        bv[st] = at(in, e * 4u);
      }
    }
    const int j = colc / a.nx, i = colc - j * a.nx;
    const double sd = a.ly[j] + a.lx[i] - a.eps;
    for (int mt = 0; mt < MT; ++mt) {
      const float* af = Zf + ((size_t)mt * 64 + lane) * (KP / 4);
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int st = 0; st < KP / 4; ++st)
        if (st < nst) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[st], bv[st], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 16 * mt + 4 * kq + r;
        // rows past K stay finite (zero): the inverse's zero-padded A
        // columns multiply them
        T[m * 16 + (lane & 15)] = m < a.K ? acc[r] * recip((float)(a.lz[m] + sd)) : 0.0f;
      }
    }
  }
  __syncthreads();   // every wavefront reaches it (no early return above)
  if (!live) return;
  float tv[KP / 4];
#pragma unroll
  for (int st = 0; st < KP / 4; ++st)
    if (st < nst) tv[st] = T[(4 * st + kq) * 16 + (lane & 15)];
  for (int mt = 0; mt < MT; ++mt) {
    const float* af = Zi + ((size_t)mt * 64 + lane) * (KP / 4);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int st = 0; st < KP / 4; ++st)
      if (st < nst) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[st], tv[st], acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 16 * mt + 4 * kq + r;
      if (m < a.K && col < a.ncol)
        at(out, ((uint32_t)m * (uint32_t)a.sq + (uint32_t)col) * 4u) = acc[r];
    }
  }
}

// The whole y-z part of the separable solve for one x frequency per
// workgroup (float32, matrix cores): with P the (Nz x Ny) plane of the
// x-transformed field at frequency ih (row k, column j),
//   A = P Yf^T,  B = D^-1 (Zf A),  C = Zi B,  out = C Yi^T
// (y forward, z forward with the spectral divide, z inverse, y inverse).
// The planes are the contiguous blocks in[ih ldt + k Ny + j] of the
// transposed work layout (k_dct4_tile); each is read once and written once,
// and the four products run from LDS and registers.  One wavefront per
// 16-column block j0 of the
// padded plane (Nzp = ceil16(Nz) rows, Nyp = ceil16(Ny) columns, zero
// padding); MFMA operand layouts as in k_coltrans_mfma.
//   A: B-operand fragments of Yf from Yfp (k_coltrans_mfma's packed
//      order, lane stride ky4 = col_pad(Ny)/4), A-operand from the plane in
//      LDS; the result tile mt sits in acc1[mt] (row 16mt + 4(l/16) + r in
//      register r of lane l, column j0 + l%16).
//   B, C: the z products sum over k in the order k = 16 mt + 4 q + r for
//      step 4 mt + r and lane quarter q, which is exactly where the previous
//      product left its rows: acc1[mt][r] is the B operand of step 4 mt + r
//      with no data movement.  ZfQ / ZiQ hold the matching A fragments,
//      ZfQ[(mt' 64 + l) 16 + 4 mt + r] = Zf[16 mt' + l%16][16 mt + 4(l/16) + r]
//      (four row tiles of 16 steps, zero past Nz).
//   out: C goes to LDS (over P, after every wavefront has read P), then
//      out = C Yi^T as in A.
// MT = ceil(Nz / 16) row tiles and NYS = ceil16(Ny) / 4 y steps are template
// constants: with runtime trip counts the guarded accumulating MFMAs make
// the compiler copy accumulators at every join, more moves than MFMAs.
#define YZ_MT 4   // row tiles of the largest plane (Nz <= 64)
template <int MT, int NYS>
__global__ void __launch_bounds__(512)
k_yz_plane(int ldt, int Ny, int Nz, int ky4, const float* __restrict__ Yfp,
           const float* __restrict__ Yip, const float* __restrict__ ZfQ,
           const float* __restrict__ ZiQ, const double* __restrict__ lx,
           const double* __restrict__ ly, const double* __restrict__ lz, double eps,
           const float* __restrict__ in, float* __restrict__ out) {
  // Every operand buffer is padded to compile-time bounds (YZ_MT row tiles,
  // 16 z steps, a 64-row LDS plane with zero rows past Nz), so operand loads
  // are unconditional and issue back to back ahead of their MFMAs instead
  // of each waiting out its own latency inside a runtime guard.
  extern __shared__ float pl[];          // 64 x (Nyp + 4)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int Nyp = 4 * NYS, nys = NYS;
  constexpr int S = Nyp + 4;             // padded row: conflict-free A reads
  const uint32_t base = (uint32_t)blockIdx.x * (uint32_t)ldt;   // plane ih = blockIdx.x
  // The plane: the block is 4 Nyp threads, thread (kt, jt) = (t / Nyp,
  // t % Nyp) holds rows k = 4 i + kt, i < 16 (zero past Nz or Ny).
  const int jt = threadIdx.x % Nyp, kt = threadIdx.x / Nyp;
  float v[YZ_MT * 4];
#pragma unroll
  for (int i = 0; i < YZ_MT * 4; ++i) {
    const int k = 4 * i + kt;
    v[i] = k < Nz && jt < Ny ? at(in, (base + (uint32_t)k * Ny + jt) * 4u) : 0.0f;
  }
  const int j0 = 16 * wv, q = lane >> 4, c = lane & 15;
  // this wavefront's B fragments of Yf (nys <= 32 steps)
  float yb[32];
  {
    const float* bf = Yfp + ((size_t)wv * 64 + lane) * ky4;
#pragma unroll
    for (int st = 0; st < 32; ++st)
      if (st < nys) yb[st] = bf[st];
  }
  // the z forward A fragments: 16 per lane and row tile, four 16-byte loads
  f32x4 za[YZ_MT][4], zb[YZ_MT][4];
#pragma unroll
  for (int mo = 0; mo < YZ_MT; ++mo) {
    const f32x4* p = reinterpret_cast<const f32x4*>(ZfQ + ((size_t)mo * 64 + lane) * 16);
#pragma unroll
    for (int u = 0; u < 4; ++u) za[mo][u] = p[u];
  }
  // the spectral divide's eigenvalue sums, loaded with the operands (a load
  // between the products would hold them on its latency)
  const int jc = 16 * wv + (lane & 15);
  const double sd = ly[jc < Ny ? jc : Ny - 1] + lx[blockIdx.x] - eps;
  double lzv[YZ_MT][4];
#pragma unroll
  for (int mo = 0; mo < YZ_MT; ++mo)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 16 * mo + 4 * (lane >> 4) + r;
      lzv[mo][r] = m < Nz ? lz[m] : 0.0;
    }
#pragma unroll
  for (int i = 0; i < YZ_MT * 4; ++i) pl[(4 * i + kt) * S + jt] = v[i];
  __syncthreads();
  // the z inverse fragments, in flight during the first product
#pragma unroll
  for (int mo = 0; mo < YZ_MT; ++mo) {
    const f32x4* p = reinterpret_cast<const f32x4*>(ZiQ + ((size_t)mo * 64 + lane) * 16);
#pragma unroll
    for (int u = 0; u < 4; ++u) zb[mo][u] = p[u];
  }
  f32x4 acc1[YZ_MT], acc2[YZ_MT];
#pragma unroll
  for (int mt = 0; mt < YZ_MT; ++mt) acc1[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // A = P Yf^T
#pragma unroll
  for (int st = 0; st < 32; ++st) {
    if (st >= nys) continue;
    const float* ar = pl + c * S + 4 * st + q;
    float a4[YZ_MT];
#pragma unroll
    for (int mt = 0; mt < YZ_MT; ++mt) a4[mt] = ar[16 * mt * S];
#pragma unroll
    for (int mt = 0; mt < YZ_MT; ++mt)
      if (mt < MT)
        acc1[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[mt], yb[st], acc1[mt], 0, 0, 0);
  }
  // the Yi fragments of the last product, in flight during the z products
  {
    const float* bf = Yip + ((size_t)wv * 64 + lane) * ky4;
#pragma unroll
    for (int st = 0; st < 32; ++st)
      if (st < nys) yb[st] = bf[st];
  }
  __syncthreads();   // P is no longer read: C may overwrite it
  // B = D^-1 (Zf A)
#pragma unroll
  for (int mo = 0; mo < YZ_MT; ++mo) {
    if (mo >= MT) continue;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < YZ_MT; ++mt)
      if (mt < MT) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(za[mo][mt][r], acc1[mt][r], acc, 0, 0, 0);
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 16 * mo + 4 * q + r;
      // rows past Nz stay zero: the inverse's zero-padded A columns meet them
      acc[r] = m < Nz ? acc[r] * recip((float)(lzv[mo][r] + sd)) : 0.0f;
    }
    acc2[mo] = acc;
  }
  // C = Zi B, into LDS (rows past Nzp keep the zeros of the load)
#pragma unroll
  for (int mo = 0; mo < YZ_MT; ++mo) {
    if (mo >= MT) continue;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < YZ_MT; ++mt)
      if (mt < MT) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(zb[mo][mt][r], acc2[mt][r], acc, 0, 0, 0);
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) pl[(16 * mo + 4 * q + r) * S + j0 + c] = acc[r];
  }
  __syncthreads();
  // out = C Yi^T
#pragma unroll
  for (int mt = 0; mt < YZ_MT; ++mt) acc1[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int st = 0; st < 32; ++st) {
    if (st >= nys) continue;
    const float* ar = pl + c * S + 4 * st + q;
    float a4[YZ_MT];
#pragma unroll
    for (int mt = 0; mt < YZ_MT; ++mt) a4[mt] = ar[16 * mt * S];
#pragma unroll
    for (int mt = 0; mt < YZ_MT; ++mt)
      if (mt < MT)
        acc1[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[mt], yb[st], acc1[mt], 0, 0, 0);
  }
  const int j = j0 + c;
  if (j >= Ny) return;
#pragma unroll
  for (int mt = 0; mt < YZ_MT; ++mt)
    if (mt < MT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = 16 * mt + 4 * q + r;
        if (k < Nz) at(out, (base + (uint32_t)k * Ny + j) * 4u) = acc1[mt][r];
      }
    }
}



// Orthonormal DCT-IV of each row (the x eigenbasis of the Neumann-left /
// Dirichlet-right operator) through an M-point complex FFT, N = 2M, M a
// power of two:
//   v[n] = (x[2n] + i x[N-1-2n]) e^{-i pi (4n+1)/(4N)},  V = FFT_M(v),
//   W[k] = V[k] e^{-i pi k/N},  X[2k] = sqrt(2/N) Re W[k],
//   X[N-1-2k] = -sqrt(2/N) Im W[k].
// One wavefront per row, four rows per workgroup, radix-2 decimation in time
// in LDS.  tw: pre-twiddles (M complex), post-twiddles (M complex), FFT
// twiddles e^{-2 pi i k/M} (M/2 complex), re/im interleaved.  O(N log N)
// instead of the dense N x N product.
// Work fused into the row transform inside the PCG iteration.
//   DCT_UPDATE (forward): alpha = sum Prz / sum Ppq (workgroup 0 stores it);
//     d += alpha p, r -= alpha q on the rows, then the transform of the new r;
//     partial ||r||^2 -> Pout.  Returns at once when S_DONE is set.
//   DCT_DOT (inverse): with check, workgroup 0 first applies the convergence
//     test to sum Prr (stores ||r||^2, counts the iteration, raises S_DONE);
//     then partial <r, out>_W of the rows -> Pout.
enum { DCT_PLAIN = 0, DCT_UPDATE = 1, DCT_DOT = 2 };
template <typename T>
struct DctFuse {
  T *d, *r;
  const T *p, *q;
  const double *Prz, *Ppq, *Prr, *dz;
  int npz, npq, nrr, check, ny;
  double* Pout;
  double* S;
};

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_dct4_rows(int nrows, int N, int logM, const T* __restrict__ tw,
            const T* __restrict__ in, T* __restrict__ out, DctFuse<T> f) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ double shr[4];
  const int M = N >> 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = blockIdx.x * 4 + wv;
  const bool act = row < nrows;
  T* a = reinterpret_cast<T*>(smem_raw) + (size_t)wv * 2 * M;
  const T* pre = tw;
  const T* post = tw + 2 * M;
  const T* W = tw + 4 * M;
  double acc = 0.0;
  T alpha = T(0);
  if constexpr (MODE == DCT_UPDATE) {
    if (f.S[S_DONE] != 0.0) return;   // uniform: nothing moves once converged
    const double rz = sum_partials(f.Prz, f.npz, shr);
    __syncthreads();
    const double pq = sum_partials(f.Ppq, f.npq, shr);
    __syncthreads();
    const double ad = rz / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      f.S[S_RZ] = rz;
      f.S[S_PQ] = pq;
      f.S[S_ALPHA] = ad;
    }
    alpha = (T)ad;
  }
  if constexpr (MODE == DCT_DOT) {
    if (f.check && blockIdx.x == 0) {
      const double rr = sum_partials(f.Prr, f.nrr, shr);
      if (threadIdx.x == 0 && f.S[S_DONE] == 0.0) {
        f.S[S_RR] = rr;
        f.S[S_IT] += 1.0;
        if (!(rr == rr)) f.S[S_DONE] = 2.0;
        else if (sqrt(rr) <= f.S[S_TOL]) f.S[S_DONE] = 1.0;
      }
      __syncthreads();
    }
  }
  if (act) {
    const size_t rb = (size_t)row * N;
    const T* x = in + rb;
    for (int n = lane; n < M; n += 64) {
      T re, im;
      if constexpr (MODE == DCT_UPDATE) {
        // each lane updates exactly the two entries it transforms
        const size_t e0 = rb + 2 * n, e1 = rb + N - 1 - 2 * n;
        f.d[e0] += alpha * f.p[e0];
        f.d[e1] += alpha * f.p[e1];
        re = f.r[e0] - alpha * f.q[e0];
        im = f.r[e1] - alpha * f.q[e1];
        f.r[e0] = re;
        f.r[e1] = im;
        acc += (double)re * (double)re + (double)im * (double)im;
      } else {
        re = x[2 * n];
        im = x[N - 1 - 2 * n];
      }
      const T pr = pre[2 * n], pi = pre[2 * n + 1];
      const int r = (int)(__brev((unsigned)n) >> (32 - logM));
      a[2 * r] = re * pr - im * pi;
      a[2 * r + 1] = re * pi + im * pr;
    }
  }
  __syncthreads();
  for (int st = 0; st < logM; ++st) {
    const int half = 1 << st;
    if (act) {
      for (int b = lane; b < (M >> 1); b += 64) {
        const int pos = b & (half - 1);
        const int i1 = ((b >> st) << (st + 1)) + pos;
        const int i2 = i1 + half;
        const int t = pos << (logM - 1 - st);
        const T wr = W[2 * t], wi = W[2 * t + 1];
        const T xr = a[2 * i2], xi = a[2 * i2 + 1];
        const T tr = wr * xr - wi * xi, ti = wr * xi + wi * xr;
        const T ur = a[2 * i1], ui = a[2 * i1 + 1];
        a[2 * i1] = ur + tr;
        a[2 * i1 + 1] = ui + ti;
        a[2 * i2] = ur - tr;
        a[2 * i2 + 1] = ui - ti;
      }
    }
    __syncthreads();
  }
  if (act) {
    const size_t rb = (size_t)row * N;
    T* y = out + rb;
    const T sc = T(sqrt(2.0 / N));
    const double dzk = MODE == DCT_DOT ? f.dz[row / f.ny] : 0.0;
    for (int kk = lane; kk < M; kk += 64) {
      const T vr = a[2 * kk], vi = a[2 * kk + 1];
      const T pr = post[2 * kk], pi = post[2 * kk + 1];
      const T y0 = (vr * pr - vi * pi) * sc;
      const T y1 = -((vr * pi + vi * pr) * sc);
      y[2 * kk] = y0;
      y[N - 1 - 2 * kk] = y1;
      if constexpr (MODE == DCT_DOT)
        acc += dzk * ((double)f.r[rb + 2 * kk] * (double)y0 +
                      (double)f.r[rb + N - 1 - 2 * kk] * (double)y1);
    }
  }
  if constexpr (MODE != DCT_PLAIN) {
    __syncthreads();
    const double t = block_sum_wg(acc, shr);
    if (threadIdx.x == 0) f.Pout[blockIdx.x] = t;
  }
}

// The same DCT-IV with the preconditioner's work buffer in the transposed
// layout Wt[i][row] (row = k Ny + j, leading dimension ldt), the layout
// k_yz_plane reads one x frequency i at a time as a contiguous block.
// Float32 only.  A workgroup transforms DT_ROWS consecutive rows (two per
// wavefront) and moves them between the row layout and Wt through an LDS
// tile, so every global access covers DT_ROWS consecutive rows of one
// frequency (64-byte runs) instead of one element per cache line.
//   TIN = false: rows (in, stride N) -> Wt (out); modes DCT_PLAIN, DCT_UPDATE.
//   TIN = true:  Wt (in) -> rows (out, stride N); modes DCT_PLAIN, DCT_DOT.
// The fused work (DCT_UPDATE / DCT_DOT) is that of k_dct4_rows, one partial
// per workgroup.
// Memory order: the row-layout vectors move in whole rows, lane l taking
// elements l + 64 u (every access a run of 256 bytes); the FFT's pairing of
// elements 2n and N - 1 - 2n is made in LDS (the rows are staged in the
// tile buffer).  The twiddle table is copied to LDS once per workgroup, so
// no global load sits between the FFT stages.
// DT_ROWS (8 or 16) rows per workgroup, two per wavefront (32 DT_ROWS
// threads); N <= DT_NMAX, so the tile moves in at most DT_NMAX / 32 elements
// per thread, all issued before any is used.
#define DT_NMAX 512
template <int MODE, bool TIN, int DT_ROWS>
__global__ void __launch_bounds__(32 * DT_ROWS)
k_dct4_tile(int nrows, int N, int logM, const float* __restrict__ tw,
            const float* __restrict__ in, float* __restrict__ out, int ldt,
            DctFuse<float> f) {
  constexpr int DT_THREADS = 32 * DT_ROWS, DT_RPW = 2;
  extern __shared__ float sm[];
  __shared__ double shr[2 * (DT_THREADS / 64)];
  const int M = N >> 1, LX = N + 1;       // LX: conflict-free tile row
  float* A = sm;                          // DT_ROWS x N: the complex FFT rows
  float* X = sm + DT_ROWS * N;            // DT_ROWS x LX: tile / row staging
  float* TW = X + DT_ROWS * LX;           // the 5 M twiddles
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int NE = DT_NMAX * DT_ROWS / DT_THREADS;   // tile elements per thread
  constexpr int NL = DT_NMAX / 64;                     // row elements per lane
  constexpr int NT = (5 * (DT_NMAX / 2) + DT_THREADS - 1) / DT_THREADS;
  const int row0 = blockIdx.x * DT_ROWS;
  const int ntw = 5 * M;
  double acc = 0.0;
  float alpha = 0.0f;
  if constexpr (MODE == DCT_UPDATE) {
    if (f.S[S_DONE] != 0.0) return;   // uniform: nothing moves once converged
  }
  // Every global load of the workgroup is issued first, ahead of the
  // partial sums (which hold __syncthreads) and of the LDS staging, so one
  // memory latency covers the launch: the twiddles, the tile's columns of
  // Wt (TIN: consecutive threads take consecutive rows), then the
  // wavefront's rows rl = wv DT_RPW + rr, element e = lane + 64 u.
  float twv[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int e = threadIdx.x + u * DT_THREADS;
    twv[u] = e < ntw ? tw[e] : 0.0f;
  }
  float t[NE];
  if constexpr (TIN) {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = threadIdx.x + u * DT_THREADS;
      const int i = e / DT_ROWS, row = row0 + (e - i * DT_ROWS);
      t[u] = i < N && row < nrows ? in[(size_t)i * ldt + row] : 0.0f;
    }
  }
  float va[DT_RPW][NL], qa[DT_RPW][NL], pa[DT_RPW][NL], da[DT_RPW][NL];
#pragma unroll
  for (int rr = 0; rr < DT_RPW; ++rr) {
    const int row = row0 + wv * DT_RPW + rr;
    const size_t rb = (size_t)row * N;
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = lane + 64 * u;
      if (row < nrows && e < N) {
        if constexpr (MODE == DCT_UPDATE) {
          va[rr][u] = f.r[rb + e];
          qa[rr][u] = f.q[rb + e];
          pa[rr][u] = f.p[rb + e];
          da[rr][u] = f.d[rb + e];
        } else if constexpr (MODE == DCT_DOT) {   // r for the dot of the output stage
          qa[rr][u] = f.r[rb + e];
        } else if constexpr (!TIN) {
          va[rr][u] = in[rb + e];
        }
      }
    }
  }
  if constexpr (MODE == DCT_UPDATE) {
    double rz, pq;
    sum_partials2(f.Prz, f.npz, f.Ppq, f.npq, shr, rz, pq);
    __syncthreads();
    const double ad = rz / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      f.S[S_RZ] = rz;
      f.S[S_PQ] = pq;
      f.S[S_ALPHA] = ad;
    }
    alpha = (float)ad;
  }
  if constexpr (MODE == DCT_DOT) {
    if (f.check && blockIdx.x == 0) {
      const double rr = sum_partials(f.Prr, f.nrr, shr);
      if (threadIdx.x == 0 && f.S[S_DONE] == 0.0) {
        f.S[S_RR] = rr;
        f.S[S_IT] += 1.0;
        if (!(rr == rr)) f.S[S_DONE] = 2.0;
        else if (sqrt(rr) <= f.S[S_TOL]) f.S[S_DONE] = 1.0;
      }
      __syncthreads();
    }
  }
  // twiddles and the input rows into LDS
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int e = threadIdx.x + u * DT_THREADS;
    if (e < ntw) TW[e] = twv[u];
  }
  if constexpr (TIN) {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = threadIdx.x + u * DT_THREADS;
      const int i = e / DT_ROWS, rl = e - i * DT_ROWS;
      if (i < N) X[rl * LX + i] = t[u];
    }
  } else {
#pragma unroll
    for (int rr = 0; rr < DT_RPW; ++rr) {
      const int rl = wv * DT_RPW + rr, row = row0 + rl;
      const size_t rb = (size_t)row * N;
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int e = lane + 64 * u;
        if (row < nrows && e < N) {
          float v = va[rr][u];
          if constexpr (MODE == DCT_UPDATE) {
            f.d[rb + e] = da[rr][u] + alpha * pa[rr][u];
            v = v - alpha * qa[rr][u];
            f.r[rb + e] = v;
            acc += (double)v * (double)v;
          }
          X[rl * LX + e] = v;
        }
      }
    }
  }
  __syncthreads();
  const float* pre = TW;
  const float* post = TW + 2 * M;
  const float* W = TW + 4 * M;
  // pre-twiddle of the pairs (2n, N - 1 - 2n) into bit-reversed order.
  // From here to the post-twiddle a wavefront touches only its own rows of
  // A (and of X on the input side), so a wavefront-level ordering of its
  // LDS accesses stands in for the workgroup barriers; the complex values
  // move as 8-byte pairs.
  const float2* pre2 = reinterpret_cast<const float2*>(pre);
  const float2* W2 = reinterpret_cast<const float2*>(W);
#pragma unroll
  for (int rr = 0; rr < DT_RPW; ++rr) {
    const int rl = wv * DT_RPW + rr, row = row0 + rl;
    float2* a = reinterpret_cast<float2*>(A + rl * N);
    const float* x = X + rl * LX;
#pragma unroll
    for (int u = 0; u < NL / 2; ++u) {
      const int n = lane + 64 * u;
      if (row < nrows && n < M) {
        const float re = x[2 * n], im = x[N - 1 - 2 * n];
        const float2 pw = pre2[n];
        const int r = (int)(__brev((unsigned)n) >> (32 - logM));
        a[r] = make_float2(re * pw.x - im * pw.y, re * pw.y + im * pw.x);
      }
    }
  }
  wave_lds_sync();
  for (int st = 0; st < logM; ++st) {
    const int half = 1 << st;
    for (int rr = 0; rr < DT_RPW; ++rr) {
      const int rl = wv * DT_RPW + rr;
      if (row0 + rl >= nrows) break;
      float2* a = reinterpret_cast<float2*>(A + rl * N);
      for (int b = lane; b < (M >> 1); b += 64) {
        const int pos = b & (half - 1);
        const int i1 = ((b >> st) << (st + 1)) + pos;
        const int i2 = i1 + half;
        const int t = pos << (logM - 1 - st);
        const float2 wv2 = W2[t];
        const float2 xv = a[i2];
        const float tr = wv2.x * xv.x - wv2.y * xv.y, ti = wv2.x * xv.y + wv2.y * xv.x;
        const float2 uv = a[i1];
        a[i1] = make_float2(uv.x + tr, uv.y + ti);
        a[i2] = make_float2(uv.x - tr, uv.y - ti);
      }
    }
    wave_lds_sync();
  }
  // post-twiddle into the staging rows (memory order)
  const float sc = (float)sqrt(2.0 / N);
  const float2* post2 = reinterpret_cast<const float2*>(post);
#pragma unroll
  for (int rr = 0; rr < DT_RPW; ++rr) {
    const int rl = wv * DT_RPW + rr, row = row0 + rl;
    if (row >= nrows) break;
    const float2* a = reinterpret_cast<const float2*>(A + rl * N);
#pragma unroll
    for (int u = 0; u < NL / 2; ++u) {
      const int kk = lane + 64 * u;
      if (kk >= M) break;
      const float2 av = a[kk], pv = post2[kk];
      const float vr = av.x, vi = av.y;
      const float pr = pv.x, pi = pv.y;
      const float y0 = (vr * pr - vi * pi) * sc;
      const float y1 = -((vr * pi + vi * pr) * sc);
      X[rl * LX + 2 * kk] = y0;
      X[rl * LX + N - 1 - 2 * kk] = y1;
    }
  }
  // the row output reads the wavefront's own rows; the transposed tile
  // output reads every row of the workgroup
  if constexpr (TIN) wave_lds_sync();
  else __syncthreads();
  if constexpr (TIN) {
    // whole rows out, and the dot with r in the same order
#pragma unroll
    for (int rr = 0; rr < DT_RPW; ++rr) {
      const int rl = wv * DT_RPW + rr, row = row0 + rl;
      if (row >= nrows) break;
      const size_t rb = (size_t)row * N;
      const double dzk = MODE == DCT_DOT ? f.dz[row / f.ny] : 0.0;
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int e = lane + 64 * u;
        if (e < N) {
          const float y = X[rl * LX + e];
          out[rb + e] = y;
          if constexpr (MODE == DCT_DOT) acc += dzk * ((double)qa[rr][u] * (double)y);
        }
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = threadIdx.x + u * DT_THREADS;
      const int i = e / DT_ROWS, rl = e - i * DT_ROWS, row = row0 + rl;
      if (i < N && row < nrows) out[(size_t)i * ldt + row] = X[rl * LX + i];
    }
  }
  if constexpr (MODE != DCT_PLAIN) {
    __syncthreads();
    const double tt = block_sum_wg(acc, shr);
    if (threadIdx.x == 0) f.Pout[blockIdx.x] = tt;
  }
}


// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
struct ProjHandle {
  fw_grid g;
  GridArgs ga;
  size_t n;
  int dtype;
  double eps_reg, rtol;
  int maxiter;
  size_t esz;  // element size of T
  // preconditioner (type T): Ux (Nx*Nx), Yf (Ny*Ny forward = basis^T),
  // Yi (Ny*Ny inverse = basis), Zf (Nz*Nz, D folded), Zi (Nz*Nz, D^-1 folded)
  void *Ux, *UxT, *Yf, *Yi, *Zf, *Zi;
  // float32: the y and z matrices in k_coltrans_mfma's A-fragment order,
  // zero-padded to col_pad(K) (null for float64 or K > 128: k_gemm then)
  void *Yf_p, *Yi_p, *Zf_p, *Zi_p;
  // float32, 16 <= Ny <= 128, Nz <= 64: Zf, Zi in k_yz_plane's permuted
  // A-fragment order (null otherwise: the column transforms then)
  void *ZfQ, *ZiQ;
  // x transform through k_dct4_rows when Ux is the orthonormal DCT-IV basis
  // and Nx is a power of two (twx: its twiddle table, type T)
  bool fft_x;
  int logMx;
  void* twx;
  int last_inner[8];  // iterations of the previous solve's inner PCG, per round
  int round;          // refinement round of the running solve
  double *lx, *ly, *lz;  // eigenvalues (device, float64)
  // vectors
  double *x64, *b64;
  void *r, *z, *p, *q, *d, *w1, *w2;
  void* p2;   // second direction buffer of the fused update (k_matvec_pz)
  double *part, *part_b, *S, *S_host;   // part_b: b's partial norms (k_div_rhs)
  double *Ppq, *Prr, *Prz;  // partials of the PCG iteration's reductions
  void* irho;          // 1 / max(rho, 0.01) of the running solve (type T)
  void* net;           // learned correction (unet.hip), or null
  double* Pqz;         // partials of <q, z>_W (flexible PCG)
  int learned_budget;     // inner iterations a learned round may take
  int learned_every;      // the network corrects every k-th direction (0: the first only)
  int learned_fallback;   // this solve fell back to plain PCG
  int learned_fallbacks;  // solves that fell back, since creation
  float* rs;           // r / RMS(r), the network's input
  float* qf;           // the network's density feature of the running solve
  bool have_prev;
  // float32 grids: the solution of the projection before last (x64p) and
  // whether it is there; the warm start is then extrapolated linearly from
  // the last two solutions
  double* x64p;
  bool have_prev2;
  bool tail_ran;       // the last solve_refined ran its gated tail
  long long spec_tag;  // tags of gated round ends (k_finalize_round)
  int last_iters, last_outer;
  double last_rel, last_div_max;
  int precond_enabled;
  int use_plane;       // k_yz_plane when available (else the column transforms)
  int mv_blocks;       // workgroups of the fused PCG operator (k_matvec_pz)
  int dct_rows;        // rows per k_dct4_tile workgroup (8 or 16)
  GraphCache* gc;      // captured PCG batches (fused path, non-default streams)
};

template <typename T>
static void launch_gemm(ProjHandle* h, const T* A, const T* B, T* C, int M,
                        int N, int K, int batch, long sAb, long sAm, long sAk,
                        long sBb, long sBk, long sBn, long sCb, long sCm,
                        long sCn, bool spectral, hipStream_t s) {
  GemmArgs a;
  a.M = M; a.N = N; a.K = K;
  a.sAb = sAb; a.sAm = sAm; a.sAk = sAk;
  a.sBb = sBb; a.sBk = sBk; a.sBn = sBn;
  a.sCb = sCb; a.sCm = sCm; a.sCn = sCn;
  a.lz = spectral ? h->lz : nullptr;
  a.ly = h->ly; a.lx = h->lx; a.nx = h->g.nx; a.eps = h->eps_reg;
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, batch);
  hipLaunchKernelGGL(k_gemm<T>, grid, dim3(256), 0, s, a, A, B, C);
}

// The preconditioner's work buffers w1, w2 hold the x-transformed field in
// the transposed layout Wt[i][row] (leading dimension ldt) when k_yz_plane
// runs, in the row layout otherwise.
static bool work_t(const ProjHandle* h) { return h->ZfQ && h->use_plane; }
static int work_ldt(const ProjHandle* h) { return (h->g.nz * h->g.ny + 31) & ~31; }
static size_t dct_tile_shm(int rows_per_wg, int N) {
  // FFT rows, tile rows (N + 1), twiddles (5 N / 2)
  return ((size_t)rows_per_wg * (2 * N + 1) + (size_t)5 * (N / 2)) * sizeof(float);
}
// workgroups of a k_dct4_tile launch (= its partials) at the handle's tiling
static int dct_tile_blocks(const ProjHandle* h) {
  return (h->g.nz * h->g.ny + h->dct_rows - 1) / h->dct_rows;
}
template <int MODE, bool TIN>
static void launch_dct_tile(ProjHandle* h, const float* in, float* out, const DctFuse<float>& f,
                            hipStream_t s) {
  const int rows = h->g.nz * h->g.ny, N = h->g.nx, rt = h->dct_rows;
  const dim3 grid(dct_tile_blocks(h)), block(32 * rt);
  const size_t shm = dct_tile_shm(rt, N);
  if (rt == 8)
    hipLaunchKernelGGL((k_dct4_tile<MODE, TIN, 8>), grid, block, shm, s, rows, N, h->logMx,
                       (const float*)h->twx, in, out, work_ldt(h), f);
  else
    hipLaunchKernelGGL((k_dct4_tile<MODE, TIN, 16>), grid, block, shm, s, rows, N, h->logMx,
                       (const float*)h->twx, in, out, work_ldt(h), f);
}

// x transform of every row (r -> w1 forward, w1/w2 -> z inverse): DCT-IV by
// FFT when available, else the GEMM with B = Ux (forward; Ux is symmetric
// for the DCT-IV basis) or Ux^T.
template <typename T>
static void x_transform(ProjHandle* h, const T* in, T* out, bool inverse, hipStream_t s) {
  const int Nx = h->g.nx, rows = h->g.nz * h->g.ny;
  if (work_t(h)) {
    if constexpr (sizeof(T) == 4) {
      const int ldt = work_ldt(h);
      if (h->fft_x && inverse)
        launch_dct_tile<DCT_PLAIN, true>(h, in, out, DctFuse<float>{}, s);
      else if (h->fft_x)
        launch_dct_tile<DCT_PLAIN, false>(h, in, out, DctFuse<float>{}, s);
      else if (!inverse)   // C[row][i] at i ldt + row
        launch_gemm<T>(h, in, (const T*)h->Ux, out, rows, Nx, Nx, 1, 0, Nx, 1, 0, Nx, 1, 0, 1,
                       ldt, false, s);
      else                 // A[row][i] at i ldt + row
        launch_gemm<T>(h, in, (const T*)h->UxT, out, rows, Nx, Nx, 1, 0, 1, ldt, 0, Nx, 1, 0,
                       Nx, 1, false, s);
    }
    return;
  }
  if (h->fft_x) {
    const size_t shm = 4 * (size_t)Nx * sizeof(T);
    hipLaunchKernelGGL((k_dct4_rows<T, DCT_PLAIN>), dim3((rows + 3) / 4), dim3(256), shm, s,
                       rows, Nx, h->logMx, (const T*)h->twx, in, out, DctFuse<T>{});
  } else if (!inverse) {
    launch_gemm<T>(h, in, (const T*)h->Ux, out, rows, Nx, Nx, 1, 0, Nx, 1, 0, Nx, 1, 0, Nx,
                   1, false, s);
  } else {
    launch_gemm<T>(h, in, (const T*)h->UxT, out, rows, Nx, Nx, 1, 0, Nx, 1, 0, Nx, 1, 0, Nx,
                   1, false, s);
  }
}

// The float32 column transform (k_coltrans_mfma); KP = col_pad(K).
static void launch_col(ProjHandle* h, const void* Mat, int K, long sq, int ncol, int nouter,
                       long souter, bool spectral, const float* in, float* out,
                       hipStream_t s) {
  ColArgs a;
  a.K = K; a.ncol = ncol; a.nx = h->g.nx; a.sq = sq; a.souter = souter;
  a.lz = spectral ? h->lz : nullptr; a.ly = h->ly; a.lx = h->lx; a.eps = h->eps_reg;
  const dim3 grid((ncol + 63) / 64, (K + 15) / 16, nouter);
  const float* M = (const float*)Mat;
  switch (col_pad(K)) {
    case 16: hipLaunchKernelGGL(k_coltrans_mfma<16>, grid, dim3(256), 0, s, a, M, in, out); break;
    case 32: hipLaunchKernelGGL(k_coltrans_mfma<32>, grid, dim3(256), 0, s, a, M, in, out); break;
    case 64: hipLaunchKernelGGL(k_coltrans_mfma<64>, grid, dim3(256), 0, s, a, M, in, out); break;
    default: hipLaunchKernelGGL(k_coltrans_mfma<128>, grid, dim3(256), 0, s, a, M, in, out); break;
  }
}

// The fused z pair (k_zpair_mfma) for K <= 64 (its LDS slab is KP x 16).
static bool zpair_ok(const ProjHandle* h) {
  return h->Zf_p && h->Zi_p && h->g.nz <= 64;
}
static void launch_zpair(ProjHandle* h, const float* in, float* out, hipStream_t s) {
  ColArgs a;
  const int Nz = h->g.nz;
  const long NP = (long)h->g.nx * h->g.ny;
  a.K = Nz; a.ncol = (int)NP; a.nx = h->g.nx; a.sq = NP; a.souter = 0;
  a.lz = h->lz; a.ly = h->ly; a.lx = h->lx; a.eps = h->eps_reg;
  const dim3 grid((unsigned)((NP + 63) / 64));
  const float* Zf = (const float*)h->Zf_p;
  const float* Zi = (const float*)h->Zi_p;
  switch (col_pad(Nz)) {
    case 16: hipLaunchKernelGGL(k_zpair_mfma<16>, grid, dim3(256), 0, s, a, Zf, Zi, in, out); break;
    case 32: hipLaunchKernelGGL(k_zpair_mfma<32>, grid, dim3(256), 0, s, a, Zf, Zi, in, out); break;
    default: hipLaunchKernelGGL(k_zpair_mfma<64>, grid, dim3(256), 0, s, a, Zf, Zi, in, out); break;
  }
}

// k_yz_plane over every x frequency (float32; ZfQ set; Wt layout).
static void launch_yz_plane(ProjHandle* h, const float* in, float* out, hipStream_t s) {
  const int Nx = h->g.nx, Ny = h->g.ny, Nz = h->g.nz;
  const int Nyp = (Ny + 15) & ~15, MT = (Nz + 15) / 16;
  const size_t shm = (size_t)16 * YZ_MT * (Nyp + 4) * sizeof(float);   // 64 rows
  using K = void (*)(int, int, int, int, const float*, const float*, const float*,
                     const float*, const double*, const double*, const double*, double,
                     const float*, float*);
#define YZ_ROW(m) {k_yz_plane<m, 4>, k_yz_plane<m, 8>, k_yz_plane<m, 12>, k_yz_plane<m, 16>, \
                   k_yz_plane<m, 20>, k_yz_plane<m, 24>, k_yz_plane<m, 28>, k_yz_plane<m, 32>}
  static const K table[4][8] = {YZ_ROW(1), YZ_ROW(2), YZ_ROW(3), YZ_ROW(4)};
#undef YZ_ROW
  hipLaunchKernelGGL(table[MT - 1][Nyp / 16 - 1], dim3(Nx), dim3(64 * (Nyp / 16)), shm, s,
                     work_ldt(h), Ny, Nz, col_pad(Ny) / 4, (const float*)h->Yf_p,
                     (const float*)h->Yi_p, (const float*)h->ZfQ, (const float*)h->ZiQ, h->lx,
                     h->ly, h->lz, h->eps_reg, in, out);
}

// The y and z parts of the separable solve on the x-transformed field in w1
// (w2 is scratch): y forward, z forward with the spectral divide, z
// inverse, y inverse.  Returns the buffer (w1 or w2) holding the result.
template <typename T>
static const T* yz_transforms(ProjHandle* h, T* w1, T* w2, hipStream_t s) {
  const int Nx = h->g.nx, Ny = h->g.ny, Nz = h->g.nz;
  const long NP = (long)Nx * Ny;
  const bool mf = sizeof(T) == 4;
  T* cur = w1;
  T* other = w2;
  if (mf && h->ZfQ && h->use_plane) {   // all four products in one pass
    launch_yz_plane(h, (const float*)cur, (float*)other, s);
    return other;
  }
  if (Ny > 1) {  // y forward per k-plane: w2_k = Yf (Ny x Ny) * w1_k (Ny x Nx)
    if (mf && h->Yf_p)
      launch_col(h, h->Yf_p, Ny, Nx, Nx, Nz, NP, false, (const float*)cur, (float*)other, s);
    else
      launch_gemm<T>(h, (const T*)h->Yf, cur, other, Ny, Nx, Ny, Nz, 0, Ny, 1, NP, Nx,
                     1, NP, Nx, 1, false, s);
    T* t = cur; cur = other; other = t;
  }
  if (mf && zpair_ok(h)) {   // z forward, spectral divide, z inverse: one pass
    launch_zpair(h, (const float*)cur, (float*)other, s);
    T* t = cur; cur = other; other = t;
  } else {
  // z forward with sqrt(dz) folded in, spectral divide in the epilogue
  if (mf && h->Zf_p)
    launch_col(h, h->Zf_p, Nz, NP, (int)NP, 1, 0, true, (const float*)cur, (float*)other, s);
  else
    launch_gemm<T>(h, (const T*)h->Zf, cur, other, Nz, (int)NP, Nz, 1, 0, Nz, 1, 0,
                   NP, 1, 0, NP, 1, true, s);
  { T* t = cur; cur = other; other = t; }
  // z inverse with 1/sqrt(dz) folded in
  if (mf && h->Zi_p)
    launch_col(h, h->Zi_p, Nz, NP, (int)NP, 1, 0, false, (const float*)cur, (float*)other, s);
  else
    launch_gemm<T>(h, (const T*)h->Zi, cur, other, Nz, (int)NP, Nz, 1, 0, Nz, 1, 0,
                   NP, 1, 0, NP, 1, false, s);
  { T* t = cur; cur = other; other = t; }
  }
  if (Ny > 1) {
    if (mf && h->Yi_p)
      launch_col(h, h->Yi_p, Ny, Nx, Nx, Nz, NP, false, (const float*)cur, (float*)other, s);
    else
      launch_gemm<T>(h, (const T*)h->Yi, cur, other, Ny, Nx, Ny, Nz, 0, Ny, 1, NP, Nx,
                     1, NP, Nx, 1, false, s);
    T* t = cur; cur = other; other = t;
  }
  return cur;
}

// z = M^{-1} r  (separable solve of fft_poisson_3d.py:184-228)
template <typename T>
static int precond_apply(ProjHandle* h, const T* r, T* z, hipStream_t s) {
  T* w1 = (T*)h->w1;
  T* w2 = (T*)h->w2;
  // x forward: w1[r][i'] = sum_i r[r][i] Ux[i][i']
  x_transform<T>(h, r, w1, false, s);
  const T* c2 = yz_transforms<T>(h, w1, w2, s);
  // x inverse: z[r][i] = sum_i' c2[r][i'] Ux[i][i']  (B = Ux^T)
  x_transform<T>(h, c2, z, true, s);
  FW_CHECK_LAUNCH("precond");
  return FW_OK;
}

static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// z = M^-1 r for the handle's preconditioner: the separable inverse, Jacobi
// (from the solve's 1/rho field) or none.
template <typename T>
static int apply_M(ProjHandle* h, const T* r, T* z, hipStream_t s) {
  if (h->precond_enabled == FW_PRECOND_SEPARABLE) return precond_apply<T>(h, r, z, s);
  if (h->precond_enabled == FW_PRECOND_LEARNED) {
    if constexpr (sizeof(T) != 4) {
      return fw_set_error(FW_E_CONFIG, "learned preconditioner: float32 grids only");
    } else {
      int rc = precond_apply<T>(h, r, z, s);
      if (rc) return rc;
      hipLaunchKernelGGL(k_rr_part, dim3(PJ_RED), dim3(256), 0, s, h->n, (const float*)r,
                         h->Pqz);
      hipLaunchKernelGGL(k_scale_rms, dim3(PJ_RED), dim3(256), 0, s, h->n, (const float*)r,
                         (const double*)h->Pqz, PJ_RED, h->S, h->rs);
      const float* c = nullptr;
      int nn = 0;
      if ((rc = net_apply(h->net, h->rs, h->qf, &c, &nn, s))) return rc;
      hipLaunchKernelGGL(k_add_scaled, dim3(PJ_RED), dim3(256), 0, s, h->n, c,
                         (const double*)h->S, (float*)z);
      FW_CHECK_LAUNCH("learned preconditioner");
      return FW_OK;
    }
  }
  if (h->precond_enabled == FW_PRECOND_JACOBI) {
    const Launch L = cell_launch(&h->g);
    hipLaunchKernelGGL(k_jacobi<T>, L.grid, L.block, 0, s, h->ga, (const T*)h->irho, r, z,
                       (T)h->eps_reg);
    FW_CHECK_LAUNCH("jacobi");
    return FW_OK;
  }
  FW_CHECK_HIP(hipMemcpyAsync(z, r, h->n * sizeof(T), hipMemcpyDeviceToDevice, s));
  return FW_OK;
}

static int read_scalars(ProjHandle* h, hipStream_t s) {
  FW_CHECK_HIP(hipMemcpyAsync(h->S_host, h->S, S_NS * sizeof(double),
                              hipMemcpyDeviceToHost, s));
  FW_CHECK_HIP(hipStreamSynchronize(s));
  return FW_OK;
}

// Inner PCG on A d = r (type T); r holds the right-hand side on entry and
// the recursively updated residual on exit.  Stops when ||r||_2 <= S_TOL,
// the inner tolerance k_round_init left on the device.
//
// round_end (may be null) enqueues the round's end behind every batch,
// gated on the inner convergence (k_finalize_round's round-end gate), before
// the batch's read; *ended then says whether it ran, and *iters is the
// round's iteration count either way.
template <typename T, typename TR>
static int pcg_inner(ProjHandle* h, const TR* rho, int maxit, int* iters, hipStream_t s,
                     const std::function<void(double)>* round_end = nullptr,
                     bool* ended = nullptr) {
  if (ended) *ended = false;
  const size_t n = h->n;
  T* r = (T*)h->r;
  T* z = (T*)h->z;
  T* p = (T*)h->p;
  T* p_alt = (T*)h->p2;
  T* q = (T*)h->q;
  T* d = (T*)h->d;
  // S_DONE, S_IT and S_TOL come from k_round_init, launched by the caller
  hipLaunchKernelGGL(k_zero<T>, dim3(nblk(n)), dim3(256), 0, s, n, d);
  int rc;
  (void)rho;   // the iterations use h->irho, formed from rho by solve_refined
  const T* ir = (const T*)h->irho;
  // separable preconditioner with the DCT-IV x transform: the update of d, r
  // rides in the forward row transform and <r, z> in the inverse one
  const bool fused = h->fft_x && h->precond_enabled == FW_PRECOND_SEPARABLE;
  const bool flex = h->precond_enabled == FW_PRECOND_LEARNED;
  // the row transforms in tiles (transposed work layout for k_yz_plane) or
  // four rows per workgroup; one partial per workgroup either way
  const bool tile = sizeof(T) == 4 && work_t(h);
  const int rows = h->g.nz * h->g.ny;
  const int ndct = tile ? dct_tile_blocks(h) : (rows + 3) / 4;
  const size_t shm = 4 * (size_t)h->g.nx * sizeof(T);   // k_dct4_rows
  const int npz = fused ? ndct : PJ_RED;
  DctFuse<T> fu{};
  fu.d = d; fu.r = r; fu.p = p; fu.q = q;
  fu.Prz = h->Prz; fu.Ppq = h->Ppq; fu.Prr = h->Prr; fu.dz = h->g.dz;
  // workgroups (= <p, q>_W partials) of the operator launch: the handle's
  // tiling on the fused path, PJ_RED where k_update_xr2 sums them
  const int mvb = fused ? h->mv_blocks : PJ_RED;
  fu.npz = ndct; fu.npq = mvb; fu.nrr = ndct; fu.ny = h->g.ny; fu.S = h->S;
  // the fused row transforms: MODE on rows of `in` into w1 (forward) or on
  // the work buffer `in` into z (DCT_DOT)
  auto row_transform = [&](auto mode, const T* in, T* out, const DctFuse<T>& f) {
    constexpr int MODE = decltype(mode)::value;
    if constexpr (sizeof(T) == 4) {
      if (tile) {
        launch_dct_tile<MODE, MODE == DCT_DOT>(h, (const float*)in, (float*)out, f, s);
        return;
      }
    }
    hipLaunchKernelGGL((k_dct4_rows<T, MODE>), dim3(ndct), dim3(256), shm, s, rows, h->g.nx,
                       h->logMx, (const T*)h->twx, in, out, f);
  };
  using plain_t = std::integral_constant<int, DCT_PLAIN>;
  using update_t = std::integral_constant<int, DCT_UPDATE>;
  using dot_t = std::integral_constant<int, DCT_DOT>;
  // z = M^-1 r; with fused, the inverse row transform also leaves <r, z>_W
  auto precond_dot = [&](int check) -> int {
    T* w1 = (T*)h->w1;
    T* w2 = (T*)h->w2;
    const T* c2 = yz_transforms<T>(h, w1, w2, s);
    DctFuse<T> fd = fu;
    fd.check = check;
    fd.Pout = h->Prz;
    row_transform(dot_t{}, c2, z, fd);
    FW_CHECK_LAUNCH("precond row transform");
    return FW_OK;
  };
  if (fused) {
    row_transform(plain_t{}, (const T*)r, (T*)h->w1, DctFuse<T>{});
    if ((rc = precond_dot(0))) return rc;
  } else {
    if ((rc = apply_M<T>(h, r, z, s))) return rc;
    hipLaunchKernelGGL(k_dot_rows<T>, dim3(PJ_RED), dim3(256), 0, s, h->ga, r, z, h->Prz,
                       h->Prr, h->S, 0);
  }
  hipLaunchKernelGGL(k_p_update2<T>, dim3(PJ_RED), dim3(256), 0, s, n, p, z, h->Prz, npz,
                     h->S, 1);
  FW_CHECK_LAUNCH("pcg setup");
  // iterations in batches with one host read per batch; the device flag stops
  // the Krylov updates at the first converged iteration, so the result is
  // that of a one-by-one loop.  The first batch is the previous inner
  // solve's iteration count (they change slowly from step to step), then
  // one iteration at a time.
  int queued = 0;
  const int rd = h->round < 8 ? h->round : 7;
  int batch = h->last_inner[rd] > 0 ? h->last_inner[rd] : 1;
  // A batch on the fused path is replayed from a captured graph (its
  // launches depend only on the batch length, whether it opens the round,
  // the direction buffers' order and the tiling); launch = false then only
  // advances the direction pointers as the launches would.
  auto fused_batch = [&](T*& pp, T*& pa, int q0, int count, bool launch) -> int {
    for (int b = 0; b < count; ++b) {
      if (q0 + b > 0) {
        if (launch)
          hipLaunchKernelGGL(k_matvec_pz<T>, dim3(mvb), dim3(256), 0, s, h->ga, ir, pp, z,
                             pa, q, (T)h->eps_reg, h->Prz, npz, h->S, h->Ppq);
        T* t = pp;
        pp = pa;
        pa = t;
      } else if (launch) {
        hipLaunchKernelGGL(k_matvec_rows<T>, dim3(mvb), dim3(256), 0, s, h->ga, ir, pp, q,
                           (T)h->eps_reg, h->Ppq);
      }
      if (launch) {
        DctFuse<T> fw = fu;
        fw.p = pp;
        fw.Pout = h->Prr;
        row_transform(update_t{}, (const T*)r, (T*)h->w1, fw);
        const int rc2 = precond_dot(1);
        if (rc2) return rc2;
      }
    }
    FW_CHECK_LAUNCH("pcg iteration");
    return FW_OK;
  };
  const bool graph = fused && h->gc && s && loop_graphs_on();
  for (;;) {
    const int nb_it = maxit - queued < batch ? maxit - queued : batch;
    batch = 1;
    if (graph) {
      // one graph per iteration kind (opening the round or not, direction
      // buffers' order): at most four recordings per tiling
      for (int b = 0; b < nb_it; ++b) {
        const int q0 = queued + b;
        const std::vector<uint64_t> key = {kw(p), kw(p_alt), (uint64_t)(q0 == 0),
                                           (uint64_t)mvb, (uint64_t)h->dct_rows,
                                           (uint64_t)h->use_plane};
        T* const p0 = p;
        T* const pa0 = p_alt;
        // recording: the launch helpers above issue to s, so s is the
        // recording stream for the duration
        if ((rc = h->gc->run(key, s, [&](hipStream_t cs) {
               const hipStream_t keep = s;
               s = cs;
               T* pp = p0;
               T* pa = pa0;
               const int r2 = fused_batch(pp, pa, q0, 1, true);
               s = keep;
               return r2;
             })))
          return rc;
        if ((rc = fused_batch(p, p_alt, q0, 1, false))) return rc;
      }
    } else if (fused) {
      if ((rc = fused_batch(p, p_alt, queued, nb_it, true))) return rc;
    }
    // the other preconditioners: direct launches
    for (int b = 0; b < nb_it && !fused; ++b) {
      hipLaunchKernelGGL(k_matvec_rows<T>, dim3(mvb), dim3(256), 0, s, h->ga, ir, p, q,
                         (T)h->eps_reg, h->Ppq);
      hipLaunchKernelGGL(k_update_xr2<T>, dim3(PJ_RED), dim3(256), 0, s, n, d, r, p, q,
                         h->Prz, npz, h->Ppq, h->S, h->Prr);
      // learned mode: the network corrects every learned_every-th
      // direction (its training inputs are residuals of every stage of a
      // solve), the others use the separable inverse; the flexible beta
      // keeps the directions conjugate to the last one
      const int itn = queued + b + 1;
      const bool use_net = flex && h->learned_every > 0 && itn % h->learned_every == 0;
      if ((rc = flex && !use_net ? precond_apply<T>(h, r, z, s) : apply_M<T>(h, r, z, s)))
        return rc;
      if (flex) {
        hipLaunchKernelGGL(k_dot2_rows<T>, dim3(PJ_RED), dim3(256), 0, s, h->ga, r, q, z,
                           h->Prz, h->Pqz, h->Prr, h->S, 1);
        hipLaunchKernelGGL(k_p_update_flex<T>, dim3(PJ_RED), dim3(256), 0, s, n, p, z, h->Pqz,
                           h->S);
      } else {
        hipLaunchKernelGGL(k_dot_rows<T>, dim3(PJ_RED), dim3(256), 0, s, h->ga, r, z,
                           h->Prz, h->Prr, h->S, 1);
        hipLaunchKernelGGL(k_p_update2<T>, dim3(PJ_RED), dim3(256), 0, s, n, p, z, h->Prz,
                           npz, h->S, 0);
      }
    }
    queued += nb_it;
    FW_CHECK_LAUNCH("pcg iteration");
    double tag = 0.0;
    if (round_end) {
      tag = (double)++h->spec_tag;
      (*round_end)(tag);
    }
    if ((rc = read_scalars(h, s))) return rc;
    if (h->S_host[S_DONE] == 2.0) return fw_set_error(FW_E_NONFINITE, "projection: NaN residual");
    if (tag != 0.0 && h->S_host[S_SPEC] == tag) {
      // the round ended on the device: S_IT was reset, S_ITL holds its count
      *ended = true;
      *iters = (int)h->S_host[S_ITL];
      h->last_inner[rd] = *iters;
      return FW_OK;
    }
    if (h->S_host[S_DONE] != 0.0 || queued >= maxit) break;
  }
  *iters = (int)h->S_host[S_IT];
  // a round that found the solve converged (S_DONE 3) ran no iteration and
  // says nothing about the next round's length
  if (h->S_host[S_DONE] != 3.0) h->last_inner[rd] = *iters;
  return FW_OK;
}

// Solve A x = b: x64 (warm start on entry), b64 given, ||b||^2 in S_BB on
// the device; ||b - A x||_2 <= rtol ||b||_2 with the residual recomputed in
// float64 after each inner solve.  have_residual: residual64 has already run
// on the warm start (its partials in part, r formed).
//
// The host reads the device scalars once per inner batch.  Round 0 starts
// without a read: k_round_init decides on the device whether it runs.  A
// later round is enqueued behind the previous round's residual without a
// read when the previous solve needed more rounds than have run (the usual
// case, so the read is saved); the last expected round's residual is read
// before anything else is enqueued, so a converged solve queues no idle
// round.  Iterates, iteration counts and results are those of the
// read-every-round loop.
//
// tail (may be null) enqueues work gated on the solve's convergence
// (solve_done): it is issued before the read that follows the last expected
// round's residual, so a solve that converges there costs that one read.
// h->tail_ran says whether the gated work ran.
template <typename T, typename TR>
static int solve_refined(ProjHandle* h, const TR* rho, double rtol, bool have_residual,
                         hipStream_t s, const std::function<void()>* tail = nullptr) {
  const size_t n = h->n;
  const double inner_floor = (sizeof(T) == 4) ? 2e-5 : 0.0;
  const int expected_rounds = h->last_outer;
  int total = 0, outer = 0, rc;
  hipLaunchKernelGGL((k_irho<T, TR>), dim3(nblk(n)), dim3(256), 0, s, n, rho, (T*)h->irho);
  if (h->precond_enabled == FW_PRECOND_LEARNED) {
    if constexpr (sizeof(TR) == 4) {
      if ((rc = density_feature((const float*)rho, h->qf, n, s))) return rc;
    } else {
      return fw_set_error(FW_E_CONFIG, "learned preconditioner: float32 grids only");
    }
  }
  auto residual = [&]() {
    hipLaunchKernelGGL((k_residual64<T, TR>), dim3(PJ_RED), dim3(256), 0, s, h->ga, n, rho,
                       h->b64, h->x64, (T*)h->r, h->eps_reg, h->part, (const double*)nullptr);
    hipLaunchKernelGGL(k_finalize_round, dim3(1), dim3(256), 0, s, h->part, PJ_RED, h->S,
                       rtol, inner_floor, 0, 0.0, (const double*)nullptr);
  };
  // have_residual: the caller left b's partial norms in part_b as well
  if (have_residual)
    hipLaunchKernelGGL(k_finalize_round, dim3(1), dim3(256), 0, s, h->part, PJ_RED, h->S,
                       rtol, inner_floor, 0, 0.0, (const double*)h->part_b);
  else
    residual();
  FW_CHECK_LAUNCH("residual64");
  bool known = false;   // S_host holds the outcome of the last k_round_init
  h->tail_ran = false;
  for (;;) {
    if (known) {
      if (h->S_host[S_DONE] == 4.0)
        return fw_set_error(FW_E_NONFINITE, "projection: NaN divergence");
      if (h->S_host[S_DONE] == 2.0)
        return fw_set_error(FW_E_NONFINITE, "projection: NaN residual");
      if (h->S_host[S_DONE] == 3.0) break;
    }
    if (total >= h->maxiter || outer >= 20) {
      if (!known) {   // converged or not: read before deciding
        if ((rc = read_scalars(h, s))) return rc;
        known = true;
        continue;
      }
      h->last_iters = total;
      if (h->learned_fallback) {
        h->precond_enabled = FW_PRECOND_LEARNED;
        h->learned_fallback = 0;
      }
      const double bn = sqrt(h->S_host[S_BB]);
      return fw_set_error(FW_E_NOCONV,
                          "projection PCG did not converge: ||r||/||b|| = %.3e after "
                          "%d iterations (rtol %.1e)", sqrt(h->S_host[S_TMP]) / bn, total,
                          rtol);
    }
    int it = 0;
    h->round = outer;
    // the learned mode gets an iteration budget per inner solve; a round
    // that does not converge within it hands the rest of this solve to plain
    // PCG (the separable preconditioner stays the fallback)
    const bool learned = h->precond_enabled == FW_PRECOND_LEARNED;
    int budget = h->maxiter - total;
    if (learned && budget > h->learned_budget) budget = h->learned_budget;
    // from the last expected round on, the round's end and the tail go
    // behind the first iteration batch, gated, so that one read covers the
    // batch, the residual and the tail when the batch converges
    const bool spec = !learned && outer + 1 >= expected_rounds;
    const std::function<void(double)> round_end = [&](double tag) {
      hipLaunchKernelGGL(k_x64_add<T>, dim3(nblk(n)), dim3(256), 0, s, n, h->x64,
                         (const T*)h->d, (const double*)h->S);
      hipLaunchKernelGGL((k_residual64<T, TR>), dim3(PJ_RED), dim3(256), 0, s, h->ga, n, rho,
                         h->b64, h->x64, (T*)h->r, h->eps_reg, h->part, (const double*)h->S);
      hipLaunchKernelGGL(k_finalize_round, dim3(1), dim3(256), 0, s, h->part, PJ_RED, h->S,
                         rtol, inner_floor, 1, tag, (const double*)nullptr);
      if (tail) (*tail)();
    };
    bool ended = false;
    if ((rc = pcg_inner<T, TR>(h, rho, budget, &it, s, spec ? &round_end : nullptr, &ended)))
      return rc;
    known = true;
    if (spec) h->tail_ran = tail && h->S_host[S_DONE] == 3.0 && h->S_host[S_BB] != 0.0;
    if (ended) {
      FW_CHECK_LAUNCH("gated round end");
      total += it;
      ++outer;
      continue;   // S_host holds the round's end: the loop's top decides
    }
    if (h->S_host[S_DONE] == 4.0)
      return fw_set_error(FW_E_NONFINITE, "projection: NaN divergence");
    if (h->S_host[S_DONE] == 3.0) break;   // converged before this round
    if (learned && h->S_host[S_DONE] != 1.0) {
      h->precond_enabled = FW_PRECOND_SEPARABLE;
      h->learned_fallback = 1;
    }
    total += it;
    ++outer;
    hipLaunchKernelGGL(k_x64_add<T>, dim3(nblk(n)), dim3(256), 0, s, n, h->x64,
                       (const T*)h->d, (const double*)nullptr);
    residual();
    FW_CHECK_LAUNCH("x64_add");
    if (outer < expected_rounds && total < h->maxiter && outer < 20) {
      known = false;   // the next round's first batch read will tell
    } else {
      if (tail) {
        (*tail)();
        FW_CHECK_LAUNCH("gated tail");
      }
      if ((rc = read_scalars(h, s))) return rc;
      // the gate read the scalars this read returns (the tail writes none)
      h->tail_ran = tail && h->S_host[S_DONE] == 3.0 && h->S_host[S_BB] != 0.0;
    }
  }
  const double bnorm = sqrt(h->S_host[S_BB]);
  const double rnorm = sqrt(h->S_host[S_TMP]);
  if (bnorm == 0.0) {
    // scipy's bicgstab returns x = 0 for b = 0
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(n)), dim3(256), 0, s, n, h->x64);
    FW_CHECK_LAUNCH("zero");
  }
  h->last_iters = total;
  h->last_outer = outer;
  h->last_rel = bnorm > 0.0 ? rnorm / bnorm : 0.0;
  if (h->learned_fallback) {
    h->precond_enabled = FW_PRECOND_LEARNED;
    h->learned_fallbacks += 1;
    h->learned_fallback = 0;
  }
  return FW_OK;
}

template <typename T>
static int project_impl(ProjHandle* h, void* u, void* v, void* w,
                        const void* rho, double dt, const void* target,
                        const void* u_inlet, hipStream_t s) {
  const size_t n = h->n;
  // the warm start: float32 (the production path) starts from
  // 2 p_{n-1} - p_{n-2}: the pressure moves smoothly from step to step, and
  // that starts the solve about half as far from the solution as p_{n-1}
  // (one to two PCG iterations of 14 fewer at 256x64x48); the solve still
  // runs to the same tolerance.  The float64 build keeps the reference's
  // warm start p_{n-1}.  The pointwise update rides in the divergence pass
  // (xmode 2: extrapolate, 1: keep p_{n-1} for the next step).
  int xmode = 0;
  if (!h->have_prev) {
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(n)), dim3(256), 0, s, n, h->x64);
    h->have_prev = true;
    h->have_prev2 = false;
  } else if (sizeof(T) == 4) {
    xmode = h->have_prev2 ? 2 : 1;
    h->have_prev2 = true;
  }
  // b and its partial norms (summed into S_BB by the first round set-up)
  hipLaunchKernelGGL(k_div_rhs<T>, dim3(PJ_RED), dim3(256), 0, s, h->ga, n,
                     (const T*)u, (const T*)v, (const T*)w, (const T*)u_inlet,
                     (const T*)target, 1.0 / dt, h->b64, h->part_b, h->x64, h->x64p, xmode);
  FW_CHECK_LAUNCH("div_rhs");
  // the warm start's residual, decided on together with ||b|| on the device
  hipLaunchKernelGGL((k_residual64<T, T>), dim3(PJ_RED), dim3(256), 0, s, h->ga, n,
                     (const T*)rho, h->b64, h->x64, (T*)h->r, h->eps_reg, h->part,
                     (const double*)nullptr);
  FW_CHECK_LAUNCH("residual64");
  // the correction and the divergence check (max into S_DMAX), gated on the
  // device's convergence decision when issued ahead of the host's read
  auto correct = [&](const double* gate) {
    hipLaunchKernelGGL((k_correct<T, T>), dim3(nblk(n)), dim3(256), 0, s, h->ga, n,
                       (T*)u, (T*)v, (T*)w, (const double*)h->x64, (const T*)rho, dt, gate);
    hipLaunchKernelGGL(k_div_check<T>, dim3(PJ_RED), dim3(256), 0, s, h->ga, n,
                       (const T*)u, (const T*)v, (const T*)w, (const T*)u_inlet,
                       (const T*)target, h->part, gate);
    hipLaunchKernelGGL(k_finalize_max, dim3(1), dim3(256), 0, s, h->part, PJ_RED, h->S,
                       S_DMAX, gate ? 1 : 0);
  };
  const std::function<void()> gated = [&]() { correct(h->S); };
  int rc = solve_refined<T, T>(h, (const T*)rho, h->rtol, true, s, &gated);
  if (rc) return rc;
  if (!h->tail_ran) {
    correct(nullptr);
    FW_CHECK_LAUNCH("correct");
    if ((rc = read_scalars(h, s))) return rc;
  }
  h->last_div_max = h->S_host[S_DMAX];
  return FW_OK;
}

template <typename T>
static int upload_cast(void** dst, const double* src, size_t n) {
  std::vector<T> tmp(n);
  for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
  FW_CHECK_HIP(hipMalloc(dst, n * sizeof(T)));
  FW_CHECK_HIP(hipMemcpy(*dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return FW_OK;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_proj_create(const fw_grid* g, const double* Ux, const double* Uy,
                   const double* Uz, const double* lam_x, const double* lam_y,
                   const double* lam_z, double eps_reg, double cg_rtol,
                   int maxiter, void** handle) {
  LegacyScope legacy_scope;
  FW_GRID(g);
  FW_GRID_OFFSETS32(g);
  FW_REQUIRE(g && handle && Ux && Uy && Uz && lam_x && lam_y && lam_z,
             "proj_create: null argument");
  FW_REQUIRE(g->nx >= 2 && g->nz >= 2 && g->ny >= 1, "proj_create: needs Nx, Nz >= 2");
  ProjHandle* h = new ProjHandle();
  memset(h, 0, sizeof(*h));
  h->g = *g;
  h->ga = grid_args(g);
  h->n = (size_t)g->nx * g->ny * g->nz;
  h->dtype = g->dtype;
  h->eps_reg = eps_reg;
  h->rtol = cg_rtol;
  h->maxiter = maxiter > 0 ? maxiter : 400;
  h->esz = g->dtype == FW_F64 ? 8 : 4;
  h->precond_enabled = FW_PRECOND_SEPARABLE;
  h->learned_budget = 60;
  h->learned_every = 4;
  h->use_plane = 1;
  h->mv_blocks = PJ_MV_DEFAULT;
  h->dct_rows = 16;
  h->gc = new GraphCache();
  const int Nx = g->nx, Ny = g->ny, Nz = g->nz;
  // host copies of the dz metric for folding sqrt(dz) into the z matrices
  std::vector<double> dz(Nz);
  FW_CHECK_HIP(hipMemcpy(dz.data(), g->dz, Nz * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> UxT((size_t)Nx * Nx), Yf((size_t)Ny * Ny), Zf((size_t)Nz * Nz),
      Zi((size_t)Nz * Nz);
  for (int a = 0; a < Nx; ++a)
    for (int b = 0; b < Nx; ++b) UxT[(size_t)a * Nx + b] = Ux[(size_t)b * Nx + a];
  for (int a = 0; a < Ny; ++a)
    for (int b = 0; b < Ny; ++b) Yf[(size_t)a * Ny + b] = Uy[(size_t)b * Ny + a];
  // forward: out[k'] = sum_k Uz[k][k'] sqrt(dz_k) in[k]
  // inverse: out[k]  = (1/sqrt(dz_k)) sum_k' Uz[k][k'] in[k']
  for (int kp = 0; kp < Nz; ++kp)
    for (int k = 0; k < Nz; ++k) {
      Zf[(size_t)kp * Nz + k] = Uz[(size_t)k * Nz + kp] * sqrt(dz[k]);
      Zi[(size_t)k * Nz + kp] = Uz[(size_t)k * Nz + kp] / sqrt(dz[k]);
    }
  int rc;
  if (g->dtype == FW_F64) {
    if ((rc = upload_cast<double>(&h->Ux, Ux, (size_t)Nx * Nx)) ||
        (rc = upload_cast<double>(&h->UxT, UxT.data(), (size_t)Nx * Nx)) ||
        (rc = upload_cast<double>(&h->Yf, Yf.data(), (size_t)Ny * Ny)) ||
        (rc = upload_cast<double>(&h->Yi, Uy, (size_t)Ny * Ny)) ||
        (rc = upload_cast<double>(&h->Zf, Zf.data(), (size_t)Nz * Nz)) ||
        (rc = upload_cast<double>(&h->Zi, Zi.data(), (size_t)Nz * Nz)))
      return rc;
  } else {
    if ((rc = upload_cast<float>(&h->Ux, Ux, (size_t)Nx * Nx)) ||
        (rc = upload_cast<float>(&h->UxT, UxT.data(), (size_t)Nx * Nx)) ||
        (rc = upload_cast<float>(&h->Yf, Yf.data(), (size_t)Ny * Ny)) ||
        (rc = upload_cast<float>(&h->Yi, Uy, (size_t)Ny * Ny)) ||
        (rc = upload_cast<float>(&h->Zf, Zf.data(), (size_t)Nz * Nz)) ||
        (rc = upload_cast<float>(&h->Zi, Zi.data(), (size_t)Nz * Nz)))
      return rc;
  }
  // FFT path for x when Ux is the orthonormal DCT-IV matrix and Nx = 2^p >= 4
  {
    bool pow2 = Nx >= 4 && (Nx & (Nx - 1)) == 0;
    double dmax = 0.0;
    for (int a = 0; pow2 && a < Nx; ++a)
      for (int b = 0; b < Nx; ++b) {
        const double e = sqrt(2.0 / Nx) * cos(M_PI * (a + 0.5) * (b + 0.5) / Nx);
        const double d = fabs(e - Ux[(size_t)a * Nx + b]);
        dmax = d > dmax ? d : dmax;
      }
    h->fft_x = pow2 && dmax < 1e-12 && 4 * (size_t)Nx * h->esz <= 65536;
    if (h->fft_x) {
      const int Mx = Nx / 2;
      int lg = 0;
      while ((1 << lg) < Mx) ++lg;
      h->logMx = lg;
      std::vector<double> tw(2 * Mx + 2 * Mx + Mx);
      for (int q = 0; q < Mx; ++q) {
        const double a1 = -M_PI * (4.0 * q + 1.0) / (4.0 * Nx);
        const double a2 = -M_PI * q / (double)Nx;
        tw[2 * q] = cos(a1);
        tw[2 * q + 1] = sin(a1);
        tw[2 * Mx + 2 * q] = cos(a2);
        tw[2 * Mx + 2 * q + 1] = sin(a2);
      }
      for (int q = 0; q < Mx / 2; ++q) {
        const double a3 = -2.0 * M_PI * q / (double)Mx;
        tw[4 * Mx + 2 * q] = cos(a3);
        tw[4 * Mx + 2 * q + 1] = sin(a3);
      }
      if (g->dtype == FW_F64) rc = upload_cast<double>(&h->twx, tw.data(), tw.size());
      else rc = upload_cast<float>(&h->twx, tw.data(), tw.size());
      if (rc) return rc;
    }
  }
  // zero-padded y and z matrices for the float32 column transforms
  {
    // fragment order of k_coltrans_mfma's A operand: for the 16-row tile
    // mt, lane l and step st, Mat[16 mt + l%16][4 st + l/16] (zero past K)
    auto pad = [&](void** dst, const double* src, int K) -> int {
      if (K > 128) return FW_OK;
      const int KP = col_pad(K), MT = (K + 15) / 16;
      std::vector<double> m((size_t)MT * 64 * (KP / 4), 0.0);
      for (int mt = 0; mt < MT; ++mt)
        for (int l = 0; l < 64; ++l)
          for (int st = 0; st < KP / 4; ++st) {
            const int row = 16 * mt + (l & 15), col = 4 * st + (l >> 4);
            if (row < K && col < K)
              m[((size_t)mt * 64 + l) * (KP / 4) + st] = src[(size_t)row * K + col];
          }
      return g->dtype == FW_F64 ? upload_cast<double>(dst, m.data(), m.size())
                                : upload_cast<float>(dst, m.data(), m.size());
    };
    if (g->dtype == FW_F32 &&
        ((rc = pad(&h->Yf_p, Yf.data(), Ny)) || (rc = pad(&h->Yi_p, Uy, Ny)) ||
         (rc = pad(&h->Zf_p, Zf.data(), Nz)) || (rc = pad(&h->Zi_p, Zi.data(), Nz))))
      return rc;
    // k_yz_plane's z fragments: lane l of row tile mo at step 4 mt + r holds
    // Mat[16 mo + l%16][16 mt + 4 (l/16) + r] (zero past Nz)
    // (always 4 row tiles x 16 steps, zero past Nz: the kernel's loads are
    // unconditional)
    auto quad = [&](void** dst, const double* src) -> int {
      std::vector<double> m((size_t)4 * 64 * 16, 0.0);
      for (int mo = 0; mo < 4; ++mo)
        for (int l = 0; l < 64; ++l)
          for (int mt = 0; mt < 4; ++mt)
            for (int r = 0; r < 4; ++r) {
              const int row = 16 * mo + (l & 15), col = 16 * mt + 4 * (l >> 4) + r;
              if (row < Nz && col < Nz)
                m[((size_t)mo * 64 + l) * 16 + 4 * mt + r] = src[(size_t)row * Nz + col];
            }
      return upload_cast<float>(dst, m.data(), m.size());
    };
    // (with the FFT x transform, its tile kernel takes Nx <= DT_NMAX)
    if (g->dtype == FW_F32 && Ny >= 16 && Ny <= 128 && Nz <= 64 &&
        (!h->fft_x || Nx <= DT_NMAX) &&
        ((rc = quad(&h->ZfQ, Zf.data())) || (rc = quad(&h->ZiQ, Zi.data()))))
      return rc;
  }
  FW_CHECK_HIP(hipMalloc((void**)&h->lx, Nx * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->ly, Ny * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->lz, Nz * sizeof(double)));
  FW_CHECK_HIP(hipMemcpy(h->lx, lam_x, Nx * sizeof(double), hipMemcpyHostToDevice));
  FW_CHECK_HIP(hipMemcpy(h->ly, lam_y, Ny * sizeof(double), hipMemcpyHostToDevice));
  FW_CHECK_HIP(hipMemcpy(h->lz, lam_z, Nz * sizeof(double), hipMemcpyHostToDevice));
  const size_t n = h->n;
  FW_CHECK_HIP(hipMalloc((void**)&h->x64, n * 8));
  FW_CHECK_HIP(hipMalloc((void**)&h->x64p, n * 8));
  h->have_prev2 = false;
  FW_CHECK_HIP(hipMalloc((void**)&h->b64, n * 8));
  void** vecs[] = {&h->r, &h->z, &h->p, &h->q, &h->d, &h->p2};
  for (void** v : vecs) FW_CHECK_HIP(hipMalloc(v, n * h->esz));
  // the work buffers hold the transposed layout (ldt >= Nz Ny per x
  // frequency) when k_yz_plane can run
  const size_t nw = h->ZfQ ? (size_t)work_ldt(h) * Nx : n;
  if (h->ZfQ && dct_tile_shm(16, Nx) > 65536) {
    const void* tk[] = {(const void*)k_dct4_tile<DCT_PLAIN, false, 16>,
                        (const void*)k_dct4_tile<DCT_PLAIN, true, 16>,
                        (const void*)k_dct4_tile<DCT_UPDATE, false, 16>,
                        (const void*)k_dct4_tile<DCT_DOT, true, 16>};
    for (const void* k : tk)
      FW_CHECK_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)dct_tile_shm(16, Nx)));
  }
  FW_CHECK_HIP(hipMalloc(&h->w1, nw * h->esz));
  FW_CHECK_HIP(hipMalloc(&h->w2, nw * h->esz));
  FW_CHECK_HIP(hipMalloc((void**)&h->part, PJ_RED * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->part_b, PJ_RED * sizeof(double)));
  {
    const int ndct = (g->nz * g->ny + 3) / 4;   // row-transform workgroups
    const size_t np = ndct > PJ_MV_MAX ? ndct : PJ_MV_MAX;
    FW_CHECK_HIP(hipMalloc((void**)&h->Ppq, 3 * np * sizeof(double)));
    h->Prr = h->Ppq + np;
    h->Prz = h->Ppq + 2 * np;
  }
  FW_CHECK_HIP(hipMalloc(&h->irho, n * h->esz));
  FW_CHECK_HIP(hipMalloc((void**)&h->Pqz, PJ_RED * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->S, S_NS * sizeof(double)));
  FW_CHECK_HIP(hipHostMalloc((void**)&h->S_host, S_NS * sizeof(double)));
  FW_CHECK_HIP(hipMemset(h->S, 0, S_NS * sizeof(double)));   // S_SPEC matches no tag
  FW_CHECK_HIP(hipMemset(h->x64, 0, n * 8));
  *handle = h;
  return FW_OK;
}

int fw_proj_destroy(void* handle) {
  LegacyScope legacy_scope;
  ProjHandle* h = (ProjHandle*)handle;
  if (!h) return FW_OK;
  void* bufs[] = {h->Ux, h->UxT, h->Yf, h->Yi, h->Zf, h->Zi, h->lx, h->ly, h->lz,
                  h->x64, h->x64p, h->b64, h->r, h->z, h->p, h->p2, h->q, h->d, h->w1,
                  h->w2, h->part, h->part_b, h->S, h->twx, h->Ppq, h->irho, h->Yf_p, h->Yi_p,
                  h->Zf_p, h->Zi_p, h->ZfQ, h->ZiQ, h->Pqz, h->rs, h->qf};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (h->S_host) (void)hipHostFree(h->S_host);
  delete h->gc;
  delete h;
  return FW_OK;
}

int fw_proj_project(void* handle, void* u, void* v, void* w, const void* rho,
                    double dt, const void* div_target, const void* u_inlet,
                    void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && u && v && w && rho && u_inlet && dt > 0.0,
             "proj_project: bad argument");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == FW_F64
             ? project_impl<double>(h, u, v, w, rho, dt, div_target, u_inlet, s)
             : project_impl<float>(h, u, v, w, rho, dt, div_target, u_inlet, s);
}

int fw_proj_stats(void* handle, int* iters, double* rel_resid, double* div_max) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h, "proj_stats: null handle");
  if (iters) *iters = h->last_iters;
  if (rel_resid) *rel_resid = h->last_rel;
  if (div_max) *div_max = h->last_div_max;
  return FW_OK;
}

int fw_proj_pressure(void* handle, double* p_out, void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && p_out, "proj_pressure: null argument");
  FW_CHECK_HIP(hipMemcpyAsync(p_out, h->x64, h->n * 8, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
  return FW_OK;
}

int fw_proj_set_pressure(void* handle, const double* p_in, void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && p_in, "proj_set_pressure: null argument");
  FW_CHECK_HIP(hipMemcpyAsync(h->x64, p_in, h->n * 8, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
  h->have_prev = true;
  h->have_prev2 = false;   // a new history starts from this pressure
  return FW_OK;
}

int fw_proj_set_learned(void* handle, void* net) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h, "proj_set_learned: null handle");
  if (!net) {
    h->net = nullptr;
    if (h->precond_enabled == FW_PRECOND_LEARNED) h->precond_enabled = FW_PRECOND_SEPARABLE;
    return FW_OK;
  }
  FW_REQUIRE(h->dtype == FW_F32, "proj_set_learned: float32 grids only");
  int nx, ny, nz;
  net_dims(net, &nx, &ny, &nz);
  FW_REQUIRE(nx == h->g.nx && ny == h->g.ny && nz == h->g.nz,
             "proj_set_learned: network grid %d x %d x %d != solver grid %d x %d x %d", nx, ny,
             nz, h->g.nx, h->g.ny, h->g.nz);
  if (!h->rs) FW_CHECK_HIP(hipMalloc((void**)&h->rs, h->n * sizeof(float)));
  if (!h->qf) FW_CHECK_HIP(hipMalloc((void**)&h->qf, h->n * sizeof(float)));
  h->net = net;
  return FW_OK;
}

int fw_proj_set_learned_every(void* handle, int every) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && every >= 0, "proj_set_learned_every: bad argument");
  h->learned_every = every;
  return FW_OK;
}

int fw_proj_learned_stats(void* handle, int* fallbacks) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && fallbacks, "proj_learned_stats: null argument");
  *fallbacks = h->learned_fallbacks;
  return FW_OK;
}

int fw_proj_rhs(void* handle, double* b_out, void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && b_out, "proj_rhs: null argument");
  FW_CHECK_HIP(hipMemcpyAsync(b_out, h->b64, h->n * 8, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
  return FW_OK;
}

int fw_proj_divergence(void* handle, const void* u, const void* v,
                       const void* w, const void* u_inlet, void* div_out,
                       void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && u && v && w && u_inlet && div_out, "proj_divergence: null argument");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == FW_F64)
    hipLaunchKernelGGL(k_divergence<double>, dim3(nblk(h->n)), dim3(256), 0, s, h->ga,
                       h->n, (const double*)u, (const double*)v, (const double*)w,
                       (const double*)u_inlet, (double*)div_out);
  else
    hipLaunchKernelGGL(k_divergence<float>, dim3(nblk(h->n)), dim3(256), 0, s, h->ga,
                       h->n, (const float*)u, (const float*)v, (const float*)w,
                       (const float*)u_inlet, (float*)div_out);
  FW_CHECK_LAUNCH("divergence");
  return FW_OK;
}

int fw_proj_matvec(void* handle, const void* rho, const void* x, void* y,
                   void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && rho && x && y, "proj_matvec: null argument");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == FW_F64)
    hipLaunchKernelGGL((k_matvec_pq<double, double>), dim3(PJ_RED), dim3(256), 0, s,
                       h->ga, h->n, (const double*)rho, (const double*)x, (double*)y,
                       h->eps_reg, h->part);
  else
    hipLaunchKernelGGL((k_matvec_pq<float, float>), dim3(PJ_RED), dim3(256), 0, s,
                       h->ga, h->n, (const float*)rho, (const float*)x, (float*)y,
                       h->eps_reg, h->part);
  FW_CHECK_LAUNCH("matvec");
  return FW_OK;
}

int fw_proj_precondition(void* handle, const void* r, void* z, void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && r && z, "proj_precondition: null argument");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == FW_F64 ? precond_apply<double>(h, (const double*)r, (double*)z, s)
                            : precond_apply<float>(h, (const float*)r, (float*)z, s);
}

int fw_proj_set_tiling(void* handle, int yz_plane, int matvec_blocks, int dct_rows,
                       int* yz_active) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h, "proj_set_tiling: null handle");
  FW_REQUIRE(matvec_blocks == 0 || (matvec_blocks >= 64 && matvec_blocks <= PJ_MV_MAX),
             "proj_set_tiling: matvec_blocks must be 0 (default) or 64..%d", PJ_MV_MAX);
  FW_REQUIRE(dct_rows == 0 || dct_rows == 8 || dct_rows == 16,
             "proj_set_tiling: dct_rows must be 0 (default), 8 or 16");
  h->use_plane = yz_plane ? 1 : 0;
  h->mv_blocks = matvec_blocks ? matvec_blocks : PJ_MV_DEFAULT;
  h->dct_rows = dct_rows ? dct_rows : 16;
  if (yz_active) *yz_active = h->use_plane && h->ZfQ ? 1 : 0;
  return FW_OK;
}

int fw_proj_set_preconditioner(void* handle, int enabled) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h, "proj_set_preconditioner: null handle");
  FW_REQUIRE(enabled == FW_PRECOND_NONE || enabled == FW_PRECOND_SEPARABLE ||
                 enabled == FW_PRECOND_JACOBI || enabled == FW_PRECOND_LEARNED,
             "proj_set_preconditioner: mode must be 0 (none), 1 (separable), 2 (Jacobi) "
             "or 3 (learned)");
  FW_REQUIRE(enabled != FW_PRECOND_LEARNED || h->net,
             "proj_set_preconditioner: attach a network with fw_proj_set_learned first");
  h->precond_enabled = enabled;
  return FW_OK;
}

int fw_proj_solve(void* handle, const void* rho, const double* b, double* x,
                  double rtol, int* iters, double* rel_resid, void* stream) {
  ProjHandle* h = (ProjHandle*)handle;
  FW_REQUIRE(h && rho && b && x, "proj_solve: null argument");
  hipStream_t s = (hipStream_t)stream;
  FW_CHECK_HIP(hipMemcpyAsync(h->b64, b, h->n * 8, hipMemcpyDeviceToDevice, s));
  FW_CHECK_HIP(hipMemcpyAsync(h->x64, x, h->n * 8, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_sumsq, dim3(PJ_RED), dim3(256), 0, s, h->n,
                     (const double*)h->b64, h->part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, h->part, PJ_RED, h->S,
                     S_BB, F_STORE);
  FW_CHECK_LAUNCH("solve setup");
  h->have_prev = true;
  h->have_prev2 = false;
  int rc = h->dtype == FW_F64
               ? solve_refined<double, double>(h, (const double*)rho, rtol, false, s)
               : solve_refined<float, float>(h, (const float*)rho, rtol, false, s);
  if (iters) *iters = h->last_iters;
  if (rel_resid) *rel_resid = h->last_rel;
  if (rc) return rc;
  FW_CHECK_HIP(hipMemcpyAsync(x, h->x64, h->n * 8, hipMemcpyDeviceToDevice, s));
  FW_CHECK_HIP(hipStreamSynchronize(s));
  return FW_OK;
}

}  // extern "C"
