// Shared definitions for the firewave HIP kernels (gfx950 / CDNA4).
//
// Field layout: every 3-D field is a dense (Nz, Ny, Nx) C-order array, flat
// index (k*Ny + j)*Nx + i, x fastest, so a wavefront of 64 lanes reads 64
// consecutive x cells (one 256 B / 512 B coalesced line group for f32 / f64).
// Storage type is selected per grid (fw_grid.dtype); every stencil kernel is
// a template over that type and is instantiated for float and double.
#pragma once
#include <hip/hip_runtime.h>
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

inline Launch cell_launch(const fw_grid* g) {
  int bx = 64;
  while (bx < g->nx && bx < 256) bx *= 2;
  Launch L;
  L.block = dim3(bx, 1, 1);
  L.grid = dim3((g->nx + bx - 1) / bx, g->ny, g->nz);
  return L;
}

// Plain-value view of fw_grid passed to kernels by value.
struct GridArgs {
  int nx, ny, nz, n_z_bed;
  double dx, dy;
  const double* dz;
  const double* d_above;
  const double* d_below;
  const double* z_face;
};

inline GridArgs grid_args(const fw_grid* g) {
  GridArgs a;
  a.nx = g->nx; a.ny = g->ny; a.nz = g->nz; a.n_z_bed = g->n_z_bed;
  a.dx = g->dx; a.dy = g->dy;
  a.dz = g->dz; a.d_above = g->d_above; a.d_below = g->d_below;
  a.z_face = g->z_face;
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

// minmod limiter (muscl_3d.py:36-45): smaller-magnitude argument when both
// have the same sign, zero otherwise.
template <typename T>
__device__ __forceinline__ T minmod(T a, T b) {
  if (a * b <= T(0)) return T(0);
  if (a > T(0)) return a < b ? a : b;
  return a > b ? a : b;
}

// MUSCL face value at face i+1/2 from the upwind side (muscl_3d.py:48-64).
// The upwind side is chosen from the sign of `uf` (the caller passes the
// cell-centre velocity, as the reference does).
template <typename T>
__device__ __forceinline__ T muscl_face(T pm1, T p0, T pp1, T pp2, T uf) {
  if (uf >= T(0)) {
    T s = minmod<T>(p0 - pm1, pp1 - p0);
    return p0 + T(0.5) * s;
  }
  T s = minmod<T>(pp1 - p0, pp2 - pp1);
  return pp1 - T(0.5) * s;
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

// Dispatch a templated implementation on the grid's storage type.
#define FW_DISPATCH(g, fn, ...)                                               \
  ((g)->dtype == FW_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))
