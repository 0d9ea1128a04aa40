// Discrete-ordinates radiation (dom_3d.py:98-473 of the reference).
//
// Sweep design.  For ordinate n the upwind step-differencing update
//   I(i) = (kB + ax I(i-1) + ay I_y + az I_z) / (k + ax + ay + az)
// is a first-order linear recurrence along x once the y and z inflows of the
// row are known.  A row (n, j, k) is therefore one wavefront: each lane folds
// its run of cells into an affine map x -> a x + b, a shuffle scan across the
// 64 lanes composes the maps, and each lane then re-evaluates its cells with
// the reference formula from the exact upwind value.  Rows whose
// sweep-ordered (j', k') lie on the same anti-diagonal j' + k' = d are
// independent, so one launch processes diagonal d for all M ordinates at
// once; Ny + Nz - 1 launches make one full source iteration.
//
// Periodic y uses the reference's stale wrap (dom_3d.py:189-202): the
// inflow of row j' = 0 is row j' = Ny-1 of the same plane, which lies on a
// later diagonal and still holds the previous iteration's intensities (for
// Ny = 1 it is the row itself, read before it is written).  I_set persists
// across solves (warm start, dom_3d.py:282).
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <vector>
#include "fw_common.h"
#include "graph.h"

namespace fw {


// The solve's device state (DomHandle::S): the largest |G - G_prev| and the
// largest G of the last completed iteration, the converged flag (set by
// k_dom_fin; every kernel of a source iteration returns at once when it is
// set, so iterations can be enqueued in batches and the host reads the
// state once per batch), the count of completed iterations (its parity
// names the G buffer the next iteration writes) and the non-finite flag.
enum { DS_DIFF = 0, DS_GMAX, DS_DONE, DS_ITERS, DS_NAN, DS_EPOCH, DS_STALL, DS_N };

// The dataflow sweep (k_dom_sweep_flow): per XCD a queue counter and a
// give-up word, then per (ordinate, block) the epoch its block was stored in.
// Each XCD's words, and each ordinate's flags, sit in 256-byte lines of
// their own: no cache line is written from two XCDs' L2s.
constexpr int FLOW_XCD = 8;        // XCDs of an MI355X (the XCC_ID register's range)
constexpr int FLOW_LINE = 64;      // ints per 256-byte line
__host__ __device__ constexpr int flow_ctr(int x) { return FLOW_LINE * x; }
__host__ __device__ constexpr int flow_gave_up(int x) { return FLOW_LINE * x + 1; }
constexpr int FLOW_MAXD = 127;     // block diagonals a plan holds
constexpr int FLOW_SPIN = 1 << 20; // polls before a workgroup gives up (about 1 s)
struct FlowPlan {
  int TJ, TK, ndiag;
  int fstride;              // ints per ordinate's flags (TJ TK rounded up to a line)
  int cum[FLOW_MAXD + 1];   // blocks on the diagonals before D
};
// rows the dataflow sweep takes: 64 PER cells (PER 1 .. 8) whose 4 x 4
// blocks, with the kernel's own two words, fit 64 KB of LDS
static inline bool flow_rows(int nx, size_t esz) {
  const bool pow2 = nx == 64 || nx == 128 || nx == 256 || nx == 512;
  return pow2 && 16 * (size_t)nx * esz < 65536;
}

struct DomHandle {
  fw_grid g;
  GridArgs ga;
  int M, maxiter;
  double eps_ground, tol, relax, kappa_gas_max;
  size_t n, esz;
  double* Om;  // device M x 3
  double* w;   // device M
  std::vector<double> Om_h, w_h;
  void *I, *G, *Gp, *kap, *kap_s, *kap_safe, *B;
  void *Ig_out, *Ig_emit, *q_in;  // (Ny, Nx)
  double *part, *part2, *S, *S_host;   // S: DS_N solve-state slots
  int last_iters;
  GraphCache* gc;   // captured source iterations (non-default streams, timer off)
  int tile;         // requested sweep tile side (0: the default 4)
  // the stream a step's DOM solve runs on beside the step (dom_fork), and
  // the events that order it; created on first use
  hipStream_t side;
  hipEvent_t ev_fork, ev_join;
  // dataflow sweep: queue counters and give-up words (a line per XCD), the
  // blocks' epochs (a run of lines per ordinate); flow: 1 in use, 0 not built
  // for this grid, -1 switched off
  int *flow_q, *flow_flags;
  int flow;
  int flow_wg;      // dataflow sweep workgroups per XCD
  int flow_spin;    // polls before a waiting workgroup gives up
  FlowPlan plan;
};

template <typename T>
__global__ void __launch_bounds__(256)
k_dom_props(size_t n, const T* __restrict__ Ts, const T* __restrict__ Tg,
            const T* __restrict__ alpha, const T* __restrict__ omega,
            const T* __restrict__ Yw, const T* __restrict__ rho,
            const T* __restrict__ Mloc, T sigma_sav, T T_amb, T kgmax,
            T* __restrict__ kap, T* __restrict__ kap_s, T* __restrict__ kap_safe,
            T* __restrict__ B) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  T ks = sigma_sav * alpha[c];
  if (Mloc) ks = ks * (T(1) + T(5.0) * Mloc[c]);
  T of = omega[c] / T(1.0e-3);
  of = of < T(0) ? T(0) : (of > T(1) ? T(1) : of);
  T tf = (Tg[c] - T(600.0)) / T(400.0);
  tf = tf < T(0) ? T(0) : (tf > T(1) ? T(1) : tf);
  T kg = kgmax * (of > tf ? of : tf);
  if (Yw && rho) kg = kg + T(30.0) * rho[c] * Yw[c];
  const T k = ks + kg;
  const T ksafe = k > T(1.0e-3) ? k : T(1.0e-3);
  T Tr4;
  if (k > T(1.0e-9)) {
    const T a = Ts[c], b = Tg[c];
    Tr4 = (ks * ((a * a) * (a * a)) + kg * ((b * b) * (b * b))) / ksafe;
  } else {
    Tr4 = (T_amb * T_amb) * (T_amb * T_amb);
  }
  kap[c] = k;
  kap_s[c] = ks;
  kap_safe[c] = ksafe;
  B[c] = T(5.67e-8) * Tr4 / T(M_PI);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_dom_ground_init(int ncol, const T* __restrict__ Tsoil, T eps, T I_amb,
                  T* __restrict__ emit, T* __restrict__ out, double* __restrict__ S,
                  int* __restrict__ flow_q) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c == 0) {   // a new solve: not converged, no iteration done
    S[DS_DONE] = 0.0;
    S[DS_ITERS] = 0.0;
    S[DS_NAN] = 0.0;
    S[DS_STALL] = 0.0;
    S[DS_EPOCH] += 1.0;   // the first sweep's epoch
    if (flow_q)
      for (int x = 0; x < FLOW_XCD; ++x) flow_q[flow_ctr(x)] = flow_q[flow_gave_up(x)] = 0;
  }
  if (c >= ncol) return;
  T e;
  if (Tsoil) {
    const T t = Tsoil[c];
    e = eps * T(5.67e-8) * ((t * t) * (t * t)) / T(M_PI);
  } else {
    e = eps * I_amb;
  }
  emit[c] = e;
  out[c] = e;
}

// Cross-lane moves for the wavefront scans, as DPP modifiers of a VALU move
// (no LDS round trip).  Lanes whose source lies outside the pattern, or
// whose row is masked off, keep `old`.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_mov(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v),
                                                    CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_mov(double old, double v) {
  const unsigned long long o = __double_as_longlong(old), x = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)x,
                                                            CTRL, ROWS, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32),
                                                            (int)(unsigned)(x >> 32), CTRL,
                                                            ROWS, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// (A, Bc) o (pa, pb): x -> A (pa x + pb) + Bc, with the lane CTRL names as
// the earlier map; lanes without one compose with the identity (1, 0).
template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ void compose_from(T& A, T& Bc) {
  const T pa = dpp_mov<CTRL, ROWS>(T(1), A);
  const T pb = dpp_mov<CTRL, ROWS>(T(0), Bc);
  Bc = A * pb + Bc;
  A = A * pa;
}

// Inclusive scan of affine maps over the 64 lanes: shifts of 1, 2, 4, 8
// inside each row of 16 lanes, then lane 15 of rows 0 and 2 into rows 1 and
// 3, then lane 31 into rows 2 and 3.  On return (eA, eB) is the exclusive
// prefix (the identity on lane 0).
template <typename T>
__device__ __forceinline__ void affine_scan64(T& A, T& Bc, T& eA, T& eB) {
  compose_from<0x111, 0xF>(A, Bc);   // row_shr:1
  compose_from<0x112, 0xF>(A, Bc);   // row_shr:2
  compose_from<0x114, 0xF>(A, Bc);   // row_shr:4
  compose_from<0x118, 0xF>(A, Bc);   // row_shr:8
  compose_from<0x142, 0xA>(A, Bc);   // row_bcast:15 into rows 1, 3
  compose_from<0x143, 0xC>(A, Bc);   // row_bcast:31 into rows 2, 3
  eA = dpp_mov<0x138, 0xF>(T(1), A); // wave_shr:1
  eB = dpp_mov<0x138, 0xF>(T(0), Bc);
}

// One anti-diagonal d of the row wavefront, all ordinates.  One wavefront
// per (row, ordinate), four ordinates per workgroup.  Lane l owns the PER
// consecutive cells [l PER, (l+1) PER) of the row in sweep order: it folds
// them into an affine map x -> A x + Bc, the maps are composed across the
// wavefront with a shuffle scan (no LDS, no barriers), and each lane then
// re-evaluates its cells with the reference formula from the exact upwind
// value.
template <typename T, int PER>
__global__ void __launch_bounds__(256)
k_dom_sweep_diag(GridArgs g, int d, int jlo, int M, const double* __restrict__ Om,
                 T* __restrict__ Iset, const T* __restrict__ kap,
                 const T* __restrict__ B, const T* __restrict__ Ig_out, T I_amb,
                 const double* __restrict__ S) {
  if (S[DS_DONE] != 0.0) return;   // uniform: the solve has converged
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const int lane = threadIdx.x & 63;
  const int n = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (n >= M) return;                       // whole wavefront exits together
  const int jp = jlo + blockIdx.x;
  const int kp = d - jp;
  const double xi = Om[3 * n], eta = Om[3 * n + 1], mu = Om[3 * n + 2];
  const int j = eta >= 0.0 ? jp : Ny - 1 - jp;
  const int k = mu >= 0.0 ? kp : Nz - 1 - kp;
  const size_t plane = (size_t)Nx * Ny;
  T* In = Iset + (size_t)n * plane * Nz;
  const T ax = T(fabs(xi) / g.dx);
  const T ay = T(fabs(eta) / g.dy);
  const T az = T(fabs(mu) / g.dz[k]);
  const int jy = eta >= 0.0 ? (j > 0 ? j - 1 : Ny - 1) : (j < Ny - 1 ? j + 1 : 0);
  const bool zb = mu >= 0.0 ? (k == 0) : (k == Nz - 1);
  const int kz = mu >= 0.0 ? k - 1 : k + 1;
  const size_t row = ((size_t)k * Ny + j) * Nx;
  const size_t rowy = ((size_t)k * Ny + jy) * Nx;
  const size_t rowz = zb ? 0 : ((size_t)kz * Ny + j) * Nx;
  const bool fwd = xi >= 0.0;
  // pass 1: fold the lane's cells (sweep order) into x -> A x + Bc; keep the
  // y and z inflows for the exact re-evaluation in pass 2
  T Iys[PER], Izs[PER];
  T A = T(1), Bc = T(0);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int s = lane * PER + q;
    if (s < Nx) {
      const int i = fwd ? s : Nx - 1 - s;
      const size_t c = row + i;
      const T kc = kap[c];
      const T Iy = In[rowy + i];
      const T Iz = zb ? (mu >= 0.0 ? Ig_out[(size_t)j * Nx + i] : I_amb) : In[rowz + i];
      const T den = kc + ax + ay + az;
      const T a = ax / den;
      const T b = (kc * B[c] + ay * Iy + az * Iz) / den;
      Bc = a * Bc + b;  // compose: x -> a (A x + Bc) + b
      A = a * A;
      Iys[q] = Iy;
      Izs[q] = Iz;
    }
  }
  // scan of the maps over the wavefront; the exclusive prefix applied to
  // the x inflow I_left = I_right = I_amb (dom_3d.py:376, :424)
  T eA, eB;
  affine_scan64(A, Bc, eA, eB);
  T Iprev = eA * I_amb + eB;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int s = lane * PER + q;
    if (s < Nx) {
      const int i = fwd ? s : Nx - 1 - s;
      const size_t c = row + i;
      const T kc = kap[c];
      const T num = kc * B[c] + ax * Iprev + ay * Iys[q] + az * Izs[q];
      const T den = kc + ax + ay + az;
      const T v = num / den;
      In[c] = v;
      Iprev = v;
    }
  }
}

// Tiled wavefront: one workgroup (TS wavefronts) sweeps a TS x TS block of
// rows (sweep-ordered j' in [TS tj, TS tj + TS), k' in [TS tk, TS tk + TS))
// of one ordinate.  Blocks on the block anti-diagonal tj + tk = D are
// independent, so Ny/TS + Nz/TS - 1 launches make a source iteration.
// Wavefront w owns the column j' = TS tj + w and walks k' upward: at stage
// s (0 .. 2 TS - 2) it computes row (w, s - w), so the
// row diagonals of the block run in order with every wavefront busy on the
// middle stages.  The z inflow of a row is the same wavefront's previous row
// (registers); the y inflow is wavefront w - 1's row of the previous stage
// (LDS, one slot per row).  Inflows from outside the block come from global
// memory: rows written by earlier launches, the ground or top boundary, or
// the stale periodic wrap.  All global inputs are read before any row of the
// block is stored, so the wrap of small Ny still reads the previous
// iteration.  One reciprocal per cell: a = ax r, b = (...) r, I = num r with
// r = 1 / (kappa + ax + ay + az).
//
// Everything of a row that does not depend on its inflows is computed once,
// before the first stage, while the other rows' loads are in flight: the
// reciprocals r, kappa B, and the multiplicative half of the row's affine
// scan (the lane's fold product A and its value at each scan level, which
// the additive half multiplies by).  A stage is then the additive half only:
// the inflow terms, the lane's fold of b, six DPP steps and the
// re-evaluation.  The arithmetic is the same operation for operation
// as folding and scanning (A, Bc) together, so the intensities are too.
// The PER cells of one lane, consecutive in memory, as one aligned vector:
// a row whose length is 64 PER moves in 16-byte (or 8-byte) accesses that
// cover whole cache lines, instead of PER strided scalar accesses.
template <typename T, int N>
struct alignas(N * sizeof(T) >= 16 ? 16 : N * sizeof(T)) LaneCells {
  T v[N];
};

// v = *(V*)(base + off) read at device scope (the vector L1 cache is not
// consulted): rows another workgroup of the same XCD stored during the
// running kernel come from that XCD's L2, where the stores went
template <typename V>
__device__ __forceinline__ V load_l2(const V* base, uint32_t off) {
  const char* p = reinterpret_cast<const char*>(base) + off;
  V v;
  if constexpr (sizeof(V) % 8 == 0) {
    unsigned long long u[sizeof(V) / 8];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(V) / 8); ++i)
      u[i] = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p) + i,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_memcpy(&v, u, sizeof(V));
  } else {
    static_assert(sizeof(V) == 4, "4- or 8k-byte values");
    const unsigned u = __hip_atomic_load(reinterpret_cast<const unsigned*>(p),
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_memcpy(&v, &u, 4);
  }
  return v;
}

// One TS x TS block (tj, tk) of ordinate n, the body of both sweep kernels.
// FLOW (the dataflow kernel): the block's inflow-independent part first,
// then wait() (every thread; false: give up), then the inflows, read from
// L2; without FLOW every input is read up front.
template <typename T, int PER, bool FULL, int TS, bool FLOW, typename Wait>
__device__ __forceinline__ bool sweep_block(const GridArgs& g, int n, int tj, int tk,
                                            const double* __restrict__ Om,
                                            T* __restrict__ Iset, const T* __restrict__ kap,
                                            const T* __restrict__ B,
                                            const T* __restrict__ Ig_out, T I_amb, T* sh,
                                            Wait&& wait) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int jp = TS * tj + w;
  const bool col = jp < Ny;                 // uniform per wavefront
  const double xi = Om[3 * n], eta = Om[3 * n + 1], mu = Om[3 * n + 2];
  const int j = eta >= 0.0 ? jp : Ny - 1 - jp;
  T* In = Iset + (size_t)n * Nx * Ny * Nz;
  const bool fwd = xi >= 0.0;
  const T ax = T(fabs(xi) * g.inv_dx);
  const T ay = T(fabs(eta) * g.inv_dy);
  const int jy = eta >= 0.0 ? (j > 0 ? j - 1 : Ny - 1) : (j < Ny - 1 ? j + 1 : 0);
  // x index of each of the lane's cells (sweep position s = lane PER + q);
  // byte offsets within one ordinate's field fit 32 bits (field < 4 GiB)
  constexpr uint32_t E = sizeof(T);
  uint32_t ix[PER];
  bool ok[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int sq = lane * PER + q;
    ok[q] = FULL || sq < Nx;
    ix[q] = (uint32_t)(fwd ? sq : Nx - 1 - sq);
  }
  // whole rows (FULL): the lane's cells are x in [i0, i0 + PER), sweep
  // position q at element q (forward) or PER - 1 - q (backward)
  using LC = LaneCells<T, PER>;
  const uint32_t i0 = (uint32_t)(fwd ? lane * PER : Nx - PER * (lane + 1));
  auto load_cells = [&](const T* p, uint32_t rowoff, T* out) {
    if constexpr (FULL) {
      const LC c = at(reinterpret_cast<const LC*>(p), (rowoff + i0) * E);
#pragma unroll
      for (int q = 0; q < PER; ++q) out[q] = fwd ? c.v[q] : c.v[PER - 1 - q];
    } else {
#pragma unroll
      for (int q = 0; q < PER; ++q)
        if (ok[q]) out[q] = at(p, (rowoff + ix[q]) * E);
    }
  };
  // the inflow rows: from L2 in the dataflow kernel
  auto load_inflow = [&](const T* p, uint32_t rowoff, T* out) {
    if constexpr (!FLOW) {
      load_cells(p, rowoff, out);
    } else if constexpr (FULL) {
      const LC c = load_l2(reinterpret_cast<const LC*>(p), (rowoff + i0) * E);
#pragma unroll
      for (int q = 0; q < PER; ++q) out[q] = fwd ? c.v[q] : c.v[PER - 1 - q];
    } else {
#pragma unroll
      for (int q = 0; q < PER; ++q)
        if (ok[q]) out[q] = load_l2(p, (rowoff + ix[q]) * E);
    }
  };
  // global inputs read up front: kappa and B of the column's TS rows, the y
  // inflow of the column's rows (w == 0), the z inflow of its first row
  // (kk == 0) and each row's z coefficient.  A stage then issues no global
  // load (on gfx9 the wait for a load also waits for every earlier store).
  T Iyg[TS][PER], Iz[PER], kB[TS][PER], rc[TS][PER], azs[TS];
  T Al[TS][7];   // per row: A at the six scan levels, then the exclusive prefix
  // the z inflow of the column's first row: the block below, the ground
  // (upward ordinates) or the top boundary
  auto load_z_inflow = [&]() {
    const int kp0 = TS * tk;
    const int k0 = mu >= 0.0 ? kp0 : Nz - 1 - kp0;
    const bool zb = mu >= 0.0 ? (k0 == 0) : (k0 == Nz - 1);
    const int kz = mu >= 0.0 ? k0 - 1 : k0 + 1;
    if (!zb) {
      load_inflow(In, ((uint32_t)kz * Ny + j) * Nx, Iz);
    } else if (mu >= 0.0) {
      load_cells(Ig_out, (uint32_t)j * Nx, Iz);
    } else {
#pragma unroll
      for (int q = 0; q < PER; ++q) Iz[q] = I_amb;
    }
  };
  if (col) {
    T kvs[TS][PER];
#pragma unroll
    for (int kk = 0; kk < TS; ++kk) {
      const int kp = TS * tk + kk;
      azs[kk] = T(0);
      if (kp < Nz) {
        const int k = mu >= 0.0 ? kp : Nz - 1 - kp;
        azs[kk] = T(fabs(mu) * g.inv_dz[k]);
        const uint32_t row = ((uint32_t)k * Ny + j) * Nx;
        load_cells(kap, row, kvs[kk]);
        load_cells(B, row, kB[kk]);
        if constexpr (!FLOW)
          if (w == 0) load_cells(In, ((uint32_t)k * Ny + jy) * Nx, Iyg[kk]);
      }
    }
    if constexpr (!FLOW) load_z_inflow();
    // the inflow-independent part of each row
#pragma unroll
    for (int kk = 0; kk < TS; ++kk) {
      if (TS * tk + kk < Nz) {
        T A = T(1);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          if (ok[q]) {
            const T kv = kvs[kk][q];
            rc[kk][q] = recip(kv + ax + ay + azs[kk]);
            kB[kk][q] = kv * kB[kk][q];
            const T a = ax * rc[kk][q];
            A = a * A;
          }
        }
        Al[kk][0] = A;
        A = A * dpp_mov<0x111, 0xF>(T(1), A);   // row_shr:1
        Al[kk][1] = A;
        A = A * dpp_mov<0x112, 0xF>(T(1), A);   // row_shr:2
        Al[kk][2] = A;
        A = A * dpp_mov<0x114, 0xF>(T(1), A);   // row_shr:4
        Al[kk][3] = A;
        A = A * dpp_mov<0x118, 0xF>(T(1), A);   // row_shr:8
        Al[kk][4] = A;
        A = A * dpp_mov<0x142, 0xA>(T(1), A);   // row_bcast:15 into rows 1, 3
        Al[kk][5] = A;
        A = A * dpp_mov<0x143, 0xC>(T(1), A);   // row_bcast:31 into rows 2, 3
        Al[kk][6] = dpp_mov<0x138, 0xF>(T(1), A);   // wave_shr:1: exclusive prefix
      }
    }
  }
  if constexpr (FLOW) {
    // the blocks this one reads from have been stored, then its inflows
    if (!wait()) return false;
    if (col) {
#pragma unroll
      for (int kk = 0; kk < TS; ++kk) {
        const int kp = TS * tk + kk;
        if (w == 0 && kp < Nz) {
          const int k = mu >= 0.0 ? kp : Nz - 1 - kp;
          load_inflow(In, ((uint32_t)k * Ny + jy) * Nx, Iyg[kk]);
        }
      }
      load_z_inflow();
    }
  }
  // every global inflow is read before any row is stored: the wait for all
  // of the wavefront's loads (s_waitcnt vmcnt(0); the compiler's wait
  // tracking sees it, so no later wait holds a stage on its own stores),
  // then the workgroup barrier
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  // Stage st ends at the workgroup's (st + 1)-th barrier, and wavefront w
  // computes its row r at stage w + r: it passes w barriers, then does its
  // rows in order with a barrier after each, then passes the rest, so every
  // wavefront passes 2 TS - 1 barriers.  The row loop is unrolled, so each
  // row's registers are addressed statically.
  for (int b = 0; b < w; ++b) __syncthreads();
#pragma unroll
  for (int r = 0; r < TS; ++r) {
    const int kp = TS * tk + r;
    if (col && kp < Nz) {
      const int k = mu >= 0.0 ? kp : Nz - 1 - kp;
      const uint32_t row = ((uint32_t)k * Ny + j) * Nx;
      const T* shy = sh + ((w - 1) * TS + r) * Nx;   // wavefront w - 1's row
      T* sho = sh + (w * TS + r) * Nx;
      const T az = azs[r];
      T num0[PER], Iy[PER];
      if (w == 0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) Iy[q] = Iyg[r][q];
      } else if constexpr (FULL) {
        const LC c = *reinterpret_cast<const LC*>(shy + lane * PER);
#pragma unroll
        for (int q = 0; q < PER; ++q) Iy[q] = c.v[q];
      } else {
#pragma unroll
        for (int q = 0; q < PER; ++q)
          if (ok[q]) Iy[q] = shy[lane * PER + q];
      }
      T Bc = T(0);
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (ok[q]) {
          num0[q] = kB[r][q] + ay * Iy[q] + az * Iz[q];
          const T a = ax * rc[r][q];
          const T b = num0[q] * rc[r][q];
          Bc = a * Bc + b;
        }
      }
      // the additive half of the affine scan: x -> A x + Bc composed with
      // the earlier lane's map; lanes without one add A * 0
      Bc = Al[r][0] * dpp_mov<0x111, 0xF>(T(0), Bc) + Bc;
      Bc = Al[r][1] * dpp_mov<0x112, 0xF>(T(0), Bc) + Bc;
      Bc = Al[r][2] * dpp_mov<0x114, 0xF>(T(0), Bc) + Bc;
      Bc = Al[r][3] * dpp_mov<0x118, 0xF>(T(0), Bc) + Bc;
      Bc = Al[r][4] * dpp_mov<0x142, 0xA>(T(0), Bc) + Bc;
      Bc = Al[r][5] * dpp_mov<0x143, 0xC>(T(0), Bc) + Bc;
      const T eB = dpp_mov<0x138, 0xF>(T(0), Bc);
      T Iprev = Al[r][6] * I_amb + eB;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (ok[q]) {
          const T v = (num0[q] + ax * Iprev) * rc[r][q];
          Iz[q] = v;                      // z inflow of the next row up
          Iprev = v;
        }
      }
      // the row to LDS (y inflow of wavefront w + 1) and to the intensities
      if constexpr (FULL) {
        LC c, m;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          c.v[q] = Iz[q];
          m.v[q] = fwd ? Iz[q] : Iz[PER - 1 - q];
        }
        *reinterpret_cast<LC*>(sho + lane * PER) = c;
        at(reinterpret_cast<LC*>(In), (row + i0) * E) = m;
      } else {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          if (ok[q]) {
            sho[lane * PER + q] = Iz[q];
            at(In, (row + ix[q]) * E) = Iz[q];
          }
        }
      }
    }
    __syncthreads();
  }
  for (int b = w; b < TS - 1; ++b) __syncthreads();
  return true;
}

template <typename T, int PER, bool FULL, int TS>
__global__ void __launch_bounds__(64 * TS)
k_dom_sweep_tile(GridArgs g, int D, int tjlo, int M, const double* __restrict__ Om,
                 T* __restrict__ Iset, const T* __restrict__ kap,
                 const T* __restrict__ B, const T* __restrict__ Ig_out, T I_amb,
                 const double* __restrict__ S) {
  if (S[DS_DONE] != 0.0) return;   // uniform: the solve has converged
  (void)M;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sh = reinterpret_cast<T*>(smem_raw);   // TS*TS rows in sweep order: slot (w, kk)
  const int tj = tjlo + blockIdx.x;
  sweep_block<T, PER, FULL, TS, false>(g, blockIdx.y, tj, D - tj, Om, Iset, kap, B, Ig_out,
                                       I_amb, sh, [] { return true; });
}

// G = sum_n w_n I_n (ordinates summed in order), one cell per thread; per
// workgroup partial max |G - G_prev| and max G.  The cells of the ground
// layer (c < ncol) also update the ground from the same loads (dom_3d.py's
// q_in_ground / I_ground_out step): q_in = sum over the downward ordinates of
// w |mu| I, in order, and I_ground_out relaxed toward emit + (1 - eps) q_in / pi.
template <typename T, int V>
__global__ void __launch_bounds__(256)
k_dom_G(size_t n, int M, const T* __restrict__ Iset, const double* __restrict__ w,
        T* __restrict__ G0, T* __restrict__ G1, double* __restrict__ pdiff,
        double* __restrict__ pmax, const double* __restrict__ S,
        const double* __restrict__ Om, int ncol, const T* __restrict__ emit,
        T* __restrict__ Ig_out, T* __restrict__ q_in, T eps, T relax) {
  if (S[DS_DONE] != 0.0) return;   // uniform: the solve has converged
  // this iteration writes G0 or G1 by its parity; the other holds G_prev
  const bool odd = ((long long)S[DS_ITERS] & 1) != 0;
  T* __restrict__ G = odd ? G1 : G0;
  const T* __restrict__ Gp = odd ? G0 : G1;
  __shared__ double sh1[256], sh2[256];
  double md = 0.0, mg = -1e300;
  // V consecutive cells per thread, one V-wide vector per ordinate (the host
  // launches V > 1 only when n and ncol are multiples of V)
  using LC = LaneCells<T, V>;
  const size_t c0 = ((size_t)blockIdx.x * 256 + threadIdx.x) * V;
  if (c0 < n) {
    T s[V], qg[V];
#pragma unroll
    for (int v = 0; v < V; ++v) s[v] = qg[v] = T(0);
    const bool ground = c0 < (size_t)ncol;
    // one ordinate's terms, in ordinate order per cell (the ground sum over
    // the downward ordinates only)
    auto add = [&](int m, const LC& iv) {
      const T wm = T(w[m]);
#pragma unroll
      for (int v = 0; v < V; ++v) s[v] += wm * iv.v[v];
      if (ground) {
        const double mu = Om[3 * m + 2];
        if (mu < 0.0) {
#pragma unroll
          for (int v = 0; v < V; ++v) qg[v] += T(w[m]) * T(fabs(mu)) * iv.v[v];
        }
      }
    };
    const LC* p = reinterpret_cast<const LC*>(Iset + c0);
    const size_t stride = n / V;   // one ordinate's field, in vectors
    int q = 0;
    for (; q + 4 <= M; q += 4) {
      const LC i0 = p[(size_t)q * stride], i1 = p[(size_t)(q + 1) * stride];
      const LC i2 = p[(size_t)(q + 2) * stride], i3 = p[(size_t)(q + 3) * stride];
      add(q, i0);
      add(q + 1, i1);
      add(q + 2, i2);
      add(q + 3, i3);
    }
    for (; q < M; ++q) add(q, p[(size_t)q * stride]);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const size_t c = c0 + v;
      G[c] = s[v];
      if (ground) {
        q_in[c] = qg[v];
        const T Inew = emit[c] + (T(1) - eps) / T(M_PI) * qg[v];
        Ig_out[c] = relax * Inew + (T(1) - relax) * Ig_out[c];
      }
      const double dd = fabs((double)s[v] - (double)Gp[c]);
      if (dd > md || dd != dd) md = dd;
      if ((double)s[v] > mg) mg = (double)s[v];
    }
  }
  const int t = threadIdx.x;
  sh1[t] = md;
  sh2[t] = mg;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      if (sh1[t + st] > sh1[t] || sh1[t + st] != sh1[t + st]) sh1[t] = sh1[t + st];
      if (sh2[t + st] > sh2[t]) sh2[t] = sh2[t + st];
    }
    __syncthreads();
  }
  if (t == 0) { pdiff[blockIdx.x] = sh1[0]; pmax[blockIdx.x] = sh2[0]; }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_dom_out(GridArgs g, size_t n, const T* __restrict__ G, const T* __restrict__ kap,
          const T* __restrict__ kap_s, const T* __restrict__ kap_safe,
          const T* __restrict__ B, T* __restrict__ qs, T* __restrict__ qg) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int k = (int)(c / ((size_t)g.nx * g.ny));
  const T k_ = kap[c];
  const T net = k_ * (G[c] - T(4.0 * M_PI) * B[c]);
  const T net_h = net * T(g.dz[k]);
  T fs = T(0), fg = T(0);
  if (k_ > T(1.0e-9)) {
    fs = kap_s[c] / kap_safe[c];
    fg = (k_ - kap_s[c]) / kap_safe[c];
  }
  qs[c] = net_h * fs;
  qg[c] = net_h * fg;
}

template <typename T>
__global__ void k_scale_copy(int ncol, const T* __restrict__ a, T s,
                             T* __restrict__ out) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c < ncol) out[c] = s * a[c];
}

__global__ void __launch_bounds__(256)
k_dom_fin(const double* __restrict__ pd, const double* __restrict__ pm, int np,
          double* __restrict__ S, int first, double tol, int* __restrict__ flow_q,
          int flow_blocks, int M) {
  if (S[DS_DONE] != 0.0) return;   // uniform: the solve has converged
  __shared__ double a[256], b[256];
  const int t = threadIdx.x;
  double x = 0.0, y = -1e300;
  for (int i = t; i < np; i += 256) {
    if (pd[i] > x || pd[i] != pd[i]) x = pd[i];
    if (pm[i] > y) y = pm[i];
  }
  a[t] = x;
  b[t] = y;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      if (a[t + s] > a[t] || a[t + s] != a[t + s]) a[t] = a[t + s];
      if (b[t + s] > b[t]) b[t] = b[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    // the iteration is complete; from the second on, the change of G
    // relative to its largest value decides convergence
    S[DS_ITERS] += 1.0;
    if (!first) {
      S[DS_DIFF] = a[0];
      S[DS_GMAX] = b[0];
      const double gmax = b[0] > 1e-12 ? b[0] : 1e-12;
      const double err = a[0] / gmax;
      if (!(err == err)) {
        S[DS_NAN] = 1.0;
        S[DS_DONE] = 1.0;
      } else if (err < tol) {
        S[DS_DONE] = 1.0;
      }
    }
    if (flow_q) {
      // the dataflow sweep of this iteration: every XCD's queue handed out
      // all of its blocks and no workgroup gave up; then the next sweep's
      // queues and epoch
      bool stalled = false;
      for (int x = 0; x < FLOW_XCD; ++x) {
        const int items = ((M - x + FLOW_XCD - 1) / FLOW_XCD) * flow_blocks;
        if (flow_q[flow_ctr(x)] < items || flow_q[flow_gave_up(x)] != 0) stalled = true;
      }
      if (stalled) {
        S[DS_STALL] = 1.0;
        S[DS_DONE] = 1.0;
      }
      for (int x = 0; x < FLOW_XCD; ++x) flow_q[flow_ctr(x)] = 0;
      S[DS_EPOCH] += 1.0;
    }
  }
}

// XCD (XCC) of the calling wavefront: the XCC_ID hardware register
__device__ __forceinline__ int xcc_id() {
  // s_getreg_b32 hwreg(HW_REG_XCC_ID = 20, offset 0, 4 bits)
  return __builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & (FLOW_XCD - 1);
}

// Dataflow sweep: one launch per source iteration instead of one per block
// diagonal.  Ordinate n belongs to XCD n % 8, and the blocks of the
// ordinates of one XCD form a queue in diagonal order (its counter).
// A workgroup reads the XCD it runs on, takes the next block of that XCD's
// queue, does its inflow-independent work, waits until the blocks it reads
// from (y and z neighbours, one diagonal earlier) have stored their rows,
// sweeps the block and publishes it: the epoch of this sweep into the
// block's flag once its wavefronts' stores are complete.
//
// Producer and consumer share the XCD and so its L2: stores pass the
// write-through L1 to the L2, and the flags and inflow rows are read at
// device scope (past the L1), so no L2 write-back or invalidation is
// needed.  A block waits only on blocks earlier in its queue, which
// workgroups already running have taken, so the sweep cannot deadlock
// however many workgroups are resident.  Every wait is bounded: after
// spin_limit polls (at most FLOW_SPIN), or once another workgroup of the XCD
// has given up, a workgroup raises the XCD's give-up word and leaves;
// k_dom_fin turns that, or a queue not handed out in full, into DS_STALL.
template <typename T, int PER, int TS>
__global__ void __launch_bounds__(64 * TS)
k_dom_sweep_flow(GridArgs g, int M, const double* __restrict__ Om, T* __restrict__ Iset,
                 const T* __restrict__ kap, const T* __restrict__ B,
                 const T* __restrict__ Ig_out, T I_amb, const double* __restrict__ S,
                 int* __restrict__ flow_q, int* __restrict__ flags, FlowPlan plan,
                 int spin_limit) {
  if (S[DS_DONE] != 0.0) return;   // uniform: the solve has converged
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sh = reinterpret_cast<T*>(smem_raw);
  __shared__ int s_item, s_ok;
  const int x = xcc_id();
  const int epoch = (int)S[DS_EPOCH];
  const int nord = (M - x + FLOW_XCD - 1) / FLOW_XCD;   // ordinates x, x + 8, ...
  const int items = nord * plan.cum[plan.ndiag];
  const int TJ = plan.TJ;
  int D = 0;
  for (;;) {
    if (threadIdx.x == 0)
      s_item = __hip_atomic_fetch_add(&flow_q[flow_ctr(x)], 1, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int it = s_item;
    __syncthreads();   // s_item read by all before the next take
    if (it >= items) return;
    while (it >= nord * plan.cum[D + 1]) ++D;   // items come in increasing order
    const int nt = plan.cum[D + 1] - plan.cum[D];
    const int r = it - nord * plan.cum[D];
    const int o = r / nt;
    const int tjlo = D - (plan.TK - 1) > 0 ? D - (plan.TK - 1) : 0;
    const int tj = tjlo + (r - o * nt), tk = D - tj;
    const int n = x + FLOW_XCD * o;
    int* fl = flags + (size_t)n * plan.fstride;
    auto wait = [&]() -> bool {
      if (threadIdx.x == 0) {
        // one load per poll (the y flag, then the z flag), the give-up word
        // every 256 polls: fewer requests queue in the L2 ahead of the
        // blocks' own loads and stores (0.32 -> 0.23 ms per source iteration
        // at 256 x 64 x 48, S8, against two flags and the word every poll)
        int ok = 1;
        bool ry = tj == 0, rz = tk == 0;
        for (int spin = 0;; ++spin) {
          if (!ry)
            ry = __hip_atomic_load(&fl[tk * TJ + tj - 1], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) >= epoch;
          else if (!rz)
            rz = __hip_atomic_load(&fl[(tk - 1) * TJ + tj], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) >= epoch;
          if (ry && rz) break;
          if (spin >= spin_limit ||
              ((spin & 255) == 255 && __hip_atomic_load(&flow_q[flow_gave_up(x)],
                                                        __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT) != 0)) {
            __hip_atomic_store(&flow_q[flow_gave_up(x)], 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        s_ok = ok;
      }
      __syncthreads();
      return s_ok != 0;
    };
    if (!sweep_block<T, PER, true, TS, true>(g, n, tj, tk, Om, Iset, kap, B, Ig_out, I_amb, sh,
                                             wait))
      return;   // uniform: s_ok
    // every wavefront's stores of the block have reached the L2 (on gfx9
    // vmcnt counts stores too), then the block's flag
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(&fl[tk * TJ + tj], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

static inline unsigned nb(size_t n) { return (unsigned)((n + 255) / 256); }

#define TRY_DOM(call)        \
  do {                       \
    int _rc = (call);        \
    if (_rc) return _rc;     \
  } while (0)

// Tiled sweep launch for block diagonal D (TS*TS rows of LDS per workgroup).
template <typename T, int PER, int TS>
static void launch_tile(DomHandle* h, int D, int tjlo, int ntiles, T I_amb, hipStream_t s) {
  const size_t shm = TS * TS * (size_t)h->g.nx * sizeof(T);
  const dim3 grid(ntiles, h->M);
  // the event-recording launch costs more host time than a plain one: use
  // it only for the launches a timer asked for
  const KernelEvents ev = timer_next(FW_TIMER_DOM_SWEEP);
  auto kern = h->g.nx == 64 * PER ? k_dom_sweep_tile<T, PER, true, TS>
                                  : k_dom_sweep_tile<T, PER, false, TS>;
  if (ev.start)
    hipExtLaunchKernelGGL(kern, grid, dim3(64 * TS), shm, s, ev.start, ev.stop, 0, h->ga, D,
                          tjlo, h->M, (const double*)h->Om, (T*)h->I, (const T*)h->kap_safe,
                          (const T*)h->B, (const T*)h->Ig_out, I_amb, (const double*)h->S);
  else
    hipLaunchKernelGGL(kern, grid, dim3(64 * TS), shm, s, h->ga, D, tjlo, h->M,
                       (const double*)h->Om, (T*)h->I, (const T*)h->kap_safe,
                       (const T*)h->B, (const T*)h->Ig_out, I_amb, (const double*)h->S);
}

// Tile side: 4 x 4 rows when their LDS rows fit 64 KB, else 0 (the per-row
// kernel).  8 x 8 tiles halve the launches but not the time: at 256 x 64 x
// 48, S8, a source iteration is VALU-throughput-bound across the tiles of a
// diagonal (13 launches of 29.2 us against 27 of 13.6 us).
// The handle's requested side (fw_dom_set_tiling: 2, 4 or 8; 0 = 4) is used
// when its TS x TS rows fit 64 KB of LDS (and, past 512 cells a row, only
// the side 4 has a kernel); the per-cell arithmetic does not depend on the
// side, so every side gives the same intensities.
template <typename T>
static int tile_side(const DomHandle* h) {
  const size_t row = (size_t)h->g.nx * sizeof(T);
  const int want = h->tile ? h->tile : 4;
  if (want != 4 && h->g.nx <= 512 && (size_t)want * want * row <= 65536) return want;
  return 16 * row <= 65536 ? 4 : 0;
}

// Whether a (cells per lane, tile side) pair has a kernel: the LDS rows of an
// 8 x 8 tile fit 64 KB only for rows of <= 1 KB (float32 Nx <= 256,
// float64 Nx <= 128), which is what tile_side allows; the other pairs are
// never launched and are not instantiated.
template <typename T, int PER, int TS>
constexpr bool tile_kernel() {
  return TS != 8 || PER * sizeof(T) <= 16;
}

// One source iteration of the tiled sweep (cells per lane as below; a float64
// row of more than 512 cells does not fit the LDS and never gets here).
template <typename T, int TS>
static int sweep_tiled_ts(DomHandle* h, T I_amb, hipStream_t s) {
  const int Nx = h->g.nx;
  const int TJ = (h->g.ny + TS - 1) / TS, TK = (h->g.nz + TS - 1) / TS;
  const int per = Nx <= 64 ? 1 : Nx <= 128 ? 2 : Nx <= 256 ? 4 : Nx <= 512 ? 8 : 16;
  auto go = [&](auto PC, int D, int tjlo, int nt) -> int {
    constexpr int PER = decltype(PC)::value;
    if constexpr (tile_kernel<T, PER, TS>() && (PER < 16 || (sizeof(T) == 4 && TS == 4))) {
      launch_tile<T, PER, TS>(h, D, tjlo, nt, I_amb, s);
      return FW_OK;
    } else {
      return fw_set_error(FW_E_CONFIG, "dom sweep: no %dx%d tile kernel for rows of %d cells",
                          TS, TS, Nx);
    }
  };
  for (int D = 0; D <= TJ + TK - 2; ++D) {
    const int tjlo = D - (TK - 1) > 0 ? D - (TK - 1) : 0;
    const int tjhi = D < TJ - 1 ? D : TJ - 1;
    const int nt = tjhi - tjlo + 1;
    int rc;
    switch (per) {
      case 1: rc = go(std::integral_constant<int, 1>{}, D, tjlo, nt); break;
      case 2: rc = go(std::integral_constant<int, 2>{}, D, tjlo, nt); break;
      case 4: rc = go(std::integral_constant<int, 4>{}, D, tjlo, nt); break;
      case 8: rc = go(std::integral_constant<int, 8>{}, D, tjlo, nt); break;
      default: rc = go(std::integral_constant<int, 16>{}, D, tjlo, nt); break;
    }
    if (rc) return rc;
  }
  return FW_OK;
}

template <typename T>
int tile_side_of(const DomHandle* h) { return tile_side<T>(h); }

// Whether the dataflow sweep runs this handle's source iterations: built for
// the grid (whole rows of 64 PER cells, 4 x 4 blocks in LDS, a plan that
// fits), not switched off, and the tile side 4 in use.
template <typename T>
bool flow_active(const DomHandle* h) {
  return h->flow == 1 && tile_side<T>(h) == 4 && flow_rows(h->g.nx, sizeof(T));
}

// One source iteration's sweep as one dataflow launch: 128 workgroups per
// XCD (four per CU), each taking blocks from its XCD's queue.
template <typename T>
static int sweep_flow(DomHandle* h, T I_amb, hipStream_t s) {
  const int Nx = h->g.nx;
  const size_t shm = 16 * (size_t)Nx * sizeof(T);
  const dim3 grid(FLOW_XCD * h->flow_wg);
  const KernelEvents ev = timer_next(FW_TIMER_DOM_SWEEP);
  auto go = [&](auto kern) {
    if (ev.start)
      hipExtLaunchKernelGGL(kern, grid, dim3(256), shm, s, ev.start, ev.stop, 0, h->ga, h->M,
                            (const double*)h->Om, (T*)h->I, (const T*)h->kap_safe,
                            (const T*)h->B, (const T*)h->Ig_out, I_amb, (const double*)h->S,
                            h->flow_q, h->flow_flags, h->plan, h->flow_spin);
    else
      hipLaunchKernelGGL(kern, grid, dim3(256), shm, s, h->ga, h->M, (const double*)h->Om,
                         (T*)h->I, (const T*)h->kap_safe, (const T*)h->B,
                         (const T*)h->Ig_out, I_amb, (const double*)h->S, h->flow_q,
                         h->flow_flags, h->plan, h->flow_spin);
  };
  switch (Nx) {
    case 64: go(k_dom_sweep_flow<T, 1, 4>); break;
    case 128: go(k_dom_sweep_flow<T, 2, 4>); break;
    case 256: go(k_dom_sweep_flow<T, 4, 4>); break;
    case 512: go(k_dom_sweep_flow<T, 8, 4>); break;
    default: return fw_set_error(FW_E_CONFIG, "dom: no dataflow sweep for rows of %d cells", Nx);
  }
  FW_CHECK_LAUNCH("dom_sweep_flow");
  return FW_OK;
}

template <typename T>
static int sweep_tiled(DomHandle* h, T I_amb, hipStream_t s) {
  if (flow_active<T>(h)) return sweep_flow<T>(h, I_amb, s);
  switch (tile_side<T>(h)) {
    case 2: return sweep_tiled_ts<T, 2>(h, I_amb, s);
    case 8: return sweep_tiled_ts<T, 8>(h, I_amb, s);
    default: return sweep_tiled_ts<T, 4>(h, I_amb, s);
  }
}

// Cells per lane: ceil(Nx / 64) rounded up to a power of two (Nx <= 2048).
template <typename T>
static int launch_sweep(DomHandle* h, int d, int jlo, int nrows, T I_amb, hipStream_t s) {
  const int Nx = h->g.nx, M = h->M;
  const dim3 grid(nrows, (M + 3) / 4), blk(256);
  const double* Om = (const double*)h->Om;
  T* I = (T*)h->I;
  const T* kap = (const T*)h->kap_safe;
  const T* B = (const T*)h->B;
  const T* Ig = (const T*)h->Ig_out;
  if (Nx <= 64)
    hipLaunchKernelGGL((k_dom_sweep_diag<T, 1>), grid, blk, 0, s, h->ga, d, jlo, M, Om, I, kap, B, Ig, I_amb, h->S);
  else if (Nx <= 128)
    hipLaunchKernelGGL((k_dom_sweep_diag<T, 2>), grid, blk, 0, s, h->ga, d, jlo, M, Om, I, kap, B, Ig, I_amb, h->S);
  else if (Nx <= 256)
    hipLaunchKernelGGL((k_dom_sweep_diag<T, 4>), grid, blk, 0, s, h->ga, d, jlo, M, Om, I, kap, B, Ig, I_amb, h->S);
  else if (Nx <= 512)
    hipLaunchKernelGGL((k_dom_sweep_diag<T, 8>), grid, blk, 0, s, h->ga, d, jlo, M, Om, I, kap, B, Ig, I_amb, h->S);
  else if (Nx <= 1024)
    hipLaunchKernelGGL((k_dom_sweep_diag<T, 16>), grid, blk, 0, s, h->ga, d, jlo, M, Om, I, kap, B, Ig, I_amb, h->S);
  else
    hipLaunchKernelGGL((k_dom_sweep_diag<T, 32>), grid, blk, 0, s, h->ga, d, jlo, M, Om, I, kap, B, Ig, I_amb, h->S);
  return FW_OK;
}

template <typename T>
static int dom_solve_impl(DomHandle* h, const void* Ts, const void* Tg,
                          const void* alpha, const void* omega, double sigma_sav,
                          double T_amb, void* qs, void* qg, const void* Tsoil,
                          void* q_soil, const void* Yw, const void* rho,
                          const void* Mloc, hipStream_t s) {
  const size_t n = h->n;
  const int Nx = h->g.nx, Ny = h->g.ny, Nz = h->g.nz, M = h->M;
  const int ncol = Nx * Ny;
  const T I_amb = T(5.67e-8 * (T_amb * T_amb) * (T_amb * T_amb) / M_PI);
  hipLaunchKernelGGL(k_dom_props<T>, dim3(nb(n)), dim3(256), 0, s, n, (const T*)Ts,
                     (const T*)Tg, (const T*)alpha, (const T*)omega, (const T*)Yw,
                     (const T*)rho, (const T*)Mloc, T(sigma_sav), T(T_amb),
                     T(h->kappa_gas_max), (T*)h->kap, (T*)h->kap_s,
                     (T*)h->kap_safe, (T*)h->B);
  hipLaunchKernelGGL(k_dom_ground_init<T>, dim3(nb(ncol)), dim3(256), 0, s, ncol,
                     (const T*)Tsoil, T(h->eps_ground), I_amb, (T*)h->Ig_emit,
                     (T*)h->Ig_out, h->S, h->flow_q);
  FW_CHECK_LAUNCH("dom_props");
  // One source iteration: the sweep, G with its change and the ground
  // update, and the change's final sum with the convergence test.  Every kernel of
  // it returns at once once the solve has converged.  Its launches depend
  // only on whether it is the first, so off the legacy default stream and
  // with the sweep timer off it is replayed from a captured graph.
  const bool vec_G = n % 4 == 0 && ncol % 4 == 0;
  auto iteration = [&](int first) -> int {
    if (tile_side<T>(h) > 0) {
      TRY_DOM(sweep_tiled<T>(h, I_amb, s));
    } else {
      for (int d = 0; d <= Ny + Nz - 2; ++d) {
        const int jlo = d - (Nz - 1) > 0 ? d - (Nz - 1) : 0;
        const int jhi = d < Ny - 1 ? d : Ny - 1;
        const int nrows = jhi - jlo + 1;
        TRY_DOM(launch_sweep<T>(h, d, jlo, nrows, I_amb, s));
      }
    }
    FW_CHECK_LAUNCH("dom_sweep");
    // four cells per thread in 16-byte (float) vectors when the fields allow
    if (vec_G) {
      hipLaunchKernelGGL((k_dom_G<T, 4>), dim3(nb(n / 4)), dim3(256), 0, s, n, M,
                         (const T*)h->I, (const double*)h->w, (T*)h->G, (T*)h->Gp,
                         h->part, h->part2, (const double*)h->S, (const double*)h->Om, ncol,
                         (const T*)h->Ig_emit, (T*)h->Ig_out, (T*)h->q_in, T(h->eps_ground),
                         T(h->relax));
    } else {
      hipLaunchKernelGGL((k_dom_G<T, 1>), dim3(nb(n)), dim3(256), 0, s, n, M,
                         (const T*)h->I, (const double*)h->w, (T*)h->G, (T*)h->Gp,
                         h->part, h->part2, (const double*)h->S, (const double*)h->Om, ncol,
                         (const T*)h->Ig_emit, (T*)h->Ig_out, (T*)h->q_in, T(h->eps_ground),
                         T(h->relax));
    }
    hipLaunchKernelGGL(k_dom_fin, dim3(1), dim3(256), 0, s, h->part, h->part2,
                       (int)nb(vec_G ? n / 4 : n), h->S, first, h->tol,
                       flow_active<T>(h) ? h->flow_q : (int*)nullptr, h->plan.cum[h->plan.ndiag],
                       M);
    FW_CHECK_LAUNCH("dom_G");
    return FW_OK;
  };
  const bool graph = h->gc && s && !timer_on(FW_TIMER_DOM_SWEEP) && loop_graphs_on();
  auto run_iteration = [&](int first) -> int {
    if (!graph) return iteration(first);
    // I_amb rides in the key: the launches take it by value; recording:
    // iteration() issues to s, the recording stream meanwhile
    return h->gc->run({(uint64_t)first, kw((double)I_amb)}, s, [&](hipStream_t cs) {
      const hipStream_t keep = s;
      s = cs;
      const int r2 = iteration(first);
      s = keep;
      return r2;
    });
  };
  // Iterations in batches with one host read per batch: the previous solve's
  // count first (the count changes slowly from solve to solve; the first
  // iteration never converges, so at least two), then one at a time.  The
  // device flag stops every kernel after the converging iteration, so the
  // result is that of a one-by-one loop.  An iteration queued past
  // convergence costs its launches, which return at once (a few us); one
  // queued short costs a read and the host's turnaround.
  int queued = 0;
  int batch = h->last_iters > 2 ? h->last_iters : 2;
  for (;;) {
    for (int b = 0; b < batch && queued < h->maxiter; ++b, ++queued)
      TRY_DOM(run_iteration(queued == 0));
    FW_CHECK_HIP(hipMemcpyAsync(h->S_host, h->S, DS_N * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    FW_CHECK_HIP(hipStreamSynchronize(s));
    if (h->S_host[DS_NAN] != 0.0) return fw_set_error(FW_E_NONFINITE, "dom: non-finite G");
    if (h->S_host[DS_STALL] != 0.0) {
      h->flow = -1;   // the per-diagonal launches from now on
      return fw_set_error(FW_E_HIP, "dom: the dataflow sweep stalled (switched off for this "
                                    "handle; the solve's intensities are invalid)");
    }
    if (h->S_host[DS_DONE] != 0.0 || queued >= h->maxiter) break;
    batch = 1;
  }
  h->last_iters = (int)h->S_host[DS_ITERS];
  hipLaunchKernelGGL(k_dom_out<T>, dim3(nb(n)), dim3(256), 0, s, h->ga, n,
                     (const T*)(h->last_iters > 0 && ((h->last_iters - 1) & 1) ? h->Gp : h->G),
                     (const T*)h->kap, (const T*)h->kap_s,
                     (const T*)h->kap_safe, (const T*)h->B, (T*)qs, (T*)qg);
  if (q_soil)
    hipLaunchKernelGGL(k_scale_copy<T>, dim3(nb(ncol)), dim3(256), 0, s, ncol,
                       (const T*)h->q_in, T(h->eps_ground), (T*)q_soil);
  FW_CHECK_LAUNCH("dom_out");
  return FW_OK;
}

// Sum over forward ordinates (xi > 0) of w |xi| I_n in the masked cells
// (flame_front_3d.py:594-630); mask is uint8 (Nz, Ny, Nx).
template <typename T>
__global__ void __launch_bounds__(256)
k_dom_fwd(size_t n, int M, const double* __restrict__ Om, const double* __restrict__ w,
          const T* __restrict__ Iset, const unsigned char* __restrict__ mask,
          T* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  T s = T(0);
  if (!mask || mask[c]) {
    for (int q = 0; q < M; ++q) {
      const double xi = Om[3 * q];
      if (xi > 0.0) s += T(w[q]) * T(xi) * Iset[(size_t)q * n + c];
    }
  }
  out[c] = s;
}

int dom_fork(void* dom, hipStream_t s, hipStream_t* side) {
  DomHandle* h = (DomHandle*)dom;
  FW_REQUIRE(h && side, "dom_fork: null argument");
  if (!h->side) {
    LegacyScope legacy_scope;
    FW_CHECK_HIP(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    FW_CHECK_HIP(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    FW_CHECK_HIP(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
  }
  FW_CHECK_HIP(hipEventRecord(h->ev_fork, s));
  FW_CHECK_HIP(hipStreamWaitEvent(h->side, h->ev_fork, 0));
  *side = h->side;
  return FW_OK;
}

int dom_join_record(void* dom) {
  DomHandle* h = (DomHandle*)dom;
  FW_CHECK_HIP(hipEventRecord(h->ev_join, h->side));
  return FW_OK;
}

int dom_join(void* dom, hipStream_t s) {
  DomHandle* h = (DomHandle*)dom;
  FW_CHECK_HIP(hipStreamWaitEvent(s, h->ev_join, 0));
  return FW_OK;
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_dom_create(const fw_grid* g, int M, const double* Omega, const double* weights,
                  double eps_w_ground, int max_source_iter, double tol_source_iter,
                  double omega_relax, double kappa_gas_max, void** handle) {
  LegacyScope legacy_scope;
  FW_GRID(g);
  FW_GRID_OFFSETS32(g);
  FW_REQUIRE(g && Omega && weights && handle && M > 0, "dom_create: bad argument");
  FW_REQUIRE(g->nx <= 64 * 32, "dom_create: Nx > 2048 not supported by the row scan");
  DomHandle* h = new DomHandle();
  h->gc = new GraphCache();
  h->g = *g;
  h->ga = grid_args(g);
  h->M = M;
  h->maxiter = max_source_iter;
  h->eps_ground = eps_w_ground;
  h->tol = tol_source_iter;
  h->relax = omega_relax;
  h->kappa_gas_max = kappa_gas_max;
  h->n = (size_t)g->nx * g->ny * g->nz;
  h->esz = g->dtype == FW_F64 ? 8 : 4;
  h->Om_h.assign(Omega, Omega + 3 * M);
  h->w_h.assign(weights, weights + M);
  FW_CHECK_HIP(hipMalloc((void**)&h->Om, 3 * M * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->w, M * sizeof(double)));
  FW_CHECK_HIP(hipMemcpy(h->Om, Omega, 3 * M * sizeof(double), hipMemcpyHostToDevice));
  FW_CHECK_HIP(hipMemcpy(h->w, weights, M * sizeof(double), hipMemcpyHostToDevice));
  const size_t n = h->n, ncol = (size_t)g->nx * g->ny;
  FW_CHECK_HIP(hipMalloc(&h->I, (size_t)M * n * h->esz));
  FW_CHECK_HIP(hipMemset(h->I, 0, (size_t)M * n * h->esz));
  void** f[] = {&h->G, &h->Gp, &h->kap, &h->kap_s, &h->kap_safe, &h->B};
  for (void** p : f) {
    FW_CHECK_HIP(hipMalloc(p, n * h->esz));
    FW_CHECK_HIP(hipMemset(*p, 0, n * h->esz));
  }
  void** c2[] = {&h->Ig_out, &h->Ig_emit, &h->q_in};
  for (void** p : c2) FW_CHECK_HIP(hipMalloc(p, ncol * h->esz));
  const size_t npart = (n + 255) / 256;
  FW_CHECK_HIP(hipMalloc((void**)&h->part, npart * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->part2, npart * sizeof(double)));
  FW_CHECK_HIP(hipMalloc((void**)&h->S, DS_N * sizeof(double)));
  FW_CHECK_HIP(hipMemset(h->S, 0, DS_N * sizeof(double)));
  FW_CHECK_HIP(hipHostMalloc((void**)&h->S_host, DS_N * sizeof(double)));
  // the dataflow sweep's plan and state: 4 x 4 blocks of whole rows
  const int TJ = (g->ny + 3) / 4, TK = (g->nz + 3) / 4;
  if (flow_rows(g->nx, h->esz) && TJ + TK - 1 <= FLOW_MAXD) {
    h->plan.TJ = TJ;
    h->plan.TK = TK;
    h->plan.ndiag = TJ + TK - 1;
    h->plan.cum[0] = 0;
    for (int D = 0; D < h->plan.ndiag; ++D) {
      const int lo = D - (TK - 1) > 0 ? D - (TK - 1) : 0, hi = D < TJ - 1 ? D : TJ - 1;
      h->plan.cum[D + 1] = h->plan.cum[D] + hi - lo + 1;
    }
    h->plan.fstride = (TJ * TK + FLOW_LINE - 1) / FLOW_LINE * FLOW_LINE;
    const size_t nflags = (size_t)M * h->plan.fstride;
    FW_CHECK_HIP(hipMalloc((void**)&h->flow_q, FLOW_LINE * FLOW_XCD * sizeof(int)));
    FW_CHECK_HIP(hipMemset(h->flow_q, 0, FLOW_LINE * FLOW_XCD * sizeof(int)));
    FW_CHECK_HIP(hipMalloc((void**)&h->flow_flags, nflags * sizeof(int)));
    FW_CHECK_HIP(hipMemset(h->flow_flags, 0, nflags * sizeof(int)));
    const char* e = getenv("FIREWAVE_DOM_FLOW");
    h->flow = (e && e[0] == '0') ? -1 : 1;
    // the queues assume the MI355X's 8 XCDs (a partitioned device shows fewer)
    int dev = 0, nxcc = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess ||
        nxcc != FLOW_XCD)
      h->flow = 0;   // not available on this device (fw_dom_set_flow cannot turn it on)
    const char* wg = getenv("FIREWAVE_DOM_FLOW_WG");
    h->flow_wg = wg ? atoi(wg) : 128;
    if (h->flow_wg < 8) h->flow_wg = 8;
    if (h->flow_wg > 512) h->flow_wg = 512;
    // a smaller poll limit makes the give-up path testable from the host
    const char* sp = getenv("FIREWAVE_DOM_FLOW_SPIN");
    h->flow_spin = sp ? atoi(sp) : FLOW_SPIN;
    if (h->flow_spin < 1 || h->flow_spin > FLOW_SPIN) h->flow_spin = FLOW_SPIN;
  }
  *handle = h;
  return FW_OK;
}

int fw_dom_destroy(void* handle) {
  LegacyScope legacy_scope;
  DomHandle* h = (DomHandle*)handle;
  if (!h) return FW_OK;
  void* b[] = {h->Om, h->w, h->I, h->G, h->Gp, h->kap, h->kap_s, h->kap_safe, h->B,
               h->Ig_out, h->Ig_emit, h->q_in, h->part, h->part2, h->S, h->flow_q,
               h->flow_flags};
  for (void* p : b)
    if (p) (void)hipFree(p);
  if (h->S_host) (void)hipHostFree(h->S_host);
  delete h->gc;
  if (h->side) {
    (void)hipStreamSynchronize(h->side);
    (void)hipEventDestroy(h->ev_fork);
    (void)hipEventDestroy(h->ev_join);
    (void)hipStreamDestroy(h->side);
  }
  delete h;
  return FW_OK;
}

int fw_dom_solve(void* handle, const void* T_s, const void* T_g, const void* alpha_s,
                 const void* omega_comb, double sigma_sav, double T_amb,
                 void* q_rad_solid_out, void* q_rad_gas_out, const void* T_soil_surface,
                 void* q_in_soil_out, const void* Y_H2O, const void* rho,
                 const void* bed_moisture, void* stream) {
  DomHandle* h = (DomHandle*)handle;
  FW_REQUIRE(h && T_s && T_g && alpha_s && omega_comb && q_rad_solid_out && q_rad_gas_out,
             "dom_solve: null argument");
  hipStream_t s = (hipStream_t)stream;
  const void* Yw = (Y_H2O && rho) ? Y_H2O : nullptr;
  const void* rh = (Y_H2O && rho) ? rho : nullptr;
  return h->g.dtype == FW_F64
             ? dom_solve_impl<double>(h, T_s, T_g, alpha_s, omega_comb, sigma_sav, T_amb,
                                      q_rad_solid_out, q_rad_gas_out, T_soil_surface,
                                      q_in_soil_out, Yw, rh, bed_moisture, s)
             : dom_solve_impl<float>(h, T_s, T_g, alpha_s, omega_comb, sigma_sav, T_amb,
                                     q_rad_solid_out, q_rad_gas_out, T_soil_surface,
                                     q_in_soil_out, Yw, rh, bed_moisture, s);
}

int fw_dom_set_tiling(void* handle, int tile_side, int* active) {
  DomHandle* h = (DomHandle*)handle;
  FW_REQUIRE(h, "dom_set_tiling: null handle");
  FW_REQUIRE(tile_side == 0 || tile_side == 2 || tile_side == 4 || tile_side == 8,
             "dom_set_tiling: tile side must be 0 (default), 2, 4 or 8");
  h->tile = tile_side;
  if (active)
    *active = h->g.dtype == FW_F64 ? tile_side_of<double>(h) : tile_side_of<float>(h);
  return FW_OK;
}

int fw_dom_set_flow(void* handle, int enable, int* active) {
  DomHandle* h = (DomHandle*)handle;
  FW_REQUIRE(h, "dom_set_flow: null handle");
  if (h->flow != 0) h->flow = enable ? 1 : -1;
  if (active)
    *active = h->g.dtype == FW_F64 ? (int)flow_active<double>(h) : (int)flow_active<float>(h);
  return FW_OK;
}

int fw_dom_sweep_launches(void* handle, int* launches) {
  DomHandle* h = (DomHandle*)handle;
  FW_REQUIRE(h && launches, "dom_sweep_launches: null argument");
  const int ts = h->g.dtype == FW_F64 ? tile_side_of<double>(h) : tile_side_of<float>(h);
  const bool fl = h->g.dtype == FW_F64 ? flow_active<double>(h) : flow_active<float>(h);
  if (fl)
    *launches = 1;
  else if (ts > 0)
    *launches = (h->g.ny + ts - 1) / ts + (h->g.nz + ts - 1) / ts - 1;
  else
    *launches = h->g.ny + h->g.nz - 1;
  return FW_OK;
}

int fw_dom_stats(void* handle, int* iters) {
  DomHandle* h = (DomHandle*)handle;
  FW_REQUIRE(h, "dom_stats: null handle");
  if (iters) *iters = h->last_iters;
  return FW_OK;
}

int fw_dom_intensity(void* handle, void** I_set) {
  DomHandle* h = (DomHandle*)handle;
  FW_REQUIRE(h && I_set, "dom_intensity: null argument");
  *I_set = h->I;
  return FW_OK;
}

int fw_dom_forward_flux(void* handle, const unsigned char* mask, void* out, void* stream) {
  DomHandle* h = (DomHandle*)handle;
  FW_REQUIRE(h && out, "dom_forward_flux: null argument");
  hipStream_t s = (hipStream_t)stream;
  if (h->g.dtype == FW_F64)
    hipLaunchKernelGGL(k_dom_fwd<double>, dim3(nb(h->n)), dim3(256), 0, s, h->n, h->M,
                       (const double*)h->Om, (const double*)h->w, (const double*)h->I,
                       mask, (double*)out);
  else
    hipLaunchKernelGGL(k_dom_fwd<float>, dim3(nb(h->n)), dim3(256), 0, s, h->n, h->M,
                       (const double*)h->Om, (const double*)h->w, (const float*)h->I,
                       mask, (float*)out);
  FW_CHECK_LAUNCH("dom_forward_flux");
  return FW_OK;
}

}  // extern "C"
