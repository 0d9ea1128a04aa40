// Native driver of one outer time step (spread_3d.py:1897-3683 of the
// reference), for the k-epsilon / EDC / DOM configuration with a Lagrangian
// (particle) or Eulerian fuel bed and a passive or an active level set.
//
// The whole step is issued from here, so the host cost per kernel is one
// hipLaunchKernelGGL instead of an interpreter round trip.  Host-side
// decisions are the reference's scalar ones: dt from the previous step's
// fused reduction, the projection's divergence loop, the turbulent
// diffusion sub-step count from max nu_t, the level-set front rows after
// each evolve or reinitialisation, and the exits.  Ping-pong fields are swapped by
// exchanging pointers inside fw_spread_fields; the caller re-reads them.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <thread>

#include "fw_common.h"
#include "graph.h"
#include "gas_solid.h"
#include "edc_cell.h"

namespace fw {

template <typename T>
__global__ void __launch_bounds__(256) k_clamp_max(size_t n, T* __restrict__ f, T hi, T lo) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  T v = f[c];
  if (v > hi) v = hi;   // np.minimum(T_g, 1e4), then np.maximum(T_g, T_amb)
  if (v < lo) v = lo;   // (NaN passes through both)
  f[c] = v;
}

// apply_turbulent_diffusion's sub-step count (turbulence_3d.py:879-918);
// 0 = skip.
static int turb_substeps(double nu_t_max, double sc_t, double dt, double dx, double dy,
                         double dz_min, double* dt_sub) {
  const double h_min2 = fmin(dx * dx, fmin(dy * dy, dz_min * dz_min));
  const double D = nu_t_max / sc_t;
  *dt_sub = 0.0;
  if (D <= 1.0e-12) return 0;
  if (!isfinite(D)) return 0;
  const double dt_max = 0.4 * h_min2 / D;
  if (dt_max <= 0.0) return 0;
  const double target = dt / dt_max;
  if (target > 1000.0) return 0;
  int n = (int)ceil(target);
  if (n < 1) n = 1;
  *dt_sub = dt / n;
  return n;
}

#define SWAP(a, b)     \
  do {                 \
    void* _t = (a);    \
    (a) = (b);         \
    (b) = _t;          \
  } while (0)

#define TRY(call)            \
  do {                       \
    int _rc = (call);        \
    if (_rc) return _rc;     \
  } while (0)

// The step's DOM solve, run beside drag .. k-epsilon on the DOM handle's own
// stream.  The solve waits on the host for each source iteration's residual,
// so it runs from a worker thread and the step's thread keeps issuing the
// flow work meanwhile.  Its inputs (T_s, T_g, alpha_s, omega, Y_H2O, rho,
// M_local, T_soil) are final once the bed has stepped, and nothing before
// the soil conduction reads what it writes (q_rad, q_rad_gas, q_in_soil and
// the intensities), so the results are those of the solve issued in line
// after k-epsilon.  FIREWAVE_DOM_BESIDE=0 issues it in line instead.
static bool dom_beside_on() {
  const char* e = getenv("FIREWAVE_DOM_BESIDE");
  return !(e && e[0] == '0');
}

// The next sub-step's EDC fused into the temperature pass (same results as
// its own launch); FIREWAVE_FUSE_EDC=0 launches it on its own.
static bool fuse_edc_on() {
  const char* e = getenv("FIREWAVE_FUSE_EDC");
  return !(e && e[0] == '0');
}

struct DomBeside {
  std::thread th;
  void* dom = nullptr;
  int rc = FW_OK;
  char msg[1024] = "";

  template <typename Solve>
  int start(void* dom_handle, hipStream_t s, Solve solve) {
    hipStream_t side = nullptr;
    int dev = 0;
    FW_CHECK_HIP(hipGetDevice(&dev));
    TRY(dom_fork(dom_handle, s, &side));
    dom = dom_handle;
    void* timers = timer_share();
    th = std::thread([this, dev, side, timers, solve]() {
      timer_adopt(timers);
      int e = hipSetDevice(dev) == hipSuccess
                  ? solve(side)
                  : fw_set_error(FW_E_HIP, "spread_step: DOM worker cannot select device %d",
                                 dev);
      if (e == FW_OK) e = dom_join_record(dom);
      if (e != FW_OK) snprintf(msg, sizeof(msg), "%s", fw_last_error());
      rc = e;
      timer_adopt(nullptr);
    });
    return FW_OK;
  }
  // s waits for the solve; the worker's error becomes this thread's
  int join(hipStream_t s) {
    if (!th.joinable()) return FW_OK;
    th.join();
    if (rc != FW_OK) return fw_set_error(rc, "%s", msg);
    return dom_join(dom, s);
  }
  ~DomBeside() {
    if (th.joinable()) th.join();
  }
};

static const double SC_T = 0.7, PR_T = 0.85;
static const double K_SOLID_GRASS = 0.20;     // solid_conduction_3d.py:52
static const double RHO_SOLID = 500.0, CP_SOLID = 1300.0;  // coupling_3d.py:50-51
static const double Q_DRIP_PER_AREA = 30000.0, F_DRIP_TO_SOLID = 0.80;
static const double Q_RAD_MAX = 1.0e5, Q_PULSE = 240000.0, T_CAP = 10000.0;
static const double T_BOOTSTRAP_S = 2.0;
static const int N_SUB = 10, REINIT_INTERVAL = 10, REINIT_SUBSTEPS = 5;

__global__ void k_set_step_scalars(double* __restrict__ sb, double dt_sub, double relax,
                                   double dsc, double dpr) {
  if (threadIdx.x == 0) {
    sb[SB_DT_SUB] = dt_sub;
    sb[SB_EDC_RELAX] = relax;
    sb[SB_DSC] = dsc;
    sb[SB_DPR] = dpr;
  }
}

// The largest interior nu_t, for the sub-step plan, read without a stall:
// the partial maxima are reduced on the device into red_out[0] (scratch until
// the end-of-step reduction) and copied into the pinned red_host[0] behind an
// event as soon as k-epsilon has written nu_t; the host waits on the event
// only when it plans the sub-steps, after issuing the work in between.
__global__ void __launch_bounds__(256)
k_max_partials(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double sh[4];
  double m = 0.0;   // the host's fold over the partials started from 0 as well
  for (int b = threadIdx.x; b < np; b += 256) {
    const double x = part[b];
    if (x > m) m = x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(m, off, 64);
    if (o > m) m = o;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0];
    for (int q = 1; q < 4; ++q)
      if (sh[q] > a) a = sh[q];
    out[0] = a;
  }
}

struct NuMaxRead {
  hipEvent_t ev = nullptr;
  int dev = -1;
};
static thread_local NuMaxRead g_numax;

static int nu_max_issue(const fw_grid* g, fw_spread_fields* f, hipStream_t s) {
  int dev = 0;
  FW_CHECK_HIP(hipGetDevice(&dev));
  if (g_numax.ev && g_numax.dev != dev) {
    (void)hipEventDestroy(g_numax.ev);
    g_numax.ev = nullptr;
  }
  if (!g_numax.ev) {
    FW_CHECK_HIP(hipEventCreateWithFlags(&g_numax.ev, hipEventDisableTiming));
    g_numax.dev = dev;
  }
  TRY(nu_t_max_partials(g, f->nu_t, f->nu_work, s));
  hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, s, (const double*)f->nu_work,
                     FW_REDUCE_BLOCKS, f->red_out);
  FW_CHECK_LAUNCH("nu_t max");
  FW_CHECK_HIP(hipMemcpyAsync(f->red_host, f->red_out, sizeof(double), hipMemcpyDeviceToHost,
                              s));
  FW_CHECK_HIP(hipEventRecord(g_numax.ev, s));
  return FW_OK;
}

static int nu_max_wait(const fw_spread_fields* f, double* nu_max) {
  FW_CHECK_HIP(hipEventSynchronize(g_numax.ev));
  *nu_max = f->red_host[0];
  return FW_OK;
}

// What the sub-step loop of a step depends on besides the fields.
struct SubstepPlan {
  int nsc, npr;              // turbulent diffusion passes (species, T_g)
  double dt_sub, dsc, dpr;   // sub-step and pass lengths
  double drip_gas, drip_solid;
  bool fuse_summary;
  bool first_edc_done;       // the first sub-step's EDC ran in the gas-energy pass
};

// The N_SUB chemistry / transport / coupling sub-steps (spread_3d.py:2678-3200)
// on the fields of f, swapping f's ping-pong pointers as it goes.  launch =
// false only advances the pointers: the replay of a captured graph of the
// same launches.  sb (may be null): the kernels read dt_sub, the EDC
// relaxation and the pass lengths from the step scalar block (the particle
// path only; the Eulerian coupling takes them by value).
static int substeps(const fw_grid* g, fw_spread_fields* f, const fw_spread_params* p,
                    const SubstepPlan& P, const double* sb, bool launch, hipStream_t s) {
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const int nzb = g->n_z_bed, s0 = p->i_src0, s1 = p->i_src1;
  const double dt_sub = P.dt_sub;
  void** Y[3] = {&f->Y_fuel, &f->Y_O2, &f->Y_H2O};
  const double y_in[3] = {0.0, 0.232, 0.0}, y_max[3] = {1.0, 0.232, 1.0};
  // the last species diffusion pass of a sub-step leaves the next EDC's
  // segment summaries (nothing between them changes the species) when the
  // launch rows are whole 256-cell segments
  bool summary_ready = false;
  // the EDC of sub-step sub + 1 rides in sub-step sub's last temperature
  // pass when that pass exists and the species pass left the summaries:
  // pointwise per cell (and per segment for the carried flag), it reads the
  // new temperature and the species, which nothing in between changes
  // Measured: the fused pass is faster at 256 x 64 x 48 (1.46 -> 1.43 ms per
  // step) and slower at 512 x 128 x 64 (6.22 -> 6.34 ms), where the workgroup
  // barriers of the EDC hold the stencil's loads; so only up to 2^21 cells
  const bool fuse_edc = P.fuse_summary && P.npr > 0 && fuse_edc_on() && n <= ((size_t)1 << 21);
  const EdcFuseArgs edc_args{f->rho, f->k, f->eps, nullptr, nullptr, nullptr, f->omega,
                             (const int*)f->edc_work,
                             edc_params(0.34, 1100.0, dt_sub, 1, 0, 1.3, 17.0e6)};
  bool edc_done = P.first_edc_done;
  for (int sub = 0; sub < N_SUB; ++sub) {
    if (launch && !edc_done)
      TRY(edc_launch(g, f->rho, f->T_g, f->Y_fuel, f->Y_O2, f->k, f->eps, f->Y_H2O, f->omega,
                     0.34, 1100.0, dt_sub, 1, 0, 1.3, 17.0e6, (int*)f->edc_work, summary_ready,
                     s, sb));
    // species with the dilution sources; the Y_O2 cap of :3059 rides in
    // y_max, the [0, 1] clip is the kernel's own
    {
      const void* yin[3] = {*Y[0], *Y[1], *Y[2]};
      void* yout[3] = {f->tmp1, f->tmp2, f->tmp3};
      if (launch)
        TRY(species3_launch(g, yin, yout, f->rho, f->u, f->v, f->w, f->src[0], f->src[1],
                            dt_sub, 1.0e-5, y_in, y_max, s, sb));
      SWAP(*Y[0], f->tmp1);
      SWAP(*Y[1], f->tmp2);
      SWAP(*Y[2], f->tmp3);
    }
    // the particle path's gas update and solid conduction (below) ride in
    // the last species diffusion pass when there is one: pointwise on the
    // cell, they read its new Y_H2O; the solid temperature goes to T_s_old
    // (free scratch during the sub-steps), which no fused workgroup reads
    const bool fuse_gs = !p->eulerian && P.nsc > 0;
    GasSolidParams gsp{};
    if (fuse_gs) {
      gsp.T_g = f->T_g;
      gsp.T_s = f->T_s;
      gsp.T_s_out = f->T_s_old;
      gsp.alpha_s = f->alpha_s;
      gsp.rho = f->rho;
      gsp.S_pyro_part = f->src[0];
      gsp.S_dry_part = f->src[1];
      gsp.Q_g_conv = f->src[7];
      gsp.q_rad_gas = f->q_rad_gas;
      gsp.Q_comb_out = sub == N_SUB - 1 ? f->Q_comb : nullptr;
      gsp.n_z_bed = nzb;
      gsp.i_src0 = s0;
      gsp.i_src1 = s1;
      gsp.drip_W_m3 = P.drip_gas;
      gsp.dt_sub = dt_sub;
      gsp.T_amb = p->T_amb;
      gsp.k_solid = K_SOLID_GRASS;
      gsp.rho_solid = RHO_SOLID;
      gsp.cp_solid = CP_SOLID;
      gsp.T_s_lo = p->T_amb;
      gsp.T_s_hi = T_CAP;
    }
    // turbulent diffusion; the clips of :3082-3084 fold into the last pass
    // (with no pass they are the identity: each field is already in range)
    for (int m = 0; m < P.nsc; ++m) {
      const bool last = m == P.nsc - 1;
      const double lo[3] = {last ? 0.0 : 1.0, last ? 0.0 : 1.0, last ? 0.0 : 1.0};
      const double hi[3] = {last ? y_max[0] : 0.0, last ? y_max[1] : 0.0,
                            last ? y_max[2] : 0.0};
      const void* fin[3] = {*Y[0], *Y[1], *Y[2]};
      void* fout[3] = {f->tmp1, f->tmp2, f->tmp3};
      if (launch)
        TRY(turb3_launch(g, fin, fout, f->nu_t, SC_T, P.dsc, lo, hi,
                         P.fuse_summary && last ? (int*)f->edc_work : nullptr, s, sb,
                         fuse_gs && last ? &gsp : nullptr));
      SWAP(*Y[0], f->tmp1);
      SWAP(*Y[1], f->tmp2);
      SWAP(*Y[2], f->tmp3);
    }
    summary_ready = P.fuse_summary;
    if (p->eulerian) {
      // drip-torch split, q_rad_gas into Q_comb and the gas-solid coupling
      // (:2916-2956, :3094-3099, :3148-3156); the solid share accumulates in
      // Q_pyro once per sub-step, as the reference subtracts it in the loop
      if (launch) {
        TRY(coupling_launch(g, f->T_g, f->T_s, f->rho, f->u, f->v, f->w, f->alpha_s,
                            p->sigma_sav, f->q_rad, f->Q_pyro, nullptr, f->q_rad_gas,
                            sub == N_SUB - 1 ? f->Q_comb : nullptr, f->eu_m_water, p->eu_L_v,
                            dt_sub, p->T_amb, 0, p->h_conv_mult, nzb, s0, s1, P.drip_gas,
                            P.drip_solid, s));
        TRY(fw_solid_conduction_vertical(g, f->T_s, f->tmp1, f->alpha_s, K_SOLID_GRASS,
                                         RHO_SOLID, CP_SOLID, dt_sub, p->T_amb, T_CAP, s));
      }
    } else if (launch && !fuse_gs) {
      // gas-energy update and vertical solid conduction with the T_s clamps
      // of :3195-3196, in one pass
      TRY(gas_solid_launch(g, f->T_g, f->T_s, f->tmp1, f->alpha_s, f->rho, f->Y_H2O,
                           f->src[0], f->src[1], f->src[7], f->q_rad_gas,
                           sub == N_SUB - 1 ? f->Q_comb : nullptr, nzb, s0, s1, P.drip_gas,
                           dt_sub, p->T_amb, K_SOLID_GRASS, RHO_SOLID, CP_SOLID, p->T_amb,
                           T_CAP, s, sb));
    }
    if (fuse_gs) SWAP(f->T_s, f->T_s_old);
    else SWAP(f->T_s, f->tmp1);
    // T_g turbulent diffusion, the clamps of :3193-3194 in the last pass
    const bool edc_next = fuse_edc && sub < N_SUB - 1;
    for (int m = 0; m < P.npr; ++m) {
      const bool last = m == P.npr - 1;
      EdcFuseArgs ea = edc_args;
      ea.Y_fuel = f->Y_fuel;
      ea.Y_O2 = f->Y_O2;
      ea.Y_H2O = f->Y_H2O;
      if (launch)
        TRY(turb_pass_launch(g, f->T_g, f->tmp1, f->nu_t, PR_T, P.dpr,
                             last ? p->T_amb : 1.0, last ? T_CAP : 0.0, s, sb,
                             edc_next && last ? &ea : nullptr));
      SWAP(f->T_g, f->tmp1);
    }
    edc_done = edc_next;
    if (P.npr == 0 && launch) {
      // the particle path's gas update already floors T_g at T_amb; the
      // coupling does not
      const double lo = p->eulerian ? p->T_amb : -INFINITY;
      if (g->dtype == FW_F64)
        hipLaunchKernelGGL(k_clamp_max<double>, dim3((n + 255) / 256), dim3(256), 0, s, n,
                           (double*)f->T_g, T_CAP, lo);
      else
        hipLaunchKernelGGL(k_clamp_max<float>, dim3((n + 255) / 256), dim3(256), 0, s, n,
                           (float*)f->T_g, (float)T_CAP, (float)lo);
      FW_CHECK_LAUNCH("clamp");
    }
  }
  return FW_OK;
}

// The graph key of a sub-step loop: the plan and every pointer it touches.
static std::vector<uint64_t> substep_key(const fw_grid* g, const fw_spread_fields* f,
                                         const SubstepPlan& P) {
  return {kw(g), (uint64_t)g->dtype, (uint64_t)P.nsc, (uint64_t)P.npr,
          (uint64_t)P.fuse_summary, (uint64_t)fuse_edc_on(), (uint64_t)P.first_edc_done,
          kw(P.drip_gas), kw(P.drip_solid),
          kw(f->rho), kw(f->T_g), kw(f->T_s), kw(f->Y_fuel), kw(f->Y_O2), kw(f->Y_H2O),
          kw(f->k), kw(f->eps), kw(f->omega), kw(f->nu_t), kw(f->u), kw(f->v), kw(f->w),
          kw(f->src[0]), kw(f->src[1]), kw(f->src[7]), kw(f->q_rad_gas), kw(f->Q_comb),
          kw(f->alpha_s), kw(f->tmp1), kw(f->tmp2), kw(f->tmp3), kw(f->T_s_old),
          kw(f->edc_work),
          kw(f->step_scalars)};
}

}  // namespace fw

using namespace fw;

extern "C" {

int fw_graphs_create(void** graphs) {
  FW_REQUIRE(graphs, "graphs_create: null argument");
  *graphs = new GraphCache();
  return FW_OK;
}

int fw_graphs_destroy(void* graphs) {
  LegacyScope legacy_scope;
  delete (GraphCache*)graphs;
  return FW_OK;
}

int fw_graphs_stats(void* graphs, long long* replays, long long* captures) {
  FW_REQUIRE(graphs, "graphs_stats: null handle");
  const GraphCache* gc = (const GraphCache*)graphs;
  if (replays) *replays = gc->hits;
  if (captures) *captures = gc->captures;
  return FW_OK;
}

int fw_spread_front(const fw_grid* g, const fw_spread_fields* f, fw_spread_state* st,
                    void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && f && st && f->phi, "spread_front: null argument");
  const int Nx = g->nx, Ny = g->ny, Nz = g->nz;
  const int k_exit = Nz > 1 ? 1 : 0, k_mid = g->n_z_bed / 2;
  const size_t row_exit = ((size_t)k_exit * Ny + Ny / 2) * Nx;
  const size_t row_mid = ((size_t)k_mid * Ny + Ny / 2) * Nx;
  double* h = (double*)malloc(2 * (size_t)Nx * sizeof(double));
  FW_REQUIRE(h, "spread_front: out of host memory");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e1 = hipMemcpyAsync(h, f->phi + row_exit, Nx * sizeof(double),
                                 hipMemcpyDeviceToHost, s);
  hipError_t e2 = hipMemcpyAsync(h + Nx, f->phi + row_mid, Nx * sizeof(double),
                                 hipMemcpyDeviceToHost, s);
  hipError_t e3 = hipStreamSynchronize(s);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    free(h);
    return fw_set_error(FW_E_HIP, "spread_front: %s",
                        hipGetErrorString(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3)));
  }
  st->front_exit_x = front_x_row(h, Nx, g->dx);
  st->front_mid_x = front_x_row(h + Nx, Nx, g->dx);
  free(h);
  return FW_OK;
}

int fw_spread_reduce(const fw_grid* g, const fw_spread_fields* f,
                     const fw_spread_params* p, int with_sources, fw_spread_state* st,
                     void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && f && p && st, "spread_reduce: null argument");
  hipStream_t s = (hipStream_t)stream;
  TRY(fw_step_reduce(g, f->u, f->v, f->w, f->nu_t, f->T_g, f->T_s,
                     with_sources ? f->S_pyro : nullptr, with_sources ? f->Q_comb : nullptr,
                     f->omega, f->Y_fuel, f->alpha_s, p->T_ign, p->T_ign, g->n_z_bed,
                     f->red_work, f->red_out, stream));
  FW_CHECK_HIP(hipMemcpyAsync(f->red_host, f->red_out, 14 * sizeof(double),
                              hipMemcpyDeviceToHost, s));
  FW_CHECK_HIP(hipStreamSynchronize(s));
  memcpy(st->red, f->red_host, sizeof(st->red));
  return FW_OK;
}

int fw_spread_step(const fw_grid* g, const fw_grid* g64, fw_spread_fields* f,
                   const fw_spread_params* p, fw_spread_state* st, void* proj, void* dom,
                   const fw_bed_particles* particles, const fw_bed_index* index,
                   fw_bed_params* bed, void* stream) {
  FW_GRID(g);
  FW_REQUIRE(g && g64 && f && p && st && proj && dom, "spread_step: null argument");
  FW_REQUIRE(p->eulerian ? (f->eu_m_solid && f->eu_m_char && f->eu_m_water && f->Q_pyro)
                         : (particles && index && bed),
             "spread_step: the bed's buffers are missing");
  FW_REQUIRE(g->n_z_bed > 0 && p->i_src1 > p->i_src0, "spread_step: needs a fuel bed");
  hipStream_t s = (hipStream_t)stream;
  const int nzb = g->n_z_bed;
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  const size_t esz = g->dtype == FW_F64 ? 8 : 4;
  const double* r = st->red;
  st->step_no += 1;
  st->front_appended = 0;
  const double t = st->t;
  // adaptive dt and the NaN halt (spread_3d.py:1906-1941)
  double u_now = fmax(fmax(r[0], r[1]), r[2]);
  u_now = fmax(fmax(u_now, p->U_mf), fmax(sqrt(9.81 * p->h_bed * 0.5), 0.1));
  const double nu_now = fmax(r[3], 1.0e-3);
  const double hmin = fmin(fmin(g->dx, g->dy), p->dz_min);
  double dt = fmin(fmin(p->cfl_factor * hmin / u_now,
                        0.25 * p->dz_min * p->dz_min / nu_now), 0.05);
  dt = fmin(dt, st->dt * 1.5);
  if (dt < p->min_dt) dt = p->min_dt;
  if (r[5] != 0.0)
    return fw_set_error(FW_E_NONFINITE, "State NaN/Inf at t=%.3fs -- simulation exploded.",
                        t);
  st->dt = dt;
  const int flame_n = (int)r[12];
  const bool combustion = t >= p->spin_up;
  const bool ign_on = t < p->ignition_duration;
  const double Q_drip = Q_DRIP_PER_AREA / fmax(p->h_bed, 1e-3);
  const int s0 = p->i_src0, s1 = p->i_src1;

  // rho_prev, EoS, inlet O2 (:1984-2050)
  TRY(fw_step_prologue(g, f->T_g, f->rho, f->rho_prev, f->Y_O2, p->T_amb, stream));
  // bed (:2064-2215): particles, or the Eulerian pools; the sources stay
  // zero until ignition
  if (p->eulerian) {
    if (combustion) {
      TRY(fw_pyro_step(g, f->T_s, f->alpha_s, f->Y_O2, f->eu_m_solid, f->eu_m_char,
                       f->eu_m_water, p->eu_m_water_init > 0.0 ? 1 : 0, p->eu_m_water_init,
                       p->eu_A_p, p->eu_E_p, dt, 15, f->src[0], f->Q_pyro, nullptr, nullptr,
                       nullptr, stream));
    } else {
      FW_CHECK_HIP(hipMemsetAsync(f->src[0], 0, n * esz, s));
      FW_CHECK_HIP(hipMemsetAsync(f->Q_pyro, 0, n * esz, s));
    }
  } else if (combustion) {
    TRY(fw_bed_external_heat(g, f->q_rad, f->Q_ext, nzb, s0, s1,
                             ign_on ? F_DRIP_TO_SOLID * Q_drip : 0.0, Q_RAD_MAX, stream));
    bed->dt = dt;
    TRY(fw_bed_step(g, particles, index, bed, f->T_g, f->Y_O2, f->Q_ext, f->src, f->contrib,
                    f->counters, nullptr, nullptr, stream));
    TRY(fw_bed_aggregate(g, particles, index, f->T_s, nullptr, stream));
    FW_CHECK_HIP(hipMemcpyAsync(f->T_s_old, f->T_s, n * esz, hipMemcpyDeviceToDevice, s));
    TRY(fw_bed_horizontal_conduction(g, particles, index, f->T_s_old, f->T_s, f->alpha_s,
                                     K_SOLID_GRASS, p->rho_solid_true, p->cp_solid, nzb, dt,
                                     f->dT_work, stream));
  }
  // DOM every dom_every steps (:2455-2505), issued here beside the flow work
  // or in line after k-epsilon
  const bool dom_now = st->step_no % (p->dom_every > 0 ? p->dom_every : 1) == 0;
  const bool beside = dom_now && dom_beside_on();
  const fw_spread_fields fd = *f;   // the solve's buffers: the flow work swaps others
  auto dom_solve = [dom, fd, p](hipStream_t ds) {
    return fw_dom_solve(dom, fd.T_s, fd.T_g, fd.alpha_s, fd.omega, p->sigma_sav, p->T_amb,
                        fd.q_rad, fd.q_rad_gas, fd.T_soil, fd.q_in_soil, fd.Y_H2O, fd.rho,
                        p->eulerian ? nullptr : fd.M_local, ds);
  };
  DomBeside dom_side;
  if (dom_now) {
    // the particles' moisture scales kappa_solid; the Eulerian bed has none
    if (!p->eulerian) TRY(fw_bed_aggregate(g, particles, index, nullptr, f->M_local, stream));
    if (beside) TRY(dom_side.start(dom, s, dom_solve));
  }
  // drag, tentative momentum (:2217-2277)
  TRY(fw_drag_force(g, f->u, f->v, f->w, f->rho, f->alpha_s, p->sigma_sav, p->canopy_C_d,
                    f->Fx, f->Fy, f->Fz, stream));
  TRY(fw_tentative_velocity(g, f->u, f->v, f->w, f->tmp1, f->tmp2, f->tmp3, f->rho, f->T_g,
                            f->Fx, f->Fy, f->Fz, f->u_inlet, f->v_inlet, f->w_inlet, dt,
                            p->T_amb, stream));
  SWAP(f->u, f->tmp1);
  SWAP(f->v, f->tmp2);
  SWAP(f->w, f->tmp3);
  // S_pyro, div_target and the projection loop (:2147, :2293-2360)
  TRY(fw_div_target(g, f->src[0], f->src[1], f->rho, f->rho_prev, dt, f->S_pyro,
                    f->div_target, stream));
  st->proj_n = 0;
  st->proj_div = 0.0;
  for (int it = 0; it < p->proj_max_iter; ++it) {
    TRY(fw_proj_project(proj, f->u, f->v, f->w, f->rho, dt, f->div_target, f->u_inlet,
                        stream));
    int ni = 0;
    double rel = 0.0, dmax = 0.0;
    TRY(fw_proj_stats(proj, &ni, &rel, &dmax));
    st->pcg_iters = ni;
    st->proj_n = it + 1;
    st->proj_div = dmax;
    if (dmax < p->proj_div_tol) break;
  }
  // outflow sponge (:2368-2372)
  TRY(fw_outflow_sponge(g, f->u, f->u_inlet, f->sigma_x, f->Y_fuel, 1.0e-3, dt, stream));
  // k-epsilon (:2408-2433)
  TRY(fw_k_epsilon(g, f->k, f->eps, f->tmp1, f->tmp2, f->nu_t, f->u, f->v, f->w, f->T_g,
                   f->alpha_s, f->u_inlet, f->k_wall, f->eps_wall, p->sigma_sav, dt, p->T_amb,
                   p->canopy_beta_p, p->canopy_beta_d, stream));
  SWAP(f->k, f->tmp1);
  SWAP(f->eps, f->tmp2);
  // the sub-step plan's max nu_t (nothing writes nu_t between here and the
  // sub-steps), read back while the GPU runs the work below
  TRY(nu_max_issue(g, f, s));
  // the DOM's results, then the soil every step (:2505-2510)
  if (beside) TRY(dom_side.join(s));
  else if (dom_now) TRY(dom_solve(s));
  TRY(fw_soil_conduction(g, p->n_soil, f->T_soil, f->q_in_soil, f->soil_dz,
                         f->soil_d_above, f->soil_d_below, dt, p->soil_k, p->soil_rho,
                         p->soil_cp, p->soil_eps, p->T_amb, stream));
  // ignition pulse on the cached q_rad (:2528-2532)
  if (ign_on)
    TRY(fw_add_patch(g, f->q_rad, nzb - 1, s0, s1, Q_PULSE * p->ignition_q_mult, stream));
  // gas-energy advection (:2613-2622); with whole 256-cell rows (up to 2^21
  // cells, as for the later sub-steps' fused EDC) the first sub-step's EDC
  // rides in the same pass: it reads the cell's new temperature and the
  // species, which nothing between the two changes
  const bool first_edc =
      fuse_edc_on() && g->nx % EDC_SEG == 0 && n <= ((size_t)1 << 21) && f->edc_work;
  if (first_edc) {
    TRY(edc_summary_launch(g, f->Y_fuel, f->Y_O2, f->Y_H2O, (int*)f->edc_work, s));
    const EdcFuseArgs ea{f->rho, f->k, f->eps, f->Y_H2O, f->Y_fuel, f->Y_O2, f->omega,
                         (const int*)f->edc_work,
                         edc_params(0.34, 1100.0, dt / N_SUB, 1, 0, 1.3, 17.0e6)};
    TRY(gas_energy_launch(g, f->T_g, f->tmp1, f->u, f->v, f->w, dt, 2.0e-5, p->T_amb, s,
                          &ea));
  } else {
    TRY(fw_gas_energy(g, f->T_g, f->tmp1, f->u, f->v, f->w, dt, 2.0e-5, p->T_amb, stream));
  }
  SWAP(f->T_g, f->tmp1);
  // N_SUB chemistry / transport / coupling sub-steps (:2678-3200)
  double nu_max = 0.0;
  TRY(nu_max_wait(f, &nu_max));
  SubstepPlan plan;
  plan.dt_sub = dt / N_SUB;
  plan.dsc = plan.dpr = 0.0;
  plan.nsc = turb_substeps(nu_max, SC_T, plan.dt_sub, g->dx, g->dy, p->dz_min, &plan.dsc);
  plan.npr = turb_substeps(nu_max, PR_T, plan.dt_sub, g->dx, g->dy, p->dz_min, &plan.dpr);
  plan.drip_gas = (combustion && ign_on) ? (1.0 - F_DRIP_TO_SOLID) * Q_drip : 0.0;
  plan.drip_solid = (combustion && ign_on && p->eulerian) ? F_DRIP_TO_SOLID * Q_drip : 0.0;
  plan.fuse_summary = g->nx % EDC_SEG == 0 && plan.nsc > 0;
  plan.first_edc_done = first_edc;
  if (f->graphs && f->step_scalars && s && !p->eulerian) {
    // one graph launch for the whole loop: the step's scalars go to the
    // device block first, the pointers advance as the launches would
    hipLaunchKernelGGL(k_set_step_scalars, dim3(1), dim3(64), 0, s, f->step_scalars,
                       plan.dt_sub, edc_relax(plan.dt_sub, 1), plan.dsc, plan.dpr);
    FW_CHECK_LAUNCH("step scalars");
    GraphCache* gc = (GraphCache*)f->graphs;
    const fw_spread_fields entry = *f;
    TRY(gc->run(substep_key(g, f, plan), s, [&](hipStream_t cs) {
      fw_spread_fields rec = entry;
      return substeps(g, &rec, p, plan, entry.step_scalars, true, cs);
    }));
    TRY(substeps(g, f, p, plan, nullptr, false, s));
  } else {
    TRY(substeps(g, f, p, plan, nullptr, true, s));
  }
  // level set (:3204-3300).  Active: the ahead band of the step's level set
  // and the DOM forward flux into it (:2555-2565, :2591; the intensities and
  // phi do not change between there and here), the front speed with the
  // empirical blend (:3227-3256) and the evolve (:3260).  Passive: v_n = 0,
  // the evolve is the identity.  Both: reinitialise every 10th step, float64.
  st->lset_step += 1;
  bool moved = false;
  if (p->lset_active) {
    FW_REQUIRE(f->lset_mask && f->q_dom_fwd && f->v_n, "spread_step: active level set "
               "needs lset_mask, q_dom_fwd and v_n");
    TRY(fw_ahead_band_mask(g64, f->phi, p->band_m, nzb, f->lset_mask, stream));
    TRY(fw_dom_forward_flux(dom, f->lset_mask, f->q_dom_fwd, stream));
    TRY(fw_lset_speed(g, f->q_dom_fwd, f->M_local, p->E_sens, p->E_lat_coef, p->emp_w,
                      p->emp_ros, f->v_n, stream));
    TRY(fw_levelset_update(g64, f->phi, nullptr, f->v_n, f->phi_tmp, dt, 0, stream));
    double* tp = f->phi;
    f->phi = f->phi_tmp;
    f->phi_tmp = tp;
    moved = true;
  }
  if (st->lset_step % REINIT_INTERVAL == 0) {
    const size_t nb = n * sizeof(double);
    FW_CHECK_HIP(hipMemcpyAsync(f->phi0, f->phi, nb, hipMemcpyDeviceToDevice, s));
    const double dtau = 0.5 * fmin(fmin(g->dx, g->dy), p->dz_min);
    for (int m = 0; m < REINIT_SUBSTEPS; ++m) {
      TRY(fw_levelset_update(g64, f->phi, f->phi0, nullptr, f->phi_tmp, dtau, 1, stream));
      double* tp = f->phi;
      f->phi = f->phi_tmp;
      f->phi_tmp = tp;
    }
    moved = true;
  }
  if (moved) TRY(fw_spread_front(g64, f, st, stream));
  // dynamic seeding behind the empirical-ROS front, every 50th step
  // (:3271-3300); front_exit_x is lset.front_x(k=1, j=Ny/2) after the evolve
  if (p->lset_active && p->emp_w > 0.0 && !p->eulerian) {
    st->seed_count += 1;
    if (st->seed_count >= 50) {
      st->seed_count = 0;
      TRY(fw_bed_seed(particles, p->x_lo_bed, st->front_exit_x, p->T_ign, p->seed_T_s,
                      stream));
    }
  }
  // exits (:3414-3436), before the EoS as in the reference
  if (flame_n == 0 && t > p->ignition_duration + T_BOOTSTRAP_S) {
    st->extinct_count += 1;
    if (st->extinct_count > 50) {
      st->stop = 1;
      return FW_OK;
    }
  } else {
    st->extinct_count = 0;
  }
  if (st->front_exit_x > p->bed_x_end - g->dx) {
    st->stop = 2;
    return FW_OK;
  }
  // EoS, reductions, front tracking, diagnostics (:3440-3512)
  TRY(fw_eos(g, f->T_g, f->rho, p->T_amb, stream));
  TRY(fw_spread_reduce(g, f, p, 1, st, stream));
  double new_front;
  if (isfinite(st->front_mid_x)) {
    if (st->front_mid_x > st->front_x_last) {
      st->front_appended = 1;
      st->front_t_new = t;
      st->front_x_new = st->front_mid_x;
      st->front_x_last = st->front_mid_x;
    }
    new_front = st->front_mid_x;
  } else {
    const int ib = (int)st->red[13];
    new_front = ib >= 0 ? (ib + 0.5) * p->x_step : st->front_x_last;
  }
  if (new_front > st->last_front) {
    st->last_front = new_front;
    st->last_advance_t = t;
  }
  st->t = t + dt;
  // exits after the snapshot (:3673-3683): domain end, then the stall of an
  // active level set
  if (st->front_x_last > p->Lx * 0.85) {
    st->stop = 3;
  } else if (p->lset_active && st->t - st->last_advance_t > p->stall_window &&
             st->t > p->ignition_duration) {
    st->stop = 4;
  }
  return FW_OK;
}

int fw_spread_run(const fw_grid* g, const fw_grid* g64, fw_spread_fields* f,
                  const fw_spread_params* p, fw_spread_state* st, void* proj, void* dom,
                  const fw_bed_particles* particles, const fw_bed_index* index,
                  fw_bed_params* bed, void* stream, int max_steps, fw_spread_state* hist,
                  int* n_done) {
  FW_REQUIRE(n_done && (max_steps <= 0 || hist), "spread_run: null argument");
  *n_done = 0;
  for (int i = 0; i < max_steps; ++i) {
    TRY(fw_spread_step(g, g64, f, p, st, proj, dom, particles, index, bed, stream));
    hist[i] = *st;
    *n_done = i + 1;
    if (st->stop != 0) break;
  }
  return FW_OK;
}

}  // extern "C"
