/* firewave — C ABI of the MI355X (gfx950) Tier-3 spread_3d time step.
 *
 * Every entry point is `extern "C"`, takes plain device pointers plus sizes
 * and scalars, and returns 0 on success or a negative FW_E_* code; the
 * message of the last failure on the calling thread is fw_last_error().
 * Field pointers address dense (Nz, Ny, Nx) C-order device arrays of the
 * grid's storage type (fw_grid.dtype: FW_F32 or FW_F64).  `stream` is a
 * hipStream_t (NULL = the default stream).  No entry point allocates device
 * memory unless its name says so (*_create); scratch comes from a handle or
 * from caller-supplied work pointers.  Output arrays never alias inputs
 * unless stated.
 *
 * Reference call surface replaced (paths relative to the FireWebApp tree):
 * src/model_outdoor/physics_3d/*.py module functions and the three stateful
 * classes ProjectionSolver3D, DOMRadiationSolver, LevelSetFront3D.
 */
#ifndef FIREWAVE_H
#define FIREWAVE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FW_OK = 0,
  FW_E_CONFIG = -1,    /* bad argument / configuration  (reference: ValueError) */
  FW_E_NOCONV = -2,    /* Krylov or DOM did not converge (reference: RuntimeError) */
  FW_E_NONFINITE = -3, /* NaN/Inf in state             (reference: RuntimeError) */
  FW_E_OVERFLOW = -4,  /* particle buffer overflow      (reference: ValueError) */
  FW_E_HIP = -5        /* HIP runtime failure */
};

enum { FW_F32 = 0, FW_F64 = 1 };

/* Grid metrics (Grid3D, spread_3d.py:160-424).  The 1-D arrays are device
 * pointers to float64 arrays: dz, d_above, d_below of length nz, z_face of
 * length nz+1, and inv_dz, inv_d_above, inv_d_below = the elementwise
 * float64 reciprocals 1.0 / dz etc. (all required).  y is periodic (the only
 * y boundary of the hot path). */
typedef struct fw_grid {
  int32_t nx, ny, nz;
  int32_t dtype;   /* FW_F32 or FW_F64: storage and arithmetic type of fields */
  double dx, dy;
  const double* dz;
  const double* d_above;
  const double* d_below;
  const double* z_face;
  int32_t n_z_bed;
  int32_t reserved;
  const double* inv_dz;
  const double* inv_d_above;
  const double* inv_d_below;
} fw_grid;

const char* fw_last_error(void);
int fw_version(void);

/* ---- transport stencils ------------------------------------------------- */

/* muscl_3d.advect_3d_scalar_muscl (muscl_3d.py:67-171): rhs -= (u.grad)phi.
 * rhs is updated in place (read-modify-write). */
int fw_advect_scalar_muscl(const fw_grid* g, const void* phi, const void* u,
                           const void* v, const void* w, void* rhs,
                           double phi_inlet, void* stream);

/* species_3d.step_species_transport (species_3d.py:40-175), out of place:
 * Y_out = clip(Y + dt*(adv + D lap Y + S/rho), 0, 1), then clip to y_max.
 * y_max = 1 reproduces the kernel; the driver passes 0.232 for O2 to fold the
 * caller-side clip (spread_3d.py:3059-3060). */
int fw_species_transport(const fw_grid* g, const void* Y, void* Y_out,
                         const void* rho, const void* u, const void* v,
                         const void* w, const void* S, double dt, double D,
                         double Y_inlet, double y_max, void* stream);

/* step_species_transport with the multi-species dilution source of the
 * time step (spread_3d.py:2990-3038) computed in the kernel: with
 * S_tot = S_pyro_part + S_dry_part, kind 1 (fuel) S = S_pyro_part - Y S_tot,
 * kind 2 (O2) S = -Y S_tot, kind 3 (H2O) S = S_dry_part - Y S_tot. */
int fw_species_transport_dilution(const fw_grid* g, const void* Y, void* Y_out,
                                  const void* rho, const void* u, const void* v,
                                  const void* w, const void* S_pyro_part,
                                  const void* S_dry_part, int kind, double dt,
                                  double D, double Y_inlet, double y_max,
                                  void* stream);

/* The three species of a sub-step in one pass (fw_species_transport_dilution
 * with kind 1, 2, 3 for Y[0], Y[1], Y[2]): rho, u, v, w and the particle
 * sources are read once.  Y, Y_out: 3 field pointers each (no aliasing);
 * Y_inlet, y_max: 3 host doubles each. */
int fw_species_transport3(const fw_grid* g, const void* const* Y, void* const* Y_out,
                          const void* rho, const void* u, const void* v, const void* w,
                          const void* S_pyro_part, const void* S_dry_part, double dt,
                          double D, const double* Y_inlet, const double* y_max,
                          void* stream);

/* Vertically adjacent cells per thread of fw_species_transport3 for grids of
 * g's shape and type (no counterpart in the reference): 1, 2 or 4; 0 restores
 * the default 2.  Every choice computes each cell identically.  Process
 * wide; picked per shape by the autotuner (firewebapp_amd/autotune.py). */
int fw_tune_species_cz(const fw_grid* g, int cz);

/* _adv_gas_energy (spread_3d.py:3827-3897): MUSCL advection with inlet ghost
 * T_amb + alpha_th FV diffusion, then clamp to [T_amb, 1e4]; out of place. */
int fw_gas_energy(const fw_grid* g, const void* T_g, void* T_out,
                  const void* u, const void* v, const void* w, double dt,
                  double alpha_th, double T_amb, void* stream);

/* One sub-step of turbulence_3d.apply_turbulent_diffusion
 * (turbulence_3d.py:920-949): f_out = f + nu_t/sc_t * lap(f) * dt_sub with
 * zero-flux ghosts.  If clip_lo <= clip_hi the result is clipped to that
 * range (the caller-side clips of spread_3d.py:3079-3084); pass lo > hi for
 * the bare reference kernel. */
int fw_turbulent_diffusion_pass(const fw_grid* g, const void* f, void* f_out,
                                const void* nu_t, double sc_t, double dt_sub,
                                double clip_lo, double clip_hi, void* stream);

/* fw_turbulent_diffusion_pass for three fields sharing nu_t; f, f_out: 3
 * field pointers each (no aliasing); clip_lo, clip_hi: 3 host doubles each. */
int fw_turbulent_diffusion_pass3(const fw_grid* g, const void* const* f,
                                 void* const* f_out, const void* nu_t, double sc_t,
                                 double dt_sub, const double* clip_lo,
                                 const double* clip_hi, void* stream);

/* fw_turbulent_diffusion_pass3 on (fuel, O2, H2O) that also writes the EDC
 * segment summaries of its output, as the step driver's last species pass of
 * a sub-step does (no counterpart in the reference; it saves fw_edc's own
 * summary pass): for each 256-cell segment s of the flat index,
 * edc_summary[s] = -1 if no cell of it has Y_F > 1e-9 and Y_O2 > 1e-9, else
 * the wet-bulb flag (Y_H2O > 0 and 1 - Y_H2O/0.18 < 0.5) of the last such
 * cell (chemistry_closures/edc.py:223).  Needs Nx % 256 == 0
 * (FW_E_CONFIG otherwise); edc_summary: fw_edc_workspace_ints(g) int32. */
int fw_turbulent_diffusion_pass3_summary(const fw_grid* g, const void* const* f,
                                         void* const* f_out, const void* nu_t, double sc_t,
                                         double dt_sub, const double* clip_lo,
                                         const double* clip_hi, int* edc_summary,
                                         void* stream);

/* max(nu_t) over interior cells (turbulence_3d.py:890-897); synchronises the
 * stream and writes the host scalar.  work: FW_REDUCE_BLOCKS (256) doubles. */
int fw_nu_t_max_interior(const fw_grid* g, const void* nu_t, double* work,
                         double* out_host, void* stream);

/* solid_conduction_3d.step_solid_conduction_vertical
 * (solid_conduction_3d.py:56-127), out of place; optional clamp as above
 * (the T_s clamps of spread_3d.py:3195-3196). */
int fw_solid_conduction_vertical(const fw_grid* g, const void* T_s,
                                 void* T_out, const void* alpha_s,
                                 double k_solid, double rho_solid,
                                 double cp_solid, double dt, double clip_lo,
                                 double clip_hi, void* stream);

/* drag_3d.step_drag_force (drag_3d.py:99-144). */
int fw_drag_force(const fw_grid* g, const void* u, const void* v,
                  const void* w, const void* rho, const void* alpha_s,
                  double sigma_sav, double C_D, void* Fx, void* Fy, void* Fz,
                  void* stream);

/* momentum_3d.step_tentative_velocity (momentum_3d.py:41-276), out of place.
 * u_inlet/v_inlet/w_inlet: (Nz, Ny). */
int fw_tentative_velocity(const fw_grid* g, const void* u, const void* v,
                          const void* w, void* u_out, void* v_out,
                          void* w_out, const void* rho, const void* T_g,
                          const void* Fx, const void* Fy, const void* Fz,
                          const void* u_inlet, const void* v_inlet,
                          const void* w_inlet, double dt, double T_amb,
                          void* stream);

/* momentum_3d.apply_outflow_sponge (momentum_3d.py:292-317), in place.
 * u_target: (Nz, Ny); sigma_x: (Nx). */
int fw_outflow_sponge(const fw_grid* g, void* u, const void* u_target,
                      const void* sigma_x, const void* Y_F, double Y_F_skip,
                      double dt, void* stream);

/* rho = P0 / (R_air max(T_g, T_amb))  (spread_3d.py:2042, :3440). */
int fw_eos(const fw_grid* g, const void* T_g, void* rho, double T_amb,
           void* stream);

/* soil_3d.step_soil_conduction (soil_3d.py:105-153), in place, Gauss-Seidel
 * in k.  T_soil: (n_soil, Ny, Nx); q_in_surface: (Ny, Nx); soil_* float64
 * device arrays of length n_soil. */
int fw_soil_conduction(const fw_grid* g, int n_soil, void* T_soil,
                       const void* q_in_surface, const double* soil_dz,
                       const double* soil_d_above, const double* soil_d_below,
                       double dt, double k_s, double rho_s, double cp_s,
                       double eps_s, double T_amb, void* stream);

/* ---- turbulence -------------------------------------------------------- */

/* turbulence_3d.step_k_epsilon (turbulence_3d.py:481-858) with
 * _strain_and_vorticity_squared (:391-478) fused; out of place.
 * u_inlet: (Nz, Ny); k/eps wall ghosts: (Ny, Nx). */
int fw_k_epsilon(const fw_grid* g, const void* k_turb, const void* eps_turb,
                 void* k_out, void* eps_out, void* nu_t_out, const void* u,
                 const void* v, const void* w, const void* T_g,
                 const void* alpha_s, const void* u_inlet,
                 const void* k_wall_ghost, const void* eps_wall_ghost,
                 double sigma_sav, double dt, double T_amb, double beta_p,
                 double beta_d, void* stream);

/* ---- pressure projection (ProjectionSolver3D, projection_3d.py:214-948;
 *      SeparableLaplacian3D, fft_poisson_3d.py:45-228) -------------------- */

/* Create a projection handle (replaces ProjectionSolver3D.__init__ with
 * method='fft_pcg', projection_3d.py:214-280, and SeparableLaplacian3D.__init__,
 * fft_poisson_3d.py:70-125).  Host float64 inputs: Ux (Nx x Nx, row-major,
 * columns = x eigenvectors of the Neumann/Dirichlet L_x), Uy (Ny x Ny, columns
 * = real orthonormal periodic eigenvectors), Uz (Nz x Nz, columns =
 * eigenvectors of sqrt(dz) L_z sqrt(dz)^-1), and their eigenvalues.  Allocates
 * the solver state (warm-start pressure, Krylov vectors). */
int fw_proj_create(const fw_grid* g, const double* Ux, const double* Uy,
                   const double* Uz, const double* lam_x, const double* lam_y,
                   const double* lam_z, double eps_reg, double cg_rtol,
                   int maxiter, void** handle);
int fw_proj_destroy(void* handle);

/* ProjectionSolver3D.project (projection_3d.py:677-844): rhs = (div(u) -
 * div_target)/dt in float64, solve A p = rhs to ||r||_2 <= cg_rtol ||rhs||_2
 * (PCG with the separable preconditioner, float64 refinement residual,
 * warm-started from the previous p), then correct u, v, w in place.
 * div_target may be NULL.  u_inlet: (Nz, Ny).  FW_E_NOCONV if the Krylov
 * solve fails (reference: RuntimeError, projection_3d.py:785-790). */
int fw_proj_project(void* handle, void* u, void* v, void* w, const void* rho,
                    double dt, const void* div_target, const void* u_inlet,
                    void* stream);

/* Iterations and final relative residual of the last solve, and
 * max|div(u_new) - div_target| after the last correction (float64). */
int fw_proj_stats(void* handle, int* iters, double* rel_resid, double* div_max);

/* Copy the solver's float64 pressure (the warm start, N values) out / in. */
int fw_proj_pressure(void* handle, double* p_out, void* stream);
int fw_proj_set_pressure(void* handle, const double* p_in, void* stream);

/* ProjectionSolver3D.divergence (projection_3d.py:847-948), alpha_g = None. */
int fw_proj_divergence(void* handle, const void* u, const void* v,
                       const void* w, const void* u_inlet, void* div_out,
                       void* stream);

/* y = A x for the operator of _fill_var_density_data (projection_3d.py:65-175)
 * built from rho (matrix-free). */
int fw_proj_matvec(void* handle, const void* rho, const void* x, void* y,
                   void* stream);

/* z = SeparableLaplacian3D.solve(r) (fft_poisson_3d.py:184-228). */
int fw_proj_precondition(void* handle, const void* r, void* z, void* stream);

/* Launch shapes of the inner PCG (no counterpart in the reference, whose
 * solve is NumPy/SciPy calls); results are identical up to float32
 * rounding of the preconditioner.  yz_plane = 1 (the default) runs the
 * preconditioner's y and z transforms, all four products per x frequency,
 * in one pass (k_yz_plane: float32, 16 <= Ny <= 128, Nz <= 64), 0 the
 * per-direction column transforms; *yz_active (may be NULL) says whether
 * the one-pass kernel runs.  matvec_blocks: workgroups of the fused
 * operator launch (64..4096; 0 = 1024).  dct_rows: rows per workgroup of
 * the x transforms that feed the one-pass kernel (8 or 16; 0 = 16).  The
 * autotuner (firewebapp_amd/autotune.py) picks them per grid. */
#define FW_PROJ_MV_MAX 4096
int fw_proj_set_tiling(void* handle, int yz_plane, int matvec_blocks, int dct_rows,
                       int* yz_active);

/* Preconditioner of the PCG: FW_PRECOND_SEPARABLE (default; the reference's
 * SeparableLaplacian3D, fft_poisson_3d.py:184-228), FW_PRECOND_NONE, or
 * FW_PRECOND_JACOBI (diag(A)^-1; not in the reference, a comparison row for
 * the time-to-residual measurement of the plain PCG). */
enum { FW_PRECOND_NONE = 0, FW_PRECOND_SEPARABLE = 1, FW_PRECOND_JACOBI = 2,
       FW_PRECOND_LEARNED = 3 };
int fw_proj_set_preconditioner(void* handle, int mode);

/* FW_PRECOND_LEARNED (no counterpart in the reference): the preconditioned
 * direction is z = M_sep r + s N(r / s, RHO_REF / rho - 1), s the RMS of r,
 * on every k-th inner PCG direction (fw_proj_set_learned_every, default 4;
 * the other directions use M_sep r alone); N is the attached network
 * (fw_net_create on the same grid, float32); flexible PCG
 * (beta = <z+, r+ - r>_W / <z, r>_W).
 * net = null detaches it (the mode falls back to separable).  The network
 * stays owned by the caller and must outlive its use. */
int fw_proj_set_learned(void* handle, void* net);

/* In FW_PRECOND_LEARNED an inner PCG round has a budget of 60 iterations;
 * a round that does not converge within it hands the rest of that solve to
 * the separable preconditioner.  fallbacks = the solves that did, since
 * the handle was created. */
int fw_proj_learned_stats(void* handle, int* fallbacks);

/* The network corrects every k-th direction of the learned mode (k >= 1;
 * k = 0: the first direction of each inner round only). */
int fw_proj_set_learned_every(void* handle, int every);

/* Copy the right-hand side b = (div - target)/dt of the last fw_proj_project
 * (float64, N values) out: the system the reference hands to its Krylov
 * solver (projection_3d.py:742-791). */
int fw_proj_rhs(void* handle, double* b_out, void* stream);

/* Solve A x = b (float64 b, x; x is the warm start) to ||b-Ax||_2 <=
 * rtol ||b||_2 with the handle's precision policy; rho in the grid's type. */
int fw_proj_solve(void* handle, const void* rho, const double* b, double* x,
                  double rtol, int* iters, double* rel_resid, void* stream);

/* ---- radiation (DOMRadiationSolver, dom_3d.py:227-473) ------------------ */

/* Replaces DOMRadiationSolver.__init__ (dom_3d.py:235-284): M ordinates
 * Omega (host, M x 3) and weights (host, M); allocates I_set (M x N, zero:
 * the reference's initial warm start), G and the property scratch. */
int fw_dom_create(const fw_grid* g, int M, const double* Omega,
                  const double* weights, double eps_w_ground,
                  int max_source_iter, double tol_source_iter,
                  double omega_relax, double kappa_gas_max, void** handle);
int fw_dom_destroy(void* handle);

/* DOMRadiationSolver.solve (dom_3d.py:286-473).  Y_H2O/rho, T_soil_surface,
 * q_in_soil_out and bed_moisture may be NULL (reference: None). */
int fw_dom_solve(void* handle, const void* T_s, const void* T_g,
                 const void* alpha_s, const void* omega_comb, double sigma_sav,
                 double T_amb, void* q_rad_solid_out, void* q_rad_gas_out,
                 const void* T_soil_surface, void* q_in_soil_out,
                 const void* Y_H2O, const void* rho, const void* bed_moisture,
                 void* stream);

/* Source iterations of the last solve. */
int fw_dom_stats(void* handle, int* iters);

/* Tile side of the sweep (no counterpart in the reference, whose sweep is a
 * Numba loop): 2, 4 or 8 rows square per workgroup (0 = 4), used when those
 * rows fit 64 KB of LDS (and only 4 past 512 cells a row).  The intensities
 * are the same for every side.  *active (may be NULL): the side a float32 /
 * float64 handle of this grid then uses (0: the per-row kernel).  Picked
 * per grid by the autotuner (firewebapp_amd/autotune.py). */
int fw_dom_set_tiling(void* handle, int tile_side, int* active);

/* Dataflow sweep on or off (no counterpart in the reference): one launch per
 * source iteration in which each 4 x 4 block waits on the two blocks it
 * reads from, instead of one launch per block diagonal.  Built for rows of
 * 64, 128, 256 or 512 cells (float64: up to 256) and used with tile side 4; on
 * by default (environment FIREWAVE_DOM_FLOW=0: off).  The intensities are
 * the same either way.  *active (may be NULL): whether it now runs. */
int fw_dom_set_flow(void* handle, int enable, int* active);

/* Sweep launches per source iteration with the current settings (1 for the
 * dataflow sweep). */
int fw_dom_sweep_launches(void* handle, int* launches);

/* Device pointer of I_set (M, Nz, Ny, Nx) in the grid's type (read by
 * flame_front_3d.compute_q_dom_fwd_at_band, flame_front_3d.py:594-630). */
int fw_dom_intensity(void* handle, void** I_set);

/* out = sum over ordinates with xi > 0 of w |xi| I_n where mask != 0
 * (flame_front_3d.compute_q_dom_fwd_at_band, flame_front_3d.py:594-630);
 * mask: uint8 field or NULL (all cells). */
int fw_dom_forward_flux(void* handle, const unsigned char* mask, void* out,
                        void* stream);

/* ---- Lagrangian bed (lagrangian_bed_3d.py:193-1070) --------------------- */

/* The particle buffer dict of allocate_bed_particle_buffers
 * (lagrangian_bed_3d.py:193-218) as device float64 SoA + int8 alive. */
typedef struct fw_bed_particles {
  double *x, *y, *z, *m_solid, *m_water, *m_char, *T_s, *m_water_0, *sav,
      *m_char_max;
  int8_t* alive;
  int32_t n;
  int32_t reserved;
  /* Optional per-particle top of the fuel above the particle's column (m),
   * for the canopy attenuation exp(-kappa (h_top - z)) of the geometric view
   * factor; NULL: the scalar h_bed of fw_bed_params for every particle (the
   * reference's uniform bed).  Set by a seeded depth field or depth jitter. */
  double* h_top;
} fw_bed_particles;

/* Particle -> cell map and its inverse: cell[p] (flat (k,j,i) index or -1)
 * and, for cells 0..n_seg_cells-1, the particles of each cell in slot order
 * (seg_idx[seg_ptr[c] .. seg_ptr[c+1]-1]).  Cells >= n_seg_cells hold no
 * particle. */
typedef struct fw_bed_index {
  const int32_t* cell;
  const int32_t* seg_ptr;
  const int32_t* seg_idx;
  int32_t n_seg_cells;
  int32_t reserved;
} fw_bed_index;

/* Scalar arguments of step_bed_particles (lagrangian_bed_3d.py:593-676). */
typedef struct fw_bed_params {
  double dt, h_conv, rho_solid_true, cp_solid, eps_solid, T_amb, view_factor,
      h_bed, kappa_bed_eff, char_ox_flux_cap, char_ox_ash_exp;
  int32_t n_per_cell_for_split, view_factor_geometric, do_drying,
      do_pyrolysis, do_char_ox, do_smolder, drying_mode, reserved;
} fw_bed_params;

/* Cell of every particle, float64, in both of the reference's conventions:
 * cell_step = locate_cell (lagrangian_particles_3d.py:111-138, x < 0 -> out
 * before truncation), cell_agg = int() then range test (the aggregators and
 * the conduction scatter, lagrangian_bed_3d.py:490-497, :528-534). */
int fw_bed_locate(const fw_grid* g, int n, const double* x, const double* y,
                  const double* z, int32_t* cell_step, int32_t* cell_agg,
                  void* stream);

/* step_bed_particles (lagrangian_bed_3d.py:593-1070).  index: built from
 * cell_step.  sources[8]: S_pyro, S_drying, Q_pyro, Q_drying, Y_F_source,
 * Q_char, Q_smold, Q_g_conv (grid type); cells 0..n_seg_cells-1 are written,
 * the caller keeps the rest zero.  contrib_work: 10*n doubles.
 * counters: device int[2] = n_alive, n_burned.  diag_work (15*n doubles) and
 * diag_out (device double[16]) may be NULL; when given, diag_out receives
 * the reference's diag_max_out. */
int fw_bed_step(const fw_grid* g, const fw_bed_particles* particles,
                const fw_bed_index* index, const fw_bed_params* params,
                const void* T_g, const void* Y_O2, const void* Q_solid_ext,
                void* const* sources, double* contrib_work, int32_t* counters,
                double* diag_work, double* diag_out, void* stream);

/* aggregate_particles_to_T_s_grid (lagrangian_bed_3d.py:500-546) and/or
 * aggregate_particles_to_M_local_grid (:549-590); either output may be NULL.
 * agg_index: built from cell_agg. */
int fw_bed_aggregate(const fw_grid* g, const fw_bed_particles* particles,
                     const fw_bed_index* agg_index, void* T_s_grid,
                     void* M_local_grid, void* stream);

/* step_horizontal_solid_conduction_scatter (lagrangian_bed_3d.py:427-497):
 * T_s_grid (in/out) updated on alpha_s > 0 bed cells from T_s_old (a copy
 * of T_s_grid on entry); every alive particle of a bed cell gets the cell's
 * dT.  dT_work: N doubles. */
int fw_bed_horizontal_conduction(const fw_grid* g,
                                 const fw_bed_particles* particles,
                                 const fw_bed_index* agg_index,
                                 const void* T_s_old, void* T_s_grid,
                                 const void* alpha_s, double k_solid,
                                 double rho_solid_true, double cp_solid,
                                 int n_z_bed, double dt, double* dT_work,
                                 void* stream);

/* ---- Eulerian fuel bed (lagrangian_bed_enable = false) ------------------ */

/* pyrolysis_3d.py:270-614 on the bed pools (float64 fields m_solid = the
 * reference's m_hemi, m_char, m_water), float64 arithmetic.  `processes` is
 * a mask of the kernels run, in this order, cell by cell:
 *   1 step_drying (:270-309; skipped when has_moisture == 0, Q_dry = 0),
 *   2 step_pyrolysis_md2004 (:311-462; m_water_init > 0 enables the
 *     moisture gate; A_p, E_p the thermal Arrhenius pair),
 *   4 step_char_oxidation (:465-564) of m_char,
 *   8 step_smoldering_oxidation (:567-614) of m_char.
 * With bit 2 set, Q_pyro_out = Q_md + Q_dry - Q_char - Q_smold over the
 * processes run (spread_3d.py:2183-2215) and S_pyro_out the volatile
 * source; Q_dry_out, Q_char_out, Q_smold_out (may be NULL) receive the
 * single heats.  Grid-type fields T_s, alpha_s, Y_O2 and outputs. */
int fw_pyro_step(const fw_grid* g, const void* T_s, const void* alpha_s, const void* Y_O2,
                 double* m_solid, double* m_char, double* m_water, int has_moisture,
                 double m_water_init, double A_p, double E_p, double dt, int processes,
                 void* S_pyro_out, void* Q_pyro_out, void* Q_dry_out, void* Q_char_out,
                 void* Q_smold_out, void* stream);

/* coupling_3d.step_gas_solid_coupling (coupling_3d.py:76-192), in place on
 * T_g, T_s (grid type) and m_water (float64): gas heating by Q_comb,
 * Ranz-Marshall convection to the solid, absorbed radiation, ground loss,
 * moisture evaporation from the solid's heat budget and the Q_pyro sink.
 * sigma_sav <= 0 leaves everything unchanged, as the reference does. */
int fw_gas_solid_coupling(const fw_grid* g, void* T_g, void* T_s, const void* rho,
                          const void* u, const void* v, const void* w, const void* alpha_s,
                          double sigma_sav, const void* q_rad_in, const void* Q_pyro,
                          const void* Q_comb, double* m_water, double L_v, double dt,
                          double T_amb, int q_loss_enable, double h_conv_mult, void* stream);

/* ---- time-step glue (spread_3d.py:1897-3683) ----------------------------- */

/* rho_prev <- rho; rho <- P0/(R max(T_g, T_amb)); Y_O2[:, :, 0] <- 0.232
 * (spread_3d.py:1984-1987, :2042, :2049-2050). */
int fw_step_prologue(const fw_grid* g, const void* T_g, void* rho,
                     void* rho_prev, void* Y_O2, double T_amb, void* stream);

/* Q_solid_ext on bed layers k < n_z_bed (spread_3d.py:2078-2103): drip_W_m3
 * on columns [i_src0, i_src1) plus clip(q_rad/dz, +-q_rad_max). */
int fw_bed_external_heat(const fw_grid* g, const void* q_rad, void* Q_ext,
                         int n_z_bed, int i_src0, int i_src1, double drip_W_m3,
                         double q_rad_max, void* stream);

/* S_pyro = S_pyro_part + S_dry_part; div_target = (S_pyro - (rho -
 * rho_prev)/dt)/max(rho, 0.1)  (spread_3d.py:2147-2154, :2300-2317). */
int fw_div_target(const fw_grid* g, const void* S_pyro_part,
                  const void* S_dry_part, const void* rho, const void* rho_prev,
                  double dt, void* S_pyro, void* div_target, void* stream);

/* Sub-step gas energy update of the particle path (spread_3d.py:2916-2957,
 * :3104-3105, :3129-3146), in place on T_g; Q_comb_out may be NULL. */
int fw_gas_energy_update(const fw_grid* g, void* T_g, const void* T_s,
                         const void* rho, const void* Y_H2O,
                         const void* S_pyro_part, const void* S_dry_part,
                         const void* Q_g_conv, const void* q_rad_gas,
                         void* Q_comb_out, int n_z_bed, int i_src0, int i_src1,
                         double drip_W_m3, double dt_sub, double T_amb,
                         void* stream);

/* fw_gas_energy_update followed by fw_solid_conduction_vertical in one pass
 * (spread_3d.py:3104-3146 then :3168-3177): T_g in place, T_s -> T_s_out
 * clamped to [T_s_lo, T_s_hi]; both read the T_s of before the conduction. */
int fw_gas_update_solid_conduction(const fw_grid* g, void* T_g, const void* T_s,
                                   void* T_s_out, const void* alpha_s, const void* rho,
                                   const void* Y_H2O, const void* S_pyro_part,
                                   const void* S_dry_part, const void* Q_g_conv,
                                   const void* q_rad_gas, void* Q_comb_out, int n_z_bed,
                                   int i_src0, int i_src1, double drip_W_m3, double dt_sub,
                                   double T_amb, double k_solid, double rho_solid,
                                   double cp_solid, double T_s_lo, double T_s_hi,
                                   void* stream);

/* field[k, :, i0:i1] += add  (ignition pulse, spread_3d.py:2528-2532). */
int fw_add_patch(const fw_grid* g, void* field, int k, int i0, int i1,
                 double add, void* stream);

/* Flame-body mask = (omega > 1e-3) | (T_g > 1000 & Y_F > 1e-3), identical to
 * compute_phi_flame_from_state(...) <= 0 (flame_front_3d.py:756-813); and
 * update_cell_age (flame_front_3d.py:816-835) when cell_age != NULL.
 * mask: uint8 field or NULL. */
int fw_flame_mask_age(const fw_grid* g, const void* omega, const void* T_g,
                      const void* Y_fuel, void* cell_age, uint8_t* mask,
                      double dt, void* stream);

/* All per-step scalars in one pass (dt inputs, NaN check, diagnostics of
 * spread_3d.py:1906-1941, :3503-3512, front tracking :3743-3777).
 * out (device double[14]): max|u|, max|v|, max|w|, max nu_t, max T_g,
 * T_g non-finite flag, max T_s, max S_pyro, max Q_comb, max omega,
 * max Y_fuel, count(T_s >= T_ign), count of flame-active cells
 * ((omega > 1e-3) | (T_g > 1000 & Y_F > 1e-3): the flame-body mask of the
 * next step, flame_front_3d.py:756-813), last burning column i (T_s >=
 * T_front on alpha_s > 0 in k < n_z_bed; -1 if none).  S_pyro and Q_comb
 * may be NULL.  work: 14 * 256 doubles. */
int fw_step_reduce(const fw_grid* g, const void* u, const void* v,
                   const void* w, const void* nu_t, const void* T_g,
                   const void* T_s, const void* S_pyro, const void* Q_comb,
                   const void* omega, const void* Y_fuel, const void* alpha_s,
                   double T_ign, double T_front, int n_z_bed, double* work,
                   double* out, void* stream);

/* LevelSetFront3D (flame_front_3d.py:113-360): mode 0 = evolve,
 * phi_out = phi - step v_n |grad phi|_Godunov; mode 1 = one Sussman reinit
 * substep, phi_out = phi - step sign(phi0)(|grad phi| - 1); mode 2 =
 * godunov_grad_norm (:113-188), phi_out = |grad phi|; mode 3 =
 * reinit_godunov_grad (:191-263), phi_out = the sign(phi0)-aware |grad phi|.
 * step is unused by modes 2 and 3, v_n by 1-3, phi0 by 0 and 2. */
int fw_levelset_update(const fw_grid* g, const void* phi, const void* phi0,
                       const void* v_n, void* phi_out, double step, int mode,
                       void* stream);

/* ---- flame front (flame_front_3d.py) ------------------------------------ */

/* LevelSetFront3D.initialize_source_patch (flame_front_3d.py:304-330) on the
 * float64 phi: -L_burnout/2 on i_start <= i < i_end, k <= k_top_bed; else
 * d = x_mid[i] - (x_mid[i_end-1] + dx/2), kept if positive, floored at
 * -L_burnout/2.  x_mid: device float64 (Nx).  Synchronises the stream. */
int fw_lset_init_patch(const fw_grid* g, double* phi, const double* x_mid, int i_start,
                       int i_end, int k_top_bed, double L_burnout, void* stream);

/* mask = (lo_strict ? phi > lo : phi >= lo) && phi <= hi on k < n_z_top, 0
 * above (n_z_top >= Nz: every layer).  phi: float64 field; mask: uint8.
 * LevelSetFront3D.burned_mask (flame_front_3d.py:379-381: lo = -inf, hi = 0),
 * flame_body_mask (:383-385: lo = -L_burnout, hi = 0), ahead_band_mask
 * (:387-389: lo = 0 strict, hi = band_m), flame_body_mask_from_phi_flame
 * (:803-813: lo = -inf, hi = band_m). */
int fw_phi_window_mask(const fw_grid* g, const double* phi, double lo, int lo_strict,
                       double hi, int n_z_top, uint8_t* mask, void* stream);

/* LevelSetFront3D.front_x(k, j) (flame_front_3d.py:391-402): the linear zero
 * crossing of row (k, j) of the float64 phi, +inf if none; synchronises the
 * stream and writes the host scalar. */
int fw_lset_front_x(const fw_grid* g, const double* phi, int k, int j, double* out_host,
                    void* stream);

/* LevelSetFront3D.enforce_z_uniformity (flame_front_3d.py:362-377): every
 * column of the float64 phi set to its z mean, in place. */
int fw_phi_z_mean(const fw_grid* g, double* phi, void* stream);

/* update_cell_age (flame_front_3d.py:816-835), in place on cell_age (grid
 * type): mask cells age by dt (from 0 if inf), the others become inf. */
int fw_update_cell_age(const fw_grid* g, const uint8_t* mask, void* cell_age, double dt,
                       void* stream);

/* compute_q_in_at_front_3d (flame_front_3d.py:440-471): out = mask ?
 * q_frankman + q_dom_fwd : 0, plus q_burst_2d (Ny, Nx) on the highest band
 * cell of each column.  q_frankman (NULL = zero) and q_burst_2d (NULL = none)
 * are optional; grid type throughout, mask uint8. */
int fw_q_in_at_front_3d(const fw_grid* g, const void* q_frankman, const void* q_dom_fwd,
                        const uint8_t* mask, const void* q_burst_2d, void* out, void* stream);

/* ---- combustion ---------------------------------------------------------- */

/* combustion_3d.step_o2_supply_rate (combustion_3d.py:184-260): on interior
 * cells (1 <= k <= Nz-2, 1 <= j <= Ny-2, 1 <= i <= Nx-2) omega_O2_out = the
 * upwind O2 mass inflow over the six faces / s_stoich; other cells are left
 * as the caller filled them (the reference fills 1e30).  The EDC closure of
 * the step does not read it (chemistry_closures/edc.py:133-376), so the
 * step driver skips it. */
int fw_o2_supply_rate(const fw_grid* g, const void* rho, const void* u, const void* v,
                      const void* w, const void* Y_O2, void* omega_O2_out, double s_stoich,
                      void* stream);


/* Number of int32 workspace entries fw_edc needs for grid g. */
int fw_edc_workspace_ints(const fw_grid* g);

/* chemistry_closures.run('edc') -> step_chemistry_ode_edc
 * (chemistry_closures/edc.py:133-376; dispatch __init__.py:46-62).
 * In place on T_g, Y_fuel, Y_O2; writes omega_out (time-averaged rate).
 * Reproduces the reference's carried wet-bulb flag (edc.py:223) exactly
 * (deterministic prefix scan in (k, j, i) order).  work: device int32 array
 * of fw_edc_workspace_ints(g) entries. */
int fw_edc(const fw_grid* g, const void* rho, void* T_g, void* Y_fuel,
           void* Y_O2, const void* k_turb, const void* eps_turb,
           const void* Y_H2O, void* omega_out, double chi_rad, double cp_g,
           double dt, int n_substeps, int extinction_enable, double s_stoich,
           double hoc_J, int* work, void* stream);

/* fw_edc without the workspace argument.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * segment summaries live in a library-owned static device array of
 * FW_EDC_STATIC_MAX_CELLS / 256 int32 (1 MiB), one copy per GPU, held until
 * the process ends; the entry point calls no allocator.  Grids of more than
 * FW_EDC_STATIC_MAX_CELLS cells are rejected with FW_E_CONFIG.  `stream`
 * must belong to the current device.  Every call on a GPU shares that GPU's
 * array: calls that overlap in time on one GPU (several streams or host
 * threads) race.  The array needs no clearing: each call writes every
 * summary it reads. */
#define FW_EDC_STATIC_MAX_CELLS 67108864
int fw_edc_static(const fw_grid* g, const void* rho, void* T_g, void* Y_fuel,
                  void* Y_O2, const void* k_turb, const void* eps_turb,
                  const void* Y_H2O, void* omega_out, double chi_rad, double cp_g,
                  double dt, int n_substeps, int extinction_enable, double s_stoich,
                  double hoc_J, void* stream);

/* ---- the outer time step (run_3d_spread loop, spread_3d.py:1897-3683) -- */

/* Device state of one run.  Grid-type (Nz, Ny, Nx) fields unless noted.
 * fw_spread_step exchanges ping-pong pointers (u/v/w, k/eps, T_g, T_s, the
 * species, phi) with the tmp* fields in place; read the struct back after
 * each step to know which buffer holds which field. */
typedef struct fw_spread_fields {
  void *rho, *rho_prev, *u, *v, *w, *T_g, *T_s, *Y_fuel, *Y_O2, *Y_H2O, *alpha_s;
  void *Fx, *Fy, *Fz, *k, *eps, *nu_t, *omega, *S_pyro, *div_target, *Q_comb;
  void *q_rad, *q_rad_gas, *Q_ext, *M_local, *T_s_old, *tmp1, *tmp2, *tmp3;
  /* per-cell particle sources: S_pyro, S_drying, Q_pyro, Q_drying,
   * Y_F_source, Q_char, Q_smold, Q_g_conv */
  void* src[8];
  void *u_inlet, *v_inlet, *w_inlet; /* (Nz, Ny) */
  void* sigma_x;                     /* (Nx) outflow sponge profile */
  void *k_wall, *eps_wall;           /* (Ny, Nx) wall ghosts */
  void *T_soil, *q_in_soil;          /* (n_soil, Ny, Nx), (Ny, Nx) */
  const double *soil_dz, *soil_d_above, *soil_d_below; /* (n_soil) */
  double *phi, *phi_tmp, *phi0;      /* float64 level set (Nz, Ny, Nx) */
  double* red_work;                  /* 14 * 256 */
  double* red_out;                   /* 14, device */
  double* red_host;                  /* 14, pinned host */
  double* nu_work;                   /* 256 */
  double* contrib;                   /* 10 * n_particles */
  double* dT_work;                   /* N */
  int32_t* counters;                 /* 2 */
  int32_t* edc_work;                 /* fw_edc_workspace_ints(g) */
  /* active level set only (may be NULL when lset_active == 0): */
  uint8_t* lset_mask;                /* N, the ahead band of the step */
  void* q_dom_fwd;                   /* N, grid type: DOM forward flux in the band */
  double* v_n;                       /* N, float64 front speed */
  /* Eulerian bed only (may be NULL when eulerian == 0): float64 pools and
   * the grid-type solid heat sink */
  double *eu_m_solid, *eu_m_char, *eu_m_water;
  void* Q_pyro;
  /* hipGraph replay of the sub-step loop (may be NULL: direct launches):
   * graphs from fw_graphs_create, step_scalars FW_STEP_SCALARS device
   * doubles.  Used only on a stream other than the legacy default stream
   * and for the Lagrangian bed. */
  void* graphs;
  double* step_scalars;
} fw_spread_fields;

/* Cache of instantiated hipGraphs for one case's step loops (no counterpart
 * in the reference: its step is interpreted Python).  A loop whose launch
 * sequence is fixed by its trip counts, flags and field pointers is recorded
 * once per distinct combination by stream capture and replayed with one
 * hipGraphLaunch; the per-step scalars it needs are read by its kernels from
 * fw_spread_fields.step_scalars.  Results are bit-identical to direct
 * launches.  A cache serves one case (one host thread) at a time. */
#define FW_STEP_SCALARS 8
int fw_graphs_create(void** graphs);
int fw_graphs_destroy(void* graphs);
/* replays and captures since creation */
int fw_graphs_stats(void* graphs, long long* replays, long long* captures);

/* Run constants resolved from the deck (spread_3d.py:959-1896). */
typedef struct fw_spread_params {
  double T_amb, h_bed, sigma_sav, canopy_C_d, canopy_beta_p, canopy_beta_d;
  double cfl_factor, min_dt, U_mf, dz_min, T_ign, ignition_duration, spin_up;
  double ignition_q_mult, bed_x_end, Lx, x_step, proj_div_tol;
  double rho_solid_true, cp_solid;   /* bed particles (horizontal conduction) */
  double soil_k, soil_rho, soil_cp, soil_eps;
  int32_t n_soil, i_src0, i_src1, proj_max_iter, dom_every, reserved;
  /* active level set (level_set_passive = false, spread_3d.py:3227-3300) */
  double band_m;        /* ahead band width max(dx, DX_VN_BAND_M) (:1601) */
  double E_sens;        /* rho_b cp_s h_bed (T_ign - T_amb) (flame_front_3d.py:503) */
  double E_lat_coef;    /* rho_b h_bed L_vap f_dry (:507), times M_local */
  double emp_w;         /* empirical-ROS blend weight (0 = off) */
  double emp_ros;       /* empirical ROS [m/s] */
  double seed_T_s;      /* T_s of the dynamic seeding (solid-phase ignition T) */
  double x_lo_bed;      /* x of the bed's leading face */
  double stall_window;  /* s without front advance before the stall exit */
  int32_t lset_active, reserved2;
  /* Eulerian bed (lagrangian_bed_enable = false, spread_3d.py:2183-2215,
   * :3148-3156) */
  double eu_m_water_init; /* rho_b M (0: no moisture, no drying, no gate) */
  double eu_A_p, eu_E_p;  /* MD2004 thermal Arrhenius pair */
  double eu_L_v;          /* latent heat of the coupling's evaporation (0 when dry) */
  double h_conv_mult;     /* Ranz-Marshall h_p scale */
  int32_t eulerian, reserved3;
} fw_spread_params;

/* Host-side scalars carried from step to step. */
typedef struct fw_spread_state {
  double t, dt;
  double front_exit_x, front_mid_x;  /* level-set front rows (k=1 / k=n_z_bed/2, j=Ny/2) */
  double last_front, last_advance_t, front_x_last;
  double front_t_new, front_x_new;   /* the front point appended this step */
  double proj_div;                   /* max|div(u) - target| after the projection loop */
  double red[14];                    /* end-of-step fw_step_reduce slots */
  int32_t step_no, lset_step, extinct_count;
  int32_t stop;                      /* 0 running, 1 extinct, 2 front exit, 3 domain end,
                                      4 stalled (active level set) */
  int32_t front_appended, proj_n, pcg_iters, reserved;
  int32_t seed_count, reserved2;     /* steps since the last dynamic seeding */
} fw_spread_state;

/* One outer step of run_3d_spread for the Lagrangian-bed / k-epsilon / EDC /
 * DOM configuration, passive or active level set (p->lset_active).  g: the
 * fields' grid; g64: the same grid with dtype FW_F64 (level set).  proj /
 * dom: handles from fw_proj_create / fw_dom_create.  The dt, NaN halt,
 * projection loop, exits, front tracking and diagnostics of the reference
 * live in `st`.  Returns FW_E_NONFINITE where the reference raises on
 * NaN/Inf. */
int fw_spread_step(const fw_grid* g, const fw_grid* g64, fw_spread_fields* f,
                   const fw_spread_params* p, fw_spread_state* st, void* proj,
                   void* dom, const fw_bed_particles* particles,
                   const fw_bed_index* index, fw_bed_params* bed, void* stream);

/* Up to max_steps calls of fw_spread_step in one call, the driver loop of
 * run_3d_spread (spread_3d.py:1897-3683) without a return to the caller
 * between steps: hist[i] receives the state after step i; the loop stops
 * after the first step that sets st->stop or returns an error (that step's
 * error is returned; *n_done counts the steps that completed). */
int fw_spread_run(const fw_grid* g, const fw_grid* g64, fw_spread_fields* f,
                  const fw_spread_params* p, fw_spread_state* st, void* proj,
                  void* dom, const fw_bed_particles* particles,
                  const fw_bed_index* index, fw_bed_params* bed, void* stream,
                  int max_steps, fw_spread_state* hist, int* n_done);

/* The end-of-step reduction into st->red (with_sources = 0 skips S_pyro and
 * Q_comb: the state before the first step). */
int fw_spread_reduce(const fw_grid* g, const fw_spread_fields* f,
                     const fw_spread_params* p, int with_sources,
                     fw_spread_state* st, void* stream);

/* Benchmark timing: while enabled, fw_spread_step records a HIP event pair
 * on its stream around every k-th sub-step k_species3 launch (at most
 * 4096 pairs).  fw_spread_timer(k), k >= 1, resets and starts with stride
 * k, fw_spread_timer(0) stops; fw_spread_timer_read waits for the events and
 * returns the summed milliseconds and the number of launches covered.  The timer
 * belongs to the calling host thread: it covers the steps that thread drives,
 * and threads stepping other cases concurrently keep their own. */
int fw_spread_timer(int enable);
int fw_spread_timer_read(double* total_ms, int* launches);

/* The same timing for other kernels (no counterpart in the reference):
 * which = FW_TIMER_SPECIES3 (the sub-steps' species kernel; what
 * fw_spread_timer drives) or FW_TIMER_DOM_SWEEP (every launch of the DOM's
 * tiled sweep, 27 per source iteration at 256 x 64 x 48).  stride >= 1
 * resets and starts the timer, 0 stops it; at most 4096 launches are
 * recorded.  Per calling host thread. */
enum { FW_TIMER_SPECIES3 = 0, FW_TIMER_DOM_SWEEP = 1, FW_TIMER_KINDS = 2 };
int fw_kernel_timer(int which, int stride);
int fw_kernel_timer_read(int which, double* total_ms, int* launches);

/* LevelSetFront3D.front_x (flame_front_3d.py:391-402) at (k=1, j=Ny/2) and
 * (k=n_z_bed/2, j=Ny/2) into st->front_exit_x / st->front_mid_x. */
int fw_spread_front(const fw_grid* g64, const fw_spread_fields* f,
                    fw_spread_state* st, void* stream);

/* ---- snapshot fields (snap_%04d.npz, spread_3d.py:3537-3670) ------------ */

/* compute_phi_flame_from_state (flame_front_3d.py:756-800): signed exact
 * Euclidean distance between cell centres to the active-flame set
 * (omega > 1e-3 or (T_g > 1000 and Y_fuel > 1e-3)), negative inside, with
 * sampling (dz_eff, dy, dx); +1e6 everywhere when no cell is active, -1e6
 * when every cell is.  dz_eff is the caller's mean of dz_arr (:788).
 * phi and work: float64 fields of the grid's size.  Axes up to 2048 cells. */
int fw_phi_flame(const fw_grid* g, const void* omega, const void* T_g,
                 const void* Y_fuel, double dz_eff, double* phi, double* work,
                 void* stream);

/* LevelSetFront3D.ahead_band_mask (flame_front_3d.py:387-389) cleared above
 * the bed top (spread_3d.py:2555-2565): mask = (0 < phi <= band_m) for
 * k < n_z_top, 0 elsewhere.  phi: the float64 level set. */
int fw_ahead_band_mask(const fw_grid* g, const double* phi, double band_m,
                       int n_z_top, uint8_t* mask, void* stream);

/* q_in_at_front_2d (spread_3d.py:3227-3233): the z sum of
 * compute_q_in_at_front_3d (flame_front_3d.py:440-471) with no Frankman flux
 * and no burst term, i.e. of q_dom_fwd where mask != 0 (mask may be NULL).
 * out: float64 (Ny, Nx). */
int fw_q_in_at_front_2d(const fw_grid* g, const void* q_dom_fwd,
                        const uint8_t* mask, double* out, void* stream);

/* tau_mix = max(k / max(eps, 1e-12), tau_buoy) (spread_3d.py:2454-2455),
 * tau_buoy = sqrt(2 max(dx, dy, dz) / g) (:1560-1561). */
int fw_tau_mix(const fw_grid* g, const void* k, const void* eps,
               double tau_buoy, void* out, void* stream);

/* Per-cell aggregates of the live bed particles (spread_3d.py:3569-3610):
 * float32 count, m_solid, m_water, m_char sums and the (m_solid + m_water +
 * m_char)-weighted T_s (0 where that mass is 0), over each cell's particles
 * of the index in slot order.  Outputs: float32 grid-sized fields. */
int fw_bed_snapshot(const fw_grid* g, const fw_bed_particles* particles,
                    const fw_bed_index* index, float* count, float* m_solid,
                    float* m_water, float* m_char, float* T_s_avg, void* stream);

/* ---- active level set (spread_3d.py:3227-3300) --------------------------- */

/* compute_v_n_3d (flame_front_3d.py:474-514) on q_in = q_dom_fwd (the
 * Frankman flux is zero and there is no burst term), then the empirical
 * blend (spread_3d.py:3252-3255):
 *   E = max(E_sens + E_lat_coef M_local, 1); v = max(q / E, 0);
 *   v = v (1 - emp_w) + emp_w emp_ros   when emp_w > 0.
 * q_dom_fwd, M_local: grid type; v_n: float64. */
int fw_lset_speed(const fw_grid* g, const void* q_dom_fwd, const void* M_local,
                  double E_sens, double E_lat_coef, double emp_w, double emp_ros,
                  double* v_n, void* stream);

/* v_n_2d of the snapshot (spread_3d.py:3645): the z mean of v_n, float64
 * (Ny, Nx). */
int fw_v_n_2d(const fw_grid* g, const double* v_n, double* out, void* stream);

/* Dynamic bed seeding behind the empirical-ROS front (spread_3d.py:3271-3300):
 * T_s = T_seed for particles with alive == 1, x_lo <= x <= x_front,
 * T_s < T_ign and m_solid > 1e-9. */
int fw_bed_seed(const fw_bed_particles* particles, double x_lo, double x_front,
                double T_ign, double T_seed, void* stream);

/* ---- learned preconditioner (no counterpart in the reference) ----------- */
/* The reference preconditions its Krylov solve with the separable inverse
 * only (projection_3d.py:756-791).  These entry points add a learned
 * correction: a 4-level 3-D U-Net (16/32/64/64 channels, 3x3x3 kernels, y
 * periodic, zero padding in x and z, grid dims multiples of 8) with bfloat16
 * weights and activations and float32 accumulation on the matrix cores,
 * trained on the solver's own projection systems, applied inside flexible
 * PCG (FW_PRECOND_LEARNED).  Layers (level, input channels -> output):
 * e0 L0 [r, q] 2 -> 16; e1 L1 pool(e0) 16 -> 32; e2 L2 pool(e1) 32 -> 64;
 * e3 L3 pool(e2) 64 -> 64; d2 L2 [e2, up(e3)] 128 -> 64; d1 L1 [e1, up(d2)]
 * 96 -> 32; d0 L0 [e0, up(d1)] 48 -> 16; c L0 d0 16 -> 1 (linear), relu
 * elsewhere.  Parameters are one flat float32 vector of fw_net_nparams()
 * values (per layer: weights (27, ci, co) then biases (co)). */
int fw_net_nparams(void);
/* q = RHO_REF / max(rho, 0.01) - 1 for n float32 values (the network's
 * density input, as the learned preconditioner forms it). */
int fw_net_density_feature(const void* rho, void* q, int64_t n, void* stream);
int fw_net_create(int nx, int ny, int nz, const float* theta_host, void** handle);
int fw_net_destroy(void* handle);
int fw_net_get_params(void* handle, float* theta_host);
/* Replaces the parameters and zeroes the Adam moments. */
int fw_net_set_params(void* handle, const float* theta_host);
/* c = N(r, q) (device pointers, N float32 values each; q = RHO_REF / rho - 1,
 * RHO_REF = 101325 / (8.314 / 0.029 * 300)). */
int fw_net_forward(void* handle, const void* r, const void* q, void* c_out, void* stream);
/* One training step on the sample (r, q, t): forward, the loss
 * L = sum dz_k (N(r, q) - t)^2 / sum dz_k t^2 (dz: device float64, Nz values),
 * backward, then an Adam step (lr, step number it >= 1; lr = 0 leaves the
 * parameters unchanged).  Returns L of the parameters before the step and
 * synchronises the stream. */
int fw_net_train_step(void* handle, const void* r, const void* q, const void* t,
                      const double* dz, double lr, int it, double* loss, void* stream);
/* Copies the gradient of the last fw_net_train_step to the host. */
int fw_net_grad_copy(void* handle, float* grad_host);

#ifdef __cplusplus
}
#endif
#endif /* FIREWAVE_H */
