/*
 * cdft.h - C ABI of the MI355X (gfx950) free-energy + gradient library.
 *
 * This is the drop-in boundary for the hot path of migduroli/classicalDFT:
 * the functions below replace, one for one, the reference C++ entry points
 * cited beside them.  Plain pointers and sizes only: host arrays are borrowed
 * for the duration of a call; `_device` variants take device pointers that
 * the caller owns (allocated with hipMalloc on the context's device).
 *
 * Conventions (all mirror the reference):
 *   - real fields are double, length N = Nx*Ny*Nz, idx = iz + Nz*(iy + Ny*ix)
 *     (include/dft/grid.hpp:24-37);
 *   - forces are delta(beta Omega)/delta rho * dV (functionals.hpp:55-63);
 *   - mu and V_ext are in kT units;
 *   - every function returns CDFT_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is cdft_last_error().
 *     CDFT_EINVAL is the reference's std::invalid_argument, CDFT_ERUNTIME its
 *     std::runtime_error (for example FIRE's "exceeded max uphill steps",
 *     include/dft/algorithms/fire.hpp:106-108).
 *   - a context is bound to one device and one HIP stream and is not
 *     thread-safe; calls on one context must be serialised by the caller.
 */
#ifndef CDFT_H
#define CDFT_H

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define CDFT_API __attribute__((visibility("default")))
#else
#define CDFT_API
#endif

#define CDFT_OK 0
#define CDFT_EINVAL (-1)
#define CDFT_ERUNTIME (-2)
#define CDFT_EHIP (-3)
#define CDFT_EUNSUPPORTED (-4)

typedef struct cdft_ctx cdft_ctx;

/* dft::make_grid(dx, box) (include/dft/grid.hpp:87-108): each box length must
 * be an integer multiple of dx within 1e-10. */
typedef struct {
  double dx;
  double box[3];
} cdft_grid;

/* physics::Interaction (include/dft/physics/interactions.hpp:23-30) with its
 * potential (include/dft/physics/potentials.hpp:53-368). */
typedef struct {
  int species_i, species_j;
  int potential;     /* 0 Lennard-Jones, 1 ten Wolde-Frenkel, 2 Wang-Ramirez-Dobnikar-Frenkel */
  double sigma, epsilon, r_cutoff, alpha;
  int split;         /* 0 Weeks-Chandler-Andersen, 1 Barker-Henderson */
  int weight_scheme; /* 0 zero, 1 linear-E, 2 linear-F, 3 quadratic-F (default), 4 Gauss-E, 5 Gauss-F */
} cdft_interaction;

#define CDFT_FMT_NONE (-1)
#define CDFT_FMT_ROSENFELD 0
#define CDFT_FMT_RSLT 1
#define CDFT_FMT_WHITEBEAR_I 2
#define CDFT_FMT_WHITEBEAR_II 3
#define CDFT_FMT_ESFMT 4

#define CDFT_PRECISION_FP64 0
#define CDFT_PRECISION_MIXED 1

/* Kernel schedule of an evaluation.  AUTO picks the fused six-kernel
 * schedule when the grid has kernels for it (y and z extents 64 or 128, x
 * extent 64, 128 or 256; in FP64 an x extent of 256 needs a y extent of 64)
 * and the general per-axis schedule otherwise; GENERAL forces the latter.
 * Both give the same results to rounding. */
#define CDFT_SCHEDULE_AUTO 0
#define CDFT_SCHEDULE_GENERAL 1
#define CDFT_SCHEDULE_FUSED 2

typedef struct {
  int fmt_model;          /* fmt::FMTModel (include/dft/functionals/fmt/models.hpp:363-419) */
  double esfmt_A, esfmt_B; /* esFMT(A, B), models.hpp:314-358 */
  int precision;          /* CDFT_PRECISION_FP64 or CDFT_PRECISION_MIXED */
  int device;             /* HIP device ordinal */
  int schedule;           /* CDFT_SCHEDULE_AUTO or CDFT_SCHEDULE_GENERAL */
} cdft_options;

/* Replaces functionals::make_weights + make_functional
 * (include/dft/functionals/functionals.hpp:27-33, functional.hpp:93-107):
 * builds the FMT and mean-field weights for a physics::Model.  n_species is
 * 1..8.  A context of more than one species (a mixture; up to 36 interactions
 * between any species pairs) is evaluated by cdft_total_species, in FP64, or
 * in mixed precision (FP32 real-space fields) on grids the fused transforms
 * cover (y and z extents 64 or 128, x extent 64, 128 or 256; elsewhere
 * CDFT_EUNSUPPORTED); the other evaluation and algorithm entry points act on
 * one species and return CDFT_EUNSUPPORTED for it. */
CDFT_API int cdft_ctx_create(const cdft_grid* grid, const double* diameters, int n_species,
                    const cdft_interaction* interactions, int n_interactions, double kT,
                    const cdft_options* options, cdft_ctx** out);
CDFT_API void cdft_ctx_destroy(cdft_ctx* ctx);
CDFT_API const char* cdft_last_error(void);

/* Grid::shape (grid.hpp:26-33). */
CDFT_API int cdft_grid_shape(const cdft_ctx* ctx, long shape[3]);
/* The schedule the context evaluates with: CDFT_SCHEDULE_FUSED or
 * CDFT_SCHEDULE_GENERAL (no counterpart in the reference, which has one
 * CPU schedule). */
CDFT_API int cdft_schedule(const cdft_ctx* ctx, int* schedule);
/* InteractionWeight::a_vdw of interaction k (mean_field.hpp:23-28). */
CDFT_API int cdft_a_vdw(const cdft_ctx* ctx, int k, double* a_vdw);
/* BulkThermodynamics::chemical_potential for a one-component fluid with the
 * grid-synchronised a_vdw (bulk/thermodynamics.hpp:222-237, functional.hpp:93-102). */
CDFT_API int cdft_bulk_chemical_potential(const cdft_ctx* ctx, double rho, double* mu);

/* functionals::Result plus the per-term split of free_energy
 * (functionals.hpp:64-95, types.hpp:21-32). */
typedef struct {
  double free_energy;
  double grand_potential;
  double ideal;
  double external;
  double hard_sphere;
  double mean_field;
  double mass; /* sum rho dV */
} cdft_result;

/* functionals::total(model, state, weights) (functionals.hpp:64): one
 * F + delta F / delta rho evaluation.  rho, vext (nullable: zero field) and
 * forces (nullable) are host arrays of N doubles; mu is beta*mu. */
CDFT_API int cdft_total(cdft_ctx* ctx, const double* rho, const double* vext, double mu, cdft_result* result,
               double* forces);
/* Same with device-resident arrays; returns when the result is complete. */
CDFT_API int cdft_total_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu, cdft_result* result,
                      double* d_forces);
/* Enqueues `count` evaluations back to back on the context stream without
 * host synchronisation (throughput measurement); the result of the last is
 * returned. */
CDFT_API int cdft_total_device_repeat(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu, int count,
                             cdft_result* result, double* d_forces);

/* physics::Model::species.size() (physics/model.hpp:15-22). */
CDFT_API int cdft_species_count(const cdft_ctx* ctx, int* n_species);
/* functionals::total for a state of any number of species
 * (functionals.hpp:64-95 with hard_sphere.hpp:174-254 and
 * mean_field.hpp:233-281): rho, vext (nullable) and forces (nullable) are host
 * arrays of n_species * N doubles, species-major; mu holds n_species values of
 * beta*mu.  result->mass is the total sum rho dV over species.  On a
 * one-species context this is cdft_total with mu[0]. */
CDFT_API int cdft_total_species(cdft_ctx* ctx, const double* rho, const double* vext, const double* mu,
                                cdft_result* result, double* forces);
/* Same with device-resident arrays; returns when the result is complete. */
CDFT_API int cdft_total_species_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext, const double* mu,
                                       cdft_result* result, double* d_forces);
/* The terms of functionals::total separately, for any species count:
 * ideal_gas (ideal_gas.hpp:18-38), external_field (external_field.hpp:16-31),
 * hard_sphere (hard_sphere.hpp:263-298) and mean_field (mean_field.hpp:233-281)
 * forces, each nullable, n_species * N doubles species-major; result holds
 * the energy of every term as cdft_total_species gives it.  The hard-sphere
 * forces are the total less the three pointwise terms. */
CDFT_API int cdft_contributions(cdft_ctx* ctx, const double* rho, const double* vext, const double* mu,
                                cdft_result* result, double* f_ideal, double* f_external, double* f_hard_sphere,
                                double* f_mean_field);
/* bulk::BulkThermodynamics (bulk/thermodynamics.hpp:200-255) for the context's
 * species, FMT model and grid-synchronised a_vdw: rho and mu hold n_species
 * positive bulk densities and the chemical potentials (kT units).  As in the
 * reference, the free-energy density counts a cross pair as a*rho_i*rho_j
 * while the chemical potentials give it half (thermodynamics.hpp:206-237).
 * The grand-potential density is -pressure. */
CDFT_API int cdft_bulk_free_energy_density(const cdft_ctx* ctx, const double* rho, double* f);
CDFT_API int cdft_bulk_chemical_potentials(const cdft_ctx* ctx, const double* rho, double* mu);
CDFT_API int cdft_bulk_pressure(const cdft_ctx* ctx, const double* rho, double* p);
/* The same equation of state for n_states bulk states at once on the device,
 * one thread per state (for phase-diagram and coexistence scans;
 * bulk/coexistence.hpp and phase_diagram.hpp evaluate it state by state):
 * rho holds n_states * n_species densities (state-major); f and p receive
 * n_states values and mu n_states * n_species; each output nullable. */
CDFT_API int cdft_bulk_batch(const cdft_ctx* ctx, long n_states, const double* rho, double* f, double* mu,
                             double* p);
/* GPU time of `count` back-to-back mixture evaluations (HIP events on the
 * context stream; no counterpart in the reference). */
CDFT_API int cdft_time_total_species_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext,
                                            const double* mu, int count, double* d_forces, double* ms);

/* Measurement hooks (no counterpart in the reference, which has no timers):
 * GPU time of `count` back-to-back evaluations measured with HIP events on the
 * context stream, and the per-stage (per-kernel) average time and
 * algorithmic HBM bytes of one evaluation; at most CDFT_MAX_STAGES stages. */
#define CDFT_MAX_STAGES 11
CDFT_API int cdft_time_total_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu, int count,
                                    double* d_forces, double* ms);
CDFT_API int cdft_profile_stages(cdft_ctx* ctx, const double* d_rho, const double* d_vext, double mu,
                                 double* d_forces, int reps, double* stage_ms, double* stage_bytes, int* n_stages);
CDFT_API const char* cdft_stage_name(const cdft_ctx* ctx, int i);

/* hard_sphere.hpp:50-67 (convolve_weights) and mean_field.hpp:259: the 12
 * weighted-density fields eta, n2, n_v2[x,y,z], T[xx,xy,xz,yy,yz,zz] and the
 * mean-field convolution [w * rho] (host arrays, 12*N doubles). */
CDFT_API int cdft_weighted_densities(cdft_ctx* ctx, const double* rho, double* out);

/* cdft_weighted_densities without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory during
 * a call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The fields are staged in a library-wide device workspace
 * of CDFT_NOALLOC_SCRATCH_BYTES per device, held until the process ends; a
 * context whose n_fields * N doubles do not fit is refused with CDFT_EINVAL.
 * Every call on one device shares that workspace, so calls that overlap on a
 * device race: serialise them. */
#define CDFT_NOALLOC_SCRATCH_BYTES (64L << 20)
CDFT_API int cdft_weighted_densities_noalloc(cdft_ctx* ctx, const double* rho, double* out);

/* math::FourierTransform::forward / backward (src/math/fourier.cpp:97-103):
 * unnormalised 3D r2c / c2r in the reference half-complex layout
 * idx = kz + (Nz/2+1)*(iy + Ny*ix), complex as interleaved (re, im). */
CDFT_API int cdft_fft_r2c(cdft_ctx* ctx, const double* in, double* out_half);
CDFT_API int cdft_fft_c2r(cdft_ctx* ctx, const double* in_half, double* out);

/* fire::Fire (include/dft/algorithms/fire.hpp:33-44).  force_max_tolerance
 * has no counterpart in the reference: when positive, the Minimizer runs
 * also stop once max |delta(beta Omega)/delta rho| (density space, without
 * the dV factor, of the forces FIRE relaxes: after the homogeneous-boundary
 * average and the fixed-mass multiplier) falls below it, checked after every
 * step; zero (the default of a zero-initialised struct) leaves only the
 * reference's energy monitor. */
typedef struct {
  double dt, dt_max, dt_min, alpha_start, f_inc, f_dec, f_alpha;
  int n_delay, max_uphill;
  double force_tolerance;
  int max_steps;
  double force_max_tolerance;
} cdft_fire;

/* minimization::Parametrization (minimization.hpp:20-39): kind 0 Unbounded
 * (rho = rho_min + x^2), kind 1 Bounded (rho_min + (rho_max-rho_min) x^2/(1+x^2)). */
typedef struct {
  int kind;
  double rho_min, rho_max;
} cdft_parametrization;

/* minimization::Result (minimization.hpp:114-121) plus the final FIRE dt, alpha. */
typedef struct {
  double free_energy, grand_potential, rms_force;
  int iterations, converged;
  double dt, alpha;
} cdft_min_result;

/* Minimizer::grand_potential (minimization.hpp:280-353): FIRE2 in the alias
 * variable at fixed mu.  vext (nullable) adds an external field, which the
 * reference's from_profile path always zeroes (init.hpp:28-42).
 * energies (nullable, max_steps + 1 doubles) receives state.energy after
 * initialisation and after every step.  Returns CDFT_ERUNTIME when FIRE
 * exceeds max_uphill. */
CDFT_API int cdft_minimize_grand_potential(cdft_ctx* ctx, const cdft_fire* fire, const cdft_parametrization* param,
                                  int use_homogeneous_boundary, const double* rho0, const double* vext, double mu,
                                  double* rho_out, cdft_min_result* result, double* energies);

/* Minimizer::fixed_mass (minimization.hpp:182-278): FIRE2 in the alias
 * variable with the density rescaled to target_mass before every evaluation,
 * the evaluation at mu = 0 (its Omega is F), and the Lagrange multiplier
 * lambda = sum(g rho) / target_mass taken off the forces, g -= lambda dV, after
 * the homogeneous-boundary average.  rho_out receives the rescaled final
 * density; result holds its F and Omega at chemical potential mu.  energies
 * (nullable, max_steps + 1 values) receive F after initialisation and every
 * step. */
CDFT_API int cdft_minimize_fixed_mass(cdft_ctx* ctx, const cdft_fire* fire, const cdft_parametrization* param,
                             int use_homogeneous_boundary, const double* rho0, const double* vext, double mu,
                             double target_mass, double* rho_out, cdft_min_result* result, double* energies);

/* Several independent Minimizer::grand_potential runs advanced in lockstep
 * (no counterpart in the reference, which solves one state point at a time;
 * for many small state points that leave most of the device idle alone).
 * Run i minimises on context ctxs[i] from rho0[i] at mu[i] with the shared
 * FIRE settings; the contexts must be distinct and on one device, their
 * grids may differ.  One HIP graph advances every running state point by a
 * block of up to 8 FIRE steps (each run's step kernels captured on its own
 * context stream, chained per run only), so the device works on all of them
 * at once from one launch; each run follows exactly the trajectory
 * cdft_minimize_grand_potential gives it.  status[i] receives the run's
 * status (CDFT_ERUNTIME when FIRE exceeded max_uphill, which ends that run
 * only); the call itself fails on invalid arguments and device errors. */
CDFT_API int cdft_minimize_grand_potential_batch(int n, cdft_ctx* const* ctxs, const cdft_fire* fire,
                                        const cdft_parametrization* param, int use_homogeneous_boundary,
                                        const double* const* rho0, const double* mu, double* const* rho_out,
                                        cdft_min_result* results, int* status);

/* dynamics::StepConfig (include/dft/algorithms/dynamics.hpp:26-33). */
typedef struct {
  double dt, diffusion_coefficient, min_density, dt_max, fp_tolerance;
  int fp_max_iterations;
} cdft_step_config;

/* Simulation::run summary (dynamics.hpp:54-61). */
typedef struct {
  double mass_initial, mass_final, time, energy, dt;
  long evaluations;
  int steps;  /* steps taken: n_steps, or fewer when a stop callback ended the run */
} cdft_sim_result;

/* Host callbacks of the algorithm entry points.  The reference's algorithms
 * take any force callable (dynamics::ForceCallback, dynamics.hpp:21-22), a
 * boundary condition and a stop condition (Simulation::boundary and
 * stop_condition, dynamics.hpp:545-562,567) and record snapshots
 * (dynamics.hpp:43-58); Picard takes any constraint (picard.hpp:21).  Every
 * member is nullable.  Callbacks receive `user` and host arrays of N doubles
 * valid for the duration of the call; a nonzero return aborts the entry point
 * with CDFT_ERUNTIME.  The transforms, stencils and propagators stay on the
 * device; only the fields the callbacks see travel to the host.
 *   force:    the energy and the forces (delta(beta Omega)/delta rho * dV) at
 *             rho; null: the context's functional at mu with vext.
 *   project:  modifies rho in place: the boundary condition of a DDFT run
 *             (applied to the initial state and after every step), or a
 *             Picard constraint (applied after every update).
 *   stop:     after every DDFT step; sets *stop nonzero to end the run there.
 *   snapshot: receives the density after step 0 and after every
 *             snapshot_interval steps (snapshot_interval <= 0: step 0 only). */
typedef struct {
  void* user;
  int (*force)(void* user, const double* rho, double* energy, double* forces);
  int (*project)(void* user, double* rho);
  int (*stop)(void* user, int step, double time, double energy, int* stop);
  int (*snapshot)(void* user, int step, double time, double energy, const double* rho);
  int snapshot_interval;
} cdft_hooks;

/* Simulation::run with integrating_factor_step (dynamics.hpp:333-401,582-693)
 * for one species.  The force callback is hooks->force, or (hooks or its
 * force null) the context's grand_potential_callback(mu) (functional.hpp:62-73)
 * plus an optional external field; hooks (nullable) may also carry the
 * boundary condition (project), the stop condition and the snapshot receiver.
 * rho is updated in place; energies and times (each nullable, n_steps + 1
 * values) receive the energy and the elapsed time after the initial state
 * and after every step taken (SimulationResult::energies / times, recorded
 * there every log_interval steps).  A step that has to restart more than 30
 * times (dt cut by 1e30), or whose iterate is not finite, fails with
 * CDFT_ERUNTIME; the reference has no such floor and would keep dividing dt. */
CDFT_API int cdft_ddft_run(cdft_ctx* ctx, const cdft_step_config* config, int n_steps, const double* vext, double mu,
                  const cdft_hooks* hooks, double* rho, cdft_sim_result* result, double* energies, double* times);

/* Autotuner (no counterpart in the reference, which leaves FFT planning to
 * FFTW_ESTIMATE, src/math/fourier.cpp:34-57): for grids the fused kernels do
 * not cover, times the evaluation under alternative FFT radix orders and
 * launch shapes and keeps the fastest whose Omega agrees with the default's.
 * Mixture contexts (which always run the general transforms) are tuned on
 * their own evaluation.  Results are cached per grid, precision, channel set
 * and species count for the process, so later contexts of the same shape take
 * them without timing.  us_default and
 * us_tuned (nullable) receive the time of one evaluation before and after, in
 * microseconds; a fused-schedule context is timed and left as it is. */
CDFT_API int cdft_autotune(cdft_ctx* ctx, int reps, double* us_default, double* us_tuned);

/* saddle_point::EigenvalueSolver (include/dft/algorithms/saddle_point.hpp:136-155). */
typedef struct {
  double tolerance;     /* residual norm |Hv - lambda v| to stop at */
  int max_iterations;
  double hessian_eps;   /* > 0: forward-difference products; 0: exact products (FP64 contexts) */
  int fixed_boundary;   /* zero the boundary forces, as Search::find's force function does */
} cdft_eigen_config;

typedef struct {
  double eigenvalue;
  int iterations, converged, hessian_products;
} cdft_eigen_result;

/* EigenvalueSolver::solve (saddle_point.hpp:159-270): the smallest eigenpair
 * of the Hessian of the grand potential at rho (forces of cdft_total at mu,
 * vext nullable) by LOBPCG.  initial_guess (nullable, N values) is the start;
 * without it the start is the seeded standard-normal sequence of `seed`
 * (SplitMix64 + Box-Muller, standing in for arma::randn).  eigenvector
 * receives the normalised eigenvector (N values). */
CDFT_API int cdft_smallest_eigenpair(cdft_ctx* ctx, const cdft_eigen_config* config, const double* rho,
                            const double* vext, double mu, const double* initial_guess, unsigned long long seed,
                            double* eigenvector, cdft_eigen_result* result);

/* saddle_point::Search (saddle_point.hpp:313-347). */
typedef struct {
  cdft_fire fire;             /* Search::fire (dt 1e-3, dt_max 1, alpha_start 0.01, f_alpha 1, tol 1e-4) */
  cdft_eigen_config eigen;    /* Search::eigenvalue; fixed_boundary is forced on */
  cdft_parametrization param; /* Search::param (Unbounded{1e-18}) */
  double omega_tolerance, tolerance_start, tolerance_factor;
  int max_outer_iterations, eigenvalue_clear_count;
} cdft_saddle_config;

typedef struct {
  double grand_potential, eigenvalue;
  int outer_iterations, converged;
  long evaluations;   /* functional evaluations of the FIRE runs */
} cdft_saddle_result;

/* Search::find (saddle_point.hpp:367-589): the critical cluster by
 * eigenvector following at fixed mass.  Per outer iteration: the smallest
 * eigenpair of the Hessian (cdft_smallest_eigenpair with a fixed boundary;
 * the previous eigenvector as start from outer iteration
 * eigenvalue_clear_count on, else the seeded normal sequence of seed + outer),
 * then FIRE in the alias variable with the interior mass held at target_mass
 * minus the boundary mass, the boundary forces zeroed, the interior Lagrange
 * multiplier removed and the force component along the (alias-transformed)
 * eigenvector removed, stopped on its rms force at a tolerance tightened by
 * tolerance_factor each outer iteration; stops when Omega changes by less
 * than omega_tolerance.  rho_out and eigenvector (nullable) receive N values. */
CDFT_API int cdft_saddle_search(cdft_ctx* ctx, const cdft_saddle_config* config, const double* rho0,
                       const double* vext, double mu, double target_mass, unsigned long long seed, double* rho_out,
                       double* eigenvector, cdft_saddle_result* result);

/* split_operator_step (include/dft/algorithms/dynamics.hpp:405-462), n_steps
 * times: half a step of exact diffusion exp(-D k^2 dt / 2) (continuous k^2,
 * compute_k_squared), a forward-Euler step of the excess term, another half
 * step; densities clamped at min_density after each part.  The force
 * callback is hooks->force, or (hooks or its force null) the grand-potential
 * gradient at mu (vext nullable); no other hook applies here.  energy
 * (nullable) receives the energy of the final density. */
CDFT_API int cdft_ddft_split_operator(cdft_ctx* ctx, const cdft_step_config* config, int n_steps, const double* vext,
                             double mu, const cdft_hooks* hooks, double* rho, double* energy);

/* crank_nicholson_step (dynamics.hpp:464-538), n_steps times: the fixed-point
 * iteration rho1_k = ((1 - a) rho0_k + dt/2 (R0_k + R1_k)) / (1 + a),
 * a = D k^2 dt / 2, until the largest change is below fp_tolerance or
 * fp_max_iterations.  Force callback as for cdft_ddft_split_operator.
 * iterations (nullable) receives the last step's count. */
CDFT_API int cdft_ddft_crank_nicolson(cdft_ctx* ctx, const cdft_step_config* config, int n_steps, const double* vext,
                             double mu, const cdft_hooks* hooks, double* rho, double* energy, int* iterations);

/* picard::Picard (include/dft/algorithms/picard.hpp:25-33). */
typedef struct {
  double mixing, min_density, tolerance;
  int max_iterations;
} cdft_picard;

/* picard::PicardResult (picard.hpp:17-23). */
typedef struct {
  double grand_potential, residual;
  int iterations, converged;
} cdft_picard_result;

/* Picard::solve (picard.hpp:63-124): rho *= exp(-mixing f / dV) until the
 * rms force or the change of Omega is below the tolerance.  The forces are
 * hooks->force, or (hooks or its force null) the grand-potential gradient at
 * mu (vext nullable).  target_mass (nullable: no constraint) applies
 * fixed_mass_constraint (minimization.hpp:157-180) after every update, for
 * any target, positive, zero or negative, with background (nullable: zero);
 * hooks->project (nullable) is then applied as a further constraint.  rho
 * (N values) is updated in place. */
CDFT_API int cdft_picard_solve(cdft_ctx* ctx, const cdft_picard* config, const double* vext, double mu,
                      const double* target_mass, const double* background, const cdft_hooks* hooks, double* rho,
                      cdft_picard_result* result);

/* ---------------------------------------------------------------- adjoint
 * The reference has no derivative of its dynamics; the entry points below
 * extend its DDFT integrating-factor step (dynamics.hpp:333-401) with a
 * reverse-mode adjoint.  They require a context of CDFT_PRECISION_FP64.
 *
 * cdft_hessian_vector: out = H(rho) u with H = d(force)/d(rho), the Jacobian
 * of the forces of cdft_total (delta(beta Omega)/delta rho * dV): dV u / rho
 * (ideal), dV W_mf u (mean field) and dV sum_ab W_a^T Phi_ab W_b u (FMT, the
 * second derivatives of Phi by forward-mode dual numbers).  H is symmetric.
 * Host arrays of N values. */
CDFT_API int cdft_hessian_vector(cdft_ctx* ctx, const double* rho, const double* u, double* out);

/* External potential of n_wells Gaussian wells (kT units):
 *   V(r) = sum_j A_j exp(-|r - c_j|^2 / (2 w_j^2)),
 * r = (ix, iy, iz) dx, minimum-image distances in the periodic box.
 * params holds 5 values per well: A, w, cx, cy, cz.  vext: N values (host). */
CDFT_API int cdft_wells_potential(cdft_ctx* ctx, int n_wells, const double* params, double* vext);

/* The map the adjoint differentiates: n_steps integrating-factor steps with
 * a fixed dt and exactly fp_iterations fixed-point iterations per step (the
 * reference's tolerance exit, restarts and dt adaptation are branches and
 * have no derivative). */
typedef struct {
  double dt;
  int fp_iterations;        /* >= 1 */
  int n_steps;              /* K >= 1 */
  int checkpoint_interval;  /* steps between stored densities; <= 0: ceil(sqrt(K)) */
} cdft_adjoint_config;

typedef struct {
  double loss;             /* sum (rho_K - target)^2 dV */
  double min_density;      /* smallest density among all iterates of the forward map */
  long evaluations;        /* functional evaluations (forward pass and segment recomputation) */
  long hessian_products;   /* Hessian-vector products of the backward pass */
  int checkpoints;         /* densities stored by the forward pass */
  double store_bytes;      /* device bytes of the checkpoints and of one segment's step states */
} cdft_adjoint_result;

/* Forward map only: rho (host, N values) is advanced in place.  vext may be
 * null.  min_density (nullable) receives the smallest density of any iterate. */
CDFT_API int cdft_ddft_fixed_run(cdft_ctx* ctx, const cdft_adjoint_config* config, const double* vext, double mu,
                        double* rho, double* min_density);

/* Loss L = sum_r (rho_K(r) - target(r))^2 dV after the forward map from rho0
 * under the well potential, and its gradient by the adjoint, run backwards
 * through the K steps with checkpointing (densities every
 * checkpoint_interval steps; each segment is recomputed once and its step
 * states kept on the device).
 *   grad:      5 * n_wells values, dL/d(A, w, cx, cy, cz) of every well;
 *   grad_rho0: N values or null, dL/d(rho0);
 *   grad_vext: N values or null, dL/dV(r) for the potential as a free field;
 *   rho_final: N values or null, rho_K. */
CDFT_API int cdft_ddft_adjoint_wells(cdft_ctx* ctx, const cdft_adjoint_config* config, int n_wells,
                            const double* params, double mu, const double* rho0, const double* target,
                            cdft_adjoint_result* result, double* grad, double* grad_rho0, double* grad_vext,
                            double* rho_final);

#ifdef __cplusplus
}
#endif

#endif /* CDFT_H */
