/*
 * gravis.h -- C ABI of the MI355X (gfx950) force-and-integration engine.
 *
 * This is the drop-in boundary for sim-core's hot path. Every entry point
 * is extern "C", takes plain pointers and sizes, returns a gravis_status and
 * never aborts or throws across the ABI. The reference interface each entry
 * replaces is cited beside it (paths relative to the gravis repository).
 *
 * Conventions
 *   - Particle arrays are structure-of-arrays f64, as sim-core's Particles
 *     (crates/sim-core/src/particle.rs:21-64); no alignment requirement.
 *   - Plain entry points take HOST pointers borrowed for the call; the
 *     *_device variants take device pointers and run on the context's
 *     stream (gravis_set_stream) without synchronising.
 *   - Accelerations follow GravitySolver's accumulate contract
 *     (crates/sim-core/src/gravity.rs:6-10): results are ADDED to ax/ay/az;
 *     the caller zeroes them first.
 *   - A context is used by one host thread at a time (sim-core guards its
 *     Barnes-Hut tree cache with a Mutex, barnes_hut.rs:27,146).
 *   - G = 1 (crates/sim-core/src/units.rs:11).
 */
#ifndef GRAVIS_H
#define GRAVIS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gravis_status {
    GRAVIS_OK = 0,
    GRAVIS_INVALID_ARG = 1,   /* bad size, null pointer, eps < 0, dt <= 0 ... */
    GRAVIS_HIP_ERROR = 2,     /* a HIP runtime call failed (see gravis_last_error) */
    GRAVIS_OUT_OF_MEMORY = 3, /* device allocation failed */
    GRAVIS_UNSUPPORTED = 4,   /* valid request this build does not implement */
    GRAVIS_INTERNAL = 5       /* device-side capacity exceeded (e.g. walk frontier) */
} gravis_status;

/* --algorithm (crates/headless/src/main.rs:61-67,114-121) */
typedef enum gravis_algorithm {
    GRAVIS_BRUTE_FORCE = 0, /* gravity.rs:36-80 */
    GRAVIS_BARNES_HUT = 1,  /* barnes_hut.rs:21-192 (linear-octree form) */
    GRAVIS_NO_GRAVITY = 2   /* gravity.rs:13-19 */
} gravis_algorithm;

typedef enum gravis_precision {
    GRAVIS_FP64 = 0,  /* f64 throughout, as sim-core */
    GRAVIS_MIXED = 1  /* FP32 pair forces on separations formed in FP64
                         relative to each wavefront's target-box centre,
                         compensated FP32 accumulation, FP64 state */
} gravis_precision;

/* Solver description: BruteForce::new(softening) (gravity.rs:41-45) or
 * BarnesHut::new(softening, theta) (barnes_hut.rs:32-40). */
typedef struct gravis_solver {
    int32_t algorithm;   /* gravis_algorithm */
    int32_t precision;   /* gravis_precision */
    double softening;    /* Plummer softening eps >= 0 */
    double theta;        /* Barnes-Hut opening angle >= 0 */
    uint32_t leaf_size;  /* Barnes-Hut leaf capacity (0 = 16) */
    uint32_t group_size; /* Barnes-Hut target group, at most 64 (0 = the
                            context's bh_group_size, default 32) */
} gravis_solver;

/* Diagnostics (crates/sim-core/src/diagnostics.rs:5-30). */
typedef struct gravis_diag {
    double time;
    uint64_t step;
    double kinetic_energy;
    double potential_energy;
    double total_energy;
    double momentum[3];
    double angular_momentum[3];
    double center_of_mass[3];
    double center_of_mass_velocity[3];
    double virial_ratio;
    double thermal_energy;
} gravis_diag;

typedef struct gravis_ctx gravis_ctx;

/* Context: owns device buffers, the stream, the cached tree arena (the
 * analogue of BarnesHut::tree_cache, barnes_hut.rs:27) and tuning choices. */
gravis_status gravis_create(int device, gravis_ctx** out);
void gravis_destroy(gravis_ctx* ctx);
const char* gravis_last_error(const gravis_ctx* ctx);
const char* gravis_status_string(gravis_status s);
/* Stream for the *_device entry points (a hipStream_t; NULL = the
 * context's own stream). */
gravis_status gravis_set_stream(gravis_ctx* ctx, void* stream);
gravis_status gravis_synchronize(gravis_ctx* ctx);
/* Library version string. */
const char* gravis_version(void);

/* GravitySolver::compute_accelerations (gravity.rs:6-10, :47-80;
 * barnes_hut.rs:139-192). Accumulates into ax/ay/az. Host pointers. */
gravis_status gravis_accel(gravis_ctx* ctx, const gravis_solver* solver, size_t n,
                           const double* x, const double* y, const double* z,
                           const double* m, double* ax, double* ay, double* az);

/* Same with device pointers, asynchronous on the context stream. */
gravis_status gravis_accel_device(gravis_ctx* ctx, const gravis_solver* solver, size_t n,
                                  const double* x, const double* y, const double* z,
                                  const double* m, double* ax, double* ay, double* az);

/* No reference counterpart: gravis_accel_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Brute force only (FP64 or
 * MIXED; GRAVIS_NO_GRAVITY adds nothing; Barnes-Hut gives
 * GRAVIS_UNSUPPORTED), n <= 65,536 (larger n gives GRAVIS_INVALID_ARG), at
 * most 8 j chunks. Same kernels and accumulate contract as
 * gravis_accel_device; every scratch buffer lives in one static 20 MiB
 * device array per GPU, held until the process ends and shared by all
 * contexts on that GPU, so calls that overlap on one GPU race. Device
 * pointers, asynchronous on the context stream; no allocation. */
gravis_status gravis_accel_device_noalloc(gravis_ctx* ctx, const gravis_solver* solver,
                                          size_t n, const double* x, const double* y,
                                          const double* z, const double* m, double* ax,
                                          double* ay, double* az);

/* nsteps of LeapfrogKDK::step (integrator.rs:41-77) with the given solver:
 * v += a dt/2; x += v dt; a = F(x); v += a dt/2. On entry a must hold
 * a(t) (as after sim-core's initial clear + compute_accelerations,
 * headless/src/main.rs:151-152); on return it holds a(t + nsteps dt).
 * Host pointers; state is device resident across the nsteps. */
gravis_status gravis_kdk(gravis_ctx* ctx, const gravis_solver* solver, size_t n, double* x,
                         double* y, double* z, double* vx, double* vy, double* vz,
                         const double* m, double* ax, double* ay, double* az, double dt,
                         uint64_t nsteps);

/* Same with device pointers, asynchronous on the context stream. FP64
 * brute force runs every step in one launch: in one workgroup for n up to
 * "kdk_fused_max_n" (default 256, bit-identical to the step-by-step path),
 * and, for several steps of 256 < n <= "kdk_coop_max_n" (default 2040) on a
 * stream that is not being captured, in one cooperative launch with a grid
 * barrier per step (forces summed in another order, so they differ from the
 * step-by-step path in rounding); that call synchronises the stream, and
 * if the runtime refuses the cooperative launch the steps run launch by
 * launch. */
gravis_status gravis_kdk_device(gravis_ctx* ctx, const gravis_solver* solver, size_t n,
                                double* x, double* y, double* z, double* vx, double* vy,
                                double* vz, const double* m, double* ax, double* ay, double* az,
                                double dt, uint64_t nsteps);

/* diagnostics::compute / compute_fast (diagnostics.rs:37-48): FP64.
 * compute_potential = 0 gives compute_fast (U, E_pot and virial = 0).
 * u and ptype may be NULL (no gas: thermal energy 0); ptype 4 = gas
 * (particle.rs:4-10). Host pointers. */
gravis_status gravis_diagnostics(gravis_ctx* ctx, size_t n, double softening,
                                 int compute_potential, const double* x, const double* y,
                                 const double* z, const double* vx, const double* vy,
                                 const double* vz, const double* m, const double* u,
                                 const uint8_t* ptype, double time, uint64_t step,
                                 gravis_diag* out);

/* Same with device pointers; synchronises the stream to return *out. */
gravis_status gravis_diagnostics_device(gravis_ctx* ctx, size_t n, double softening,
                                        int compute_potential, const double* x,
                                        const double* y, const double* z, const double* vx,
                                        const double* vy, const double* vz, const double* m,
                                        const double* u, const uint8_t* ptype, double time,
                                        uint64_t step, gravis_diag* out);

/* ---- render view (crates/render-core/src/gpu_types.rs:5-9) --------------
 * GpuParticle: 32 bytes per particle, position f32 x3, mass f32, colour
 * f32 x4 (RGBA). Positions, masses and internal energies are rounded to
 * f32 as the sim thread's `as f32` does (crates/native-app/src/
 * sim_thread.rs:250-258); the colour is particle_color (crates/render-core/
 * src/color.rs:9-52): gas (type 4) by log10 of its internal energy, other
 * types from the fixed palette. u and ptype may be NULL (u = 0, type 0).
 * The _device form takes device pointers (out: 32 n bytes of device
 * memory a render thread can map) and runs asynchronously on the context
 * stream; gravis_render_pack takes host pointers and synchronises. */
typedef struct gravis_gpu_particle {
    float position[3];
    float mass;
    float color[4];
} gravis_gpu_particle;

gravis_status gravis_render_pack_device(gravis_ctx* ctx, size_t n, const double* x,
                                        const double* y, const double* z, const double* m,
                                        const double* u, const uint8_t* ptype, void* out);
gravis_status gravis_render_pack(gravis_ctx* ctx, size_t n, const double* x, const double* y,
                                 const double* z, const double* m, const double* u,
                                 const uint8_t* ptype, void* out);

/* ---- batched ensemble and reverse mode (no reference counterpart) -------
 * nsys independent systems stored back to back: system s owns particles
 * [offsets[s], offsets[s+1]) of the SoA arrays (offsets is a HOST array of
 * nsys+1 entries, offsets[0] = 0). Each system runs LeapfrogKDK
 * (integrator.rs:41-77) with brute-force gravity among its own particles
 * (solver->algorithm must be GRAVIS_BRUTE_FORCE; FP64 or MIXED). */

/* a = F(x) for every system (assigns, unlike gravis_accel). Host
 * pointers. */
gravis_status gravis_ensemble_accel(gravis_ctx* ctx, const gravis_solver* solver, int nsys,
                                    const uint64_t* offsets, const double* x, const double* y,
                                    const double* z, const double* m, double* ax, double* ay,
                                    double* az);
/* Same with device pointers, asynchronous. */
gravis_status gravis_ensemble_accel_device(gravis_ctx* ctx, const gravis_solver* solver,
                                           int nsys, const uint64_t* offsets, const double* x,
                                           const double* y, const double* z, const double* m,
                                           double* ax, double* ay, double* az);

/* nsteps of KDK for every system; a must hold a(t) on entry (as
 * gravis_kdk). use_graph != 0 captures one step (kick-drift + force/kick
 * launches) as a hipGraph and replays it. Host pointers. */
gravis_status gravis_ensemble_kdk(gravis_ctx* ctx, const gravis_solver* solver, int nsys,
                                  const uint64_t* offsets, double* x, double* y, double* z,
                                  double* vx, double* vy, double* vz, const double* m,
                                  double* ax, double* ay, double* az, double dt,
                                  uint64_t nsteps, int use_graph);
gravis_status gravis_ensemble_kdk_device(gravis_ctx* ctx, const gravis_solver* solver, int nsys,
                                         const uint64_t* offsets, double* x, double* y,
                                         double* z, double* vx, double* vy, double* vz,
                                         const double* m, double* ax, double* ay, double* az,
                                         double dt, uint64_t nsteps, int use_graph);

/* Reverse mode of nsteps KDK steps from (x0, v0) with masses m (a(x0) is
 * computed first). The final-state cotangent is either the gradient of
 * L = sum_i |x_i(T) - target_i|^2 (targets tx/ty/tz non-NULL; per-system
 * loss written to loss[nsys] if non-NULL) or given directly as gx/gy/gz
 * (position) and gvx/gvy/gvz (velocity, may be NULL). Outputs the
 * cotangents of the initial positions, velocities and masses; the
 * trajectory is recomputed from checkpoints every checkpoint_every steps
 * (0 = 32). final_state (6 n: x, y, z, vx, vy, vz at T) may be NULL. The
 * pair VJP runs in FP64; the recomputed forward uses solver->precision.
 * Host pointers. */
gravis_status gravis_ensemble_adjoint(gravis_ctx* ctx, const gravis_solver* solver, int nsys,
                                      const uint64_t* offsets, const double* x0,
                                      const double* y0, const double* z0, const double* vx0,
                                      const double* vy0, const double* vz0, const double* m,
                                      double dt, uint64_t nsteps, uint32_t checkpoint_every,
                                      const double* tx, const double* ty, const double* tz,
                                      const double* gx, const double* gy, const double* gz,
                                      const double* gvx, const double* gvy, const double* gvz,
                                      double* loss, double* final_state, double* grad_x,
                                      double* grad_y, double* grad_z, double* grad_vx,
                                      double* grad_vy, double* grad_vz, double* grad_m);
gravis_status gravis_ensemble_adjoint_device(
    gravis_ctx* ctx, const gravis_solver* solver, int nsys, const uint64_t* offsets,
    const double* x0, const double* y0, const double* z0, const double* vx0, const double* vy0,
    const double* vz0, const double* m, double dt, uint64_t nsteps, uint32_t checkpoint_every,
    const double* tx, const double* ty, const double* tz, const double* gx, const double* gy,
    const double* gz, const double* gvx, const double* gvy, const double* gvz, double* loss,
    double* final_state, double* grad_x, double* grad_y, double* grad_z, double* grad_vx,
    double* grad_vy, double* grad_vz, double* grad_m);

/* ---- smoothed particle hydrodynamics ------------------------------------
 * SphSolver (crates/sim-core/src/sph/solver.rs:20-53) and ForceParams
 * (sph/forces.rs:15-41). Defaults (gravis_sph_default_params) are
 * sim-core's: gamma 5/3, eta 1.2, h_max +inf, CFL 0.3, force factor 0.25,
 * beta 2, alpha in [0.1, 2], conductivity 1, switch decay sigma 0.1, and the
 * fixed-point h iteration (<= 5 passes, tolerance 1e-3) with the Wendland C2
 * kernel in FP64.
 * The remaining knobs select a physics variant with no reference
 * counterpart (gravis_sph_variant_params): cubic-spline M4 kernel,
 * Newton-Raphson h on rho(h) = m (eta/h)^3 to 1e-4 (re-query when 2 h
 * leaves the list's radius), alpha fixed (no switch), no conductivity,
 * gravity softening eps_i = h_i of the target (Barnes-Hut only), and FP32
 * pair arithmetic with FP64 per-particle sums. */
typedef struct gravis_sph_params {
    double gamma;        /* adiabatic index */
    double eta;          /* h = eta (m / rho)^(1/3) */
    double h_max;        /* upper bound on h (+inf disables) */
    double cfl;          /* CFL safety factor */
    double force_factor; /* acceleration timestep factor */
    double beta_visc;    /* Monaghan beta */
    double alpha_min;    /* Morris-Monaghan switch floor */
    double alpha_max;    /* switch ceiling */
    double alpha_cond;   /* Price conductivity coefficient (0 disables) */
    double sigma_visc;   /* switch decay parameter */
    double h_tol;        /* h iteration tolerance |dh|/h (1e-3) */
    int32_t kernel;      /* 0 Wendland C2 (sim-core), 1 cubic spline M4 */
    int32_t h_solver;    /* 0 fixed point (sim-core), 1 Newton-Raphson */
    int32_t max_h_iter;  /* h iterations (5) */
    int32_t alpha_switch;/* 1 Morris-Monaghan switch (sim-core), 0 alpha fixed */
    int32_t pair_fp32;   /* 0 FP64 pair loops (sim-core), 1 FP32 pairs, FP64 sums */
    int32_t soft_h;      /* 0 the gravity solver's softening, 1 eps_i = h_i */
} gravis_sph_params;

/* Particles (crates/sim-core/src/particle.rs:21-64): all 18 SoA arrays of n
 * particles; type 4 = gas (particle.rs:4-10). Host or device pointers per
 * entry point. */
typedef struct gravis_sph_state {
    size_t n;
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *m;
    double *rho, *p, *u, *dudt, *h, *c, *alpha;
    uint8_t* type;
} gravis_sph_state;

void gravis_sph_default_params(gravis_sph_params* out);
/* The variant above: kernel 1, h_solver 1, h_tol 1e-4, max_h_iter 30,
 * alpha_switch 0, alpha_cond 0, soft_h 1, pair_fp32 1. */
void gravis_sph_variant_params(gravis_sph_params* out);

/* SphSolver::compute (solver.rs:61-80): ball-query neighbour lists at
 * 2.2 h (sph/neighbors.rs:56-119), the density / smoothing-length iteration
 * with re-queries and the equation of state (sph/density.rs:42-157), then
 * the forces (sph/forces.rs:69-232). Writes rho, h, p, c of gas particles,
 * ADDS to ax/ay/az and dudt, and writes the ForceIntermediates div_v,
 * max_mu, dalpha_dt (n each; 0 for non-gas). omega (grad-h factors, 1 for
 * non-gas) and nneigh (final neighbour-list lengths) may be NULL. Host
 * pointers. */
gravis_status gravis_sph_compute(gravis_ctx* ctx, const gravis_sph_params* params,
                                 const gravis_sph_state* state, double* div_v, double* max_mu,
                                 double* dalpha_dt, double* omega, uint32_t* nneigh);
/* Same with device pointers, asynchronous except for the synchronisations
 * that read back the re-query count. */
gravis_status gravis_sph_compute_device(gravis_ctx* ctx, const gravis_sph_params* params,
                                        const gravis_sph_state* state, double* div_v,
                                        double* max_mu, double* dalpha_dt, double* omega,
                                        uint32_t* nneigh);

/* GravitySolver::compute_accelerations for an SPH state, as step_with_sph
 * calls it (solver.rs:180-181): accumulates into ax/ay/az with the solver's
 * softening, or with eps_i = h_i of each target when params->soft_h
 * (Barnes-Hut only; GRAVIS_UNSUPPORTED otherwise). Host pointers. */
gravis_status gravis_sph_accel(gravis_ctx* ctx, const gravis_sph_params* params,
                               const gravis_solver* gravity, gravis_sph_state* state);
/* Same with device pointers, asynchronous. */
gravis_status gravis_sph_accel_device(gravis_ctx* ctx, const gravis_sph_params* params,
                                      const gravis_solver* gravity,
                                      const gravis_sph_state* state);

/* step_with_sph (solver.rs:145-212): half kicks of v and u (floor 1e-10),
 * drift, gravity with `gravity` (NULL or GRAVIS_NO_GRAVITY: none), SPH
 * compute, half kicks of v and u, viscosity switch update, and the next
 * timestep of compute_timestep (solver.rs:90-131) in *dt_next. Host
 * pointers. */
gravis_status gravis_sph_step(gravis_ctx* ctx, const gravis_sph_params* params,
                              const gravis_solver* gravity, gravis_sph_state* state, double dt,
                              double* dt_next);

/* nsteps of headless's SPH loop (crates/headless/src/main.rs:178-187):
 * each step runs step_with_sph with the current dt, then
 * dt = min(dt_next, 1.5 dt); *dt_io holds the first dt on entry and the
 * next step's dt on return; *time_io is advanced by the steps taken
 * (may be NULL). State stays device resident across the steps. Host
 * pointers. */
gravis_status gravis_sph_run(gravis_ctx* ctx, const gravis_sph_params* params,
                             const gravis_solver* gravity, gravis_sph_state* state,
                             uint64_t nsteps, double* dt_io, double* time_io);
/* Same with device state pointers (dt_io, time_io stay host scalars). */
gravis_status gravis_sph_run_device(gravis_ctx* ctx, const gravis_sph_params* params,
                                    const gravis_solver* gravity, gravis_sph_state* state,
                                    uint64_t nsteps, double* dt_io, double* time_io);

/* ---- introspection (no reference counterpart: test and bench hooks) ---- */

/* Kernel timing: when enabled, the dominant force kernel of every call is
 * bracketed by HIP events on the stream it runs on. gravis_kernel_time
 * returns the summed milliseconds and launch count for `name`
 * ("bf_fp64", "bf_mixed", "bf_matrix", "bh_walk", "sph_query",
 * "sph_density", "sph_force", and the octree phases "morton_sort" (root
 * box, keys, radix sort) and "tree_build" (gather, levels, sums, groups),
 * which include the gaps between their launches) and resets them. */
gravis_status gravis_set_profiling(gravis_ctx* ctx, int enable);
gravis_status gravis_kernel_time(gravis_ctx* ctx, const char* name, double* ms_total,
                                 uint64_t* launches);

/* Tuning knobs (no reference counterpart): "bh_frontier_cap" = per-wave
 * breadth-first frontier capacity of the Barnes-Hut walk in nodes (default
 * 16384; a group whose frontier would exceed 4096 nodes, or this capacity if
 * smaller, is walked one target at a time, and a single target's frontier
 * must fit it); "sph_query_leaves" = overlapping tree leaves one wave of the SPH
 * neighbour search holds (1..512, default 512; a group needing more is
 * searched one target at a time, depth first); "sph_pool_init" = initial
 * neighbour-pool capacity in entries per particle (1..4096, default 96; the
 * pool grows past it, keeping the lists written, when a round runs out);
 * "kdk_fused_max_n" (0..256, default 256) and "kdk_coop_max_n" (0..2040,
 * default 2040) = the largest systems gravis_kdk(_device) steps in one
 * workgroup and in one cooperative launch with FP64 brute force (0
 * disables). */
gravis_status gravis_set_tuning(gravis_ctx* ctx, const char* key, int64_t value);

/* Launch shapes of the context (no reference counterpart). Brute force:
 * targets per lane of the mixed kernel (1 or 2: 64 or 128 targets per
 * wavefront), j records per staged tile (128 or 256), wavefronts per CU the
 * j-range split aims for; Barnes-Hut: target group size used when the
 * solver's group_size is 0. Set one at a time with gravis_set_tuning keys
 * "bf_ipt", "bf_tile", "bf_waves_per_cu", "bh_group_size", "bf_matrix"
 * (0: VALU kernel; 1 or 2: the mixed brute force forms |x_j - x_i|^2 +
 * eps^2 on the matrix cores, 64 targets per wavefront, bf_ipt unused --
 * 1 with v_mfma_f32_16x16x4f32, 2 with v_mfma_f32_16x16x32_bf16 on FP32
 * values split exactly in three bf16 pieces);
 * "bf_near_64ths" (0..64, default 1) is the near-tile fraction q, in 64ths,
 * of the mixed brute-force and ensemble kernels: a source tile whose gap to
 * the wavefront's target box and the softening are both below q times the
 * box diagonal gets FP64 pair separations (the matrix-core kernel uses the
 * stricter q' with (1 + 1/q')^2 = 2 (1 + 1/q), which keeps the same
 * per-pair error bound). */
typedef struct gravis_tuning {
    int32_t bf_ipt;
    int32_t bf_tile;
    int32_t bf_waves_per_cu;
    uint32_t bh_group_size;
    int32_t bf_matrix; /* 1, 2: mixed brute force on the matrix-core kernel */
    double best_ms;    /* gravis_autotune: mean time of one force call, chosen shape */
} gravis_tuning;

gravis_status gravis_get_tuning(gravis_ctx* ctx, gravis_tuning* out);

/* Autotuner (no reference counterpart): times the candidate launch shapes
 * for solver s on the given particles (host arrays, copied once) and keeps
 * the chosen one in the context. Mixed brute force: bf_waves_per_cu
 * 64, 128, 256, 512, 768, 1024 x (bf_ipt 2, bf_ipt 1, bf_matrix 1,
 * bf_matrix 2) x bf_tile (48 shapes); FP64 brute force: bf_waves_per_cu
 * 32..1024 (7 values). Each brute-force
 * shape is timed once (a warm-up and two calls); the four fastest are timed
 * again in three interleaved rounds of four calls and the fastest total is
 * kept, a shape later in that order (more j chunks, i.e. more partial-sum
 * memory) only when it is more than 0.5 % faster.
 * Barnes-Hut with group_size 0: group sizes 64, 32, 16, keeping 32 or 16
 * only when it is more than 2 % faster than the ones before it; each group
 * size costs five force calls. Shapes change the summation grouping,
 * so forces move in their last bits (brute force) or within the
 * opening-angle error (Barnes-Hut groups). out may be NULL. When a
 * candidate fails the call returns its error and the context keeps the
 * shapes it had before the call. */
gravis_status gravis_autotune(gravis_ctx* ctx, const gravis_solver* s, size_t n,
                              const double* x, const double* y, const double* z,
                              const double* m, gravis_tuning* out);

/* Work counters of the last Barnes-Hut force call: out[0] frontier node
 * tests, out[1] accepted (monopole) nodes, out[2] leaf particles used as
 * sources, out[3] target-source pair interactions, out[4] tree nodes. */
gravis_status gravis_bh_stats(gravis_ctx* ctx, uint64_t out[5]);

/* Work counters of the last SPH compute: out[0] target groups searched
 * depth first (LDS capacity exceeded), out[1] re-queries, out[2] query
 * rounds, out[3] neighbour-pool entries in use. */
gravis_status gravis_sph_stats(gravis_ctx* ctx, uint64_t out[4]);

/* Morton keys (63-bit, 21 bits per axis over the root cube of
 * octree.rs:14-44) and their stable sort permutation, device computed,
 * copied to host. box_out = (cx, cy, cz, half_width). */
gravis_status gravis_morton_sort(gravis_ctx* ctx, size_t n, const double* x, const double* y,
                                 const double* z, uint64_t* keys_sorted, uint32_t* perm,
                                 double* box_out);


/* ---- snapshots (host; crates/sim-core/src/snapshot.rs:5-157) ------------
 * Binary format NBODY001/002/003, little endian. gravis_snapshot_write is
 * Snapshot::capture + write_to (always v3: positions, velocities, masses,
 * particle types, internal energy, smoothing length, density).
 * gravis_snapshot_info reads the header; gravis_snapshot_read
 * (Snapshot::read_from) fills a state of that count (state->n must equal
 * it) and zeroes every field the file does not carry, accelerations
 * included: recompute them before stepping (Snapshot::initialize). Files
 * that are not snapshots give GRAVIS_INVALID_ARG. */
gravis_status gravis_snapshot_write(const char* path, double time, uint64_t step,
                                    double softening, double dt, const gravis_sph_state* state);
gravis_status gravis_snapshot_info(const char* path, int* version, double* time, uint64_t* step,
                                   double* softening, double* dt, uint64_t* count);
gravis_status gravis_snapshot_read(const char* path, gravis_sph_state* state);

/* ---- initial conditions (host; sim-core's scenario generators) ---------- */

/* PlummerSphere::generate (crates/sim-core/src/scenarios/plummer_sphere.rs:
 * 106-133): n equal masses M/n, inverse-CDF radii, Aarseth-Henon-Wielen
 * velocities, ChaCha20Rng::seed_from_u64(seed) stream, COM frame. */
gravis_status gravis_scenario_plummer(size_t n, double scale_radius, double total_mass,
                                      uint64_t seed, double* x, double* y, double* z,
                                      double* vx, double* vy, double* vz, double* m);
/* plummer_sphere.rs:135-152 */
double gravis_plummer_suggested_dt(double scale_radius, double total_mass);
double gravis_plummer_suggested_softening(size_t n, double scale_radius);
/* TwoBody::generate (crates/sim-core/src/scenarios/two_body.rs:49-92):
 * 2 particles at periapsis in the COM frame; period (two_body.rs:38-41). */
gravis_status gravis_scenario_two_body(double eccentricity, double mass_ratio,
                                       double semi_major_axis, double* x, double* y, double* z,
                                       double* vx, double* vy, double* vz, double* m);
double gravis_two_body_period(double semi_major_axis);
/* EvrardCollapse::generate (crates/sim-core/src/scenarios/
 * evrard_collapse.rs:45-57) with fill_uniform_gas_sphere and sphere_spacing
 * (scenarios/mod.rs:17-51): n gas particles (type 4) uniform in a sphere of
 * radius R by rejection from the cube [-R, R)^3 (ChaCha20Rng seeded with
 * seed), mass M/n, h = 1.5 (V/n)^(1/3), u = u0, alpha 1, at rest; every
 * other field 0 (Particles::add_gas, particle.rs:136-154). state->n must be
 * n; host arrays. The suggested softening is 0.3 (V/n)^(1/3)
 * (evrard_collapse.rs:67-70). */
gravis_status gravis_scenario_evrard(size_t n, double total_mass, double radius, double u0,
                                     uint64_t seed, gravis_sph_state* state);
double gravis_evrard_softening(size_t n, double radius);
/* Variant initial conditions (no reference counterpart; the configuration
 * with rho ~ 1/r): n gas particles in a sphere of radius R and mass M with
 * rho(r) = M / (2 pi R^2 r), i.e. r = R sqrt(U) with U = random::<f64>(),
 * cos(theta) in U[-1, 1), phi in U[0, 2 pi), drawn per particle in that
 * order from ChaCha20Rng::seed_from_u64(seed); at rest, u = u0, alpha 1,
 * m = M/n, h = eta (m / rho(r))^(1/3). Host arrays, state->n = n. */
gravis_status gravis_scenario_evrard_1r(size_t n, double total_mass, double radius, double u0,
                                        double eta, uint64_t seed, gravis_sph_state* state);
/* ColdCollapse::generate (crates/sim-core/src/scenarios/cold_collapse.rs:
 * 56-102): n particles of mass M/n uniform in a sphere of radius R at rest,
 * then the COM frame. gas = 0: rejection with random_range(-1..1) per axis
 * scaled by R, type 0; gas = 1: fill_uniform_gas_sphere with u = 0.05 and
 * h = 1.5 (V/n)^(1/3), type 4. state->n = n; host arrays. Suggested dt
 * t_ff/200 (t_ff/500 for gas), softening 0.3 (V/n)^(1/3)
 * (cold_collapse.rs:104-123). */
gravis_status gravis_scenario_cold_collapse(size_t n, double radius, double total_mass,
                                            uint64_t seed, int gas, gravis_sph_state* state);

/* GalaxyCollision's fields (crates/sim-core/src/scenarios/galaxy_collision.rs:
 * 20-68); gravis_galaxy_default_params gives Default::default(). */
typedef struct gravis_galaxy_params {
    uint64_t n_per_galaxy;
    uint64_t seed;
    double disk_fraction;
    double halo_mass_ratio;
    double baryon_mass;
    double disk_scale_radius;
    double disk_scale_height;
    double bulge_scale_radius;
    double halo_scale_radius;
    double disk_max_radius;
    double halo_max_radius;
    double separation;
    double impact_parameter;
    double approach_velocity;
} gravis_galaxy_params;

void gravis_galaxy_default_params(gravis_galaxy_params* params);
/* GalaxyCollision::generate (galaxy_collision.rs:75-157): two galaxies of
 * disk (type 1), Plummer bulge (type 2) and Hernquist halo (type 3) from one
 * ChaCha20Rng::seed_from_u64(seed) stream, placed at (-+sep/2, -+b/2, 0)
 * with x velocities +-v/2, then moved to the COM frame.
 * state->n = 2 * n_per_galaxy; host arrays; SPH fields zero. */
gravis_status gravis_scenario_galaxy_collision(const gravis_galaxy_params* params,
                                               gravis_sph_state* state);
/* suggested_dt (t_orb(R_d) / 500) and suggested_softening (0.05 R_d),
 * galaxy_collision.rs:159-185. */
double gravis_galaxy_suggested_dt(const gravis_galaxy_params* params);
double gravis_galaxy_suggested_softening(const gravis_galaxy_params* params);
/* Batched-ensemble layout (no reference counterpart): counts[s] =
 * lo + (u64 draw) % (hi - lo + 1) from ChaCha20Rng::seed_from_u64(seed),
 * in system order; and loss targets U[-1, 1)^3 drawn x, y, z per particle. */
gravis_status gravis_scenario_ensemble_counts(uint64_t seed, int nsys, uint32_t lo, uint32_t hi,
                                              uint32_t* counts);
gravis_status gravis_scenario_targets(uint64_t seed, size_t n, double* tx, double* ty,
                                      double* tz);
/* Two Plummer spheres (n_each particles, scale radius a, mass M each;
 * seeds seed_a, seed_b) whose centres are `separation` apart along x with
 * offset `impact` along y, on a parabolic approach: relative speed
 * sqrt(2 G 2M / sqrt(separation^2 + impact^2)), centre-of-mass frame.
 * Sphere A fills indices [0, n_each), B [n_each, 2 n_each). The layout
 * follows GalaxyCollision (crates/sim-core/src/scenarios/
 * galaxy_collision.rs:134-154), with Plummer spheres in place of disks. */
gravis_status gravis_scenario_plummer_pair(size_t n_each, double scale_radius,
                                           double total_mass, uint64_t seed_a, uint64_t seed_b,
                                           double separation, double impact, double* x,
                                           double* y, double* z, double* vx, double* vy,
                                           double* vz, double* m);

#ifdef __cplusplus
}
#endif

#endif /* GRAVIS_H */
