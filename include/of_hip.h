/*
 * of_hip.h -- C ABI of the MI355X (gfx950) incompressible projection step.
 *
 * This is the boundary a host program binds to: plain C types, device pointers and sizes,
 * no C++ or torch types. Every entry point names the reference interface it replaces
 * (paths relative to the zfifteen/solver checkout). The C++ host layer in
 * solver_amd/cpp/ restates those C++ interfaces on top of this ABI, and solver_amd/__init__.py
 * binds it with ctypes.
 *
 * Field storage. Every field argument is a device pointer to FP64 data in the reference's
 * ghosted storage layout (core/field_layout.hpp:69-290, core/grid.hpp:57-136): one ghost
 * layer, i fastest, flat index i + sx*(j + sy*k), with
 *   cell-centred  (sx,sy,sz) = (nx+2, ny+2, nz+2)
 *   u (x faces)   (sx,sy,sz) = (nx+3, ny+2, nz+2)   (the periodic face is stored twice)
 *   v (y faces)   (sx,sy,sz) = (nx+2, ny+3, nz+2)
 *   w (z faces)   (sx,sy,sz) = (nx+2, ny+2, nz+3)
 * so a reference StructuredField::data() buffer can be copied to the device unchanged.
 *
 * Errors. Every function returns OF_OK (0) or a negative status. The message of the last
 * failure on the calling thread is returned by of_last_error(); it carries the same text
 * as the reference's exception (e.g. "Taylor-Green pressure solve did not converge at step
 * 3 (iterations=1, relative_residual=..., tolerance=...)"). OF_INVALID_ARGUMENT maps to the
 * reference's std::invalid_argument, OF_RUNTIME_ERROR to std::runtime_error.
 *
 * Streams. Work is queued on the context's stream (hipStream_t passed as void*; NULL means
 * the context creates its own default-priority stream). Entry points that return
 * diagnostics synchronise that stream before returning; the others only enqueue.
 */
#ifndef OF_HIP_H
#define OF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OF_OK = 0,
  OF_INVALID_ARGUMENT = -1,
  OF_RUNTIME_ERROR = -2,
  OF_HIP_ERROR = -3,
  OF_NOT_CONVERGED = -4
};

/* core/grid.hpp:57-89 (ghost_layers is fixed at 1, as in every reference driver) */
typedef struct of_grid {
  int nx, ny, nz;
  double dx, dy, dz;
} of_grid;

/* core/field_layout.hpp:149-154 */
enum { OF_CELL = 0, OF_FACE_X = 1, OF_FACE_Y = 2, OF_FACE_Z = 3 };

/* bc/boundary_conditions.hpp:10-22 */
enum { OF_NO_SLIP_WALL = 0, OF_PRESCRIBED_VELOCITY = 1, OF_SYMMETRY = 2, OF_FIXED_PRESSURE = 3, OF_PERIODIC = 4 };
enum { OF_PRESSURE_NEUMANN = 0, OF_PRESSURE_DIRICHLET = 1, OF_PRESSURE_PERIODIC = 2 };

/* bc/boundary_conditions.hpp:24-44; faces ordered x_min, x_max, y_min, y_max, z_min, z_max */
typedef struct of_bc {
  int type;
  double velocity[3];
  double pressure;
  double pressure_gradient;
} of_bc;
typedef struct of_bc_set {
  of_bc faces[6];
} of_bc_set;

/* bc/boundary_conditions.hpp:46-62 */
typedef struct of_pressure_bc {
  int type;
  double value;
  double gradient;
} of_pressure_bc;
typedef struct of_pressure_bc_set {
  of_pressure_bc faces[6];
} of_pressure_bc_set;

/* solver/projection.hpp:15-20 */
typedef struct of_projection_options {
  double dt;
  double density;
  int poisson_max_iterations;
  double poisson_tolerance;
} of_projection_options;

/* Precision of the pressure solve. FP64: the reference's MGPCG in FP64 throughout.
 * MIXED: FP32 multigrid V-cycle (smoothing, residual, transfer) inside the FP64 outer CG
 * (residual, dot products, pressure update); the outer CG iterates to the same relative
 * residual tolerance. */
enum { OF_PRECISION_FP64 = 0, OF_PRECISION_MIXED = 1 };

/* solver/projection.hpp:22-39 (strings are fixed: mgpcg / geometric_multigrid / v_cycle /
 * damped_jacobi; residual_history is returned through a caller buffer) */
typedef struct of_poisson_diagnostics {
  int iterations;
  double initial_residual_l2;
  double final_residual_l2;
  double relative_residual;
  double removed_rhs_mean;
  int multigrid_levels;
  int coarse_unknowns;
  int pre_smoothing_steps;
  int post_smoothing_steps;
  int converged;
  int zero_mean_enforced;
  int history_length; /* number of residual_history entries (may exceed the buffer) */
} of_poisson_diagnostics;

/* solver/projection.hpp:41-47 (+ the L-infinity divergence, which the reference omits) */
typedef struct of_projection_diagnostics {
  double rhs_l2;
  double divergence_l2_before;
  double divergence_l2_after;
  double divergence_linf_after;
  double pressure_mean;
  of_poisson_diagnostics pressure_solve;
} of_projection_diagnostics;

/* solver/momentum_terms.hpp:27-33 */
typedef struct of_cfl_diagnostics {
  double dt, max_cfl, max_u, max_v, max_w;
} of_cfl_diagnostics;

typedef struct of_context of_context;

/* ---------------------------------------------------------------- context and errors */
/* No reference counterpart (the reference is single-threaded CPU code). Creates the device
 * context (kernel scratch, reduction buffers, stream) on HIP device `device`. */
int of_context_create(int device, void* stream, of_context** out);
/* Drops the caller's reference. Every stepper and adjoint run created on the context holds a
 * reference of its own, so the native context (and its stream) is released with the last of
 * them: handles may be destroyed in any order. A HIP failure while releasing is returned as
 * OF_HIP_ERROR by the destroy call that ran it (and clears HIP's last error), never left for
 * a later call. */
int of_context_destroy(of_context* ctx);
int of_context_synchronize(of_context* ctx);
const char* of_last_error(void);
/* library version / build string; no reference counterpart */
const char* of_build_info(void);
/* Live timing of one kernel by name (no reference counterpart): while set, every launch of
 * that kernel on the context stream (finest multigrid level only for the V-cycle kernels) is
 * bracketed with HIP events; of_profile_read returns the count and total/min/max duration
 * and restarts the record. NULL or "" switches timing off. */
int of_profile_kernel(of_context* ctx, const char* kernel_name);
int of_profile_read(of_context* ctx, int* launches, double* total_ms, double* min_ms, double* max_ms);
/* The same record launch by launch: the first `capacity` durations in launch order (out_ms may
 * be NULL), the count in *launches; restarts the record. */
int of_profile_durations(of_context* ctx, double* out_ms, int capacity, int* launches);
/* Brackets only every stride-th launch of the timed kernel (default 1): each bracketing event
 * record leaves the device idle for a few microseconds, so a throughput run samples. */
int of_profile_stride(of_context* ctx, int stride);

/* No reference counterpart: sqrt(mean((a - b)^2)) over the active entries of two fields of one
 * location (a deterministic device reduction; the manufactured-solution check uses it). */
int of_active_rms_difference(of_context* ctx, const of_grid* grid, int location, const double* a, const double* b,
                             double* out);

/* ---------------------------------------------------------------- boundary conditions */
/* solver/projection.cpp:966-971 apply_velocity_boundary_conditions */
int of_apply_velocity_boundary_conditions(of_context* ctx, const of_grid* grid, const of_bc_set* bcs, double* u,
                                          double* v, double* w);
/* solver/projection.cpp:973-984 apply_pressure_boundary_conditions */
int of_apply_pressure_boundary_conditions(of_context* ctx, const of_grid* grid, const of_pressure_bc_set* bcs,
                                          double* p);
/* solver/projection.cpp:986-1026 apply_total_pressure_boundary_conditions */
int of_apply_total_pressure_boundary_conditions(of_context* ctx, const of_grid* grid, const of_bc_set* bcs,
                                                const double* src_u, const double* src_v, const double* src_w,
                                                double* p_total);
/* solver/projection.cpp:934-964 derive_pressure_correction_boundary_conditions (host-only) */
int of_derive_pressure_correction_boundary_conditions(const of_bc_set* bcs, of_pressure_bc_set* out);

/* ---------------------------------------------------------------- discrete operators */
/* operators/discrete_operators.hpp:7 compute_gradient */
int of_compute_gradient(of_context* ctx, const of_grid* grid, const double* p, double* gx, double* gy, double* gz);
/* operators/discrete_operators.hpp:8 compute_divergence */
int of_compute_divergence(of_context* ctx, const of_grid* grid, const double* u, const double* v, const double* w,
                          double* div);
/* operators/discrete_operators.hpp:9-10 compute_laplacian (location = OF_CELL / OF_FACE_*) */
int of_compute_laplacian(of_context* ctx, const of_grid* grid, int location, const double* in, double* out);

/* ---------------------------------------------------------------- momentum terms */
/* solver/momentum_terms.hpp:36-38 compute_advection_term; scheme 0 tvd(van Leer), 1 upwind, 2 central */
int of_compute_advection_term(of_context* ctx, const of_grid* grid, int scheme, const double* u, const double* v,
                              const double* w, double* au, double* av, double* aw);
/* solver/momentum_terms.hpp:40-42 compute_diffusion_term */
int of_compute_diffusion_term(of_context* ctx, const of_grid* grid, double viscosity, const double* u,
                              const double* v, const double* w, double* du, double* dv, double* dw);
/* solver/momentum_terms.hpp:44 compute_advective_cfl */
int of_compute_advective_cfl(of_context* ctx, const of_grid* grid, const double* u, const double* v,
                             const double* w, double dt, of_cfl_diagnostics* out);
/* solver/taylor_green.cpp:196-225 compute_factorized_correction (variant 0) and
 * solver/lid_driven_cavity.cpp:234-252 compute_factorized_correction_2d (variant 1) */
int of_compute_factorized_correction(of_context* ctx, const of_grid* grid, int variant, double alpha,
                                     const double* u, const double* v, const double* w, double* fu, double* fv,
                                     double* fw);
/* solver/taylor_green.cpp:227-256 assemble_predictor_rhs (variant 0) and
 * solver/lid_driven_cavity.cpp:489-518 assemble_lid_driven_cavity_predictor_rhs (variant 1).
 * u,v,w carry BC ghosts. prev_* may be NULL (first step: forward-Euler advection). The
 * advection of u,v,w is written to adv_*; adv_* may alias prev_* (read before written). */
int of_assemble_predictor_rhs(of_context* ctx, const of_grid* grid, int variant, int scheme, double viscosity,
                              double dt, const double* u, const double* v, const double* w, const double* p_total,
                              const double* prev_u, const double* prev_v, const double* prev_w, double* adv_u,
                              double* adv_v, double* adv_w, double* rhs_u, double* rhs_v, double* rhs_w);

/* ---------------------------------------------------------------- Helmholtz predictor */
/* solver/projection.hpp:49-52 solve_predictor_adi: (1-a dxx)(1-a dyy)(1-a dzz) u* = rhs,
 * solved exactly by batched line solves. line_solves as counted by the reference. */
int of_solve_predictor_adi(of_context* ctx, const of_grid* grid, const of_bc_set* bcs, double alpha,
                           const double* rhs_u, const double* rhs_v, const double* rhs_w, double* out_u,
                           double* out_v, double* out_w, int* line_solves);

/* The CG alternative to the line solves, on grids periodic on every axis: unpreconditioned
 * conjugate gradients on the factorised operator per component, started from the right-hand
 * side, stopping at sqrt(r.r / r0.r0) <= tolerance (metal/taylor_green_backend.mm:546-621, the
 * reference's Metal backend; its FP64 tolerance is 1e-12 with at most 80 iterations). The result
 * carries BC ghosts like of_solve_predictor_adi. iterations (may be NULL): the largest count of
 * the three components; relative_residual (may be NULL): the largest final relative residual.
 * Fails with a message when a component does not converge or a boundary is not periodic. */
int of_solve_predictor_cg(of_context* ctx, const of_grid* grid, const of_bc_set* bcs, double alpha,
                          const double* rhs_u, const double* rhs_v, const double* rhs_w, double* out_u,
                          double* out_v, double* out_w, double tolerance, int max_iterations, int* iterations,
                          double* relative_residual);

/* ---------------------------------------------------------------- projection */
/* solver/projection.hpp:54-57 build_pressure_rhs */
int of_build_pressure_rhs(of_context* ctx, const of_grid* grid, const of_bc_set* bcs,
                          const of_projection_options* opts, const double* u, const double* v, const double* w,
                          double* rhs);
/* solver/projection.hpp:59-63 correct_velocity */
int of_correct_velocity(of_context* ctx, const of_grid* grid, const of_bc_set* bcs, const of_projection_options* opts,
                        const double* u, const double* v, const double* w, const double* phi, double* out_u,
                        double* out_v, double* out_w);
/* solver/projection.hpp:65-70 project_velocity (pressure_rhs may be NULL) */
int of_project_velocity(of_context* ctx, const of_grid* grid, const of_bc_set* bcs, const of_projection_options* opts,
                        int precision, const double* u, const double* v, const double* w, double* phi,
                        double* out_u, double* out_v, double* out_w, double* pressure_rhs,
                        of_projection_diagnostics* diag, double* residual_history, int history_capacity);

/* ---------------------------------------------------------------- pressure Poisson (MGPCG) */
/* linsolve/poisson_solver.hpp:34-37 solve_pressure_poisson. p is the warm start and the
 * result. residual_history may be NULL. */
int of_solve_pressure_poisson(of_context* ctx, const of_grid* grid, const of_pressure_bc_set* bcs,
                              const of_projection_options* opts, int precision, const double* rhs, double* p,
                              of_poisson_diagnostics* diag, double* residual_history, int history_capacity);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP stream capture, tight serving loops, harnesses that forbid hipMalloc): the solve of
 * of_solve_pressure_poisson with every device array (multigrid hierarchy, CG vectors,
 * reduction partials, scalars, residual history, coarse factors) carved from the library's
 * fixed scratch block of OF_NOALLOC_SCRATCH_BYTES per device, so that the call allocates and
 * frees no device or pinned host memory. It runs on `stream` (NULL: the null stream) of HIP
 * device `device`, reads the stop flag back with a copy and a stream synchronisation after
 * each iteration, and returns when the solve is done. Grids of more than OF_NOALLOC_MAX_CELLS
 * cells, and solves whose arrays do not fit the block (e.g. a very large coarsest system),
 * return OF_INVALID_ARGUMENT. Every call on one device uses that device's one block, which
 * keeps the last call's contents: calls that overlap on one device race. */
#define OF_NOALLOC_MAX_CELLS 262144
#define OF_NOALLOC_SCRATCH_BYTES (32u << 20)
int of_solve_pressure_poisson_noalloc(int device, void* stream, const of_grid* grid, const of_pressure_bc_set* bcs,
                                      const of_projection_options* opts, int precision, const double* rhs, double* p,
                                      of_poisson_diagnostics* diag);
/* linsolve/poisson_solver.hpp:30-32 build_poisson_rhs_from_pressure */
int of_build_poisson_rhs_from_pressure(of_context* ctx, const of_grid* grid, const of_pressure_bc_set* bcs,
                                       const double* p, double* rhs);
/* linsolve/poisson_solver.cpp:244-269 apply_poisson_operator (internal in the reference;
 * exported for the operator tests) */
int of_apply_poisson_operator(of_context* ctx, const of_grid* grid, const of_pressure_bc_set* bcs, const double* in,
                              double* out);

/* ---------------------------------------------------------------- device-resident stepper */
/* solver/taylor_green.hpp:18-57 TaylorGreenConfig / TaylorGreenState,
 * solver/lid_driven_cavity.hpp:12-25,88-97 (case 1) and solver/channel_flow.hpp:10-28 (cases 2
 * and 3). The stepper owns the device state (u, v, w, p_total, A_prev x3,
 * has_previous_advection, metrics) and runs run_taylor_green_steps
 * (solver/taylor_green.cpp:540-634), run_lid_driven_cavity_steps
 * (solver/lid_driven_cavity.cpp:524-632) or the step loop of run_channel_flow
 * (solver/channel_flow.cpp:468-523) on the device. */
enum {
  OF_CASE_TAYLOR_GREEN = 0,
  OF_CASE_LID_DRIVEN_CAVITY = 1,
  OF_CASE_CHANNEL_COUETTE = 2,    /* periodic x, moving upper wall (ChannelFlowCase::couette) */
  OF_CASE_CHANNEL_POISEUILLE = 3  /* fixed-pressure x ends, walls in y (ChannelFlowCase::poiseuille) */
};

typedef struct of_stepper_config {
  int flow_case;
  int nx, ny, nz;
  double viscosity;           /* TG, channel: viscosity; cavity: lid_velocity / reynolds */
  double dt;                  /* from of_taylor_green_dt / of_lid_driven_cavity_dt / of_channel_flow_dt */
  int scheme;                 /* advection scheme, 0 = tvd */
  int poisson_max_iterations;
  double poisson_tolerance;
  int precision;              /* OF_PRECISION_* */
  double lid_velocity;        /* cavity: lid speed; Couette: top_velocity */
  double pressure_drop;       /* Poiseuille: pressure at x_min (0 at x_max) */
} of_stepper_config;

/* solver/taylor_green.hpp:32-42 TaylorGreenStepMetrics (+ L-infinity divergence) */
typedef struct of_step_metrics {
  int step;
  double time;
  double dt;
  double max_cfl;
  double max_velocity_change;
  double divergence_l2;
  double max_divergence_l2;
  int pressure_iterations;
  double pressure_relative_residual;
  double divergence_linf;
  double max_divergence_linf;
} of_step_metrics;

typedef struct of_stepper of_stepper;

/* solver/taylor_green.cpp:519-525 taylor_green_dt and :497-508 make_taylor_green_grid */
int of_taylor_green_dt(int nx, int ny, int nz, double cfl_limit, double final_time, double* dt, int* steps);
/* solver/lid_driven_cavity.cpp:95-98,386-388 lid_driven_cavity_dt */
int of_lid_driven_cavity_dt(int nx, int ny, double lid_velocity, double cfl_limit, double* dt);
/* solver/channel_flow.cpp:142-158 fixed_dt: cfl_limit / (max_u / dx) with max_u = top_velocity
 * (Couette) or pressure_drop / (8 viscosity) (Poiseuille); dx = 1 / nx. */
int of_channel_flow_dt(int flow_case, int nx, double viscosity, double top_velocity, double pressure_drop,
                       double cfl_limit, double* dt);

int of_stepper_create(of_context* ctx, const of_stepper_config* cfg, of_stepper** out);
int of_stepper_destroy(of_stepper* s);
/* solver/taylor_green.cpp:527-538 initialize_taylor_green_state on the device.
 * convention 0: the reference's fields (u=-cos x sin y [cos z]); 1: shifted by (pi/2, pi/2)
 * (u = sin x cos y [cos z], v = -cos x sin y [cos z], p = +1/4(..)).
 * For the cavity this is initialize_lid_driven_cavity_state (solver/lid_driven_cavity.cpp:390-409). */
int of_stepper_initialize(of_stepper* s, int convention);
/* Host <-> device state transfer (host pointers, reference storage layout). adv_* may be
 * NULL on upload (zeros). */
int of_stepper_upload(of_stepper* s, const double* u, const double* v, const double* w, const double* p_total,
                      const double* adv_u, const double* adv_v, const double* adv_w, int has_previous_advection,
                      const of_step_metrics* metrics);
int of_stepper_download(of_stepper* s, double* u, double* v, double* w, double* p_total, double* adv_u,
                        double* adv_v, double* adv_w, int* has_previous_advection, of_step_metrics* metrics);
/* Device pointers of the state fields (u, v, w, p_total, adv_u, adv_v, adv_w) for zero-copy
 * access by a caller sharing the stream. */
int of_stepper_device_fields(of_stepper* s, double** fields7);
/* Runs `steps` fractional steps. Throws the reference's non-convergence / non-finite errors
 * as OF_RUNTIME_ERROR with the reference's message. */
int of_stepper_run(of_stepper* s, int steps, of_step_metrics* metrics);
/* Same, without reading metrics back after every step (diagnostics are still accumulated on
 * the device and checked once at the end). Used for throughput runs. */
int of_stepper_run_batched(of_stepper* s, int steps, of_step_metrics* metrics);
/* Kinetic energy of the current state (solver/taylor_green.cpp:269-288). */
int of_stepper_kinetic_energy(of_stepper* s, double* out);
/* Body force f (device face arrays, velocity layout; copied) added to the Taylor-Green
 * predictor right-hand side as rhs += dt f after the reference's terms
 * (solver/taylor_green.cpp:227-256 has no such term). NULL arrays clear it. */
int of_stepper_set_forcing(of_stepper* s, const double* fu, const double* fv, const double* fw);

/* Helmholtz predictor of the stepper's steps: OF_PREDICTOR_ADI (default: the reference's exact
 * line solves) or OF_PREDICTOR_CG (of_solve_predictor_cg with the given tolerance and iteration
 * limit; periodic Taylor-Green only). No counterpart in the reference's CPU path; the CG form is
 * its Metal backend's. */
#define OF_PREDICTOR_ADI 0
#define OF_PREDICTOR_CG 1
int of_stepper_set_predictor(of_stepper* s, int kind, double tolerance, int max_iterations);

/* ---------------------------------------------------------------- discrete adjoint */
/* The reference has no derivative code. These entry points add the discrete adjoint and the
 * tangent-linear model of the device step above for the three-dimensional periodic
 * Taylor-Green case. All field pointers are device pointers in the reference storage layout.
 *
 * Per-operator transposes (periodic boundaries only; `bcs` must be all periodic). Outputs
 * are overwritten over the whole ghosted storage.
 *   of_apply_velocity_boundary_conditions  (solver/projection.cpp:233-300)  -> in place
 *   of_apply_pressure_boundary_conditions  (solver/projection.cpp:207-231)  -> in place
 *   of_compute_gradient    (operators/discrete_operators.cpp:64-96)   gbar -> pbar
 *   of_compute_divergence  (operators/discrete_operators.cpp:98-113)  dbar -> ubar
 *   of_solve_predictor_adi (solver/projection.hpp:49-52)              ybar -> rbar
 * and the linearised advection (solver/momentum_terms.cpp:88-312, branches frozen at u):
 *   tangent: out = dA(u)[du];  adjoint: bar = dA(u)^T[abar]. */
int of_apply_velocity_boundary_conditions_adjoint(of_context* ctx, const of_grid* grid, const of_bc_set* bcs,
                                                  double* u, double* v, double* w);
int of_apply_pressure_boundary_conditions_adjoint(of_context* ctx, const of_grid* grid,
                                                  const of_pressure_bc_set* bcs, double* p);
int of_compute_gradient_adjoint(of_context* ctx, const of_grid* grid, const double* gx, const double* gy,
                                const double* gz, double* p);
int of_compute_divergence_adjoint(of_context* ctx, const of_grid* grid, const double* d, double* u, double* v,
                                  double* w);
int of_solve_predictor_adi_adjoint(of_context* ctx, const of_grid* grid, const of_bc_set* bcs, double alpha,
                                   const double* yu, const double* yv, const double* yw, double* ru, double* rv,
                                   double* rw);
int of_compute_advection_tangent(of_context* ctx, const of_grid* grid, int scheme, const double* u, const double* v,
                                 const double* w, const double* du, const double* dv, const double* dw, double* au,
                                 double* av, double* aw);
int of_compute_advection_adjoint(of_context* ctx, const of_grid* grid, int scheme, const double* u, const double* v,
                                 const double* w, const double* abar_u, const double* abar_v, const double* abar_w,
                                 double* ubar, double* vbar, double* wbar);

/* Checkpointed adjoint run: the forward pass stores the state every `checkpoint_every` steps;
 * the reverse sweep recomputes one window at a time. */
typedef struct of_adjoint of_adjoint;
int of_adjoint_create(of_context* ctx, const of_stepper_config* cfg, int checkpoint_every, of_adjoint** out);
int of_adjoint_destroy(of_adjoint* a);
/* forcing as for of_stepper_set_forcing; NULL clears it */
int of_adjoint_set_forcing(of_adjoint* a, const double* fu, const double* fv, const double* fw);
/* Runs `steps` steps from (u0, v0, w0, p0); p0 may be NULL (zero). Only active entries are
 * read. The final kinetic energy and metrics are returned. */
int of_adjoint_forward(of_adjoint* a, const double* u0, const double* v0, const double* w0, const double* p0,
                       int steps, double* kinetic_energy, of_step_metrics* metrics);
/* Copies the final velocity (ghosted storage). */
int of_adjoint_final_velocity(of_adjoint* a, double* u, double* v, double* w);
/* Reverse sweep. seed_*: adjoint of the final velocity storage; all three NULL selects the
 * gradient of the final kinetic energy. Outputs (NULL to skip): gradients w.r.t. the active
 * entries of u0, v0, w0, p0 and the forcing field (inactive entries zero). */
int of_adjoint_backward(of_adjoint* a, const double* seed_u, const double* seed_v, const double* seed_w,
                        double* gu, double* gv, double* gw, double* gp, double* gfu, double* gfv, double* gfw);
/* Tangent-linear run along (du0, dv0, dw0, dp0, dfu, dfv, dfw) (NULL: zero) from the initial
 * state of the last forward pass. Outputs: the final velocity perturbation (NULL to skip) and
 * the directional derivative of the final kinetic energy (NULL to skip). */
int of_adjoint_tangent(of_adjoint* a, const double* du0, const double* dv0, const double* dw0, const double* dp0,
                       const double* dfu, const double* dfv, const double* dfw, double* du, double* dv, double* dw,
                       double* dkinetic_energy);

#ifdef __cplusplus
}
#endif

#endif /* OF_HIP_H */
