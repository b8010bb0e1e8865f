// Host-side declarations shared by the translation units of the gfx950 library.
#pragma once

#include "common.hpp"

#include <functional>
#include <string>
#include <vector>

namespace ofh {

// Runs the body of an extern "C" entry point: exceptions become status codes plus the message
// of_last_error returns, and a release failure recorded meanwhile is reported (capi.hip).
int guarded_call(const std::function<void()>& body);

// ------------------------------------------------------------------ device context
// Device memory handed out from one block the caller provides (nothing is allocated): 256-byte
// aligned pieces, never freed one by one. take() throws InvalidArgument when the block is used up.
struct DeviceArena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~(size_t)255;
    if(at + bytes > cap)
      throw InvalidArgument("the solve needs more than the " + std::to_string(cap >> 20) +
                            " MiB of no-allocation scratch");
    used = at + bytes;
    return base + at;
  }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool borrowed = false;  // partials and scal belong to someone else: the destructor frees nothing
  double* partials = nullptr;   // kMaxRed x kMaxPartials reduction partials
  double* scal = nullptr;       // device scalars
  double* host_scal = nullptr;  // pinned mirror for read-back
  // grow-only scratch pool for the per-operator entry points
  std::vector<std::pair<void*, size_t>> scratch;
  void* scratch_get(int slot, size_t bytes);
  // live timing of one kernel: launches of the kernel named `prof_name` are bracketed with
  // events on the context stream (finest-level launches only for the multigrid kernels)
  std::string prof_name;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
  size_t prof_used = 0;
  int prof_stride = 1;    // bracket every prof_stride-th launch of the kernel
  size_t prof_seen = 0;   // launches of the kernel since timing was switched on
  // true for the launches to time: an event record leaves the device idle for a few
  // microseconds, so a timed run samples every prof_stride-th launch
  bool prof_on(const char* name) { return !prof_name.empty() && prof_name == name && prof_seen++ % prof_stride == 0; }
  void prof_begin(const char* name, hipEvent_t* end);
  ~Context();
};

// Sets up a context on `device`: its stream (`stream`, or a new non-blocking one of default
// priority that the context owns), reduction partials and scalars. Throws HipError on failure;
// the context's destructor frees whatever was set up.
void context_init(Context& x, int device, hipStream_t stream);

// Brackets one launch with timing events when `name` is the profiled kernel.
struct ProfScope {
  Context& c;
  hipEvent_t end = nullptr;
  ProfScope(Context& ctx, const char* name, bool eligible = true) : c(ctx) {
    if(eligible && c.prof_on(name)) c.prof_begin(name, &end);
  }
  ~ProfScope() {
    if(end) (void)hipEventRecord(end, c.stream);
  }
};

// Fixed slots in Context::scal
enum ScalarSlot : int {
  kScalRed0 = 0,      // 8 values: generic reduction results
  kScalCfl = 8,       // 5 values: cfl dt, max cfl, max u, max v, max w
  kScalStep = 16,     // step diagnostics (32 values)
  kScalPoisson = 64,  // MGPCG scalars (see poisson.hip)
  kScalCount = 256
};

// Launch shape for a "k-column" kernel over an active box ex*ey*ez: blocks of 64x4 threads,
// each thread walking up to kz consecutive k so that the grid stays within kMaxPartials.
struct ColumnLaunch {
  dim3 grid, block;
  int kz;
};
ColumnLaunch column_launch(int ex, int ey, int ez, int min_kz = 1);

// ------------------------------------------------------------------ fields.hip
void launch_velocity_bc(Context& c, const of_grid& g, const of_bc_set& b, double* u, double* v, double* w);
void launch_component_bc(Context& c, const of_grid& g, const of_bc_set& b, int comp, double* f);
// Ghost fill of the components selected by bit c of `mask` (the others may be null).
void launch_component_bc_mask(Context& c, const of_grid& g, const of_bc_set& b, double* u, double* v, double* w,
                              int mask);
void launch_pressure_bc(Context& c, const of_grid& g, const of_pressure_bc_set& b, double* p);
void launch_total_pressure_bc(Context& c, const of_grid& g, const of_bc_set& b, const double* su, const double* sv,
                              const double* sw, double* p);
void launch_gradient(Context& c, const of_grid& g, const double* p, double* gx, double* gy, double* gz);
void launch_divergence(Context& c, const of_grid& g, const double* u, const double* v, const double* w, double* d);
void launch_laplacian(Context& c, const of_grid& g, int loc, double scale, const double* in, double* out);
void launch_advection(Context& c, const of_grid& g, int scheme, const double* u, const double* v, const double* w,
                      double* au, double* av, double* aw);
void launch_cfl(Context& c, const of_grid& g, const double* u, const double* v, const double* w, double dt,
                double* out5_device);
void launch_factorized_correction(Context& c, const of_grid& g, int variant, double alpha, const double* u,
                                  const double* v, const double* w, double* fu, double* fv, double* fw);
void launch_predictor_rhs(Context& c, const of_grid& g, int variant, int scheme, double nu, double dt,
                          const double* u, const double* v, const double* w, const double* p, const double* pu,
                          const double* pv, const double* pw, double* au, double* av, double* aw, double* ru,
                          double* rv, double* rw);
// sum(x^2) and max|x| over an active box, into scal[slot], scal[slot+1]
void launch_norms(Context& c, const Lay& L, const double* x, int slot);
// scaled copy of the active range: out = scale * in
void launch_scale_active(Context& c, const Lay& L, double scale, const double* in, double* out);
// y = y + a * x on the active range
void launch_axpy_active(Context& c, const Lay& L, double a, const double* x, double* y);

// ------------------------------------------------------------------ helmholtz.hip
// scratch: three work fields of the velocity layout the solve may overwrite (null: context
// scratch); they may be the rhs arrays themselves when the caller no longer needs them.
// keep_storage: the output starts as a copy of the whole rhs storage (edge ghosts included), as
// the reference's does; a caller that never reads the output's edge ghosts passes false.
int predictor_adi(Context& c, const of_grid& g, const of_bc_set& b, double alpha, const double* ru,
                  const double* rv, const double* rw, double* ou, double* ov, double* ow,
                  double* const* scratch = nullptr, bool keep_storage = true);
// Conjugate-gradient alternative to predictor_adi for grids periodic on every axis (work arrays
// of one grid in a HelmholtzCG). Returns the largest iteration count of the three components;
// throws when one did not reach tol within max_iterations. worst_rel (may be null) receives the
// largest final relative residual.
struct HelmholtzCG;
HelmholtzCG* helmholtz_cg_create(Context& c, const of_grid& g);
void helmholtz_cg_destroy(HelmholtzCG* h);
int predictor_cg(HelmholtzCG& S, const of_grid& g, const of_bc_set& b, double alpha, const double* ru,
                 const double* rv, const double* rw, double* ou, double* ov, double* ow, double tol,
                 int max_iterations, double* worst_rel);
// Transpose of predictor_adi (rhs -> out) for periodic boundaries; rbar may alias ybar.
void predictor_adi_adjoint(Context& c, const of_grid& g, const of_bc_set& b, double alpha, const double* const* ybar,
                           double* const* rbar);

// ------------------------------------------------------------------ poisson.hip
struct PoissonResult {
  of_poisson_diagnostics diag;
  std::vector<double> history;
};
class PoissonSolver;  // MGPCG on the device, reused across solves of one grid/BC set
PoissonSolver* poisson_create(Context& c, const of_grid& g, const of_pressure_bc_set& b, int precision);
// The same, with every device array carved from `arena` and no allocation anywhere (the solve
// reads its stop flag back by a copy and a stream synchronisation per iteration).
PoissonSolver* poisson_create_in(Context& c, const of_grid& g, const of_pressure_bc_set& b, int precision,
                                 DeviceArena& arena);
void poisson_destroy(PoissonSolver* s);
// Solves -lap(p) = rhs (reference sign conventions); rhs and p in ghosted cell storage.
// p holds the warm start on entry (zero_guess: the warm start is zero and p is not read). Only
// the active range of p is written. With deferred_mean, a problem with a null space leaves the
// final mean of p in it instead of subtracting it: *deferred_mean points at that mean in device
// memory (null when there is none), valid until the solver's next solve, and the caller
// subtracts it on read (p - mean is then bitwise the value the subtraction would have stored).
PoissonResult poisson_solve(PoissonSolver* s, const double* rhs, double* p, int max_iterations, double tol,
                            bool want_history, bool zero_guess = false, const double** deferred_mean = nullptr);
void launch_poisson_apply(Context& c, const of_grid& g, const of_pressure_bc_set& b, bool homogeneous,
                          const double* in, double* out);

// ------------------------------------------------------------------ projection (fields.hip)
// div(u) * scale into a ghosted cell field (active range)
void launch_divergence_scaled(Context& c, const of_grid& g, const double* u, const double* v, const double* w,
                              double scale, double* out);
// out = pred - scale * grad(phi_with_ghosts) on every active face
void launch_correct(Context& c, const of_grid& g, const double* u, const double* v, const double* w,
                    const double* phi, double scale, double* ou, double* ov, double* ow);

}  // namespace ofh
