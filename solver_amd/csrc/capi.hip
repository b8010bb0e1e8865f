// extern "C" entry points (include/of_hip.h). Each wraps the device launches of one
// reference interface, translating C++ exceptions into status codes plus a thread-local
// message that carries the reference's exception text.

#include "adjoint.hpp"
#include "stepper.hpp"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace ofh {

void* Context::scratch_get(int slot, size_t bytes) {
  if((int)scratch.size() <= slot) scratch.resize(slot + 1, {nullptr, 0});
  auto& s = scratch[slot];
  if(s.second < bytes) {
    if(s.first) OFH_CHECK(hipStreamSynchronize(stream));
    if(s.first) OFH_CHECK(hipFree(s.first));
    OFH_CHECK(hipMalloc(&s.first, bytes));
    s.second = bytes;
  }
  return s.first;
}

void Context::prof_begin(const char*, hipEvent_t* end) {
  if(prof_used == prof_events.size()) {
    hipEvent_t a, b;
    OFH_CHECK(hipEventCreate(&a));
    OFH_CHECK(hipEventCreate(&b));
    prof_events.emplace_back(a, b);
  }
  auto& e = prof_events[prof_used++];
  OFH_CHECK(hipEventRecord(e.first, stream));
  *end = e.second;
}

Context::~Context() {
  if(borrowed) return;
  for(auto& e : prof_events) {
    OFH_RELEASE(hipEventDestroy(e.first));
    OFH_RELEASE(hipEventDestroy(e.second));
  }
  for(auto& s : scratch)
    if(s.first) OFH_RELEASE(hipFree(s.first));
  if(partials) OFH_RELEASE(hipFree(partials));
  if(scal) OFH_RELEASE(hipFree(scal));
  if(host_scal) OFH_RELEASE(hipHostFree(host_scal));
  if(own_stream && stream) OFH_RELEASE(hipStreamDestroy(stream));
}

namespace {
thread_local std::string g_release_error;
}

void release_note(hipError_t e, const char* call) {
  if(e != hipSuccess && g_release_error.empty())
    g_release_error = std::string(call) + " failed: " + hipGetErrorString(e);
}

std::string release_take() {
  (void)hipGetLastError();  // a failed release call is reported through the record, not later
  std::string r;
  r.swap(g_release_error);
  return r;
}

int launch_check_level() {
  static const int level = [] {
    const char* v = std::getenv("OFH_CHECK_LAUNCHES");
    return v ? std::atoi(v) : 0;
  }();
  return level;
}

void launch_check(const char* kernel) {
  hipError_t e = hipGetLastError();
  if(e == hipSuccess && launch_check_level() > 1) e = hipDeviceSynchronize();
  if(e != hipSuccess) throw HipError(std::string("launch of ") + kernel + " failed: " + hipGetErrorString(e));
}

bool axis_periodic(const of_bc_set& b, int axis);
bool pressure_axis_periodic(const of_pressure_bc_set& b, int axis);

}  // namespace ofh

using namespace ofh;

// A context is shared by the caller's handle and every stepper or adjoint run created on it:
// of_context_destroy drops the caller's reference and the native state goes with the last one,
// so the order in which a caller (or a garbage collector) releases handles does not matter.
struct of_context {
  Context ctx;
  std::atomic<int> refs{1};
};

namespace {

void context_retain(of_context* c) { c->refs.fetch_add(1, std::memory_order_relaxed); }

// Drops one reference; the last one waits for the context stream and frees the context.
void context_release(of_context* c) {
  if(c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  OFH_RELEASE(hipSetDevice(c->ctx.device));
  OFH_RELEASE(hipStreamSynchronize(c->ctx.stream));
  delete c;
}

}  // namespace

struct of_adjoint {
  AdjointRun* a;
  of_context* owner;
  int device;
};
struct of_stepper {
  Stepper* s;
  of_context* owner;
  int device;
};

namespace {

thread_local std::string g_error;

// The body's own error is the one reported; a release failure during its unwinding is dropped.
int failed(int code, const char* what) {
  g_error = what;
  (void)release_take();
  return code;
}

// Runs an entry point's body: C++ exceptions become status codes plus the message, and a
// release failure recorded while the body ran is reported as the call's HIP error.
template <typename F>
int guarded(F&& f) {
  try {
    f();
    std::string rel = release_take();
    if(!rel.empty()) {
      g_error = rel;
      return OF_HIP_ERROR;
    }
    return OF_OK;
  } catch(const InvalidArgument& e) {
    return failed(OF_INVALID_ARGUMENT, e.what());
  } catch(const std::invalid_argument& e) {
    return failed(OF_INVALID_ARGUMENT, e.what());
  } catch(const HipError& e) {
    return failed(OF_HIP_ERROR, e.what());
  } catch(const std::exception& e) {
    return failed(OF_RUNTIME_ERROR, e.what());
  } catch(...) {
    return failed(OF_RUNTIME_ERROR, "unknown error");
  }
}

Context& C(of_context* c) {
  if(c == nullptr) throw InvalidArgument("context must not be null");
  OFH_CHECK(hipSetDevice(c->ctx.device));
  return c->ctx;
}

void need(const void* p, const char* what) {
  if(p == nullptr) throw InvalidArgument(std::string(what) + " must not be null");
}

of_poisson_diagnostics to_diag(const PoissonResult& r, double* hist, int cap) {
  of_poisson_diagnostics d = r.diag;
  if(hist != nullptr)
    for(int i = 0; i < std::min(cap, (int)r.history.size()); ++i) hist[i] = r.history[i];
  return d;
}

void check_projection_options(const of_projection_options* o, const char* fn) {
  need(o, "options");
  if(!(o->dt > 0.0)) throw InvalidArgument(std::string(fn) + " expects dt > 0");
  if(!(o->density > 0.0)) throw InvalidArgument(std::string(fn) + " expects density > 0");
}

size_t fsize(const of_grid& g, int loc) { return make_lay(g, loc).size() * sizeof(double); }

void need_periodic(const of_bc_set* b) {
  need(b, "bcs");
  for(int a = 0; a < 3; ++a)
    if(!axis_periodic(*b, a)) throw InvalidArgument("adjoint boundary conditions support periodic faces only");
}

AdjointRun& A(of_adjoint* a) {
  need(a, "adjoint");
  OFH_CHECK(hipSetDevice(a->device));
  return *a->a;
}

}  // namespace

void ofh::context_init(Context& x, int device, hipStream_t stream) {
  x.device = device;
  OFH_CHECK(hipSetDevice(device));
  if(stream) {
    x.stream = stream;
  } else {
    OFH_CHECK(hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking));
    x.own_stream = true;
  }
  OFH_CHECK(hipMalloc(&x.partials, sizeof(double) * kMaxRed * kMaxPartials));
  OFH_CHECK(hipMalloc(&x.scal, sizeof(double) * kScalCount));
  OFH_CHECK(hipMemsetAsync(x.scal, 0, sizeof(double) * kScalCount, x.stream));
  OFH_CHECK(hipHostMalloc(&x.host_scal, sizeof(double) * kScalCount, hipHostMallocDefault));
}

int ofh::guarded_call(const std::function<void()>& body) { return guarded(body); }

extern "C" {

const char* of_last_error(void) { return g_error.c_str(); }

const char* of_build_info(void) {
  return "solver_amd gfx950 projection step (HIP " __DATE__ ")";
}

int of_context_create(int device, void* stream, of_context** out) {
  return guarded([&] {
    need(out, "out");
    auto* c = new of_context();
    try {
      context_init(c->ctx, device, static_cast<hipStream_t>(stream));
    } catch(...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int of_context_destroy(of_context* c) {
  return guarded([&] {
    if(c == nullptr) return;
    context_release(c);
  });
}

int of_profile_kernel(of_context* c, const char* name) {
  return guarded([&] {
    Context& x = C(c);
    OFH_CHECK(hipStreamSynchronize(x.stream));
    x.prof_name = name ? name : "";
    x.prof_used = 0;
    x.prof_seen = 0;
  });
}

int of_profile_stride(of_context* c, int stride) {
  return guarded([&] {
    if(stride < 1) throw InvalidArgument("the profiling stride must be at least 1");
    C(c).prof_stride = stride;
  });
}

int of_profile_read(of_context* c, int* launches, double* total_ms, double* min_ms, double* max_ms) {
  return guarded([&] {
    Context& x = C(c);
    OFH_CHECK(hipStreamSynchronize(x.stream));
    double tot = 0.0, mn = 1e300, mx = 0.0;
    for(size_t i = 0; i < x.prof_used; ++i) {
      float ms = 0.0f;
      OFH_CHECK(hipEventElapsedTime(&ms, x.prof_events[i].first, x.prof_events[i].second));
      tot += ms;
      mn = std::min(mn, (double)ms);
      mx = std::max(mx, (double)ms);
    }
    if(launches) *launches = (int)x.prof_used;
    if(total_ms) *total_ms = tot;
    if(min_ms) *min_ms = x.prof_used ? mn : 0.0;
    if(max_ms) *max_ms = mx;
    x.prof_used = 0;
  });
}

int of_profile_durations(of_context* c, double* out_ms, int capacity, int* launches) {
  return guarded([&] {
    Context& x = C(c);
    OFH_CHECK(hipStreamSynchronize(x.stream));
    for(size_t i = 0; i < x.prof_used; ++i) {
      float ms = 0.0f;
      OFH_CHECK(hipEventElapsedTime(&ms, x.prof_events[i].first, x.prof_events[i].second));
      if(out_ms && (int)i < capacity) out_ms[i] = ms;
    }
    if(launches) *launches = (int)x.prof_used;
    x.prof_used = 0;
  });
}

int of_active_rms_difference(of_context* c, const of_grid* g, int location, const double* a, const double* b,
                             double* out) {
  return guarded([&] {
    validate_grid(g);
    if(location < OF_CELL || location > OF_FACE_Z) throw InvalidArgument("unknown field location");
    need(a, "a"), need(b, "b"), need(out, "out");
    Context& x = C(c);
    const Lay L = make_lay(*g, location);
    double* diff = static_cast<double*>(x.scratch_get(15, L.size() * sizeof(double)));
    launch_scale_active(x, L, 1.0, a, diff);
    launch_axpy_active(x, L, -1.0, b, diff);
    launch_norms(x, L, diff, kScalRed0);
    OFH_CHECK(hipMemcpyAsync(x.host_scal, x.scal + kScalRed0, sizeof(double), hipMemcpyDeviceToHost, x.stream));
    OFH_CHECK(hipStreamSynchronize(x.stream));
    *out = std::sqrt(x.host_scal[0] / ((double)L.ax * (double)L.ay * (double)L.az));
  });
}

int of_context_synchronize(of_context* c) {
  return guarded([&] { OFH_CHECK(hipStreamSynchronize(C(c).stream)); });
}

int of_apply_velocity_boundary_conditions(of_context* c, const of_grid* g, const of_bc_set* b, double* u, double* v,
                                          double* w) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    need(u, "u"), need(v, "v"), need(w, "w");
    launch_velocity_bc(C(c), *g, *b, u, v, w);
  });
}

int of_apply_pressure_boundary_conditions(of_context* c, const of_grid* g, const of_pressure_bc_set* b, double* p) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    need(p, "p");
    launch_pressure_bc(C(c), *g, *b, p);
  });
}

int of_apply_total_pressure_boundary_conditions(of_context* c, const of_grid* g, const of_bc_set* b, const double* su,
                                                const double* sv, const double* sw, double* p) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    need(p, "p_total");
    launch_total_pressure_bc(C(c), *g, *b, su, sv, sw, p);
  });
}

int of_derive_pressure_correction_boundary_conditions(const of_bc_set* b, of_pressure_bc_set* out) {
  return guarded([&] {
    need(b, "bcs");
    need(out, "out");
    for(int a = 0; a < 3; ++a) (void)axis_periodic(*b, a);
    *out = derive_pressure_bcs(*b);
  });
}

int of_compute_gradient(of_context* c, const of_grid* g, const double* p, double* gx, double* gy, double* gz) {
  return guarded([&] {
    validate_grid(g);
    launch_gradient(C(c), *g, p, gx, gy, gz);
  });
}

int of_compute_divergence(of_context* c, const of_grid* g, const double* u, const double* v, const double* w,
                          double* d) {
  return guarded([&] {
    validate_grid(g);
    launch_divergence(C(c), *g, u, v, w, d);
  });
}

int of_compute_laplacian(of_context* c, const of_grid* g, int loc, const double* in, double* out) {
  return guarded([&] {
    validate_grid(g);
    if(loc < OF_CELL || loc > OF_FACE_Z) throw InvalidArgument("compute_laplacian: unknown field location");
    launch_laplacian(C(c), *g, loc, 1.0, in, out);
  });
}

int of_compute_advection_term(of_context* c, const of_grid* g, int scheme, const double* u, const double* v,
                              const double* w, double* au, double* av, double* aw) {
  return guarded([&] {
    validate_grid(g);
    if(scheme < 0 || scheme > 2) throw InvalidArgument("compute_advection_term: unknown advection scheme");
    launch_advection(C(c), *g, scheme, u, v, w, au, av, aw);
  });
}

int of_compute_diffusion_term(of_context* c, const of_grid* g, double nu, const double* u, const double* v,
                              const double* w, double* du, double* dv, double* dw) {
  return guarded([&] {
    validate_grid(g);
    Context& x = C(c);
    const double* in[3] = {u, v, w};
    double* out[3] = {du, dv, dw};
    for(int a = 0; a < 3; ++a) {
      // compute_laplacian then scale_active(viscosity): lap * nu
      launch_laplacian(x, *g, OF_FACE_X + a, 1.0, in[a], out[a]);
      launch_scale_active(x, make_lay(*g, OF_FACE_X + a), nu, out[a], out[a]);
    }
  });
}

int of_compute_advective_cfl(of_context* c, const of_grid* g, const double* u, const double* v, const double* w,
                             double dt, of_cfl_diagnostics* out) {
  return guarded([&] {
    validate_grid(g);
    need(out, "out");
    if(dt < 0.0) throw InvalidArgument("compute_advective_cfl expects a non-negative dt");
    Context& x = C(c);
    launch_cfl(x, *g, u, v, w, dt, x.scal + kScalCfl);
    OFH_CHECK(hipMemcpyAsync(x.host_scal, x.scal + kScalCfl, 5 * sizeof(double), hipMemcpyDeviceToHost, x.stream));
    OFH_CHECK(hipStreamSynchronize(x.stream));
    out->dt = x.host_scal[0];
    out->max_cfl = x.host_scal[1];
    out->max_u = x.host_scal[2];
    out->max_v = x.host_scal[3];
    out->max_w = x.host_scal[4];
  });
}

int of_compute_factorized_correction(of_context* c, const of_grid* g, int variant, double alpha, const double* u,
                                     const double* v, const double* w, double* fu, double* fv, double* fw) {
  return guarded([&] {
    validate_grid(g);
    launch_factorized_correction(C(c), *g, variant, alpha, u, v, w, fu, fv, fw);
  });
}

int of_assemble_predictor_rhs(of_context* c, const of_grid* g, int variant, int scheme, double nu, double dt,
                              const double* u, const double* v, const double* w, const double* p, const double* pu,
                              const double* pv, const double* pw, double* au, double* av, double* aw, double* ru,
                              double* rv, double* rw) {
  return guarded([&] {
    validate_grid(g);
    Context& x = C(c);
    const bool prev = pu != nullptr;
    if(prev && (pv == nullptr || pw == nullptr)) throw InvalidArgument("previous advection must be all or none");
    // predictor_rhs = current_velocity (whole storage), then active-range updates
    const double* in[3] = {u, v, w};
    double* out[3] = {ru, rv, rw};
    for(int a = 0; a < 3; ++a)
      OFH_CHECK(hipMemcpyAsync(out[a], in[a], fsize(*g, OF_FACE_X + a), hipMemcpyDeviceToDevice, x.stream));
    launch_predictor_rhs(x, *g, variant, scheme, nu, dt, u, v, w, p, pu, pv, pw, au, av, aw, ru, rv, rw);
  });
}

int of_solve_predictor_adi(of_context* c, const of_grid* g, const of_bc_set* b, double alpha, const double* ru,
                           const double* rv, const double* rw, double* ou, double* ov, double* ow, int* line_solves) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    const int n = predictor_adi(C(c), *g, *b, alpha, ru, rv, rw, ou, ov, ow);
    if(line_solves) *line_solves = n;
  });
}

int of_solve_predictor_cg(of_context* c, const of_grid* g, const of_bc_set* b, double alpha, const double* ru,
                          const double* rv, const double* rw, double* ou, double* ov, double* ow, double tol, int maxit,
                          int* iterations, double* relative_residual) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    if(!(tol > 0.0)) throw InvalidArgument("solve_predictor_cg expects tolerance > 0");
    Context& x = C(c);
    HelmholtzCG* h = helmholtz_cg_create(x, *g);
    try {
      double rel = 0.0;
      const int n = predictor_cg(*h, *g, *b, alpha, ru, rv, rw, ou, ov, ow, tol, maxit, &rel);
      if(iterations) *iterations = n;
      if(relative_residual) *relative_residual = rel;
    } catch(...) {
      helmholtz_cg_destroy(h);
      throw;
    }
    helmholtz_cg_destroy(h);
  });
}

int of_build_pressure_rhs(of_context* c, const of_grid* g, const of_bc_set* b, const of_projection_options* o,
                          const double* u, const double* v, const double* w, double* rhs) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    check_projection_options(o, "build_pressure_rhs");
    Context& x = C(c);
    // working = predicted + BC ghosts (solver/projection.cpp:1060-1063)
    double* wk[3];
    const double* in[3] = {u, v, w};
    for(int a = 0; a < 3; ++a) {
      wk[a] = static_cast<double*>(x.scratch_get(a, fsize(*g, OF_FACE_X + a)));
      OFH_CHECK(hipMemcpyAsync(wk[a], in[a], fsize(*g, OF_FACE_X + a), hipMemcpyDeviceToDevice, x.stream));
    }
    launch_velocity_bc(x, *g, *b, wk[0], wk[1], wk[2]);
    launch_divergence_scaled(x, *g, wk[0], wk[1], wk[2], o->density / o->dt, rhs);
  });
}

int of_correct_velocity(of_context* c, const of_grid* g, const of_bc_set* b, const of_projection_options* o,
                        const double* u, const double* v, const double* w, const double* phi, double* ou, double* ov,
                        double* ow) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    check_projection_options(o, "correct_velocity");
    Context& x = C(c);
    const of_pressure_bc_set pb = derive_pressure_bcs(*b);
    double* pg = static_cast<double*>(x.scratch_get(3, fsize(*g, OF_CELL)));
    OFH_CHECK(hipMemcpyAsync(pg, phi, fsize(*g, OF_CELL), hipMemcpyDeviceToDevice, x.stream));
    launch_pressure_bc(x, *g, pb, pg);
    const double* in[3] = {u, v, w};
    double* out[3] = {ou, ov, ow};
    for(int a = 0; a < 3; ++a)
      OFH_CHECK(hipMemcpyAsync(out[a], in[a], fsize(*g, OF_FACE_X + a), hipMemcpyDeviceToDevice, x.stream));
    launch_correct(x, *g, u, v, w, pg, -o->dt / o->density, out[0], out[1], out[2]);
    launch_velocity_bc(x, *g, *b, ou, ov, ow);
  });
}

int of_solve_pressure_poisson(of_context* c, const of_grid* g, const of_pressure_bc_set* b,
                              const of_projection_options* o, int precision, const double* rhs, double* p,
                              of_poisson_diagnostics* diag, double* hist, int cap) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    need(o, "options");
    if(o->poisson_max_iterations <= 0) throw InvalidArgument("solve_pressure_poisson expects iterations > 0");
    if(!(o->poisson_tolerance > 0.0)) throw InvalidArgument("solve_pressure_poisson expects tolerance > 0");
    Context& x = C(c);
    PoissonSolver* s = poisson_create(x, *g, *b, precision);
    PoissonResult r;
    try {
      r = poisson_solve(s, rhs, p, o->poisson_max_iterations, o->poisson_tolerance, hist != nullptr);
    } catch(...) {
      poisson_destroy(s);
      throw;
    }
    poisson_destroy(s);
    const of_poisson_diagnostics d = to_diag(r, hist, cap);
    if(diag) *diag = d;
  });
}

int of_build_poisson_rhs_from_pressure(of_context* c, const of_grid* g, const of_pressure_bc_set* b, const double* p,
                                       double* rhs) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    Context& x = C(c);
    double* tmp = static_cast<double*>(x.scratch_get(4, fsize(*g, OF_CELL)));
    launch_poisson_apply(x, *g, *b, false, p, tmp);
    launch_scale_active(x, make_lay(*g, OF_CELL), -1.0, tmp, rhs);
  });
}

int of_apply_poisson_operator(of_context* c, const of_grid* g, const of_pressure_bc_set* b, const double* in,
                              double* out) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    launch_poisson_apply(C(c), *g, *b, false, in, out);
  });
}

int of_project_velocity(of_context* c, const of_grid* g, const of_bc_set* b, const of_projection_options* o,
                        int precision, const double* u, const double* v, const double* w, double* phi, double* ou,
                        double* ov, double* ow, double* prhs_out, of_projection_diagnostics* diag, double* hist,
                        int cap) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    check_projection_options(o, "project_velocity");
    Context& x = C(c);
    const of_pressure_bc_set pb = derive_pressure_bcs(*b);
    const Lay Lc = make_lay(*g, OF_CELL);
    double* wk[3];
    const double* in[3] = {u, v, w};
    for(int a = 0; a < 3; ++a) {
      wk[a] = static_cast<double*>(x.scratch_get(10 + a, fsize(*g, OF_FACE_X + a)));
      OFH_CHECK(hipMemcpyAsync(wk[a], in[a], fsize(*g, OF_FACE_X + a), hipMemcpyDeviceToDevice, x.stream));
    }
    launch_velocity_bc(x, *g, *b, wk[0], wk[1], wk[2]);
    double* div = static_cast<double*>(x.scratch_get(13, fsize(*g, OF_CELL)));
    double* rhs = static_cast<double*>(x.scratch_get(14, fsize(*g, OF_CELL)));
    OFH_CHECK(hipMemsetAsync(div, 0, fsize(*g, OF_CELL), x.stream));
    OFH_CHECK(hipMemsetAsync(rhs, 0, fsize(*g, OF_CELL), x.stream));
    of_projection_diagnostics d{};
    launch_divergence(x, *g, wk[0], wk[1], wk[2], div);
    launch_norms(x, Lc, div, kScalRed0);
    launch_divergence_scaled(x, *g, wk[0], wk[1], wk[2], o->density / o->dt, rhs);
    launch_norms(x, Lc, rhs, kScalRed0 + 2);
    OFH_CHECK(hipMemcpyAsync(x.host_scal, x.scal + kScalRed0, 4 * sizeof(double), hipMemcpyDeviceToHost, x.stream));
    OFH_CHECK(hipStreamSynchronize(x.stream));
    const double ncell = (double)g->nx * g->ny * g->nz;
    d.divergence_l2_before = std::sqrt(x.host_scal[0] / ncell);
    d.rhs_l2 = std::sqrt(x.host_scal[2] / ncell);
    PoissonSolver* s = poisson_create(x, *g, pb, precision);
    PoissonResult r;
    try {
      r = poisson_solve(s, rhs, phi, o->poisson_max_iterations, o->poisson_tolerance, hist != nullptr);
    } catch(...) {
      poisson_destroy(s);
      throw;
    }
    poisson_destroy(s);
    d.pressure_solve = to_diag(r, hist, cap);
    // correct_velocity(working, phi, ...)
    double* pg = static_cast<double*>(x.scratch_get(3, fsize(*g, OF_CELL)));
    OFH_CHECK(hipMemcpyAsync(pg, phi, fsize(*g, OF_CELL), hipMemcpyDeviceToDevice, x.stream));
    launch_pressure_bc(x, *g, pb, pg);
    double* out[3] = {ou, ov, ow};
    for(int a = 0; a < 3; ++a)
      OFH_CHECK(hipMemcpyAsync(out[a], wk[a], fsize(*g, OF_FACE_X + a), hipMemcpyDeviceToDevice, x.stream));
    launch_correct(x, *g, wk[0], wk[1], wk[2], pg, -o->dt / o->density, out[0], out[1], out[2]);
    launch_velocity_bc(x, *g, *b, ou, ov, ow);
    launch_divergence(x, *g, ou, ov, ow, div);
    launch_norms(x, Lc, div, kScalRed0);
    OFH_CHECK(hipMemcpyAsync(x.host_scal, x.scal + kScalRed0, 2 * sizeof(double), hipMemcpyDeviceToHost, x.stream));
    OFH_CHECK(hipStreamSynchronize(x.stream));
    d.divergence_l2_after = std::sqrt(x.host_scal[0] / ncell);
    d.divergence_linf_after = x.host_scal[1];
    {
      // mean of phi over the active range
      std::vector<double> hp(make_lay(*g, OF_CELL).size());
      OFH_CHECK(hipMemcpy(hp.data(), phi, fsize(*g, OF_CELL), hipMemcpyDeviceToHost));
      double s2 = 0.0;
      for(int k = 1; k <= g->nz; ++k)
        for(int j = 1; j <= g->ny; ++j)
          for(int i = 1; i <= g->nx; ++i) s2 += hp[Lc.at(i, j, k)];
      d.pressure_mean = s2 / ncell;
    }
    if(prhs_out) {
      launch_scale_active(x, Lc, 1.0, rhs, prhs_out);
      OFH_CHECK(hipStreamSynchronize(x.stream));
    }
    if(diag) *diag = d;
  });
}

int of_taylor_green_dt(int nx, int ny, int nz, double cfl, double final_time, double* dt, int* steps) {
  return guarded([&] {
    if(nx <= 0 || ny <= 0 || nz <= 0) throw InvalidArgument("Taylor-Green config expects positive nx, ny, and nz");
    if(!(cfl > 0.0)) throw InvalidArgument("Taylor-Green config expects cfl_limit > 0");
    if(!(final_time > 0.0)) throw InvalidArgument("Taylor-Green config expects final_time > 0");
    const double L = 2.0 * std::acos(-1.0);
    const double dx = L / (double)nx, dy = L / (double)ny, dz = nz > 1 ? L / (double)nz : 1.0;
    const double inv = 1.0 / dx + 1.0 / dy + (nz == 1 ? 0.0 : 1.0 / dz);
    const double upper = cfl / inv;
    const int n = std::max(1, (int)std::ceil(final_time / upper));
    if(dt) *dt = final_time / (double)n;
    if(steps) *steps = n;
  });
}

int of_lid_driven_cavity_dt(int nx, int ny, double lid, double cfl, double* dt) {
  return guarded([&] {
    if(nx <= 0 || ny <= 0) throw InvalidArgument("lid-driven cavity config expects positive nx and ny");
    const double dx = 1.0 / (double)nx, dy = 1.0 / (double)ny;
    *dt = cfl / (lid / dx + lid / dy);
  });
}

int of_channel_flow_dt(int flow_case, int nx, double nu, double top, double drop, double cfl, double* dt) {
  return guarded([&] {  // solver/channel_flow.cpp:142-158
    need(dt, "dt");
    if(nx <= 0) throw InvalidArgument("channel flow config expects positive nx and ny");
    double max_u;
    if(flow_case == OF_CASE_CHANNEL_COUETTE) max_u = top;
    else if(flow_case == OF_CASE_CHANNEL_POISEUILLE) max_u = drop / (8.0 * nu);
    else throw InvalidArgument("unsupported channel flow case");
    const double dx = 1.0 / (double)nx;
    *dt = cfl / (max_u / dx);
  });
}

int of_stepper_create(of_context* c, const of_stepper_config* cfg, of_stepper** out) {
  return guarded([&] {
    need(cfg, "config");
    need(out, "out");
    Context& x = C(c);
    auto* s = new of_stepper{nullptr, c, x.device};
    try {
      s->s = new Stepper(x, *cfg);
    } catch(...) {
      delete s;
      throw;
    }
    context_retain(c);
    *out = s;
  });
}

int of_stepper_destroy(of_stepper* s) {
  return guarded([&] {
    if(!s) return;
    OFH_RELEASE(hipSetDevice(s->device));
    delete s->s;
    context_release(s->owner);
    delete s;
  });
}

int of_stepper_initialize(of_stepper* s, int convention) {
  return guarded([&] {
    need(s, "stepper");
    OFH_CHECK(hipSetDevice(s->device));
    s->s->initialize(convention);
  });
}

int of_stepper_upload(of_stepper* s, const double* u, const double* v, const double* w, const double* p,
                      const double* au, const double* av, const double* aw, int has_prev, const of_step_metrics* m) {
  return guarded([&] {
    need(s, "stepper");
    Stepper& S = *s->s;
    OFH_CHECK(hipSetDevice(s->device));
    const double* vin[3] = {u, v, w};
    const double* ain[3] = {au, av, aw};
    for(int a = 0; a < 3; ++a) need(vin[a], "velocity");
    need(p, "p_total");
    // the copies go on the stepper's stream (non-blocking, so the legacy stream would not
    // order them against its kernels) and the call returns when they are done
    OFH_CHECK(hipStreamSynchronize(S.c.stream));
    std::vector<double> active_only;
    for(int a = 0; a < 3; ++a) {
      const size_t n = fsize(S.g, OF_FACE_X + a);
      OFH_CHECK(hipMemcpyAsync(S.vel[a], vin[a], n, hipMemcpyHostToDevice, S.c.stream));
      // The reference carries the state's edge and corner ghosts unchanged from step to step
      // (every copy in run_taylor_green_steps takes them along and no BC fill writes them, so
      // the factorisation correction keeps reading the injected values). The step writes its new
      // velocity into the second buffer set and swaps, so both sets start with the same storage.
      OFH_CHECK(hipMemcpyAsync(S.corr[a], S.vel[a], n, hipMemcpyDeviceToDevice, S.c.stream));
      if(ain[a]) {
        // the reference's A_prev after a step carries the ghosts of its advection work array,
        // which are never written (zero); the step never reads them, so they are zero here too
        const Lay L = make_lay(S.g, OF_FACE_X + a);
        active_only.assign(L.size(), 0.0);
        for(int k = 1; k <= L.az; ++k)
          for(int j = 1; j <= L.ay; ++j)
            for(int i = 1; i <= L.ax; ++i) active_only[L.at(i, j, k)] = ain[a][L.at(i, j, k)];
        OFH_CHECK(hipMemcpyAsync(S.adv[a], active_only.data(), n, hipMemcpyHostToDevice, S.c.stream));
        OFH_CHECK(hipStreamSynchronize(S.c.stream));  // active_only is refilled for the next component
      } else {
        OFH_CHECK(hipMemsetAsync(S.adv[a], 0, n, S.c.stream));
      }
    }
    OFH_CHECK(hipMemcpyAsync(S.p_total, p, fsize(S.g, OF_CELL), hipMemcpyHostToDevice, S.c.stream));
    OFH_CHECK(hipStreamSynchronize(S.c.stream));
    S.has_prev = has_prev ? 1 : 0;
    if(m) S.metrics = *m;
    else {
      std::memset(&S.metrics, 0, sizeof S.metrics);
      S.metrics.dt = S.cfg.dt;
    }
  });
}

int of_stepper_download(of_stepper* s, double* u, double* v, double* w, double* p, double* au, double* av, double* aw,
                        int* has_prev, of_step_metrics* m) {
  return guarded([&] {
    need(s, "stepper");
    Stepper& S = *s->s;
    OFH_CHECK(hipSetDevice(s->device));
    OFH_CHECK(hipStreamSynchronize(S.c.stream));
    double* vo[3] = {u, v, w};
    double* ao[3] = {au, av, aw};
    for(int a = 0; a < 3; ++a) {
      if(vo[a]) OFH_CHECK(hipMemcpy(vo[a], S.vel[a], fsize(S.g, OF_FACE_X + a), hipMemcpyDeviceToHost));
      if(ao[a]) OFH_CHECK(hipMemcpy(ao[a], S.adv[a], fsize(S.g, OF_FACE_X + a), hipMemcpyDeviceToHost));
    }
    if(p) OFH_CHECK(hipMemcpy(p, S.p_total, fsize(S.g, OF_CELL), hipMemcpyDeviceToHost));
    if(has_prev) *has_prev = S.has_prev;
    if(m) *m = S.metrics;
  });
}

int of_stepper_device_fields(of_stepper* s, double** f) {
  return guarded([&] {
    need(s, "stepper");
    need(f, "fields");
    Stepper& S = *s->s;
    f[0] = S.vel[0];
    f[1] = S.vel[1];
    f[2] = S.vel[2];
    f[3] = S.p_total;
    f[4] = S.adv[0];
    f[5] = S.adv[1];
    f[6] = S.adv[2];
  });
}

int of_stepper_run(of_stepper* s, int steps, of_step_metrics* m) {
  return guarded([&] {
    need(s, "stepper");
    if(steps < 0) throw InvalidArgument("run_taylor_green_steps expects a non-negative step count");
    Stepper& S = *s->s;
    OFH_CHECK(hipSetDevice(s->device));
    for(int i = 0; i < steps; ++i) S.step();
    if(m) *m = S.metrics;
  });
}

int of_stepper_run_batched(of_stepper* s, int steps, of_step_metrics* m) { return of_stepper_run(s, steps, m); }

int of_stepper_kinetic_energy(of_stepper* s, double* out) {
  return guarded([&] {
    need(s, "stepper");
    need(out, "out");
    OFH_CHECK(hipSetDevice(s->device));
    *out = s->s->kinetic_energy();
  });
}

int of_stepper_set_forcing(of_stepper* s, const double* fu, const double* fv, const double* fw) {
  return guarded([&] {
    need(s, "stepper");
    OFH_CHECK(hipSetDevice(s->device));
    const double* f[3] = {fu, fv, fw};
    s->s->set_forcing(fu || fv || fw ? f : nullptr);
  });
}

int of_stepper_set_predictor(of_stepper* s, int kind, double tol, int maxit) {
  return guarded([&] {
    need(s, "stepper");
    OFH_CHECK(hipSetDevice(s->device));
    s->s->set_predictor(kind, tol, maxit);
  });
}

// ------------------------------------------------------------------ discrete adjoint
int of_apply_velocity_boundary_conditions_adjoint(of_context* c, const of_grid* g, const of_bc_set* b, double* u,
                                                  double* v, double* w) {
  return guarded([&] {
    validate_grid(g);
    need_periodic(b);
    need(u, "u"), need(v, "v"), need(w, "w");
    double* f[3] = {u, v, w};
    launch_velocity_bc_adjoint(C(c), *g, f);
    OFH_CHECK(hipGetLastError());
  });
}

int of_apply_pressure_boundary_conditions_adjoint(of_context* c, const of_grid* g, const of_pressure_bc_set* b,
                                                  double* p) {
  return guarded([&] {
    validate_grid(g);
    need(b, "bcs");
    for(int a = 0; a < 3; ++a)
      if(!pressure_axis_periodic(*b, a)) throw InvalidArgument("adjoint boundary conditions support periodic faces only");
    need(p, "p");
    launch_pressure_bc_adjoint(C(c), *g, p);
    OFH_CHECK(hipGetLastError());
  });
}

int of_compute_gradient_adjoint(of_context* c, const of_grid* g, const double* gx, const double* gy, const double* gz,
                                double* p) {
  return guarded([&] {
    validate_grid(g);
    need(gx, "gx"), need(gy, "gy"), need(gz, "gz"), need(p, "p");
    Context& x = C(c);
    OFH_CHECK(hipMemsetAsync(p, 0, fsize(*g, OF_CELL), x.stream));
    const double* gb[3] = {gx, gy, gz};
    launch_gradient_adjoint(x, *g, 1.0, gb, p);
    OFH_CHECK(hipGetLastError());
  });
}

int of_compute_divergence_adjoint(of_context* c, const of_grid* g, const double* d, double* u, double* v, double* w) {
  return guarded([&] {
    validate_grid(g);
    need(d, "d"), need(u, "u"), need(v, "v"), need(w, "w");
    Context& x = C(c);
    double* ub[3] = {u, v, w};
    for(int a = 0; a < 3; ++a) OFH_CHECK(hipMemsetAsync(ub[a], 0, fsize(*g, OF_FACE_X + a), x.stream));
    launch_divergence_adjoint(x, *g, 1.0, d, ub);
    OFH_CHECK(hipGetLastError());
  });
}

int of_solve_predictor_adi_adjoint(of_context* c, const of_grid* g, const of_bc_set* b, double alpha, const double* yu,
                                   const double* yv, const double* yw, double* ru, double* rv, double* rw) {
  return guarded([&] {
    validate_grid(g);
    need_periodic(b);
    need(yu, "yu"), need(yv, "yv"), need(yw, "yw"), need(ru, "ru"), need(rv, "rv"), need(rw, "rw");
    if(alpha < 0.0) throw InvalidArgument("solve_predictor_adi expects a non-negative alpha");
    const double* y[3] = {yu, yv, yw};
    double* r[3] = {ru, rv, rw};
    predictor_adi_adjoint(C(c), *g, *b, alpha, y, r);
  });
}

int of_compute_advection_tangent(of_context* c, const of_grid* g, int scheme, const double* u, const double* v,
                                 const double* w, const double* du, const double* dv, const double* dw, double* au,
                                 double* av, double* aw) {
  return guarded([&] {
    validate_grid(g);
    if(scheme < 0 || scheme > 2) throw InvalidArgument("unknown advection scheme");
    need(u, "u"), need(v, "v"), need(w, "w"), need(du, "du"), need(dv, "dv"), need(dw, "dw");
    need(au, "au"), need(av, "av"), need(aw, "aw");
    Context& x = C(c);
    double* out[3] = {au, av, aw};
    for(int a = 0; a < 3; ++a) OFH_CHECK(hipMemsetAsync(out[a], 0, fsize(*g, OF_FACE_X + a), x.stream));
    const double* vel[3] = {u, v, w};
    const double* dvel[3] = {du, dv, dw};
    launch_advection_tangent(x, *g, scheme, vel, dvel, out);
    OFH_CHECK(hipGetLastError());
  });
}

int of_compute_advection_adjoint(of_context* c, const of_grid* g, int scheme, const double* u, const double* v,
                                 const double* w, const double* bu, const double* bv, const double* bw, double* ubar,
                                 double* vbar, double* wbar) {
  return guarded([&] {
    validate_grid(g);
    if(scheme < 0 || scheme > 2) throw InvalidArgument("unknown advection scheme");
    need(u, "u"), need(v, "v"), need(w, "w"), need(bu, "abar_u"), need(bv, "abar_v"), need(bw, "abar_w");
    need(ubar, "ubar"), need(vbar, "vbar"), need(wbar, "wbar");
    Context& x = C(c);
    double* out[3] = {ubar, vbar, wbar};
    for(int a = 0; a < 3; ++a) OFH_CHECK(hipMemsetAsync(out[a], 0, fsize(*g, OF_FACE_X + a), x.stream));
    const double* vel[3] = {u, v, w};
    const double* ab[3] = {bu, bv, bw};
    launch_advection_adjoint(x, *g, scheme, vel, ab, out);
    OFH_CHECK(hipGetLastError());
  });
}

int of_adjoint_create(of_context* c, const of_stepper_config* cfg, int every, of_adjoint** out) {
  return guarded([&] {
    need(cfg, "config");
    need(out, "out");
    Context& x = C(c);
    auto* h = new of_adjoint{nullptr, c, x.device};
    try {
      h->a = adjoint_create(x, *cfg, every);
    } catch(...) {
      delete h;
      throw;
    }
    context_retain(c);
    *out = h;
  });
}

int of_adjoint_destroy(of_adjoint* a) {
  return guarded([&] {
    if(!a) return;
    OFH_RELEASE(hipSetDevice(a->device));
    adjoint_destroy(a->a);
    context_release(a->owner);
    delete a;
  });
}

int of_adjoint_set_forcing(of_adjoint* a, const double* fu, const double* fv, const double* fw) {
  return guarded([&] {
    AdjointRun& r = A(a);
    const double* f[3] = {fu, fv, fw};
    adjoint_set_forcing(&r, fu || fv || fw ? f : nullptr);
  });
}

int of_adjoint_forward(of_adjoint* a, const double* u0, const double* v0, const double* w0, const double* p0,
                       int steps, double* ke, of_step_metrics* m) {
  return guarded([&] {
    AdjointRun& r = A(a);
    const double* u[3] = {u0, v0, w0};
    const double e = adjoint_forward(&r, u, p0, steps, m);
    if(ke) *ke = e;
  });
}

int of_adjoint_final_velocity(of_adjoint* a, double* u, double* v, double* w) {
  return guarded([&] {
    AdjointRun& r = A(a);
    double* out[3] = {u, v, w};
    adjoint_final_velocity(&r, out);
  });
}

int of_adjoint_backward(of_adjoint* a, const double* su, const double* sv, const double* sw, double* gu, double* gv,
                        double* gw, double* gp, double* gfu, double* gfv, double* gfw) {
  return guarded([&] {
    AdjointRun& r = A(a);
    if((su || sv || sw) && !(su && sv && sw)) throw InvalidArgument("the seed needs all three components or none");
    const double* seed[3] = {su, sv, sw};
    double* g[3] = {gu, gv, gw};
    double* gf[3] = {gfu, gfv, gfw};
    adjoint_backward(&r, su ? seed : nullptr, g, gp, gf);
  });
}

int of_adjoint_tangent(of_adjoint* a, const double* du0, const double* dv0, const double* dw0, const double* dp0,
                       const double* dfu, const double* dfv, const double* dfw, double* du, double* dv, double* dw,
                       double* dke) {
  return guarded([&] {
    AdjointRun& r = A(a);
    if((dfu || dfv || dfw) && !(dfu && dfv && dfw))
      throw InvalidArgument("the forcing direction needs all three components or none");
    const double* d0[3] = {du0, dv0, dw0};
    const double* df[3] = {dfu, dfv, dfw};
    double* out[3] = {du, dv, dw};
    adjoint_tangent(&r, d0, dp0, dfu ? df : nullptr, out, dke);
  });
}

}  // extern "C"
