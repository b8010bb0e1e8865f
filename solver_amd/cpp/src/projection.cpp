// Helmholtz predictor, projection and pressure Poisson solve of the C++ interface (reference
// solver/projection.cpp:1028-1147, linsolve/poisson_solver.cpp:644-809), on the device.
#include "linsolve/poisson_solver.hpp"
#include "solver/hip_backend.hpp"
#include "solver/projection.hpp"

#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace solver {

namespace hb = hip_backend;

namespace {

void require_same_layout(const StructuredField& a, const StructuredField& b, const char* op) {
  if(!a.layout().same_shape_as(b.layout())) throw std::invalid_argument(std::string(op) + ": incompatible field layouts");
}

struct DevVel {
  hb::DeviceField x, y, z;
  explicit DevVel(const VelocityField& v) : x(v.x), y(v.y), z(v.z) {}
  void download(VelocityField& v) const {
    x.download(v.x);
    y.download(v.y);
    z.download(v.z);
  }
};

of_projection_options to_abi(const ProjectionOptions& o) {
  return of_projection_options{o.dt, o.density, o.poisson_max_iterations, o.poisson_tolerance};
}

PoissonSolveDiagnostics from_abi(const of_poisson_diagnostics& d, const std::vector<double>& history) {
  PoissonSolveDiagnostics out;
  out.iterations = d.iterations;
  out.initial_residual_l2 = d.initial_residual_l2;
  out.final_residual_l2 = d.final_residual_l2;
  out.relative_residual = d.relative_residual;
  out.removed_rhs_mean = d.removed_rhs_mean;
  out.multigrid_levels = d.multigrid_levels;
  out.coarse_unknowns = d.coarse_unknowns;
  out.pre_smoothing_steps = d.pre_smoothing_steps;
  out.post_smoothing_steps = d.post_smoothing_steps;
  out.converged = d.converged != 0;
  out.zero_mean_enforced = d.zero_mean_enforced != 0;
  const std::size_t n = std::min<std::size_t>(history.size(), static_cast<std::size_t>(d.history_length));
  out.residual_history.assign(history.begin(), history.begin() + static_cast<std::ptrdiff_t>(n));
  return out;
}

std::size_t history_capacity(const ProjectionOptions& o) {
  return static_cast<std::size_t>(o.poisson_max_iterations > 0 ? o.poisson_max_iterations : 0) + 1;
}

}  // namespace

HelmholtzDiagnostics solve_predictor_adi(const VelocityField& rhs, const double alpha,
                                         const BoundaryConditionSet& boundary_conditions,
                                         VelocityField& predicted_velocity) {
  if(alpha < 0.0) throw std::invalid_argument("solve_predictor_adi expects a non-negative alpha");
  require_same_layout(rhs.x, predicted_velocity.x, "solve_predictor_adi");
  require_same_layout(rhs.y, predicted_velocity.y, "solve_predictor_adi");
  require_same_layout(rhs.z, predicted_velocity.z, "solve_predictor_adi");
  const of_grid g = hb::to_abi(rhs.x.layout().grid());
  const of_bc_set b = hb::to_abi(boundary_conditions);
  DevVel in(rhs), out(predicted_velocity);
  int line_solves = 0;
  hb::check(of_solve_predictor_adi(hb::context(), &g, &b, alpha, in.x.get(), in.y.get(), in.z.get(), out.x.get(),
                                   out.y.get(), out.z.get(), &line_solves));
  out.download(predicted_velocity);
  return HelmholtzDiagnostics{line_solves};
}

void build_pressure_rhs(const VelocityField& predicted_velocity, const BoundaryConditionSet& boundary_conditions,
                        const ProjectionOptions& options, ScalarField& rhs) {
  const of_grid g = hb::to_abi(rhs.layout().grid());
  const of_bc_set b = hb::to_abi(boundary_conditions);
  const of_projection_options o = to_abi(options);
  DevVel in(predicted_velocity);
  hb::DeviceField r(rhs);
  hb::check(of_build_pressure_rhs(hb::context(), &g, &b, &o, in.x.get(), in.y.get(), in.z.get(), r.get()));
  r.download(rhs);
}

void correct_velocity(const VelocityField& predicted_velocity, const PressureField& pressure_correction,
                      const BoundaryConditionSet& boundary_conditions, const ProjectionOptions& options,
                      VelocityField& corrected_velocity) {
  if(!(options.dt > 0.0)) throw std::invalid_argument("correct_velocity expects dt > 0");
  if(!(options.density > 0.0)) throw std::invalid_argument("correct_velocity expects density > 0");
  require_same_layout(predicted_velocity.x, corrected_velocity.x, "correct_velocity");
  require_same_layout(predicted_velocity.y, corrected_velocity.y, "correct_velocity");
  require_same_layout(predicted_velocity.z, corrected_velocity.z, "correct_velocity");
  const of_grid g = hb::to_abi(pressure_correction.layout().grid());
  const of_bc_set b = hb::to_abi(boundary_conditions);
  const of_projection_options o = to_abi(options);
  DevVel in(predicted_velocity), out(corrected_velocity);
  hb::DeviceField phi(pressure_correction);
  hb::check(of_correct_velocity(hb::context(), &g, &b, &o, in.x.get(), in.y.get(), in.z.get(), phi.get(), out.x.get(),
                                out.y.get(), out.z.get()));
  out.download(corrected_velocity);
}

ProjectionDiagnostics project_velocity(const VelocityField& predicted_velocity,
                                       const BoundaryConditionSet& boundary_conditions,
                                       const ProjectionOptions& options, PressureField& pressure_correction,
                                       VelocityField& corrected_velocity, ScalarField* pressure_rhs) {
  const of_grid g = hb::to_abi(pressure_correction.layout().grid());
  const of_bc_set b = hb::to_abi(boundary_conditions);
  const of_projection_options o = to_abi(options);
  DevVel in(predicted_velocity), out(corrected_velocity);
  hb::DeviceField phi(pressure_correction);
  ScalarField rhs_host{pressure_correction.layout().grid()};
  if(pressure_rhs != nullptr) require_same_layout(*pressure_rhs, rhs_host, "project_velocity");
  hb::DeviceField rhs(pressure_rhs != nullptr ? static_cast<const StructuredField&>(*pressure_rhs) : rhs_host);
  std::vector<double> history(history_capacity(options));
  of_projection_diagnostics d{};
  hb::check(of_project_velocity(hb::context(), &g, &b, &o, hb::pressure_precision(), in.x.get(), in.y.get(),
                                in.z.get(), phi.get(), out.x.get(), out.y.get(), out.z.get(),
                                pressure_rhs != nullptr ? rhs.get() : nullptr, &d, history.data(),
                                static_cast<int>(history.size())));
  phi.download(pressure_correction);
  out.download(corrected_velocity);
  if(pressure_rhs != nullptr) rhs.download(*pressure_rhs);
  ProjectionDiagnostics pd;
  pd.rhs_l2 = d.rhs_l2;
  pd.divergence_l2_before = d.divergence_l2_before;
  pd.divergence_l2_after = d.divergence_l2_after;
  pd.pressure_mean = d.pressure_mean;
  pd.pressure_solve = from_abi(d.pressure_solve, history);
  return pd;
}

void require_converged_pressure_projection(const ProjectionDiagnostics& diagnostics, const ProjectionOptions& options,
                                           const char* solver_name, const int step_index) {
  if(diagnostics.pressure_solve.converged) return;
  std::ostringstream msg;
  msg << solver_name << " pressure solve did not converge at step " << step_index
      << " (iterations=" << diagnostics.pressure_solve.iterations
      << ", relative_residual=" << diagnostics.pressure_solve.relative_residual
      << ", tolerance=" << options.poisson_tolerance << ")";
  throw std::runtime_error(msg.str());
}

namespace linsolve {

std::string to_string(const MultigridCycle cycle) { return cycle == MultigridCycle::v_cycle ? "v_cycle" : "unknown"; }

std::string to_string(const MultigridSmoother smoother) {
  return smoother == MultigridSmoother::damped_jacobi ? "damped_jacobi" : "unknown";
}

MultigridPolicy default_multigrid_policy() { return MultigridPolicy{}; }

void build_poisson_rhs_from_pressure(const PressureField& pressure, const PressureBoundarySet& boundary_conditions,
                                     ScalarField& rhs) {
  require_same_layout(pressure, rhs, "build_poisson_rhs_from_pressure");
  const of_grid g = hb::to_abi(pressure.layout().grid());
  const of_pressure_bc_set b = hb::to_abi(boundary_conditions);
  hb::DeviceField p(pressure), r(rhs);
  hb::check(of_build_poisson_rhs_from_pressure(hb::context(), &g, &b, p.get(), r.get()));
  r.download(rhs);
}

PoissonSolveDiagnostics solve_pressure_poisson(const ScalarField& rhs, const PressureBoundarySet& boundary_conditions,
                                               const ProjectionOptions& options, PressureField& pressure) {
  if(options.poisson_max_iterations <= 0)
    throw std::invalid_argument("solve_pressure_poisson expects iterations > 0");
  if(!(options.poisson_tolerance > 0.0))
    throw std::invalid_argument("solve_pressure_poisson expects tolerance > 0");
  require_same_layout(rhs, pressure, "solve_pressure_poisson");
  const of_grid g = hb::to_abi(rhs.layout().grid());
  const of_pressure_bc_set b = hb::to_abi(boundary_conditions);
  const of_projection_options o = to_abi(options);
  hb::DeviceField r(rhs), p(pressure);
  std::vector<double> history(history_capacity(options));
  of_poisson_diagnostics d{};
  hb::check(of_solve_pressure_poisson(hb::context(), &g, &b, &o, hb::pressure_precision(), r.get(), p.get(), &d,
                                      history.data(), static_cast<int>(history.size())));
  p.download(pressure);
  return from_abi(d, history);
}

}  // namespace linsolve

}  // namespace solver
