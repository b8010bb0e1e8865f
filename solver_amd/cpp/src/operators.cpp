// MAC operators and momentum terms of the C++ interface (reference
// operators/discrete_operators.cpp, solver/momentum_terms.cpp:281-401), evaluated on the device.
// Inputs and outputs are copied whole, so storage the kernels do not write keeps the caller's
// values, as in the reference.
#include "operators/discrete_operators.hpp"
#include "solver/hip_backend.hpp"
#include "solver/momentum_terms.hpp"

#include <stdexcept>
#include <string>

namespace solver {

namespace hb = hip_backend;

namespace {

void require_same_layout(const StructuredField& a, const StructuredField& b, const char* op) {
  if(!a.layout().same_shape_as(b.layout())) throw std::invalid_argument(std::string(op) + ": incompatible field layouts");
}

void require_same_velocity_layouts(const VelocityField& a, const VelocityField& b, const char* op) {
  if(!a.x.layout().same_shape_as(b.x.layout()) || !a.y.layout().same_shape_as(b.y.layout()) ||
     !a.z.layout().same_shape_as(b.z.layout()))
    throw std::invalid_argument(std::string(op) + ": incompatible velocity layouts");
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

}  // namespace

namespace operators {

void compute_gradient(const PressureField& pressure, VelocityField& gradient) {
  if(pressure.layout().location() != FieldLocation::cell_center)
    throw std::invalid_argument("compute_gradient expects cell-centered pressure");
  const of_grid g = hb::to_abi(pressure.layout().grid());
  hb::DeviceField p(pressure);
  DevVel out(gradient);
  hb::check(of_compute_gradient(hb::context(), &g, p.get(), out.x.get(), out.y.get(), out.z.get()));
  out.download(gradient);
}

void compute_divergence(const VelocityField& velocity, ScalarField& divergence) {
  const of_grid g = hb::to_abi(divergence.layout().grid());
  DevVel in(velocity);
  hb::DeviceField d(divergence);
  hb::check(of_compute_divergence(hb::context(), &g, in.x.get(), in.y.get(), in.z.get(), d.get()));
  d.download(divergence);
}

void compute_laplacian(const StructuredField& input, StructuredField& output) {
  require_same_layout(input, output, "compute_laplacian");
  const of_grid g = hb::to_abi(input.layout().grid());
  hb::DeviceField in(input), out(output);
  hb::check(of_compute_laplacian(hb::context(), &g, static_cast<int>(input.layout().location()), in.get(), out.get()));
  out.download(output);
}

void compute_laplacian(const VelocityField& input, VelocityField& output) {
  compute_laplacian(input.x, output.x);
  compute_laplacian(input.y, output.y);
  compute_laplacian(input.z, output.z);
}

}  // namespace operators

std::string to_string(const AdvectionScheme scheme) {
  switch(scheme) {
    case AdvectionScheme::tvd: return "tvd";
    case AdvectionScheme::upwind: return "upwind";
    case AdvectionScheme::central: return "central";
  }
  return "unknown";
}

std::string to_string(const FluxLimiter limiter) { return limiter == FluxLimiter::van_leer ? "van_leer" : "unknown"; }

std::string describe(const AdvectionOptions& options) {
  return "scheme=" + to_string(options.scheme) + ", limiter=" + to_string(options.limiter);
}

void compute_advection_term(const VelocityField& velocity, const AdvectionOptions& options, VelocityField& advection) {
  require_same_velocity_layouts(velocity, advection, "compute_advection_term");
  const of_grid g = hb::to_abi(velocity.x.layout().grid());
  DevVel in(velocity), out(advection);
  hb::check(of_compute_advection_term(hb::context(), &g, static_cast<int>(options.scheme), in.x.get(), in.y.get(),
                                      in.z.get(), out.x.get(), out.y.get(), out.z.get()));
  out.download(advection);
}

void compute_diffusion_term(const VelocityField& velocity, const double viscosity, VelocityField& diffusion) {
  require_same_velocity_layouts(velocity, diffusion, "compute_diffusion_term");
  const of_grid g = hb::to_abi(velocity.x.layout().grid());
  DevVel in(velocity), out(diffusion);
  hb::check(of_compute_diffusion_term(hb::context(), &g, viscosity, in.x.get(), in.y.get(), in.z.get(), out.x.get(),
                                      out.y.get(), out.z.get()));
  out.download(diffusion);
}

CflDiagnostics compute_advective_cfl(const VelocityField& velocity, const double dt) {
  const of_grid g = hb::to_abi(velocity.x.layout().grid());
  DevVel in(velocity);
  of_cfl_diagnostics d{};
  hb::check(of_compute_advective_cfl(hb::context(), &g, in.x.get(), in.y.get(), in.z.get(), dt, &d));
  return CflDiagnostics{d.dt, d.max_cfl, d.max_u, d.max_v, d.max_w};
}

}  // namespace solver
