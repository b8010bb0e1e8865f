// Boundary conditions of the C++ interface (reference solver/projection.cpp:892-1026): the
// descriptors on the host, the ghost fills on the device.
#include "bc/boundary_conditions.hpp"

#include "solver/hip_backend.hpp"

namespace solver {

namespace hb = hip_backend;

BoundaryConditionSet BoundaryConditionSet::all(const PhysicalBoundaryType type) {
  BoundaryConditionSet s;
  for(BoundaryCondition& c : s.faces) c.type = type;
  return s;
}

BoundaryConditionSet BoundaryConditionSet::cavity() { return all(PhysicalBoundaryType::no_slip_wall); }

std::string to_string(const PhysicalBoundaryType type) {
  switch(type) {
    case PhysicalBoundaryType::no_slip_wall: return "no_slip_wall";
    case PhysicalBoundaryType::prescribed_velocity: return "prescribed_velocity";
    case PhysicalBoundaryType::symmetry: return "symmetry";
    case PhysicalBoundaryType::fixed_pressure: return "fixed_pressure";
    case PhysicalBoundaryType::periodic: return "periodic";
  }
  return "unknown";
}

std::string to_string(const PressureBoundaryType type) {
  switch(type) {
    case PressureBoundaryType::neumann: return "neumann";
    case PressureBoundaryType::dirichlet: return "dirichlet";
    case PressureBoundaryType::periodic: return "periodic";
  }
  return "unknown";
}

PressureBoundarySet derive_pressure_correction_boundary_conditions(const BoundaryConditionSet& boundary_conditions) {
  const of_bc_set in = hb::to_abi(boundary_conditions);
  of_pressure_bc_set out{};
  hb::check(of_derive_pressure_correction_boundary_conditions(&in, &out));
  PressureBoundarySet s;
  for(int f = 0; f < 6; ++f) {
    PressureBoundaryCondition& c = s.faces[static_cast<std::size_t>(f)];
    c.type = static_cast<PressureBoundaryType>(out.faces[f].type);
    c.value = out.faces[f].value;
    c.gradient = out.faces[f].gradient;
  }
  return s;
}

void apply_velocity_boundary_conditions(const BoundaryConditionSet& boundary_conditions, VelocityField& velocity) {
  const of_grid g = hb::to_abi(velocity.x.layout().grid());
  const of_bc_set b = hb::to_abi(boundary_conditions);
  hb::DeviceField u(velocity.x), v(velocity.y), w(velocity.z);
  hb::check(of_apply_velocity_boundary_conditions(hb::context(), &g, &b, u.get(), v.get(), w.get()));
  u.download(velocity.x);
  v.download(velocity.y);
  w.download(velocity.z);
}

void apply_pressure_boundary_conditions(const PressureBoundarySet& boundary_conditions, PressureField& pressure) {
  const of_grid g = hb::to_abi(pressure.layout().grid());
  const of_pressure_bc_set b = hb::to_abi(boundary_conditions);
  hb::DeviceField p(pressure);
  hb::check(of_apply_pressure_boundary_conditions(hb::context(), &g, &b, p.get()));
  p.download(pressure);
}

void apply_total_pressure_boundary_conditions(const BoundaryConditionSet& boundary_conditions,
                                              const VelocityField& normal_pressure_gradient_source,
                                              PressureField& pressure_total) {
  const of_grid g = hb::to_abi(pressure_total.layout().grid());
  const of_bc_set b = hb::to_abi(boundary_conditions);
  hb::DeviceField su(normal_pressure_gradient_source.x), sv(normal_pressure_gradient_source.y),
      sw(normal_pressure_gradient_source.z), p(pressure_total);
  hb::check(of_apply_total_pressure_boundary_conditions(hb::context(), &g, &b, su.get(), sv.get(), sw.get(), p.get()));
  p.download(pressure_total);
}

}  // namespace solver
