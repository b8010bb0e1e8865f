// Channel-flow driver of the C++ interface (reference solver/channel_flow.cpp): the .cfg
// surface, boundary sets and profile validation on the host, set-up and step loop on the device
// (of_stepper with OF_CASE_CHANNEL_COUETTE / OF_CASE_CHANNEL_POISEUILLE).
#include "solver/channel_flow.hpp"

#include "solver/hip_backend.hpp"

#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

namespace solver {

namespace hb = hip_backend;

namespace {

std::string trim(const std::string& s) {
  const std::size_t a = s.find_first_not_of(" \t\r\n");
  if(a == std::string::npos) return "";
  const std::size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

bool parse_bool(const std::string& v) {
  if(v == "true" || v == "1" || v == "yes" || v == "on") return true;
  if(v == "false" || v == "0" || v == "no" || v == "off") return false;
  throw std::runtime_error("invalid boolean value: " + v);
}

ChannelFlowCase parse_case(const std::string& v) {
  if(v == "couette") return ChannelFlowCase::couette;
  if(v == "poiseuille") return ChannelFlowCase::poiseuille;
  throw std::runtime_error("unsupported channel flow case: " + v);
}

// solver/channel_flow.cpp:47-69
void validate_config(const ChannelFlowConfig& c) {
  if(c.nx <= 0 || c.ny <= 0) throw std::invalid_argument("channel flow config expects positive nx and ny");
  if(c.viscosity <= 0.0) throw std::invalid_argument("channel flow config expects viscosity > 0");
  if(c.cfl_limit <= 0.0) throw std::invalid_argument("channel flow config expects cfl_limit > 0");
  if(c.steps <= 0) throw std::invalid_argument("channel flow config expects steps > 0");
  if(c.poisson_max_iterations <= 0 || c.poisson_tolerance <= 0.0)
    throw std::invalid_argument("channel flow config expects positive Poisson controls");
  if(c.case_kind == ChannelFlowCase::couette && c.top_velocity <= 0.0)
    throw std::invalid_argument("Couette config expects top_velocity > 0");
  if(c.case_kind == ChannelFlowCase::poiseuille && c.pressure_drop <= 0.0)
    throw std::invalid_argument("Poiseuille config expects pressure_drop > 0");
}

int abi_case(const ChannelFlowConfig& c) {
  return c.case_kind == ChannelFlowCase::couette ? OF_CASE_CHANNEL_COUETTE : OF_CASE_CHANNEL_POISEUILLE;
}

// analytic streamwise velocity at wall-normal coordinate y (solver/channel_flow.cpp:160-169)
double exact_u(const ChannelFlowConfig& c, const double y) {
  return c.case_kind == ChannelFlowCase::couette ? c.top_velocity * y
                                                 : 0.5 * c.pressure_drop * y * (1.0 - y) / c.viscosity;
}

// x-average of u on every active row of the first k plane (solver/channel_flow.cpp:258-279)
WallNormalProfile streamwise_profile(const VelocityField& v) {
  const FieldLayout& L = v.x.layout();
  const IndexRange3D r = L.active_range();
  const double count = static_cast<double>(r.i_end - r.i_begin);
  WallNormalProfile p;
  for(int j = r.j_begin; j < r.j_end; ++j) {
    double sum = 0.0;
    for(int i = r.i_begin; i < r.i_end; ++i) sum += v.x(i, j, r.k_begin);
    p.coordinate.push_back(L.coordinate_for_storage_index(Axis::y, j));
    p.value.push_back(sum / count);
  }
  return p;
}

// relative L2 distance of the profile to the analytic one (solver/channel_flow.cpp:281-301)
double profile_error(const ChannelFlowConfig& c, const WallNormalProfile& p) {
  if(p.coordinate.size() != p.value.size() || p.coordinate.empty())
    throw std::invalid_argument("channel profile validation requires matching non-empty arrays");
  double num = 0.0, den = 0.0;
  for(std::size_t n = 0; n < p.coordinate.size(); ++n) {
    const double ref = exact_u(c, p.coordinate[n]);
    const double d = p.value[n] - ref;
    num += d * d;
    den += ref * ref;
  }
  return den == 0.0 ? std::sqrt(num) : std::sqrt(num / den);
}

}  // namespace

ChannelFlowConfig default_channel_flow_config() { return ChannelFlowConfig{}; }

ChannelFlowConfig load_channel_flow_config(const std::string& path) {
  std::ifstream in(path);
  if(!in.is_open()) throw std::runtime_error("unable to open channel flow config: " + path);
  ChannelFlowConfig c = default_channel_flow_config();
  std::string line;
  int number = 0;
  while(std::getline(in, line)) {
    ++number;
    const std::size_t hash = line.find('#');
    if(hash != std::string::npos) line.erase(hash);
    line = trim(line);
    if(line.empty()) continue;
    const std::size_t eq = line.find('=');
    if(eq == std::string::npos) throw std::runtime_error("invalid config line " + std::to_string(number) + ": " + line);
    const std::string key = trim(line.substr(0, eq));
    const std::string value = trim(line.substr(eq + 1));
    if(key == "case") c.case_kind = parse_case(value);
    else if(key == "nx") c.nx = std::stoi(value);
    else if(key == "ny") c.ny = std::stoi(value);
    else if(key == "viscosity") c.viscosity = std::stod(value);
    else if(key == "top_velocity") c.top_velocity = std::stod(value);
    else if(key == "pressure_drop") c.pressure_drop = std::stod(value);
    else if(key == "cfl_limit") c.cfl_limit = std::stod(value);
    else if(key == "steps") c.steps = std::stoi(value);
    else if(key == "poisson_max_iterations") c.poisson_max_iterations = std::stoi(value);
    else if(key == "poisson_tolerance") c.poisson_tolerance = std::stod(value);
    else if(key == "validate_profile") c.validate_profile = parse_bool(value);
    else throw std::runtime_error("unsupported channel flow config key: " + key);
  }
  validate_config(c);
  return c;
}

std::string to_string(const ChannelFlowCase case_kind) {
  switch(case_kind) {
    case ChannelFlowCase::couette: return "couette";
    case ChannelFlowCase::poiseuille: return "poiseuille";
  }
  return "unknown";
}

std::string describe(const ChannelFlowConfig& c) {
  std::ostringstream s;
  s << "case=" << to_string(c.case_kind) << ", nx=" << c.nx << ", ny=" << c.ny << ", viscosity=" << c.viscosity
    << ", top_velocity=" << c.top_velocity << ", pressure_drop=" << c.pressure_drop << ", cfl_limit=" << c.cfl_limit
    << ", steps=" << c.steps << ", poisson_max_iterations=" << c.poisson_max_iterations
    << ", poisson_tolerance=" << c.poisson_tolerance
    << ", validate_profile=" << (c.validate_profile ? "true" : "false") << ", advection=" << describe(c.advection);
  return s.str();
}

// solver/channel_flow.cpp:408-430: walls in y, symmetry in z; Couette periodic in x with the
// upper wall moving at top_velocity, Poiseuille with pressure_drop at x_min and 0 at x_max.
BoundaryConditionSet make_channel_flow_boundary_conditions(const ChannelFlowConfig& c) {
  BoundaryConditionSet b = BoundaryConditionSet::all(PhysicalBoundaryType::no_slip_wall);
  b[BoundaryFace::z_min].type = PhysicalBoundaryType::symmetry;
  b[BoundaryFace::z_max].type = PhysicalBoundaryType::symmetry;
  if(c.case_kind == ChannelFlowCase::couette) {
    b[BoundaryFace::x_min].type = PhysicalBoundaryType::periodic;
    b[BoundaryFace::x_max].type = PhysicalBoundaryType::periodic;
    b[BoundaryFace::y_max].type = PhysicalBoundaryType::prescribed_velocity;
    b[BoundaryFace::y_max].velocity = {c.top_velocity, 0.0, 0.0};
  } else {
    b[BoundaryFace::x_min].type = PhysicalBoundaryType::fixed_pressure;
    b[BoundaryFace::x_min].pressure = c.pressure_drop;
    b[BoundaryFace::x_max].type = PhysicalBoundaryType::fixed_pressure;
    b[BoundaryFace::x_max].pressure = 0.0;
  }
  return b;
}

ChannelFlowValidation validate_channel_flow_profile(const ChannelFlowResult& result) {
  const ChannelFlowConfig& c = result.config;
  ChannelFlowValidation v;
  v.reference_dataset =
      c.case_kind == ChannelFlowCase::couette ? "analytic_couette_profile" : "analytic_poiseuille_profile";
  v.relative_l2_error = profile_error(c, result.streamwise_profile);
  v.divergence_l2 = result.final_step.divergence_l2;
  v.pass = v.relative_l2_error <= 5.0e-3 && v.divergence_l2 <= 1.0e-10;  // solver/channel_flow.cpp:421-430
  return v;
}

ChannelFlowResult run_channel_flow(const ChannelFlowConfig& c) {
  validate_config(c);
  const int kind = abi_case(c);
  double dt = 0.0;
  hb::check(of_channel_flow_dt(kind, c.nx, c.viscosity, c.top_velocity, c.pressure_drop, c.cfl_limit, &dt));
  of_stepper_config sc{};
  sc.flow_case = kind;
  sc.nx = c.nx;
  sc.ny = c.ny;
  sc.nz = 1;
  sc.viscosity = c.viscosity;
  sc.dt = dt;
  sc.scheme = static_cast<int>(c.advection.scheme);
  sc.poisson_max_iterations = c.poisson_max_iterations;
  sc.poisson_tolerance = c.poisson_tolerance;
  sc.precision = OF_PRECISION_FP64;
  sc.lid_velocity = c.top_velocity;
  sc.pressure_drop = c.pressure_drop;

  const Grid grid{c.nx, c.ny, 1, 1.0 / static_cast<double>(c.nx), 1.0 / static_cast<double>(c.ny), 1.0, 1};
  VelocityField velocity{grid}, advection{grid};
  PressureField pressure{grid};
  of_step_metrics m{};
  of_stepper* s = nullptr;
  hb::check(of_stepper_create(hb::context(), &sc, &s));
  try {
    hb::check(of_stepper_initialize(s, 0));
    hb::check(of_stepper_run(s, c.steps, &m));
    int has_prev = 0;
    hb::check(of_stepper_download(s, velocity.x.data(), velocity.y.data(), velocity.z.data(), pressure.data(),
                                  advection.x.data(), advection.y.data(), advection.z.data(), &has_prev, &m));
  } catch(...) {
    (void)of_stepper_destroy(s);
    throw;
  }
  hb::check(of_stepper_destroy(s));
  apply_velocity_boundary_conditions(make_channel_flow_boundary_conditions(c), velocity);

  ChannelFlowResult r;
  r.config = c;
  r.final_step.step = m.step;
  r.final_step.time = m.time;
  r.final_step.dt = m.dt;
  r.final_step.max_cfl = m.max_cfl;
  r.final_step.max_velocity_change = m.max_velocity_change;
  r.final_step.divergence_l2 = m.divergence_l2;
  r.final_step.pressure_iterations = m.pressure_iterations;
  r.final_step.pressure_relative_residual = m.pressure_relative_residual;
  r.streamwise_profile = streamwise_profile(velocity);
  if(c.validate_profile) r.validation = validate_channel_flow_profile(r);
  return r;
}

}  // namespace solver
