// Lid-driven cavity driver of the C++ interface (reference solver/lid_driven_cavity.cpp): the
// .cfg surface, centreline sampling and the Re = 100 envelope check on the host, the step loop
// on the device.
#include "solver/lid_driven_cavity.hpp"

#include "solver/hip_backend.hpp"

#include <algorithm>
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

void validate_config(const LidDrivenCavityConfig& c) {
  if(c.nx <= 0 || c.ny <= 0) throw std::invalid_argument("lid-driven cavity config expects positive nx and ny");
  if(c.reynolds <= 0.0) throw std::invalid_argument("lid-driven cavity config expects reynolds > 0");
  if(c.lid_velocity <= 0.0) throw std::invalid_argument("lid-driven cavity config expects lid_velocity > 0");
  if(c.cfl_limit <= 0.0) throw std::invalid_argument("lid-driven cavity config expects cfl_limit > 0");
  if(c.max_steps <= 0 || c.min_steps < 0)
    throw std::invalid_argument("lid-driven cavity config expects non-negative step controls");
  if(c.poisson_max_iterations <= 0 || c.poisson_tolerance <= 0.0)
    throw std::invalid_argument("lid-driven cavity config expects positive Poisson controls");
}

bool same_grid(const Grid& a, const Grid& b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx && a.dy == b.dy && a.dz == b.dz &&
         a.ghost_layers == b.ghost_layers;
}

// u along the vertical line through the middle x face / v along the horizontal line through the
// middle y face, at the first k plane (solver/lid_driven_cavity.cpp:101-135)
CenterlineProfile vertical_u(const VelocityField& v) {
  const FieldLayout& L = v.x.layout();
  const IndexRange3D r = L.active_range();
  const int ic = r.i_begin + L.active_extent().nx / 2;
  CenterlineProfile p;
  for(int j = r.j_begin; j < r.j_end; ++j) {
    p.coordinate.push_back(L.coordinate_for_storage_index(Axis::y, j));
    p.value.push_back(v.x(ic, j, r.k_begin));
  }
  return p;
}

CenterlineProfile horizontal_v(const VelocityField& v) {
  const FieldLayout& L = v.y.layout();
  const IndexRange3D r = L.active_range();
  const int jc = r.j_begin + L.active_extent().ny / 2;
  CenterlineProfile p;
  for(int i = r.i_begin; i < r.i_end; ++i) {
    p.coordinate.push_back(L.coordinate_for_storage_index(Axis::x, i));
    p.value.push_back(v.y(i, jc, r.k_begin));
  }
  return p;
}

// extrema over the interior samples (first and last excluded)
CenterlineExtrema extrema(const CenterlineProfile& u, const CenterlineProfile& v) {
  if(u.value.size() < 3 || v.value.size() < 3)
    throw std::invalid_argument("centerline extrema require at least three samples per profile");
  const auto [umin, umax] = std::minmax_element(u.value.begin() + 1, u.value.end() - 1);
  const auto [vmin, vmax] = std::minmax_element(v.value.begin() + 1, v.value.end() - 1);
  return CenterlineExtrema{*umax, *umin, *vmax, *vmin};
}

// relative distance to the nearest end of [lo, hi] (0 inside)
double miss_against_envelope(const double value, const double a, const double b) {
  const double lo = std::min(a, b), hi = std::max(a, b);
  if(value >= lo && value <= hi) return 0.0;
  const double nearest = value < lo ? lo : hi;
  return std::abs(value - nearest) / std::abs(nearest);
}

// cubic Lagrange interpolation on the four samples around `x` (linear with fewer than four)
// (solver/lid_driven_cavity.cpp:182-226)
double interpolate(const CenterlineProfile& p, const double x) {
  const std::size_t n = p.coordinate.size();
  if(n != p.value.size() || n < 2)
    throw std::invalid_argument("centerline profile interpolation requires matching sample arrays");
  if(x < p.coordinate.front() || x > p.coordinate.back())
    throw std::out_of_range("reference coordinate falls outside the sampled centerline range");
  const std::size_t up = static_cast<std::size_t>(std::lower_bound(p.coordinate.begin(), p.coordinate.end(), x) -
                                                  p.coordinate.begin());
  if(up == 0) return p.value.front();
  if(up == n) return p.value.back();
  if(p.coordinate[up] == x) return p.value[up];
  const std::size_t lo = up - 1;
  if(n >= 4) {
    const std::size_t s = std::min(lo > 0 ? lo - 1 : 0, n - 4);
    double acc = 0.0;
    for(std::size_t a = 0; a < 4; ++a) {
      double basis = 1.0;
      for(std::size_t b = 0; b < 4; ++b)
        if(a != b) basis *= (x - p.coordinate[s + b]) / (p.coordinate[s + a] - p.coordinate[s + b]);
      acc += basis * p.value[s + a];
    }
    return acc;
  }
  const double w = (x - p.coordinate[lo]) / (p.coordinate[up] - p.coordinate[lo]);
  return (1.0 - w) * p.value[lo] + w * p.value[up];
}

// Device-resident stepper for one cavity configuration.
class DeviceRun {
 public:
  explicit DeviceRun(const LidDrivenCavityConfig& c) {
    of_stepper_config sc{};
    sc.flow_case = OF_CASE_LID_DRIVEN_CAVITY;
    sc.nx = c.nx;
    sc.ny = c.ny;
    sc.nz = 1;
    sc.viscosity = lid_driven_cavity_viscosity(c);
    sc.dt = lid_driven_cavity_dt(c);
    sc.scheme = static_cast<int>(c.advection.scheme);
    sc.poisson_max_iterations = c.poisson_max_iterations;
    sc.poisson_tolerance = c.poisson_tolerance;
    sc.precision = OF_PRECISION_FP64;
    sc.lid_velocity = c.lid_velocity;
    hb::check(of_stepper_create(hb::context(), &sc, &s_));
  }
  DeviceRun(const DeviceRun&) = delete;
  DeviceRun& operator=(const DeviceRun&) = delete;
  ~DeviceRun() {
    if(s_ != nullptr) (void)of_stepper_destroy(s_);
  }
  void initialize() { hb::check(of_stepper_initialize(s_, 0)); }
  void upload(const LidDrivenCavityState& st) {
    of_step_metrics m{};
    m.step = st.metrics.step;
    m.time = st.metrics.time;
    m.dt = st.metrics.dt;
    m.max_cfl = st.metrics.max_cfl;
    m.max_velocity_change = st.metrics.max_velocity_change;
    m.divergence_l2 = st.metrics.divergence_l2;
    m.pressure_iterations = st.metrics.pressure_iterations;
    m.pressure_relative_residual = st.metrics.pressure_relative_residual;
    hb::check(of_stepper_upload(s_, st.velocity.x.data(), st.velocity.y.data(), st.velocity.z.data(),
                                st.pressure_total.data(), st.advection_previous.x.data(),
                                st.advection_previous.y.data(), st.advection_previous.z.data(),
                                st.has_previous_advection ? 1 : 0, &m));
  }
  of_step_metrics run(const int steps) {
    of_step_metrics m{};
    hb::check(of_stepper_run(s_, steps, &m));
    return m;
  }
  void download(LidDrivenCavityState& st) {
    int has_prev = 0;
    of_step_metrics m{};
    hb::check(of_stepper_download(s_, st.velocity.x.data(), st.velocity.y.data(), st.velocity.z.data(),
                                  st.pressure_total.data(), st.advection_previous.x.data(),
                                  st.advection_previous.y.data(), st.advection_previous.z.data(), &has_prev, &m));
    st.has_previous_advection = has_prev != 0;
    st.metrics.step = m.step;
    st.metrics.time = m.time;
    st.metrics.dt = m.dt;
    st.metrics.max_cfl = m.max_cfl;
    st.metrics.max_velocity_change = m.max_velocity_change;
    st.metrics.divergence_l2 = m.divergence_l2;
    st.metrics.pressure_iterations = m.pressure_iterations;
    st.metrics.pressure_relative_residual = m.pressure_relative_residual;
  }

 private:
  of_stepper* s_ = nullptr;
};

// advance_lid_driven_cavity_impl (solver/lid_driven_cavity.cpp:524-624) on one device stepper
void advance(const LidDrivenCavityConfig& c, const int step_count, const bool stop_on_steady,
             LidDrivenCavityState& st) {
  validate_config(c);
  if(step_count < 0) throw std::invalid_argument("run_lid_driven_cavity_steps expects a non-negative step count");
  if(!same_grid(st.grid, make_lid_driven_cavity_grid(c)))
    throw std::invalid_argument("lid-driven cavity state grid does not match the config");
  if(step_count == 0) return;
  DeviceRun run(c);
  run.upload(st);
  if(stop_on_steady) {
    for(int s = 0; s < step_count; ++s) {
      const of_step_metrics m = run.run(1);
      if(m.step >= c.min_steps && m.max_velocity_change <= c.steady_tolerance) break;
    }
  } else {
    run.run(step_count);
  }
  run.download(st);
}

}  // namespace

LidDrivenCavityState::LidDrivenCavityState(const Grid& grid_in)
    : grid(grid_in), velocity(grid), advection_previous(grid), pressure_total(grid) {}

LidDrivenCavityConfig default_lid_driven_cavity_config() { return LidDrivenCavityConfig{}; }

LidDrivenCavityConfig load_lid_driven_cavity_config(const std::string& path) {
  std::ifstream in(path);
  if(!in.is_open()) throw std::runtime_error("unable to open cavity config: " + path);
  LidDrivenCavityConfig c = default_lid_driven_cavity_config();
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
    if(key == "nx") c.nx = std::stoi(value);
    else if(key == "ny") c.ny = std::stoi(value);
    else if(key == "reynolds") c.reynolds = std::stod(value);
    else if(key == "lid_velocity") c.lid_velocity = std::stod(value);
    else if(key == "cfl_limit") c.cfl_limit = std::stod(value);
    else if(key == "max_steps") c.max_steps = std::stoi(value);
    else if(key == "min_steps") c.min_steps = std::stoi(value);
    else if(key == "steady_tolerance") c.steady_tolerance = std::stod(value);
    else if(key == "poisson_max_iterations") c.poisson_max_iterations = std::stoi(value);
    else if(key == "poisson_tolerance") c.poisson_tolerance = std::stod(value);
    else if(key == "validate_reference") c.validate_reference = parse_bool(value);
    else throw std::runtime_error("unsupported cavity config key: " + key);
  }
  validate_config(c);
  return c;
}

std::string describe(const LidDrivenCavityConfig& c) {
  std::ostringstream s;
  s << "nx=" << c.nx << ", ny=" << c.ny << ", Re=" << c.reynolds << ", lid_velocity=" << c.lid_velocity
    << ", cfl_limit=" << c.cfl_limit << ", max_steps=" << c.max_steps << ", min_steps=" << c.min_steps
    << ", steady_tolerance=" << c.steady_tolerance << ", poisson_max_iterations=" << c.poisson_max_iterations
    << ", poisson_tolerance=" << c.poisson_tolerance
    << ", validate_reference=" << (c.validate_reference ? "true" : "false") << ", advection=" << describe(c.advection);
  return s.str();
}

std::string to_string(const CenterlineSampleKind line) {
  return line == CenterlineSampleKind::u_vertical ? "u_vertical" : "v_horizontal";
}

BoundaryConditionSet make_lid_driven_cavity_boundary_conditions(const LidDrivenCavityConfig& c) {
  BoundaryConditionSet b = BoundaryConditionSet::cavity();
  b[BoundaryFace::y_max].type = PhysicalBoundaryType::prescribed_velocity;
  b[BoundaryFace::y_max].velocity = {c.lid_velocity, 0.0, 0.0};
  b[BoundaryFace::z_min].type = PhysicalBoundaryType::symmetry;
  b[BoundaryFace::z_max].type = PhysicalBoundaryType::symmetry;
  return b;
}

Grid make_lid_driven_cavity_grid(const LidDrivenCavityConfig& c) {
  return Grid{c.nx, c.ny, 1, 1.0 / c.nx, 1.0 / c.ny, 1.0, 1};
}

double lid_driven_cavity_viscosity(const LidDrivenCavityConfig& c) { return c.lid_velocity / c.reynolds; }

double lid_driven_cavity_dt(const LidDrivenCavityConfig& c) {
  double dt = 0.0;
  hb::check(of_lid_driven_cavity_dt(c.nx, c.ny, c.lid_velocity, c.cfl_limit, &dt));
  return dt;
}

LidDrivenCavityState initialize_lid_driven_cavity_state(const LidDrivenCavityConfig& c) {
  validate_config(c);
  LidDrivenCavityState st{make_lid_driven_cavity_grid(c)};
  DeviceRun run(c);
  run.initialize();
  run.download(st);
  st.metrics = SimulationStepMetrics{};
  st.metrics.dt = lid_driven_cavity_dt(c);
  st.advection_previous.fill(0.0);
  st.has_previous_advection = false;
  return st;
}

void run_lid_driven_cavity_steps(const LidDrivenCavityConfig& c, const int step_count, LidDrivenCavityState& st) {
  advance(c, step_count, false, st);
}

// Published Re = 100 centreline values (the reference's table,
// solver/lid_driven_cavity.cpp:411-448): label, line, coordinate, value, lower, upper.
LidDrivenCavityReference re100_centerline_reference_envelope() {
  LidDrivenCavityReference r;
  r.dataset = "Re100_centerline_literature_envelope";
  r.points[0] = {"u_vertical_y0.9766", CenterlineSampleKind::u_vertical, 0.9766, 0.84123, 0.84123, 0.84123};
  r.points[1] = {"u_vertical_y0.4531", CenterlineSampleKind::u_vertical, 0.4531, -0.21090, -0.21400, -0.21060};
  r.points[2] = {"v_horizontal_x0.2344", CenterlineSampleKind::v_horizontal, 0.2344, 0.17527, 0.17527, 0.17960};
  r.points[3] = {"v_horizontal_x0.8047", CenterlineSampleKind::v_horizontal, 0.8047, -0.24533, -0.25400, -0.24533};
  return r;
}

LidDrivenCavityValidation validate_lid_driven_cavity_re100(const LidDrivenCavityResult& result) {
  const LidDrivenCavityReference ref = re100_centerline_reference_envelope();
  LidDrivenCavityValidation v;
  v.reference_dataset = ref.dataset;
  v.divergence_l2 = result.final_step.divergence_l2;
  double worst = 0.0;
  bool pass = result.final_step.divergence_l2 <= 1.0e-10;
  for(std::size_t n = 0; n < ref.points.size(); ++n) {
    const LidDrivenCavityReferencePoint& pt = ref.points[n];
    const CenterlineProfile& prof =
        pt.line == CenterlineSampleKind::u_vertical ? result.u_vertical_centerline : result.v_horizontal_centerline;
    const double sample = interpolate(prof, pt.coordinate);
    LidDrivenCavityValidationPoint& out = v.points[n];
    out.label = pt.label;
    out.line = pt.line;
    out.coordinate = pt.coordinate;
    out.reference_value = pt.value;
    out.reference_lower_bound = pt.lower_bound;
    out.reference_upper_bound = pt.upper_bound;
    out.sample_value = sample;
    out.relative_error = miss_against_envelope(sample, pt.lower_bound, pt.upper_bound);
    worst = std::max(worst, out.relative_error);
    pass = pass && out.relative_error <= 0.02;
  }
  v.max_relative_error = worst;
  v.pass = pass;
  return v;
}

LidDrivenCavityResult finalize_lid_driven_cavity_result(const LidDrivenCavityConfig& c,
                                                        const LidDrivenCavityState& st) {
  validate_config(c);
  if(!same_grid(st.grid, make_lid_driven_cavity_grid(c)))
    throw std::invalid_argument("lid-driven cavity state grid does not match the config");
  VelocityField vel = st.velocity;
  apply_velocity_boundary_conditions(make_lid_driven_cavity_boundary_conditions(c), vel);
  LidDrivenCavityResult r;
  r.config = c;
  r.final_step = st.metrics;
  r.u_vertical_centerline = vertical_u(vel);
  r.v_horizontal_centerline = horizontal_v(vel);
  r.extrema = extrema(r.u_vertical_centerline, r.v_horizontal_centerline);
  if(c.validate_reference && std::abs(c.reynolds - 100.0) < 1.0e-12) r.validation = validate_lid_driven_cavity_re100(r);
  return r;
}

void run_lid_driven_cavity_to_steady_state(const LidDrivenCavityConfig& c, LidDrivenCavityState& st) {
  advance(c, std::max(0, c.max_steps - st.metrics.step), true, st);
}

LidDrivenCavityResult run_lid_driven_cavity(const LidDrivenCavityConfig& c) {
  LidDrivenCavityState st = initialize_lid_driven_cavity_state(c);
  run_lid_driven_cavity_to_steady_state(c, st);
  return finalize_lid_driven_cavity_result(c, st);
}

}  // namespace solver
