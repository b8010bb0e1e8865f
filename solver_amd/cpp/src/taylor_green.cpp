// Taylor-Green driver of the C++ interface (reference solver/taylor_green.cpp:257-740): the
// .cfg surface, grid/BC/dt set-up and result validation on the host; the initial state and
// the step loop on the device (device-resident stepper, one upload and one download per call).
#include "solver/taylor_green.hpp"

#include "solver/hip_backend.hpp"

#include <algorithm>
#include <chrono>
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

void validate_config(const TaylorGreenConfig& c) {
  if(c.nx <= 0 || c.ny <= 0 || c.nz <= 0)
    throw std::invalid_argument("Taylor-Green config expects positive nx, ny, and nz");
  if(c.viscosity <= 0.0) throw std::invalid_argument("Taylor-Green config expects viscosity > 0");
  if(c.cfl_limit <= 0.0) throw std::invalid_argument("Taylor-Green config expects cfl_limit > 0");
  if(c.final_time <= 0.0) throw std::invalid_argument("Taylor-Green config expects final_time > 0");
  if(c.poisson_max_iterations <= 0 || c.poisson_tolerance <= 0.0)
    throw std::invalid_argument("Taylor-Green config expects positive Poisson controls");
}

bool three_d(const TaylorGreenConfig& c) { return c.nz > 1; }

double two_pi() { return 2.0 * std::acos(-1.0); }

// exact solution factors of the decaying vortex (solver/taylor_green.cpp:257-266)
double exact_decay(const TaylorGreenConfig& c, const double t) {
  return std::exp(-(three_d(c) ? 3.0 : 2.0) * c.viscosity * t);
}
double exact_kinetic_energy(const TaylorGreenConfig& c, const double t) {
  return (three_d(c) ? 0.125 : 0.25) * std::exp(-(three_d(c) ? 6.0 : 4.0) * c.viscosity * t);
}

// mean over cells of 0.5 |u_c|^2 with face-to-centre averages (solver/taylor_green.cpp:269-288)
double kinetic_energy(const VelocityField& vel) {
  const IndexRange3D r = FieldLayout::cell_centered(vel.x.layout().grid()).active_range();
  double sum = 0.0;
  std::size_t count = 0;
  for(int k = r.k_begin; k < r.k_end; ++k)
    for(int j = r.j_begin; j < r.j_end; ++j)
      for(int i = r.i_begin; i < r.i_end; ++i) {
        const double uc = 0.5 * (vel.x(i, j, k) + vel.x(i + 1, j, k));
        const double vc = 0.5 * (vel.y(i, j, k) + vel.y(i, j + 1, k));
        const double wc = 0.5 * (vel.z(i, j, k) + vel.z(i, j, k + 1));
        sum += 0.5 * (uc * uc + vc * vc + wc * wc);
        ++count;
      }
  return sum / static_cast<double>(count);
}

// relative L2 distance to the analytic field at time t (solver/taylor_green.cpp:290-333)
double velocity_relative_l2_error(const VelocityField& vel, const TaylorGreenConfig& c, const double t) {
  const double decay = exact_decay(c, t);
  const bool d3 = three_d(c);
  double num = 0.0, den = 0.0;
  auto add = [&](const StructuredField& f, const int comp) {
    const FieldLayout& L = f.layout();
    const IndexRange3D r = L.active_range();
    for(int k = r.k_begin; k < r.k_end; ++k) {
      const double z = L.coordinate_for_storage_index(Axis::z, k);
      const double zf = d3 ? std::cos(z) : 1.0;
      for(int j = r.j_begin; j < r.j_end; ++j) {
        const double y = L.coordinate_for_storage_index(Axis::y, j);
        for(int i = r.i_begin; i < r.i_end; ++i) {
          const double x = L.coordinate_for_storage_index(Axis::x, i);
          double ex = 0.0;
          if(comp == 0) ex = -std::cos(x) * std::sin(y) * zf * decay;
          if(comp == 1) ex = std::sin(x) * std::cos(y) * zf * decay;
          const double e = f(i, j, k) - ex;
          num += e * e;
          den += ex * ex;
        }
      }
    }
  };
  add(vel.x, 0);
  add(vel.y, 1);
  add(vel.z, 2);
  return den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
}

TaylorGreenValidation build_validation(const TaylorGreenConfig& c, const TaylorGreenStepMetrics& m,
                                       const VelocityField& vel) {
  TaylorGreenValidation v;
  const double exact = exact_kinetic_energy(c, m.time);
  v.reference_dataset = three_d(c) ? "analytic_taylor_green_decay_3d" : "analytic_taylor_green_decay_2d";
  v.normalized_energy_error = std::abs(kinetic_energy(vel) - exact) / exact;
  v.velocity_relative_l2_error = velocity_relative_l2_error(vel, c, m.time);
  v.exact_kinetic_energy = exact;
  v.pass = v.normalized_energy_error <= 1.0e-2 && m.divergence_l2 <= 1.0e-10;
  return v;
}

of_step_metrics to_abi(const TaylorGreenStepMetrics& m) {
  of_step_metrics o{};
  o.step = m.step;
  o.time = m.time;
  o.dt = m.dt;
  o.max_cfl = m.max_cfl;
  o.max_velocity_change = m.max_velocity_change;
  o.divergence_l2 = m.divergence_l2;
  o.max_divergence_l2 = m.max_divergence_l2;
  o.pressure_iterations = m.pressure_iterations;
  o.pressure_relative_residual = m.pressure_relative_residual;
  return o;
}

TaylorGreenStepMetrics from_abi(const of_step_metrics& o) {
  TaylorGreenStepMetrics m;
  m.step = o.step;
  m.time = o.time;
  m.dt = o.dt;
  m.max_cfl = o.max_cfl;
  m.max_velocity_change = o.max_velocity_change;
  m.divergence_l2 = o.divergence_l2;
  m.max_divergence_l2 = o.max_divergence_l2;
  m.pressure_iterations = o.pressure_iterations;
  m.pressure_relative_residual = o.pressure_relative_residual;
  return m;
}

// Device-resident stepper for one Taylor-Green configuration.
class DeviceRun {
 public:
  explicit DeviceRun(const TaylorGreenConfig& c) {
    of_stepper_config sc{};
    sc.flow_case = OF_CASE_TAYLOR_GREEN;
    sc.nx = c.nx;
    sc.ny = c.ny;
    sc.nz = c.nz;
    sc.viscosity = c.viscosity;
    sc.dt = taylor_green_dt(c);
    sc.scheme = static_cast<int>(c.advection.scheme);
    sc.poisson_max_iterations = c.poisson_max_iterations;
    sc.poisson_tolerance = c.poisson_tolerance;
    sc.precision = c.precision == SolvePrecision::mixed ? OF_PRECISION_MIXED : OF_PRECISION_FP64;
    hb::check(of_stepper_create(hb::context(), &sc, &s_));
  }
  DeviceRun(const DeviceRun&) = delete;
  DeviceRun& operator=(const DeviceRun&) = delete;
  ~DeviceRun() {
    if(s_ != nullptr) (void)of_stepper_destroy(s_);
  }
  void initialize() { hb::check(of_stepper_initialize(s_, 0)); }
  void upload(const TaylorGreenState& st) {
    const of_step_metrics m = to_abi(st.metrics);
    hb::check(of_stepper_upload(s_, st.velocity.x.data(), st.velocity.y.data(), st.velocity.z.data(),
                                st.pressure_total.data(), st.advection_previous.x.data(),
                                st.advection_previous.y.data(), st.advection_previous.z.data(),
                                st.has_previous_advection ? 1 : 0, &m));
  }
  void run(const int steps) {
    of_step_metrics m{};
    hb::check(of_stepper_run(s_, steps, &m));
  }
  void download(TaylorGreenState& st) {
    int has_prev = 0;
    of_step_metrics m{};
    hb::check(of_stepper_download(s_, st.velocity.x.data(), st.velocity.y.data(), st.velocity.z.data(),
                                  st.pressure_total.data(), st.advection_previous.x.data(),
                                  st.advection_previous.y.data(), st.advection_previous.z.data(), &has_prev, &m));
    st.has_previous_advection = has_prev != 0;
    st.metrics = from_abi(m);
  }

 private:
  of_stepper* s_ = nullptr;
};

void require_hip_backend(const TaylorGreenConfig& c) {
  if(c.backend != ExecutionBackend::hip)
    throw std::runtime_error("Taylor-Green backend " + to_string(c.backend) +
                             " is not available in this build; backend = hip runs the step loop on the GPU");
}

}  // namespace

std::string to_string(const ExecutionBackend backend) {
  switch(backend) {
    case ExecutionBackend::cpu: return "cpu";
    case ExecutionBackend::metal: return "metal";
    case ExecutionBackend::hip: return "hip";
  }
  return "unknown";
}

ExecutionBackend parse_execution_backend(const std::string& value) {
  if(value == "cpu") return ExecutionBackend::cpu;
  if(value == "metal") return ExecutionBackend::metal;
  if(value == "hip") return ExecutionBackend::hip;
  throw std::runtime_error("invalid Taylor-Green backend value: " + value);
}

std::string to_string(const SolvePrecision precision) { return precision == SolvePrecision::mixed ? "mixed" : "fp64"; }

SolvePrecision parse_solve_precision(const std::string& value) {
  if(value == "fp64") return SolvePrecision::fp64;
  if(value == "mixed") return SolvePrecision::mixed;
  throw std::runtime_error("invalid Taylor-Green precision value: " + value);
}

TaylorGreenConfig default_taylor_green_config() { return TaylorGreenConfig{}; }

TaylorGreenConfig load_taylor_green_config(const std::string& path) {
  std::ifstream in(path);
  if(!in.is_open()) throw std::runtime_error("unable to open Taylor-Green config: " + path);
  TaylorGreenConfig c = default_taylor_green_config();
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
    else if(key == "nz") c.nz = std::stoi(value);
    else if(key == "viscosity") c.viscosity = std::stod(value);
    else if(key == "cfl_limit") c.cfl_limit = std::stod(value);
    else if(key == "final_time") c.final_time = std::stod(value);
    else if(key == "poisson_max_iterations") c.poisson_max_iterations = std::stoi(value);
    else if(key == "poisson_tolerance") c.poisson_tolerance = std::stod(value);
    else if(key == "validate_energy") c.validate_energy = parse_bool(value);
    else if(key == "backend") c.backend = parse_execution_backend(value);
    else if(key == "precision") c.precision = parse_solve_precision(value);
    else throw std::runtime_error("unsupported Taylor-Green config key: " + key);
  }
  validate_config(c);
  return c;
}

std::string describe(const TaylorGreenConfig& c) {
  std::ostringstream s;
  s << "nx=" << c.nx << ", ny=" << c.ny << ", nz=" << c.nz << ", viscosity=" << c.viscosity
    << ", cfl_limit=" << c.cfl_limit << ", final_time=" << c.final_time
    << ", poisson_max_iterations=" << c.poisson_max_iterations << ", poisson_tolerance=" << c.poisson_tolerance
    << ", validate_energy=" << (c.validate_energy ? "true" : "false") << ", backend=" << to_string(c.backend)
    << ", advection=" << describe(c.advection) << ", precision=" << to_string(c.precision);
  return s.str();
}

Grid make_taylor_green_grid(const TaylorGreenConfig& c) {
  validate_config(c);
  const double L = two_pi();
  return Grid{c.nx, c.ny, c.nz, L / c.nx, L / c.ny, three_d(c) ? L / c.nz : 1.0, 1};
}

BoundaryConditionSet make_taylor_green_boundary_conditions(const TaylorGreenConfig& c) {
  BoundaryConditionSet b = BoundaryConditionSet::all(PhysicalBoundaryType::periodic);
  if(!three_d(c)) {
    b[BoundaryFace::z_min].type = PhysicalBoundaryType::symmetry;
    b[BoundaryFace::z_max].type = PhysicalBoundaryType::symmetry;
  }
  return b;
}

double taylor_green_dt(const TaylorGreenConfig& c) {
  validate_config(c);
  double dt = 0.0;
  int steps = 0;
  hb::check(of_taylor_green_dt(c.nx, c.ny, c.nz, c.cfl_limit, c.final_time, &dt, &steps));
  return dt;
}

TaylorGreenState initialize_taylor_green_state(const TaylorGreenConfig& c) {
  validate_config(c);
  require_hip_backend(c);
  TaylorGreenState st{make_taylor_green_grid(c)};
  DeviceRun run(c);
  run.initialize();
  run.download(st);
  st.metrics = TaylorGreenStepMetrics{};
  st.metrics.dt = taylor_green_dt(c);
  st.advection_previous.fill(0.0);
  st.has_previous_advection = false;
  return st;
}

void run_taylor_green_steps(const TaylorGreenConfig& c, const int step_count, TaylorGreenState& st) {
  validate_config(c);
  if(step_count < 0) throw std::invalid_argument("run_taylor_green_steps expects a non-negative step count");
  require_hip_backend(c);
  const Grid g = make_taylor_green_grid(c);
  if(st.grid.nx != g.nx || st.grid.ny != g.ny || st.grid.nz != g.nz || st.grid.dx != g.dx || st.grid.dy != g.dy ||
     st.grid.dz != g.dz)
    throw std::invalid_argument("Taylor-Green state grid does not match the config");
  if(step_count == 0) return;
  DeviceRun run(c);
  run.upload(st);
  run.run(step_count);
  run.download(st);
}

TaylorGreenResult finalize_taylor_green_result(const TaylorGreenConfig& c, const TaylorGreenState& st) {
  validate_config(c);
  TaylorGreenResult r;
  r.config = c;
  r.backend_used = c.backend;
  r.final_step = st.metrics;
  r.initial_kinetic_energy = exact_kinetic_energy(c, 0.0);
  r.final_kinetic_energy = kinetic_energy(st.velocity);
  if(c.validate_energy) r.validation = build_validation(c, st.metrics, st.velocity);
  return r;
}

TaylorGreenResult run_taylor_green(const TaylorGreenConfig& c, TaylorGreenState* final_state) {
  validate_config(c);
  require_hip_backend(c);
  TaylorGreenState st{make_taylor_green_grid(c)};
  const double dt = taylor_green_dt(c);
  const int steps = static_cast<int>(std::ceil(c.final_time / dt));
  const auto t0 = std::chrono::steady_clock::now();
  {
    DeviceRun run(c);
    run.initialize();
    run.run(steps);
    run.download(st);
  }
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  TaylorGreenResult r = finalize_taylor_green_result(c, st);
  r.backend_used = ExecutionBackend::hip;
  r.accelerator_name = hb::device_name();
  r.backend_elapsed_seconds = elapsed;
  r.cleanup_elapsed_seconds = 0.0;  // FP64 on the device: no CPU clean-up projection is needed
  if(final_state != nullptr) *final_state = st;
  return r;
}

TaylorGreenResult run_taylor_green(const TaylorGreenConfig& c) { return run_taylor_green(c, nullptr); }

}  // namespace solver
