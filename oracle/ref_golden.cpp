// ref_golden: command-line driver around the reference's own public C++ API.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile together with the reference's CPU
// sources as they lie under the reference checkout (nothing is copied). It never leaves the
// test harness: tests/golden/make_golden.py runs it to write the committed fixtures, and
// bench.py may time it as the reference CPU baseline.
//
// Protocol: ref_golden <command> key=value ...
//   in=<file>   raw little-endian float64 input arrays, concatenated in the documented order
//   out=<file>  raw float64 output arrays, concatenated in the documented order
//   meta=<file> "key: value" lines (doubles printed with 17 significant digits)
// Grid keys: nx ny nz dx dy dz. Velocity BC keys: bc=t0,..,t5 bcvel=18 numbers bcp=6 numbers.
// Pressure BC keys: pbc=t0,..,t5 pbcval=6 numbers pbcgrad=6 numbers.

#include "io/checkpoint.hpp"
#include "io/vtk_export.hpp"
#include "linsolve/poisson_solver.hpp"
#include "operators/discrete_operators.hpp"
#include "solver/channel_flow.hpp"
#include "solver/lid_driven_cavity.hpp"
#include "solver/momentum_terms.hpp"
#include "solver/operator_verification.hpp"
#include "solver/projection.hpp"
#include "solver/taylor_green.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using Args = std::map<std::string, std::string>;

std::vector<double> parse_list(const std::string& s) {
  std::vector<double> v;
  std::stringstream ss(s);
  std::string item;
  while(std::getline(ss, item, ',')) v.push_back(std::stod(item));
  return v;
}

const std::string& need(const Args& a, const char* k) {
  auto it = a.find(k);
  if(it == a.end()) throw std::runtime_error(std::string("missing key ") + k);
  return it->second;
}
double num(const Args& a, const char* k) { return std::stod(need(a, k)); }
int inum(const Args& a, const char* k) { return std::stoi(need(a, k)); }
double num_or(const Args& a, const char* k, double d) { return a.count(k) ? std::stod(a.at(k)) : d; }

struct Reader {
  std::vector<double> data;
  size_t pos = 0;
  explicit Reader(const Args& a) {
    if(!a.count("in")) return;
    std::ifstream f(a.at("in"), std::ios::binary);
    f.seekg(0, std::ios::end);
    const size_t n = (size_t)f.tellg() / sizeof(double);
    f.seekg(0);
    data.resize(n);
    f.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(n * sizeof(double)));
  }
  void into(solver::StructuredField& fld) {
    if(pos + fld.size() > data.size()) throw std::runtime_error("input file too short");
    std::memcpy(fld.data(), data.data() + pos, fld.size() * sizeof(double));
    pos += fld.size();
  }
  void into(solver::VelocityField& v) { into(v.x); into(v.y); into(v.z); }
};

struct Writer {
  std::ofstream out;
  std::ofstream meta;
  explicit Writer(const Args& a) {
    if(a.count("out")) out.open(a.at("out"), std::ios::binary);
    if(a.count("meta")) meta.open(a.at("meta"));
    meta.precision(17);
  }
  void put(const solver::StructuredField& f) {
    out.write(reinterpret_cast<const char*>(f.data()), (std::streamsize)(f.size() * sizeof(double)));
  }
  void put(const solver::VelocityField& v) { put(v.x); put(v.y); put(v.z); }
  template <typename T>
  void kv(const char* k, const T& v) { meta << k << ": " << v << '\n'; }
};

solver::Grid grid_from(const Args& a) {
  return solver::Grid{inum(a, "nx"), inum(a, "ny"), inum(a, "nz"), num(a, "dx"), num(a, "dy"), num(a, "dz"), 1};
}

solver::BoundaryConditionSet bcs_from(const Args& a) {
  solver::BoundaryConditionSet b;
  const std::vector<double> t = parse_list(need(a, "bc"));
  const std::vector<double> vel = a.count("bcvel") ? parse_list(a.at("bcvel")) : std::vector<double>(18, 0.0);
  const std::vector<double> pr = a.count("bcp") ? parse_list(a.at("bcp")) : std::vector<double>(6, 0.0);
  for(int f = 0; f < 6; ++f) {
    b.faces[f].type = static_cast<solver::PhysicalBoundaryType>((int)t[f]);
    for(int c = 0; c < 3; ++c) b.faces[f].velocity[c] = vel[3 * f + c];
    b.faces[f].pressure = pr[f];
  }
  return b;
}

solver::PressureBoundarySet pbcs_from(const Args& a) {
  solver::PressureBoundarySet b;
  const std::vector<double> t = parse_list(need(a, "pbc"));
  const std::vector<double> val = a.count("pbcval") ? parse_list(a.at("pbcval")) : std::vector<double>(6, 0.0);
  const std::vector<double> gr = a.count("pbcgrad") ? parse_list(a.at("pbcgrad")) : std::vector<double>(6, 0.0);
  for(int f = 0; f < 6; ++f) {
    b.faces[f].type = static_cast<solver::PressureBoundaryType>((int)t[f]);
    b.faces[f].value = val[f];
    b.faces[f].gradient = gr[f];
  }
  return b;
}

solver::ProjectionOptions opts_from(const Args& a) {
  return solver::ProjectionOptions{num_or(a, "dt", 1.0), num_or(a, "density", 1.0),
                                   (int)num_or(a, "maxit", 4000), num_or(a, "tol", 1e-12)};
}

void put_poisson(Writer& w, const solver::PoissonSolveDiagnostics& d, const char* prefix = "") {
  const std::string p(prefix);
  w.kv((p + "iterations").c_str(), d.iterations);
  w.kv((p + "initial_residual_l2").c_str(), d.initial_residual_l2);
  w.kv((p + "final_residual_l2").c_str(), d.final_residual_l2);
  w.kv((p + "relative_residual").c_str(), d.relative_residual);
  w.kv((p + "removed_rhs_mean").c_str(), d.removed_rhs_mean);
  w.kv((p + "multigrid_levels").c_str(), d.multigrid_levels);
  w.kv((p + "coarse_unknowns").c_str(), d.coarse_unknowns);
  w.kv((p + "converged").c_str(), d.converged ? 1 : 0);
  w.kv((p + "zero_mean_enforced").c_str(), d.zero_mean_enforced ? 1 : 0);
  std::ostringstream h;
  h.precision(17);
  for(size_t i = 0; i < d.residual_history.size(); ++i) h << (i ? "," : "") << d.residual_history[i];
  w.kv((p + "residual_history").c_str(), h.str());
}

solver::TaylorGreenConfig tg_config_from(const Args& a) {
  solver::TaylorGreenConfig c;
  c.nx = inum(a, "nx");
  c.ny = inum(a, "ny");
  c.nz = inum(a, "nz");
  c.viscosity = num(a, "viscosity");
  c.cfl_limit = num(a, "cfl");
  c.final_time = num(a, "final_time");
  c.poisson_tolerance = num(a, "tol");
  c.poisson_max_iterations = inum(a, "maxit");
  c.validate_energy = false;
  return c;
}

solver::LidDrivenCavityConfig cavity_config_from(const Args& a) {
  solver::LidDrivenCavityConfig c;
  c.nx = inum(a, "nx");
  c.ny = inum(a, "ny");
  c.reynolds = num(a, "reynolds");
  c.lid_velocity = num(a, "lid");
  c.cfl_limit = num(a, "cfl");
  c.max_steps = (int)num_or(a, "max_steps", 8000);
  c.min_steps = (int)num_or(a, "min_steps", 0);
  c.steady_tolerance = num_or(a, "steady_tolerance", 1e-8);
  c.poisson_tolerance = num(a, "tol");
  c.poisson_max_iterations = inum(a, "maxit");
  c.validate_reference = num_or(a, "validate", 0) != 0;
  return c;
}

template <typename M>
void put_metrics(Writer& w, const M& m) {
  w.kv("step", m.step);
  w.kv("time", m.time);
  w.kv("dt", m.dt);
  w.kv("max_cfl", m.max_cfl);
  w.kv("max_velocity_change", m.max_velocity_change);
  w.kv("divergence_l2", m.divergence_l2);
  w.kv("pressure_iterations", m.pressure_iterations);
  w.kv("pressure_relative_residual", m.pressure_relative_residual);
}

int run(const std::string& cmd, const Args& a) {
  Reader in(a);
  Writer w(a);
  if(cmd == "tg_init" || cmd == "tg_steps") {
    const solver::TaylorGreenConfig c = tg_config_from(a);
    solver::TaylorGreenState s = solver::initialize_taylor_green_state(c);
    if(cmd == "tg_steps" && a.count("in")) {  // state injection through the public struct
      in.into(s.velocity);
      in.into(s.pressure_total);
      in.into(s.advection_previous);
      s.has_previous_advection = inum(a, "has_prev") != 0;
    }
    w.kv("dt", solver::taylor_green_dt(c));
    if(cmd == "tg_steps") {
      const int steps = inum(a, "steps");
      const auto t0 = std::chrono::steady_clock::now();
      try {
        solver::run_taylor_green_steps(c, steps, s);
      } catch(const std::exception& e) {
        w.kv("error", e.what());
        return 0;
      }
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      w.kv("elapsed_seconds", el);
      put_metrics(w, s.metrics);
      w.kv("max_divergence_l2", s.metrics.max_divergence_l2);
      w.kv("has_prev", s.has_previous_advection ? 1 : 0);
      const solver::TaylorGreenResult r = solver::finalize_taylor_green_result(c, s);
      w.kv("final_kinetic_energy", r.final_kinetic_energy);
    }
    w.put(s.velocity);
    w.put(s.pressure_total);
    w.put(s.advection_previous);
    return 0;
  }
  if(cmd == "cavity_init" || cmd == "cavity_steps") {
    const solver::LidDrivenCavityConfig c = cavity_config_from(a);
    solver::LidDrivenCavityState s = solver::initialize_lid_driven_cavity_state(c);
    if(cmd == "cavity_steps" && a.count("in")) {
      in.into(s.velocity);
      in.into(s.pressure_total);
      in.into(s.advection_previous);
      s.has_previous_advection = inum(a, "has_prev") != 0;
      s.metrics.step = (int)num_or(a, "step0", 0);
    }
    w.kv("dt", solver::lid_driven_cavity_dt(c));
    if(cmd == "cavity_steps") {
      try {
        solver::run_lid_driven_cavity_steps(c, inum(a, "steps"), s);
      } catch(const std::exception& e) {
        w.kv("error", e.what());
        return 0;
      }
      put_metrics(w, s.metrics);
      const solver::LidDrivenCavityResult r = solver::finalize_lid_driven_cavity_result(c, s);
      w.kv("u_vertical_max", r.extrema.u_vertical_max);
      w.kv("u_vertical_min", r.extrema.u_vertical_min);
      w.kv("v_horizontal_max", r.extrema.v_horizontal_max);
      w.kv("v_horizontal_min", r.extrema.v_horizontal_min);
    }
    if(a.count("ckpt")) solver::io::write_lid_driven_cavity_checkpoint(a.at("ckpt"), c, s);
    if(a.count("vtk")) solver::io::write_lid_driven_cavity_vtk(a.at("vtk"), s);
    w.put(s.velocity);
    w.put(s.pressure_total);
    w.put(s.advection_previous);
    return 0;
  }
  if(cmd == "cavity_run") {  // to steady state (or max_steps); out: u v w p_total A_prev(u v w)
    const solver::LidDrivenCavityConfig c = cavity_config_from(a);
    solver::LidDrivenCavityState s = solver::initialize_lid_driven_cavity_state(c);
    w.kv("dt", solver::lid_driven_cavity_dt(c));
    bool steady = false;
    try {
      // one step per call: the public API's step loop with the steady test of run_lid_driven_cavity
      while(s.metrics.step < c.max_steps) {
        solver::run_lid_driven_cavity_steps(c, 1, s);
        if(s.metrics.step >= c.min_steps && s.metrics.max_velocity_change <= c.steady_tolerance) {
          steady = true;
          break;
        }
      }
    } catch(const std::exception& e) {
      // one step per call numbers the failing step 1 in the message; run_lid_driven_cavity's
      // single loop numbers it by the state's step count, which is restored here
      std::string msg = e.what();
      const std::string local = "at step 1 (";
      const size_t at = msg.find(local);
      if(at != std::string::npos)
        msg.replace(at, local.size(), "at step " + std::to_string(s.metrics.step + 1) + " (");
      w.kv("error", msg);
    }
    w.kv("steady", steady ? 1 : 0);
    put_metrics(w, s.metrics);
    const solver::LidDrivenCavityResult r = solver::finalize_lid_driven_cavity_result(c, s);
    w.kv("u_vertical_max", r.extrema.u_vertical_max);
    w.kv("u_vertical_min", r.extrema.u_vertical_min);
    w.kv("v_horizontal_max", r.extrema.v_horizontal_max);
    w.kv("v_horizontal_min", r.extrema.v_horizontal_min);
    w.kv("validation_pass", r.validation.pass ? 1 : 0);
    w.kv("validation_max_relative_error", r.validation.max_relative_error);
    for(size_t k = 0; k < r.validation.points.size(); ++k)
      w.kv(("sample_" + std::to_string(k)).c_str(), r.validation.points[k].sample_value);
    w.put(s.velocity);
    w.put(s.pressure_total);
    w.put(s.advection_previous);
    return 0;
  }
  if(cmd == "channel") {  // out: streamwise profile (ny values)
    solver::ChannelFlowConfig c;
    c.case_kind = need(a, "case") == "poiseuille" ? solver::ChannelFlowCase::poiseuille
                                                  : solver::ChannelFlowCase::couette;
    c.nx = inum(a, "nx");
    c.ny = inum(a, "ny");
    c.viscosity = num(a, "viscosity");
    c.top_velocity = num(a, "top");
    c.pressure_drop = num(a, "drop");
    c.cfl_limit = num(a, "cfl");
    c.steps = inum(a, "steps");
    c.poisson_max_iterations = inum(a, "maxit");
    c.poisson_tolerance = num(a, "tol");
    c.validate_profile = true;
    solver::ChannelFlowResult r;
    try {
      r = solver::run_channel_flow(c);
    } catch(const std::exception& e) {
      w.kv("error", e.what());
      return 0;
    }
    put_metrics(w, r.final_step);
    w.kv("reference_dataset", r.validation.reference_dataset);
    w.kv("relative_l2_error", r.validation.relative_l2_error);
    w.kv("pass", r.validation.pass ? 1 : 0);
    w.out.write(reinterpret_cast<const char*>(r.streamwise_profile.value.data()),
                (std::streamsize)(r.streamwise_profile.value.size() * sizeof(double)));
    return 0;
  }
  if(cmd == "mms") {
    const solver::OperatorManufacturedSolutionResult r = solver::run_operator_manufactured_solution_case(inum(a, "n"));
    w.kv("gradient_error", r.gradient_error);
    w.kv("divergence_error", r.divergence_error);
    w.kv("laplacian_error", r.laplacian_error);
    return 0;
  }

  const solver::Grid g = grid_from(a);
  if(cmd == "operators") {  // in: p u v w ; out: gx gy gz div lap_p lap_u lap_v lap_w
    solver::PressureField p{g};
    solver::VelocityField v{g}, gr{g}, lv{g};
    solver::ScalarField div{g};
    solver::PressureField lp{g};
    in.into(p);
    in.into(v);
    solver::operators::compute_gradient(p, gr);
    solver::operators::compute_divergence(v, div);
    solver::operators::compute_laplacian(p, lp);
    solver::operators::compute_laplacian(v, lv);
    w.put(gr);
    w.put(div);
    w.put(lp);
    w.put(lv);
    return 0;
  }
  if(cmd == "advection") {  // in: u v w ; out: A
    solver::VelocityField v{g}, adv{g};
    in.into(v);
    solver::AdvectionOptions o;
    o.scheme = static_cast<solver::AdvectionScheme>((int)num_or(a, "scheme", 0));
    solver::compute_advection_term(v, o, adv);
    w.put(adv);
    return 0;
  }
  if(cmd == "diffusion") {
    solver::VelocityField v{g}, d{g};
    in.into(v);
    solver::compute_diffusion_term(v, num(a, "nu"), d);
    w.put(d);
    return 0;
  }
  if(cmd == "cfl") {
    solver::VelocityField v{g};
    in.into(v);
    const solver::CflDiagnostics d = solver::compute_advective_cfl(v, num(a, "dt"));
    w.kv("dt", d.dt);
    w.kv("max_cfl", d.max_cfl);
    w.kv("max_u", d.max_u);
    w.kv("max_v", d.max_v);
    w.kv("max_w", d.max_w);
    return 0;
  }
  if(cmd == "vbc") {
    solver::VelocityField v{g};
    in.into(v);
    solver::apply_velocity_boundary_conditions(bcs_from(a), v);
    w.put(v);
    return 0;
  }
  if(cmd == "pbc") {
    solver::PressureField p{g};
    in.into(p);
    solver::apply_pressure_boundary_conditions(pbcs_from(a), p);
    w.put(p);
    return 0;
  }
  if(cmd == "tpbc") {  // in: su sv sw p
    solver::VelocityField s{g};
    solver::PressureField p{g};
    in.into(s);
    in.into(p);
    solver::apply_total_pressure_boundary_conditions(bcs_from(a), s, p);
    w.put(p);
    return 0;
  }
  if(cmd == "cavity_rhs") {  // in: cur(u v w) p prev(u v w) ; out: adv rhs
    solver::VelocityField cur{g}, prev{g}, adv{g}, dif{g}, grad{g}, fc{g}, rhs{g};
    solver::PressureField p{g};
    in.into(cur);
    in.into(p);
    in.into(prev);
    solver::detail::assemble_lid_driven_cavity_predictor_rhs(cur, p, inum(a, "has_prev") ? &prev : nullptr, {},
                                                             num(a, "nu"), num(a, "dt"), adv, dif, grad, fc, rhs);
    w.put(adv);
    w.put(rhs);
    return 0;
  }
  if(cmd == "adi") {  // in: rhs ; out: predicted
    solver::VelocityField rhs{g}, pred{g};
    in.into(rhs);
    const solver::HelmholtzDiagnostics d = solver::solve_predictor_adi(rhs, num(a, "alpha"), bcs_from(a), pred);
    w.kv("line_solves", d.line_solves);
    w.put(pred);
    return 0;
  }
  if(cmd == "bprhs") {
    solver::VelocityField v{g};
    solver::ScalarField rhs{g};
    in.into(v);
    solver::build_pressure_rhs(v, bcs_from(a), opts_from(a), rhs);
    w.put(rhs);
    return 0;
  }
  if(cmd == "correct") {  // in: u v w phi ; out: corrected
    solver::VelocityField v{g}, c{g};
    solver::PressureField phi{g};
    in.into(v);
    in.into(phi);
    solver::correct_velocity(v, phi, bcs_from(a), opts_from(a), c);
    w.put(c);
    return 0;
  }
  if(cmd == "project") {  // in: u v w phi ; out: phi corrected prhs
    solver::VelocityField v{g}, c{g};
    solver::PressureField phi{g};
    solver::ScalarField prhs{g};
    in.into(v);
    in.into(phi);
    const solver::ProjectionDiagnostics d = solver::project_velocity(v, bcs_from(a), opts_from(a), phi, c, &prhs);
    w.kv("rhs_l2", d.rhs_l2);
    w.kv("divergence_l2_before", d.divergence_l2_before);
    w.kv("divergence_l2_after", d.divergence_l2_after);
    w.kv("pressure_mean", d.pressure_mean);
    put_poisson(w, d.pressure_solve);
    w.put(phi);
    w.put(c);
    w.put(prhs);
    return 0;
  }
  if(cmd == "poisson") {  // in: rhs p ; out: p
    solver::ScalarField rhs{g};
    solver::PressureField p{g};
    in.into(rhs);
    in.into(p);
    const solver::PoissonSolveDiagnostics d = solver::linsolve::solve_pressure_poisson(rhs, pbcs_from(a), opts_from(a), p);
    put_poisson(w, d);
    w.put(p);
    return 0;
  }
  if(cmd == "poisson_rhs") {  // in: p ; out: rhs
    solver::PressureField p{g};
    solver::ScalarField rhs{g};
    in.into(p);
    solver::linsolve::build_poisson_rhs_from_pressure(p, pbcs_from(a), rhs);
    w.put(rhs);
    return 0;
  }
  throw std::runtime_error("unknown command " + cmd);
}

}  // namespace

int main(int argc, char** argv) {
  if(argc < 2) {
    std::cerr << "usage: ref_golden <command> key=value...\n";
    return 2;
  }
  Args a;
  for(int i = 2; i < argc; ++i) {
    const std::string s = argv[i];
    const size_t eq = s.find('=');
    if(eq == std::string::npos) {
      std::cerr << "bad argument " << s << '\n';
      return 2;
    }
    a[s.substr(0, eq)] = s.substr(eq + 1);
  }
  try {
    return run(argv[1], a);
  } catch(const std::exception& e) {
    std::cerr << "ref_golden failed: " << e.what() << '\n';
    return 1;
  }
}
