// solver_cavity: the reference's tools/lid_driven_cavity_main.cpp command line and output
// (--steps, --restart <checkpoint>, --checkpoint-out <path>, --vtk-out <path>), with the step
// loop on the GPU.
#include "io/checkpoint.hpp"
#include "io/vtk_export.hpp"
#include "solver/lid_driven_cavity.hpp"

#include <iostream>
#include <optional>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
  try {
    std::string config_path = "solver_amd/cpp/configs/lid_driven_cavity_re100_128.cfg";
    std::optional<int> steps;
    std::string restart_path, checkpoint_out, vtk_out;
    bool have_config = false;
    for(int a = 1; a < argc; ++a) {
      const std::string arg = argv[a];
      auto value = [&]() -> std::string {
        if(a + 1 >= argc) throw std::runtime_error("missing value for " + arg);
        return argv[++a];
      };
      if(arg == "--steps") {
        steps = std::stoi(value());
      } else if(arg == "--restart") {
        restart_path = value();
      } else if(arg == "--checkpoint-out") {
        checkpoint_out = value();
      } else if(arg == "--vtk-out") {
        vtk_out = value();
      } else if(!have_config && !arg.empty() && arg.rfind("--", 0) != 0) {
        config_path = arg;
        have_config = true;
      } else {
        throw std::runtime_error("unsupported argument: " + arg);
      }
    }
    const solver::LidDrivenCavityConfig config = solver::load_lid_driven_cavity_config(config_path);
    solver::LidDrivenCavityState state = restart_path.empty()
                                             ? solver::initialize_lid_driven_cavity_state(config)
                                             : solver::io::load_lid_driven_cavity_checkpoint(restart_path, config).state;
    if(steps) {
      solver::run_lid_driven_cavity_steps(config, *steps, state);
    } else {
      // steps until min_steps have run and max |u_new - u_old| <= steady_tolerance, or max_steps
      solver::run_lid_driven_cavity_to_steady_state(config, state);
    }
    if(!checkpoint_out.empty()) solver::io::write_lid_driven_cavity_checkpoint(checkpoint_out, config, state);
    if(!vtk_out.empty()) solver::io::write_lid_driven_cavity_vtk(vtk_out, state);
    const solver::LidDrivenCavityResult r = solver::finalize_lid_driven_cavity_result(config, state);
    std::cout << "config_path: " << config_path << '\n';
    std::cout << "config: " << solver::describe(config) << '\n';
    std::cout << "steps: " << r.final_step.step << '\n';
    std::cout << "time: " << r.final_step.time << '\n';
    std::cout << "dt: " << r.final_step.dt << '\n';
    std::cout << "max_cfl: " << r.final_step.max_cfl << '\n';
    std::cout << "max_velocity_change: " << r.final_step.max_velocity_change << '\n';
    std::cout << "divergence_l2: " << r.final_step.divergence_l2 << '\n';
    std::cout << "pressure_iterations: " << r.final_step.pressure_iterations << '\n';
    std::cout << "pressure_relative_residual: " << r.final_step.pressure_relative_residual << '\n';
    std::cout << "u_vertical_max: " << r.extrema.u_vertical_max << '\n';
    std::cout << "u_vertical_min: " << r.extrema.u_vertical_min << '\n';
    std::cout << "v_horizontal_max: " << r.extrema.v_horizontal_max << '\n';
    std::cout << "v_horizontal_min: " << r.extrema.v_horizontal_min << '\n';
    std::cout << "reference_dataset: " << r.validation.reference_dataset << '\n';
    std::cout << "max_reference_relative_error: " << r.validation.max_relative_error << '\n';
    for(const solver::LidDrivenCavityValidationPoint& p : r.validation.points) {
      if(p.label.empty()) continue;
      std::cout << "reference_point[" << p.label << "]: line=" << solver::to_string(p.line)
                << ", coordinate=" << p.coordinate << ", reference=" << p.reference_value
                << ", reference_lower=" << p.reference_lower_bound << ", reference_upper=" << p.reference_upper_bound
                << ", sample=" << p.sample_value << ", relative_error=" << p.relative_error << '\n';
    }
    if(!restart_path.empty()) std::cout << "restart_path: " << restart_path << '\n';
    if(!checkpoint_out.empty()) std::cout << "checkpoint_out: " << checkpoint_out << '\n';
    if(!vtk_out.empty()) std::cout << "vtk_out: " << vtk_out << '\n';
    if(steps || !config.validate_reference) {
      std::cout << "benchmark_status: skipped" << '\n';
      return 0;
    }
    if(r.validation.reference_dataset.empty()) {
      std::cout << "benchmark_status: unavailable" << '\n';
      return 1;
    }
    std::cout << "benchmark_status: " << (r.validation.pass ? "pass" : "fail") << '\n';
    return r.validation.pass ? 0 : 1;
  } catch(const std::exception& e) {
    std::cerr << "solver_cavity failed: " << e.what() << '\n';
    return 1;
  }
}
