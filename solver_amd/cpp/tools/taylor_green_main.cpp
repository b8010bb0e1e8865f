// solver_taylor_green: the reference's tools/taylor_green_main.cpp command line and key: value
// output, running the step loop on the GPU (backend = hip).
#include "solver/taylor_green.hpp"

#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
  try {
    std::string config_path = "solver_amd/cpp/configs/taylor_green_128.cfg";
    std::string backend_override;
    bool have_config = false;
    for(int a = 1; a < argc; ++a) {
      const std::string arg = argv[a];
      if(arg == "--backend") {
        if(a + 1 >= argc) throw std::runtime_error("missing value for --backend");
        backend_override = argv[++a];
      } else if(arg == "--vtk-out") {
        throw std::runtime_error("--vtk-out: VTK export is not part of this build");
      } else if(!have_config && !arg.empty() && arg.rfind("--", 0) != 0) {
        config_path = arg;
        have_config = true;
      } else {
        throw std::runtime_error("unsupported argument: " + arg);
      }
    }
    solver::TaylorGreenConfig config = solver::load_taylor_green_config(config_path);
    if(!backend_override.empty()) config.backend = solver::parse_execution_backend(backend_override);
    solver::TaylorGreenState state{solver::make_taylor_green_grid(config)};
    const solver::TaylorGreenResult r = solver::run_taylor_green(config, &state);
    std::cout << "config_path: " << config_path << '\n';
    std::cout << "config: " << solver::describe(config) << '\n';
    std::cout << "backend_used: " << solver::to_string(r.backend_used) << '\n';
    if(!r.accelerator_name.empty()) std::cout << "accelerator_name: " << r.accelerator_name << '\n';
    if(r.backend_elapsed_seconds > 0.0) std::cout << "backend_elapsed_seconds: " << r.backend_elapsed_seconds << '\n';
    if(r.cleanup_elapsed_seconds > 0.0) std::cout << "cleanup_elapsed_seconds: " << r.cleanup_elapsed_seconds << '\n';
    std::cout << "steps: " << r.final_step.step << '\n';
    std::cout << "time: " << r.final_step.time << '\n';
    std::cout << "dt: " << r.final_step.dt << '\n';
    std::cout << "max_cfl: " << r.final_step.max_cfl << '\n';
    std::cout << "max_velocity_change: " << r.final_step.max_velocity_change << '\n';
    std::cout << "divergence_l2: " << r.final_step.divergence_l2 << '\n';
    std::cout << "max_divergence_l2: " << r.final_step.max_divergence_l2 << '\n';
    std::cout << "pressure_iterations: " << r.final_step.pressure_iterations << '\n';
    std::cout << "pressure_relative_residual: " << r.final_step.pressure_relative_residual << '\n';
    std::cout << "initial_kinetic_energy: " << r.initial_kinetic_energy << '\n';
    std::cout << "final_kinetic_energy: " << r.final_kinetic_energy << '\n';
    std::cout << "reference_dataset: " << r.validation.reference_dataset << '\n';
    std::cout << "exact_kinetic_energy: " << r.validation.exact_kinetic_energy << '\n';
    std::cout << "normalized_energy_error: " << r.validation.normalized_energy_error << '\n';
    std::cout << "velocity_relative_l2_error: " << r.validation.velocity_relative_l2_error << '\n';
    if(!config.validate_energy) {
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
    std::cerr << "solver_taylor_green failed: " << e.what() << '\n';
    return 1;
  }
}
