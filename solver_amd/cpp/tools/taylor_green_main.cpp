// solver_taylor_green: the reference's tools/taylor_green_main.cpp command line and key: value
// output (--backend, --vtk-out <path>), running the step loop on the GPU (backend = hip). This
// build adds --steps <n> (run n steps instead of up to final_time, no validation), --restart
// <checkpoint> and --checkpoint-out <path> (SLVCHKP1 Taylor-Green checkpoints).
#include "io/checkpoint.hpp"
#include "io/vtk_export.hpp"
#include "solver/taylor_green.hpp"

#include <cmath>
#include <iostream>
#include <optional>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
  try {
    std::string config_path = "solver_amd/cpp/configs/taylor_green_128.cfg";
    std::string backend_override, vtk_out, restart_path, checkpoint_out;
    std::optional<int> steps;
    bool have_config = false;
    for(int a = 1; a < argc; ++a) {
      const std::string arg = argv[a];
      auto value = [&]() -> std::string {
        if(a + 1 >= argc) throw std::runtime_error("missing value for " + arg);
        return argv[++a];
      };
      if(arg == "--backend") {
        backend_override = value();
      } else if(arg == "--vtk-out") {
        vtk_out = value();
      } else if(arg == "--steps") {
        steps = std::stoi(value());
      } else if(arg == "--restart") {
        restart_path = value();
      } else if(arg == "--checkpoint-out") {
        checkpoint_out = value();
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
    solver::TaylorGreenResult r;
    if(restart_path.empty() && !steps) {
      r = solver::run_taylor_green(config, &state);
    } else {
      if(restart_path.empty())
        state = solver::initialize_taylor_green_state(config);
      else
        state = solver::io::load_taylor_green_checkpoint(restart_path, config).state;
      const int total = static_cast<int>(std::ceil(config.final_time / solver::taylor_green_dt(config)));
      const int n = steps ? *steps : std::max(0, total - state.metrics.step);
      solver::run_taylor_green_steps(config, n, state);
      r = solver::finalize_taylor_green_result(config, state);
    }
    if(!checkpoint_out.empty()) solver::io::write_taylor_green_checkpoint(checkpoint_out, config, state);
    if(!vtk_out.empty()) solver::io::write_mac_fields_vtk(vtk_out, state.velocity, state.pressure_total);
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
    if(!restart_path.empty()) std::cout << "restart_path: " << restart_path << '\n';
    if(!checkpoint_out.empty()) std::cout << "checkpoint_out: " << checkpoint_out << '\n';
    if(!vtk_out.empty()) std::cout << "vtk_out: " << vtk_out << '\n';
    if(steps || !config.validate_energy) {
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
