// solver_channel: runs one channel-flow configuration and prints the reference tool's
// "key: value" report (tools/channel_flow_main.cpp). Exit status 0 on a passing (or skipped)
// validation, 1 otherwise.
#include "solver/channel_flow.hpp"

#include <iostream>
#include <string>

int main(int argc, char** argv) {
  try {
    const std::string path = argc > 1 ? argv[1] : "configs/channel_couette_128.cfg";
    const solver::ChannelFlowConfig config = solver::load_channel_flow_config(path);
    const solver::ChannelFlowResult r = solver::run_channel_flow(config);
    const solver::ChannelFlowStepMetrics& m = r.final_step;
    std::cout << "config_path: " << path << '\n'
              << "config: " << solver::describe(config) << '\n'
              << "steps: " << m.step << '\n'
              << "time: " << m.time << '\n'
              << "dt: " << m.dt << '\n'
              << "max_cfl: " << m.max_cfl << '\n'
              << "max_velocity_change: " << m.max_velocity_change << '\n'
              << "divergence_l2: " << m.divergence_l2 << '\n'
              << "pressure_iterations: " << m.pressure_iterations << '\n'
              << "pressure_relative_residual: " << m.pressure_relative_residual << '\n'
              << "reference_dataset: " << r.validation.reference_dataset << '\n'
              << "profile_relative_l2_error: " << r.validation.relative_l2_error << '\n';
    if(!config.validate_profile) {
      std::cout << "benchmark_status: skipped\n";
      return 0;
    }
    if(r.validation.reference_dataset.empty()) {
      std::cout << "benchmark_status: unavailable\n";
      return 1;
    }
    std::cout << "benchmark_status: " << (r.validation.pass ? "pass" : "fail") << '\n';
    return r.validation.pass ? 0 : 1;
  } catch(const std::exception& e) {
    std::cerr << "solver_channel failed: " << e.what() << '\n';
    return 1;
  }
}
