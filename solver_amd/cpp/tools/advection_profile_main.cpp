// solver_advection_profile: the reference's tools/advection_profile_main.cpp (2D Taylor-Green
// velocity, repeated compute_advection_term, timing in key: value form) on the GPU. Each repeat
// goes through the host C++ interface, so the figures include the host<->device copies.
#include "solver/momentum_terms.hpp"

#include <chrono>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
  try {
    const int n = argc > 1 ? std::stoi(argv[1]) : 256;
    const int repeats = argc > 2 ? std::stoi(argv[2]) : 200;
    if(n <= 0 || repeats <= 0) throw std::invalid_argument("resolution and repeats must be positive");
    const double h = 2.0 * std::acos(-1.0) / n;
    const solver::Grid grid{n, n, 1, h, h, 1.0, 1};
    solver::VelocityField vel{grid}, adv{grid};
    auto fill = [](solver::StructuredField& f, auto fn) {
      const solver::FieldLayout& L = f.layout();
      const solver::Extent3D s = L.storage_extent();
      for(int k = 0; k < s.nz; ++k)
        for(int j = 0; j < s.ny; ++j)
          for(int i = 0; i < s.nx; ++i)
            f(i, j, k) = fn(L.coordinate_for_storage_index(solver::Axis::x, i),
                            L.coordinate_for_storage_index(solver::Axis::y, j));
    };
    fill(vel.x, [](double x, double y) { return -std::cos(x) * std::sin(y); });
    fill(vel.y, [](double x, double y) { return std::sin(x) * std::cos(y); });
    fill(vel.z, [](double, double) { return 0.0; });
    const solver::AdvectionOptions options{};
    auto active = [](const solver::StructuredField& f) { return f.layout().active_range().extent().cell_count(); };
    const std::size_t cells = static_cast<std::size_t>(n) * static_cast<std::size_t>(n);
    const std::size_t bytes = 8ull * (active(vel.x) + active(vel.y) + active(vel.z) + active(adv.x) + active(adv.y) +
                                      active(adv.z));
    const auto t0 = std::chrono::steady_clock::now();
    for(int r = 0; r < repeats; ++r) solver::compute_advection_term(vel, options, adv);
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "kernel: advection" << '\n';
    std::cout << "nx: " << n << '\n';
    std::cout << "ny: " << n << '\n';
    std::cout << "repeats: " << repeats << '\n';
    std::cout << "elapsed_seconds: " << elapsed << '\n';
    std::cout << "average_seconds_per_repeat: " << elapsed / repeats << '\n';
    std::cout << "logical_cells: " << cells << '\n';
    std::cout << "cell_updates_per_second: " << static_cast<double>(cells) * repeats / elapsed << '\n';
    std::cout << "lower_bound_bandwidth_gbps: " << static_cast<double>(bytes) * repeats / elapsed / 1.0e9 << '\n';
    std::cout << "advection: " << solver::describe(options) << '\n';
    return 0;
  } catch(const std::exception& e) {
    std::cerr << "solver_advection_profile failed: " << e.what() << '\n';
    return 1;
  }
}
