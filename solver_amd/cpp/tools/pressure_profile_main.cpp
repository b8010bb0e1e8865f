// solver_pressure_profile: the reference's tools/pressure_profile_main.cpp (2D Dirichlet
// Poisson problem, repeated MGPCG solves, timing in key: value form) on the GPU. An optional
// third argument selects the precision of the solve: fp64 (default) or mixed.
#include "linsolve/poisson_solver.hpp"
#include "solver/hip_backend.hpp"

#include <chrono>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
  try {
    const int n = argc > 1 ? std::stoi(argv[1]) : 256;
    const int repeats = argc > 2 ? std::stoi(argv[2]) : 24;
    if(n <= 0 || repeats <= 0) throw std::invalid_argument("resolution and repeats must be positive");
    const std::string precision = argc > 3 ? argv[3] : "fp64";
    if(precision != "fp64" && precision != "mixed") throw std::invalid_argument("precision must be fp64 or mixed");
    solver::hip_backend::set_pressure_precision(precision == "mixed" ? OF_PRECISION_MIXED : OF_PRECISION_FP64);
    const double h = 1.0 / n;
    const double pi = std::acos(-1.0);
    const solver::Grid grid{n, n, 1, h, h, 1.0, 1};
    solver::ScalarField rhs{grid};
    solver::PressureField pressure{grid};
    solver::PressureBoundarySet bcs;
    for(auto& f : bcs.faces) {
      f.type = solver::PressureBoundaryType::dirichlet;
      f.value = 0.0;
    }
    const solver::FieldLayout& L = rhs.layout();
    const solver::Extent3D s = L.storage_extent();
    for(int k = 0; k < s.nz; ++k)
      for(int j = 0; j < s.ny; ++j)
        for(int i = 0; i < s.nx; ++i) {
          const double x = L.coordinate_for_storage_index(solver::Axis::x, i);
          const double y = L.coordinate_for_storage_index(solver::Axis::y, j);
          rhs(i, j, k) = 2.0 * pi * pi * std::sin(pi * x) * std::sin(pi * y);
        }
    const solver::ProjectionOptions options{1.0, 1.0, 4000, 1.0e-10};
    int iterations = 0, levels = 0;
    double residual = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    for(int r = 0; r < repeats; ++r) {
      pressure.fill(0.0);
      const solver::PoissonSolveDiagnostics d = solver::linsolve::solve_pressure_poisson(rhs, bcs, options, pressure);
      if(!d.converged) throw std::runtime_error("pressure profile benchmark encountered a non-converged solve");
      iterations += d.iterations;
      residual += d.relative_residual;
      levels = d.multigrid_levels;
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "kernel: pressure_poisson" << '\n';
    std::cout << "nx: " << n << '\n';
    std::cout << "ny: " << n << '\n';
    std::cout << "repeats: " << repeats << '\n';
    std::cout << "elapsed_seconds: " << elapsed << '\n';
    std::cout << "average_seconds_per_repeat: " << elapsed / repeats << '\n';
    std::cout << "logical_cells: " << static_cast<std::size_t>(n) * static_cast<std::size_t>(n) << '\n';
    std::cout << "unknown_updates_per_second: " << static_cast<double>(n) * n * repeats / elapsed << '\n';
    std::cout << "average_iterations: " << static_cast<double>(iterations) / repeats << '\n';
    std::cout << "average_relative_residual: " << residual / repeats << '\n';
    std::cout << "multigrid_levels: " << levels << '\n';
    std::cout << "precision: " << precision << '\n';
    return 0;
  } catch(const std::exception& e) {
    std::cerr << "solver_pressure_profile failed: " << e.what() << '\n';
    return 1;
  }
}
