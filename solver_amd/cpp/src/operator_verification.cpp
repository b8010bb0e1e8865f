// Manufactured-solution check of the MAC operators (reference
// solver/operator_verification.cpp:76-127): analytic fields on the host, operators on the device.
#include "solver/operator_verification.hpp"

#include "core/fields.hpp"
#include "operators/discrete_operators.hpp"

#include <cmath>
#include <stdexcept>

namespace solver {

namespace {

template <typename Fn>
void fill_all(StructuredField& f, Fn&& fn) {
  const FieldLayout& L = f.layout();
  const Extent3D s = L.storage_extent();
  for(int k = 0; k < s.nz; ++k) {
    const double z = L.coordinate_for_storage_index(Axis::z, k);
    for(int j = 0; j < s.ny; ++j) {
      const double y = L.coordinate_for_storage_index(Axis::y, j);
      for(int i = 0; i < s.nx; ++i) f(i, j, k) = fn(L.coordinate_for_storage_index(Axis::x, i), y, z);
    }
  }
}

template <typename Fn>
double rms_error(const StructuredField& f, Fn&& exact) {
  const FieldLayout& L = f.layout();
  const IndexRange3D r = L.active_range();
  double sum = 0.0;
  std::size_t n = 0;
  for(int k = r.k_begin; k < r.k_end; ++k) {
    const double z = L.coordinate_for_storage_index(Axis::z, k);
    for(int j = r.j_begin; j < r.j_end; ++j) {
      const double y = L.coordinate_for_storage_index(Axis::y, j);
      for(int i = r.i_begin; i < r.i_end; ++i) {
        const double e = f(i, j, k) - exact(L.coordinate_for_storage_index(Axis::x, i), y, z);
        sum += e * e;
        ++n;
      }
    }
  }
  return std::sqrt(sum / static_cast<double>(n));
}

}  // namespace

OperatorManufacturedSolutionResult run_operator_manufactured_solution_case(const int resolution) {
  if(resolution <= 0) throw std::invalid_argument("run_operator_manufactured_solution_case expects resolution > 0");
  const double L = 2.0 * std::acos(-1.0);
  const double h = L / resolution;
  const Grid grid{resolution, resolution, 1, h, h, 1.0, 1};
  PressureField p{grid};
  VelocityField vel{grid}, grad{grid};
  ScalarField div{grid};
  PressureField lap{grid};
  fill_all(p, [](double x, double y, double) { return std::sin(x) * std::cos(y); });
  fill_all(vel.x, [](double x, double y, double) { return std::sin(x) * std::cos(y); });
  fill_all(vel.y, [](double x, double y, double) { return std::cos(x) * std::sin(y); });
  fill_all(vel.z, [](double, double, double) { return 0.0; });
  operators::compute_gradient(p, grad);
  operators::compute_divergence(vel, div);
  operators::compute_laplacian(p, lap);
  const double ex = rms_error(grad.x, [](double x, double y, double) { return std::cos(x) * std::cos(y); });
  const double ey = rms_error(grad.y, [](double x, double y, double) { return -std::sin(x) * std::sin(y); });
  OperatorManufacturedSolutionResult r;
  r.resolution = resolution;
  r.domain_length = L;
  r.gradient_error = std::sqrt(0.5 * (ex * ex + ey * ey));
  r.divergence_error = rms_error(div, [](double x, double y, double) { return 2.0 * std::cos(x) * std::cos(y); });
  r.laplacian_error = rms_error(lap, [](double x, double y, double) { return -2.0 * std::sin(x) * std::cos(y); });
  return r;
}

}  // namespace solver
