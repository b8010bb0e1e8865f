// Manufactured-solution check of the MAC operators (the reference's
// run_operator_manufactured_solution_case, solver/operator_verification.cpp:76-127) done on the
// device: every field -- the manufactured inputs and the exact derivatives -- is sampled once on
// the host into a device buffer, the three operators run through the C ABI on those buffers,
// and each error is a device RMS difference over the active entries.
//
// The mathematics is fixed by the result's definition: on [0, 2 pi]^2 with N^2 cells,
// p = sin x cos y and u = (sin x cos y, cos x sin y); the exact gradient (cos x cos y,
// -sin x sin y), divergence 2 cos x cos y and Laplacian -2 sin x cos y; the gradient error
// combines its two components as sqrt((e_x^2 + e_y^2) / 2).
#include "solver/operator_verification.hpp"

#include "core/fields.hpp"
#include "solver/hip_backend.hpp"

#include <array>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace solver {

namespace {

namespace hb = hip_backend;

using Profile = double (*)(double, double);

FieldLayout layout_of(const Grid& grid, FieldLocation location) {
  return location == FieldLocation::cell_center
             ? FieldLayout::cell_centered(grid)
             : FieldLayout::face_centered(static_cast<Axis>(static_cast<int>(location) - 1), grid);
}

std::size_t storage_of(const Grid& grid, FieldLocation location) { return layout_of(grid, location).storage_size(); }

// The field of `location` on `grid` with every storage point (ghosts included) set to fn(x, y).
// The coordinates of one axis are tabulated once; the fill walks the storage in memory order.
StructuredField sampled(const Grid& grid, FieldLocation location, Profile fn) {
  StructuredField f{layout_of(grid, location)};
  const FieldLayout& L = f.layout();
  const Extent3D s = L.storage_extent();
  std::vector<double> xs(static_cast<std::size_t>(s.nx)), ys(static_cast<std::size_t>(s.ny));
  for(int i = 0; i < s.nx; ++i) xs[static_cast<std::size_t>(i)] = L.coordinate_for_storage_index(Axis::x, i);
  for(int j = 0; j < s.ny; ++j) ys[static_cast<std::size_t>(j)] = L.coordinate_for_storage_index(Axis::y, j);
  double* out = f.data();
  for(int k = 0; k < s.nz; ++k)
    for(int j = 0; j < s.ny; ++j)
      for(int i = 0; i < s.nx; ++i) *out++ = fn(xs[static_cast<std::size_t>(i)], ys[static_cast<std::size_t>(j)]);
  return f;
}

double p_exact(double x, double y) { return std::sin(x) * std::cos(y); }
double u_exact(double x, double y) { return std::sin(x) * std::cos(y); }
double v_exact(double x, double y) { return std::cos(x) * std::sin(y); }
double zero(double, double) { return 0.0; }
double dpdx_exact(double x, double y) { return std::cos(x) * std::cos(y); }
double dpdy_exact(double x, double y) { return -std::sin(x) * std::sin(y); }
double div_exact(double x, double y) { return 2.0 * std::cos(x) * std::cos(y); }
double lap_exact(double x, double y) { return -2.0 * std::sin(x) * std::cos(y); }

// A device buffer of `location` on `grid` holding fn sampled at every storage point.
hb::DeviceField device_sampled(const Grid& grid, FieldLocation location, Profile fn) {
  return hb::DeviceField(sampled(grid, location, fn));
}

double rms_difference(const of_grid& g, FieldLocation location, const hb::DeviceField& a, const hb::DeviceField& b) {
  double e = 0.0;
  hb::check(of_active_rms_difference(hb::context(), &g, static_cast<int>(location), a.get(), b.get(), &e));
  return e;
}

}  // namespace

OperatorManufacturedSolutionResult run_operator_manufactured_solution_case(const int resolution) {
  if(resolution <= 0) throw std::invalid_argument("run_operator_manufactured_solution_case expects resolution > 0");
  const double length = 2.0 * std::acos(-1.0);
  const double h = length / static_cast<double>(resolution);
  const Grid grid{resolution, resolution, 1, h, h, 1.0, 1};
  const of_grid g = hb::to_abi(grid);
  const FieldLocation cell = FieldLocation::cell_center, fx = FieldLocation::face_x, fy = FieldLocation::face_y,
                      fz = FieldLocation::face_z;

  // operator inputs
  const hb::DeviceField p = device_sampled(grid, cell, p_exact);
  const std::array<hb::DeviceField, 3> vel{device_sampled(grid, fx, u_exact), device_sampled(grid, fy, v_exact),
                                           device_sampled(grid, fz, zero)};
  // operator outputs (zero-initialised device buffers)
  const std::array<hb::DeviceField, 3> grad{hb::DeviceField(storage_of(grid, fx)),
                                            hb::DeviceField(storage_of(grid, fy)),
                                            hb::DeviceField(storage_of(grid, fz))};
  const hb::DeviceField div(storage_of(grid, cell));
  const hb::DeviceField lap(storage_of(grid, cell));
  of_context* ctx = hb::context();
  hb::check(of_compute_gradient(ctx, &g, p.get(), grad[0].get(), grad[1].get(), grad[2].get()));
  hb::check(of_compute_divergence(ctx, &g, vel[0].get(), vel[1].get(), vel[2].get(), div.get()));
  hb::check(of_compute_laplacian(ctx, &g, OF_CELL, p.get(), lap.get()));

  // errors against the exact derivatives, on the device
  const double ex = rms_difference(g, fx, grad[0], device_sampled(grid, fx, dpdx_exact));
  const double ey = rms_difference(g, fy, grad[1], device_sampled(grid, fy, dpdy_exact));
  OperatorManufacturedSolutionResult r;
  r.resolution = resolution;
  r.domain_length = length;
  r.gradient_error = std::sqrt(0.5 * (ex * ex + ey * ey));
  r.divergence_error = rms_difference(g, cell, div, device_sampled(grid, cell, div_exact));
  r.laplacian_error = rms_difference(g, cell, lap, device_sampled(grid, cell, lap_exact));
  return r;
}

}  // namespace solver
