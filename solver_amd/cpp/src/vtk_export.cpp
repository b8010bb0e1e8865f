// Legacy-VTK ASCII export (reference io/vtk_export.cpp:18-78): the active cells as
// STRUCTURED_POINTS at the cell centres, the velocity averaged from the two faces of each cell,
// values printed like a default-formatted ostream (%g, 6 significant digits). Lines are
// formatted into a 4 MB buffer and written in blocks.
#include "io/vtk_export.hpp"

#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace solver::io {

namespace {

bool same_grid(const Grid& a, const Grid& b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx && a.dy == b.dy && a.dz == b.dz &&
         a.ghost_layers == b.ghost_layers;
}

class LineBuffer {
 public:
  explicit LineBuffer(std::ofstream& out) : out_(out) { buf_.reserve(kCap + 256); }
  ~LineBuffer() { flush(); }
  template <typename... T>
  void line(const char* fmt, T... v) {
    char tmp[192];
    const int n = std::snprintf(tmp, sizeof tmp, fmt, v...);
    buf_.append(tmp, (std::size_t)n);
    if(buf_.size() >= kCap) flush();
  }
  void text(const std::string& s) { buf_ += s; }
  void flush() {
    out_.write(buf_.data(), (std::streamsize)buf_.size());
    buf_.clear();
  }

 private:
  static constexpr std::size_t kCap = std::size_t(4) << 20;
  std::ofstream& out_;
  std::string buf_;
};

}  // namespace

void write_mac_fields_vtk(const std::string& path, const VelocityField& velocity, const PressureField& pressure_total) {
  const Grid& g = pressure_total.layout().grid();
  if(!same_grid(velocity.x.layout().grid(), g) || !same_grid(velocity.y.layout().grid(), g) ||
     !same_grid(velocity.z.layout().grid(), g))
    throw std::invalid_argument("VTK export expects velocity and pressure on the same grid");
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if(!out.is_open()) throw std::runtime_error("unable to open VTK file for writing: " + path);
  const IndexRange3D r = pressure_total.layout().active_range();
  const std::size_t points = r.extent().cell_count();
  {
    LineBuffer b(out);
    b.text("# vtk DataFile Version 3.0\nsolver checkpoint state\nASCII\nDATASET STRUCTURED_POINTS\n");
    b.line("DIMENSIONS %d %d %d\n", g.nx, g.ny, g.nz);
    b.line("ORIGIN %g %g %g\n", 0.5 * g.dx, 0.5 * g.dy, 0.5 * g.dz);
    b.line("SPACING %g %g %g\n", g.dx, g.dy, g.dz);
    b.line("POINT_DATA %zu\n", points);
    b.text("VECTORS velocity double\n");
    for(int k = r.k_begin; k < r.k_end; ++k)
      for(int j = r.j_begin; j < r.j_end; ++j)
        for(int i = r.i_begin; i < r.i_end; ++i) {
          const double u = 0.5 * (velocity.x(i, j, k) + velocity.x(i + 1, j, k));
          const double v = 0.5 * (velocity.y(i, j, k) + velocity.y(i, j + 1, k));
          const double w = 0.5 * (velocity.z(i, j, k) + velocity.z(i, j, k + 1));
          b.line("%g %g %g\n", u, v, w);
        }
    b.text("SCALARS pressure double 1\nLOOKUP_TABLE default\n");
    for(int k = r.k_begin; k < r.k_end; ++k)
      for(int j = r.j_begin; j < r.j_end; ++j)
        for(int i = r.i_begin; i < r.i_end; ++i) b.line("%g\n", pressure_total(i, j, k));
  }
  if(!out.good()) throw std::runtime_error("failed while writing VTK file: " + path);
}

void write_lid_driven_cavity_vtk(const std::string& path, const LidDrivenCavityState& state) {
  write_mac_fields_vtk(path, state.velocity, state.pressure_total);
}

}  // namespace solver::io
