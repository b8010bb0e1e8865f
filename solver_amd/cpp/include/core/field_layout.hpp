// Ghosted storage layout of one field: the reference's FieldLayout (core/field_layout.hpp:69-290)
// with the same names. Storage is i fastest, flat index i + sx*(j + sy*k); face-centred
// layouts have one extra active entry along their normal axis (the periodic face is stored
// twice). Device fields use the identical layout, so buffers move between host and device
// unchanged.
#pragma once

#include "core/grid.hpp"

#include <cstddef>
#include <stdexcept>

namespace solver {

enum class FieldLocation : int { cell_center = 0, face_x = 1, face_y = 2, face_z = 3 };

struct Extent3D {
  int nx;
  int ny;
  int nz;

  [[nodiscard]] constexpr int count(const Axis axis) const noexcept {
    return axis == Axis::x ? nx : axis == Axis::y ? ny : nz;
  }
  [[nodiscard]] constexpr std::size_t cell_count() const noexcept {
    return static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny) * static_cast<std::size_t>(nz);
  }
};

struct Index3D {
  int i;
  int j;
  int k;
};

struct IndexRange3D {
  int i_begin;
  int i_end;
  int j_begin;
  int j_end;
  int k_begin;
  int k_end;

  [[nodiscard]] constexpr Extent3D extent() const noexcept {
    return Extent3D{i_end - i_begin, j_end - j_begin, k_end - k_begin};
  }
  [[nodiscard]] constexpr bool contains(const int i, const int j, const int k) const noexcept {
    return i >= i_begin && i < i_end && j >= j_begin && j < j_end && k >= k_begin && k < k_end;
  }
};

class FieldLayout {
 public:
  [[nodiscard]] static FieldLayout cell_centered(const Grid& grid) {
    return FieldLayout(FieldLocation::cell_center, grid);
  }

  [[nodiscard]] static FieldLayout face_centered(const Axis axis, const Grid& grid) {
    return FieldLayout(static_cast<FieldLocation>(1 + axis_index(axis)), grid);
  }

  [[nodiscard]] constexpr FieldLocation location() const noexcept { return location_; }
  [[nodiscard]] constexpr const Grid& grid() const noexcept { return grid_; }
  [[nodiscard]] constexpr int ghost_layers() const noexcept { return grid_.ghost_layers; }
  [[nodiscard]] constexpr Extent3D active_extent() const noexcept { return active_; }
  [[nodiscard]] constexpr Extent3D storage_extent() const noexcept {
    const int g2 = 2 * grid_.ghost_layers;
    return Extent3D{active_.nx + g2, active_.ny + g2, active_.nz + g2};
  }
  [[nodiscard]] constexpr std::size_t storage_size() const noexcept { return storage_extent().cell_count(); }
  [[nodiscard]] constexpr std::size_t storage_stride_j() const noexcept {
    return static_cast<std::size_t>(storage_extent().nx);
  }
  [[nodiscard]] constexpr std::size_t storage_stride_k() const noexcept {
    return storage_stride_j() * static_cast<std::size_t>(storage_extent().ny);
  }

  [[nodiscard]] constexpr bool same_shape_as(const FieldLayout& other) const noexcept {
    const Extent3D a = storage_extent(), b = other.storage_extent();
    return location_ == other.location_ && ghost_layers() == other.ghost_layers() && active_.nx == other.active_.nx &&
           active_.ny == other.active_.ny && active_.nz == other.active_.nz && a.nx == b.nx && a.ny == b.ny &&
           a.nz == b.nz;
  }

  [[nodiscard]] constexpr bool is_unit_stride_i() const noexcept { return true; }

  [[nodiscard]] constexpr IndexRange3D active_range() const noexcept {
    const int g = grid_.ghost_layers;
    return IndexRange3D{g, g + active_.nx, g, g + active_.ny, g, g + active_.nz};
  }

  [[nodiscard]] constexpr bool is_storage_index(const int i, const int j, const int k) const noexcept {
    const Extent3D s = storage_extent();
    return i >= 0 && i < s.nx && j >= 0 && j < s.ny && k >= 0 && k < s.nz;
  }

  [[nodiscard]] constexpr bool is_active_storage_index(const int i, const int j, const int k) const noexcept {
    return active_range().contains(i, j, k);
  }

  [[nodiscard]] Index3D storage_index_from_active(const int i, const int j, const int k) const {
    if(i < 0 || i >= active_.nx || j < 0 || j >= active_.ny || k < 0 || k >= active_.nz)
      throw std::out_of_range("active index outside field extent");
    return unchecked_storage_index_from_active(i, j, k);
  }

  [[nodiscard]] constexpr Index3D unchecked_storage_index_from_active(const int i, const int j,
                                                                      const int k) const noexcept {
    const int g = grid_.ghost_layers;
    return Index3D{i + g, j + g, k + g};
  }

  [[nodiscard]] std::size_t index(const int i, const int j, const int k) const {
    if(!is_storage_index(i, j, k)) throw std::out_of_range("storage index outside field extent");
    return unchecked_index(i, j, k);
  }

  [[nodiscard]] constexpr std::size_t unchecked_index(const int i, const int j, const int k) const noexcept {
    return static_cast<std::size_t>(i) + storage_stride_j() * static_cast<std::size_t>(j) +
           storage_stride_k() * static_cast<std::size_t>(k);
  }

  // the single active layer adjacent to `face`
  [[nodiscard]] IndexRange3D boundary_active_range(const BoundaryFace face) const {
    const Axis axis = boundary_axis(face);
    const int g = grid_.ghost_layers;
    const int first = is_lower_boundary(face) ? g : g + active_.count(axis) - 1;
    return with_axis(active_range(), axis, first, first + 1);
  }

  // ghost layer `layer` (0 = nearest the active range) beyond `face`
  [[nodiscard]] IndexRange3D ghost_range(const BoundaryFace face, const int layer = 0) const {
    const int g = grid_.ghost_layers;
    if(layer < 0 || layer >= g) throw std::out_of_range("ghost layer outside configured range");
    const Axis axis = boundary_axis(face);
    const int at = is_lower_boundary(face) ? g - 1 - layer : g + active_.count(axis) + layer;
    return with_axis(active_range(), axis, at, at + 1);
  }

  [[nodiscard]] double coordinate_at_active_index(const Axis axis, const int active_index) const {
    if(active_index < 0 || active_index >= active_.count(axis))
      throw std::out_of_range("active coordinate index outside field extent");
    return on_faces(axis) ? grid_.face_coordinate(axis, active_index) : grid_.cell_center(axis, active_index);
  }

  [[nodiscard]] double coordinate_for_storage_index(const Axis axis, const int storage_index) const {
    if(storage_index < 0 || storage_index >= storage_extent().count(axis))
      throw std::out_of_range("storage coordinate index outside field extent");
    const double shifted = static_cast<double>(storage_index - grid_.ghost_layers);
    return (on_faces(axis) ? shifted : shifted + 0.5) * grid_.spacing(axis);
  }

 private:
  FieldLayout(const FieldLocation location, const Grid& grid)
      : active_{grid.nx + (location == FieldLocation::face_x), grid.ny + (location == FieldLocation::face_y),
                grid.nz + (location == FieldLocation::face_z)},
        grid_(grid),
        location_(location) {}

  // true when coordinates along `axis` sit on faces (the normal axis of a face field)
  [[nodiscard]] constexpr bool on_faces(const Axis axis) const noexcept {
    return location_ != FieldLocation::cell_center && axis_index(axis) == static_cast<int>(location_) - 1;
  }

  static IndexRange3D with_axis(IndexRange3D r, const Axis axis, const int begin, const int end) noexcept {
    if(axis == Axis::x) {
      r.i_begin = begin;
      r.i_end = end;
    } else if(axis == Axis::y) {
      r.j_begin = begin;
      r.j_end = end;
    } else {
      r.k_begin = begin;
      r.k_end = end;
    }
    return r;
  }

  Extent3D active_;
  Grid grid_;
  FieldLocation location_;
};

}  // namespace solver
