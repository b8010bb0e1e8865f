// Host field containers of the C++ interface (the reference's core/fields.hpp:12-143 and
// core/aligned_buffer.hpp:11-107 by name): value semantics, 64-byte aligned zero-initialised
// storage in the ghosted layout. The HIP entry points copy these buffers to the device and
// back unchanged.
#pragma once

#include "core/field_layout.hpp"

#include <algorithm>
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <new>
#include <utility>

namespace solver {

template <typename T, std::size_t Alignment>
class AlignedBuffer {
 public:
  using value_type = T;
  static constexpr std::size_t alignment = Alignment;
  static_assert(Alignment >= alignof(T) && Alignment % alignof(T) == 0, "alignment must suit the value type");

  AlignedBuffer() = default;
  explicit AlignedBuffer(const std::size_t size) { resize(size); }
  AlignedBuffer(const AlignedBuffer& other) : AlignedBuffer(other.size_) {
    if(size_ != 0) std::copy_n(other.data_, size_, data_);
  }
  AlignedBuffer(AlignedBuffer&& other) noexcept
      : data_(std::exchange(other.data_, nullptr)), size_(std::exchange(other.size_, 0)) {}
  AlignedBuffer& operator=(const AlignedBuffer& other) {
    if(this != &other) {
      AlignedBuffer tmp(other);
      swap(tmp);
    }
    return *this;
  }
  AlignedBuffer& operator=(AlignedBuffer&& other) noexcept {
    if(this != &other) {
      reset();
      data_ = std::exchange(other.data_, nullptr);
      size_ = std::exchange(other.size_, 0);
    }
    return *this;
  }
  ~AlignedBuffer() { reset(); }

  void resize(const std::size_t size) {
    reset();
    if(size == 0) return;
    data_ = static_cast<T*>(::operator new[](size * sizeof(T), std::align_val_t{Alignment}));
    size_ = size;
    std::fill_n(data_, size_, T{});
  }
  void reset() noexcept {
    if(data_ != nullptr) ::operator delete[](data_, std::align_val_t{Alignment});
    data_ = nullptr;
    size_ = 0;
  }
  void swap(AlignedBuffer& other) noexcept {
    std::swap(data_, other.data_);
    std::swap(size_, other.size_);
  }
  [[nodiscard]] T* data() noexcept { return data_; }
  [[nodiscard]] const T* data() const noexcept { return data_; }
  [[nodiscard]] std::size_t size() const noexcept { return size_; }
  [[nodiscard]] T& operator[](const std::size_t i) noexcept { return data_[i]; }
  [[nodiscard]] const T& operator[](const std::size_t i) const noexcept { return data_[i]; }

 private:
  T* data_ = nullptr;
  std::size_t size_ = 0;
};

class StructuredField {
 public:
  using value_type = double;
  static constexpr std::size_t storage_alignment = 64;

  explicit StructuredField(FieldLayout layout) : layout_(layout), data_(layout_.storage_size()) {}

  [[nodiscard]] const FieldLayout& layout() const noexcept { return layout_; }
  [[nodiscard]] double* data() noexcept { return data_.data(); }
  [[nodiscard]] const double* data() const noexcept { return data_.data(); }
  [[nodiscard]] std::size_t size() const noexcept { return data_.size(); }
  [[nodiscard]] bool is_aligned() const noexcept {
    return reinterpret_cast<std::uintptr_t>(data()) % storage_alignment == 0;
  }

  void fill(const double value) { std::fill_n(data(), size(), value); }
  void fill_range(const IndexRange3D& r, const double value) {
    for(int k = r.k_begin; k < r.k_end; ++k)
      for(int j = r.j_begin; j < r.j_end; ++j)
        for(int i = r.i_begin; i < r.i_end; ++i) (*this)(i, j, k) = value;
  }
  void fill_ghost_layer(const BoundaryFace face, const int layer, const double value) {
    fill_range(layout_.ghost_range(face, layer), value);
  }

  [[nodiscard]] double& at(const int i, const int j, const int k) { return data_[layout_.index(i, j, k)]; }
  [[nodiscard]] const double& at(const int i, const int j, const int k) const { return data_[layout_.index(i, j, k)]; }
  [[nodiscard]] double& unchecked(const int i, const int j, const int k) noexcept {
    return data_[layout_.unchecked_index(i, j, k)];
  }
  [[nodiscard]] const double& unchecked(const int i, const int j, const int k) const noexcept {
    return data_[layout_.unchecked_index(i, j, k)];
  }
  [[nodiscard]] double* row_ptr(const int j, const int k) noexcept { return data() + layout_.unchecked_index(0, j, k); }
  [[nodiscard]] const double* row_ptr(const int j, const int k) const noexcept {
    return data() + layout_.unchecked_index(0, j, k);
  }
  [[nodiscard]] double& operator()(const int i, const int j, const int k) {
    assert(layout_.is_storage_index(i, j, k));
    return data_[layout_.unchecked_index(i, j, k)];
  }
  [[nodiscard]] const double& operator()(const int i, const int j, const int k) const {
    assert(layout_.is_storage_index(i, j, k));
    return data_[layout_.unchecked_index(i, j, k)];
  }

 protected:
  FieldLayout layout_;
  AlignedBuffer<double, storage_alignment> data_;
};

class ScalarField final : public StructuredField {
 public:
  explicit ScalarField(const Grid& grid) : StructuredField(FieldLayout::cell_centered(grid)) {}
};

class PressureField final : public StructuredField {
 public:
  explicit PressureField(const Grid& grid) : StructuredField(FieldLayout::cell_centered(grid)) {}
};

class FaceField final : public StructuredField {
 public:
  FaceField(const Axis normal_axis, const Grid& grid)
      : StructuredField(FieldLayout::face_centered(normal_axis, grid)), normal_axis_(normal_axis) {}
  [[nodiscard]] constexpr Axis normal_axis() const noexcept { return normal_axis_; }

 private:
  Axis normal_axis_;
};

struct VelocityField {
  FaceField x;
  FaceField y;
  FaceField z;

  explicit VelocityField(const Grid& grid) : x(Axis::x, grid), y(Axis::y, grid), z(Axis::z, grid) {}

  void fill(const double value) {
    x.fill(value);
    y.fill(value);
    z.fill(value);
  }
};

}  // namespace solver
