// Glue between the C++ interface and the C ABI (include/of_hip.h): the process-wide device
// context, device buffers mirroring host fields, descriptor conversion and status-to-exception
// mapping. No reference counterpart.
#pragma once

#include "bc/boundary_conditions.hpp"
#include "core/fields.hpp"
#include "of_hip.h"

#include <cstddef>
#include <string>

namespace solver::hip_backend {

// Device context on HIP device $SOLVER_HIP_DEVICE (default 0), created on first use.
of_context* context();
// Waits for the work queued on the context stream.
void synchronize();
// Throws std::invalid_argument (OF_INVALID_ARGUMENT) or std::runtime_error carrying the
// library's message for any status other than OF_OK.
void check(int status);
// Marketing name of the context's device.
std::string device_name();

of_grid to_abi(const Grid& grid);
of_bc_set to_abi(const BoundaryConditionSet& bcs);
of_pressure_bc_set to_abi(const PressureBoundarySet& bcs);

// FP64 device buffer holding one field's whole storage.
class DeviceField {
 public:
  explicit DeviceField(std::size_t count);
  explicit DeviceField(const StructuredField& host);  // allocates and uploads
  DeviceField(const DeviceField&) = delete;
  DeviceField& operator=(const DeviceField&) = delete;
  DeviceField(DeviceField&& other) noexcept;
  ~DeviceField();
  [[nodiscard]] double* get() const noexcept { return ptr_; }
  void upload(const StructuredField& host);
  // synchronises the context stream, then copies the whole storage back
  void download(StructuredField& host) const;

 private:
  double* ptr_ = nullptr;
  std::size_t count_ = 0;
};

}  // namespace solver::hip_backend

namespace solver::hip_backend {

// Precision of solve_pressure_poisson / project_velocity called through the C++ interface
// (their reference signatures carry none): OF_PRECISION_FP64 (default) or OF_PRECISION_MIXED.
void set_pressure_precision(int precision);
int pressure_precision();

}  // namespace solver::hip_backend
