// Device context, buffers and descriptor conversion for the C++ interface.
#include "solver/hip_backend.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <stdexcept>
#include <string>

namespace solver::hip_backend {

namespace {

// The context lives until process exit on purpose: tearing it down from a static destructor
// could run after the HIP runtime itself has been torn down.
of_context* g_ctx = nullptr;
int g_device = 0;
int g_precision = OF_PRECISION_FP64;

void hip_check(hipError_t e, const char* what) {
  if(e != hipSuccess) throw std::runtime_error(std::string(what) + " failed: " + hipGetErrorString(e));
}

}  // namespace

void check(const int status) {
  if(status == OF_OK) return;
  const char* msg = of_last_error();
  const std::string text = msg != nullptr ? msg : "unknown error";
  if(status == OF_INVALID_ARGUMENT) throw std::invalid_argument(text);
  throw std::runtime_error(text);
}

of_context* context() {
  if(g_ctx == nullptr) {
    const char* env = std::getenv("SOLVER_HIP_DEVICE");
    g_device = env != nullptr ? std::atoi(env) : 0;
    check(of_context_create(g_device, nullptr, &g_ctx));
  }
  return g_ctx;
}

void synchronize() { check(of_context_synchronize(context())); }

std::string device_name() {
  context();
  hipDeviceProp_t prop{};
  hip_check(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties");
  std::string name = prop.name;
  if(name.empty()) name = prop.gcnArchName;
  return name;
}

void set_pressure_precision(const int precision) {
  if(precision != OF_PRECISION_FP64 && precision != OF_PRECISION_MIXED)
    throw std::invalid_argument("unknown pressure-solve precision");
  g_precision = precision;
}

int pressure_precision() { return g_precision; }

of_grid to_abi(const Grid& grid) {
  if(grid.ghost_layers != 1) throw std::invalid_argument("the HIP backend supports grids with one ghost layer");
  return of_grid{grid.nx, grid.ny, grid.nz, grid.dx, grid.dy, grid.dz};
}

of_bc_set to_abi(const BoundaryConditionSet& bcs) {
  of_bc_set out{};
  for(int f = 0; f < 6; ++f) {
    const BoundaryCondition& c = bcs.faces[static_cast<std::size_t>(f)];
    out.faces[f].type = static_cast<int>(c.type);
    for(int a = 0; a < 3; ++a) out.faces[f].velocity[a] = c.velocity[static_cast<std::size_t>(a)];
    out.faces[f].pressure = c.pressure;
    out.faces[f].pressure_gradient = c.pressure_gradient;
  }
  return out;
}

of_pressure_bc_set to_abi(const PressureBoundarySet& bcs) {
  of_pressure_bc_set out{};
  for(int f = 0; f < 6; ++f) {
    const PressureBoundaryCondition& c = bcs.faces[static_cast<std::size_t>(f)];
    out.faces[f].type = static_cast<int>(c.type);
    out.faces[f].value = c.value;
    out.faces[f].gradient = c.gradient;
  }
  return out;
}

DeviceField::DeviceField(const std::size_t count) : count_(count) {
  context();
  hip_check(hipMalloc(reinterpret_cast<void**>(&ptr_), count_ * sizeof(double)), "hipMalloc");
  hip_check(hipMemset(ptr_, 0, count_ * sizeof(double)), "hipMemset");
  // the library's kernels run on a non-blocking stream, which the legacy stream's memset does
  // not order against
  hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

DeviceField::DeviceField(const StructuredField& host) : DeviceField(host.size()) { upload(host); }

DeviceField::DeviceField(DeviceField&& other) noexcept : ptr_(other.ptr_), count_(other.count_) {
  other.ptr_ = nullptr;
  other.count_ = 0;
}

DeviceField::~DeviceField() {
  // a destructor cannot report; clear HIP's last error so that a failed free does not surface
  // as the error of a later, unrelated call
  if(ptr_ != nullptr && hipFree(ptr_) != hipSuccess) (void)hipGetLastError();
}

void DeviceField::upload(const StructuredField& host) {
  if(host.size() != count_) throw std::invalid_argument("device buffer size does not match the field");
  synchronize();  // the context stream may still read or write this buffer
  hip_check(hipMemcpy(ptr_, host.data(), count_ * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
}

void DeviceField::download(StructuredField& host) const {
  if(host.size() != count_) throw std::invalid_argument("device buffer size does not match the field");
  synchronize();
  hip_check(hipMemcpy(host.data(), ptr_, count_ * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
}

}  // namespace solver::hip_backend
