// Pressure Poisson solve that allocates nothing: of_solve_pressure_poisson with every device
// array carved from a fixed per-device scratch block (include/of_hip.h,
// of_solve_pressure_poisson_noalloc). No reference counterpart: it serves callers that cannot
// allocate device memory.

#include "ofh.hpp"

#include <mutex>

namespace ofh {
namespace {

constexpr size_t kScratchBytes = OF_NOALLOC_SCRATCH_BYTES;

// The block itself. In host code this name is only a stand-in for the symbol, not a device
// address: every host-side use goes through the address hipGetSymbolAddress returns for the
// current device (scratch_on).
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (the current device), looked up once per device.
char* scratch_on(int device) {
  if(device < 0 || device >= kMaxDevices) throw InvalidArgument("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if(g_scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    OFH_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if(ptr == nullptr) throw HipError("hipGetSymbolAddress returned no address for the scratch block");
    g_scratch_ptr[device] = ptr;
  }
  return static_cast<char*>(g_scratch_ptr[device]);
}

}  // namespace
}  // namespace ofh

using namespace ofh;

extern "C" int of_solve_pressure_poisson_noalloc(int device, void* stream, const of_grid* g,
                                                 const of_pressure_bc_set* b, const of_projection_options* o,
                                                 int precision, const double* rhs, double* p,
                                                 of_poisson_diagnostics* diag) {
  return guarded_call([&] {
    validate_grid(g);
    if(b == nullptr) throw InvalidArgument("bcs must not be null");
    if(o == nullptr) throw InvalidArgument("options must not be null");
    if(rhs == nullptr || p == nullptr) throw InvalidArgument("rhs and p must not be null");
    if(o->poisson_max_iterations <= 0) throw InvalidArgument("solve_pressure_poisson expects iterations > 0");
    if(!(o->poisson_tolerance > 0.0)) throw InvalidArgument("solve_pressure_poisson expects tolerance > 0");
    const double cells = (double)g->nx * (double)g->ny * (double)g->nz;
    if(cells > (double)OF_NOALLOC_MAX_CELLS)
      throw InvalidArgument("solve_pressure_poisson_noalloc supports grids of at most " +
                            std::to_string(OF_NOALLOC_MAX_CELLS) + " cells");
    OFH_CHECK(hipSetDevice(device));
    DeviceArena arena;
    arena.base = scratch_on(device);
    arena.cap = kScratchBytes;
    // a context whose reduction buffers live in the scratch block; it frees nothing
    Context x;
    x.device = device;
    x.stream = static_cast<hipStream_t>(stream);
    x.borrowed = true;
    x.partials = static_cast<double*>(arena.take(sizeof(double) * kMaxRed * kMaxPartials));
    x.scal = static_cast<double*>(arena.take(sizeof(double) * kScalCount));
    PoissonSolver* s = poisson_create_in(x, *g, *b, precision, arena);
    PoissonResult r;
    try {
      r = poisson_solve(s, rhs, p, o->poisson_max_iterations, o->poisson_tolerance, false);
    } catch(...) {
      poisson_destroy(s);
      throw;
    }
    poisson_destroy(s);
    if(diag) *diag = r.diag;
  });
}
