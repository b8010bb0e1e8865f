// Shared device/host helpers of the gfx950 projection-step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "of_hip.h"

namespace ofh {

// ------------------------------------------------------------------ errors
struct InvalidArgument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct RuntimeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define OFH_CHECK(call)                                                                                   \
  do {                                                                                                    \
    hipError_t e_ = (call);                                                                               \
    if(e_ != hipSuccess)                                                                                  \
      throw ::ofh::HipError(std::string(#call) + " failed: " + hipGetErrorString(e_));                    \
  } while(0)

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Release paths (destructors and the *_destroy entry points) must not throw. They record the
// first HIP call that failed with OFH_RELEASE and go on releasing; release_take() then clears
// HIP's per-thread last error and hands the record to the entry point that ran them, which
// reports it. A failed release therefore never surfaces later as the error of an unrelated
// launch.
void release_note(hipError_t e, const char* call);
#define OFH_RELEASE(call) ::ofh::release_note((call), #call)
// The first recorded release failure ("" if none); resets the record and HIP's last error.
std::string release_take();

// Every kernel launch goes through OFH_LAUNCH. With OFH_CHECK_LAUNCHES=1 in the environment
// the launch status is read right after each launch, so a bad launch names its kernel; with
// OFH_CHECK_LAUNCHES=2 the device is also synchronised after each launch, so a kernel that
// fails while running is named too (slow: for finding a failure, not for timing).
int launch_check_level();
void launch_check(const char* kernel);
#define OFH_LAUNCH(kernel, ...)                                                                           \
  do {                                                                                                    \
    hipLaunchKernelGGL(kernel, __VA_ARGS__);                                                              \
    if(::ofh::launch_check_level() > 0) ::ofh::launch_check(#kernel);                                     \
  } while(0)

// ------------------------------------------------------------------ layout
// Ghosted storage of core/field_layout.hpp:69-290 with one ghost layer.
struct Lay {
  int ax, ay, az;  // active extents
  int sx, sy, sz;  // storage extents
  __host__ __device__ inline size_t at(int i, int j, int k) const {
    return (size_t)i + (size_t)sx * ((size_t)j + (size_t)sy * (size_t)k);
  }
  __host__ __device__ inline size_t size() const { return (size_t)sx * sy * sz; }
  __host__ __device__ inline int a(int d) const { return d == 0 ? ax : d == 1 ? ay : az; }
  __host__ __device__ inline int s(int d) const { return d == 0 ? sx : d == 1 ? sy : sz; }
};

__host__ __device__ inline Lay make_lay(int nx, int ny, int nz, int loc) {
  Lay L;
  L.ax = nx + (loc == OF_FACE_X);
  L.ay = ny + (loc == OF_FACE_Y);
  L.az = nz + (loc == OF_FACE_Z);
  L.sx = L.ax + 2;
  L.sy = L.ay + 2;
  L.sz = L.az + 2;
  return L;
}
__host__ __device__ inline Lay make_lay(const of_grid& g, int loc) { return make_lay(g.nx, g.ny, g.nz, loc); }

__host__ __device__ inline double spacing(const of_grid& g, int axis) {
  return axis == 0 ? g.dx : axis == 1 ? g.dy : g.dz;
}

inline void validate_grid(const of_grid* g) {
  if(g == nullptr) throw InvalidArgument("grid must not be null");
  if(g->nx <= 0 || g->ny <= 0 || g->nz <= 0) throw InvalidArgument("grid dimensions must be positive");
  if(!(g->dx > 0.0) || !(g->dy > 0.0) || !(g->dz > 0.0)) throw InvalidArgument("grid spacing must be positive");
}

// ------------------------------------------------------------------ launch shapes
// 2D thread blocks 64 (i) x 4 (j) over one k-plane per grid z: a wave reads 64 consecutive
// doubles of a row (coalesced 512-byte segments).
constexpr int kBX = 64;
constexpr int kBY = 4;
constexpr int kThreads = kBX * kBY;

inline dim3 grid_for(int ex, int ey, int ez) {
  return dim3((unsigned)((ex + kBX - 1) / kBX), (unsigned)((ey + kBY - 1) / kBY), (unsigned)ez);
}

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ------------------------------------------------------------------ deterministic reductions
// Each block of a reducing kernel stores one partial per value (store_partials); a one-block
// kernel launched next on the same stream combines them in a fixed order (k_finalize) and runs
// the reduction's epilogue. The kernel boundary orders the partial stores before the combine,
// so no device-scope fence or completion counter is needed: a device-scope release in every
// block writes back the XCD-private L2 and made streaming kernels 4x slower on this part.
// Results are bitwise reproducible for a given launch shape.
constexpr int kMaxRed = 8;           // values per reduction
constexpr int kMaxPartials = 65536;  // partial slots per value
constexpr int kFinalizeThreads = 1024;

enum RedOp : int { kSum = 0, kMax = 1 };

// DPP lane moves of a double (its two 32-bit halves moved alike).
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// The value of a double in one lane (uniform).
__device__ inline double lane_value(double v, int lane) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Sum (op kSum) or maximum (kMax) of a double over a full wave, the same value in every lane:
// butterfly exchanges within each row of 16 lanes by DPP moves (lane ^ 1, lane ^ 2, half-row
// mirror, row mirror: every lane of a row then holds the row's result, the same in each since the
// operations commute exactly), then the four row results, read by lane and combined in a fixed
// order. No LDS round trip: a shuffle tree through the LDS permute unit takes about ten times as
// long. Every lane of the wave must be active.
__device__ inline double wave_reduce_all(double v, int op) {
  auto f = [op](double a, double b) { return op == kSum ? a + b : fmax(a, b); };
  v = f(v, dpp_move<0xB1>(v));   // quad_perm [1,0,3,2]: lane ^ 1
  v = f(v, dpp_move<0x4E>(v));   // quad_perm [2,3,0,1]: lane ^ 2
  v = f(v, dpp_move<0x141>(v));  // row_half_mirror: lane i <-> 7 - i in each half-row
  v = f(v, dpp_move<0x140>(v));  // row_mirror: lane i <-> 15 - i in each row
  return f(f(lane_value(v, 0), lane_value(v, 16)), f(lane_value(v, 32), lane_value(v, 48)));
}
__device__ inline double wave_sum_all(double v) { return wave_reduce_all(v, kSum); }
// The value held by lane ^ 1, by a DPP quad permute ([1,0,3,2]): an exact move, so the value of
// __shfl_xor(v, 1, 64) without a trip through the LDS permute unit. Every lane must be active.
__device__ inline double lane_xor1(double v) { return dpp_move<0xB1>(v); }
__device__ inline float lane_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
// Shuffle-tree reductions for blocks whose last wave is partial (result valid in lane 0).
__device__ inline double wave_sum(double v) {
  for(int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for(int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Block-wide reduction of NV values, result valid in thread 0. ops[k] selects sum or max.
// Blocks of up to 1024 threads. Each wave reduces its values, then wave 0 reduces the wave
// results (one per lane, the identity in the lanes past the last wave): DPP moves when the block
// is whole waves, shuffle trees otherwise.
template <int NV>
__device__ inline void block_reduce(double (&v)[NV], const int (&ops)[NV]) {
  __shared__ double red[16][NV];
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthreads = blockDim.x * blockDim.y * blockDim.z;
  const int lane = tid & 63, wave = tid >> 6;
  const int nwaves = (nthreads + 63) >> 6;
  const bool full = (nthreads & 63) == 0;
  auto reduce = [full](double x, int op) {
    return full ? wave_reduce_all(x, op) : (op == kSum ? wave_sum(x) : wave_max(x));
  };
#pragma unroll
  for(int k = 0; k < NV; ++k) v[k] = reduce(v[k], ops[k]);
  if(lane == 0) {
#pragma unroll
    for(int k = 0; k < NV; ++k) red[wave][k] = v[k];
  }
  __syncthreads();
  if(wave == 0) {
#pragma unroll
    for(int k = 0; k < NV; ++k) {
      const double x = red[min(lane, nwaves - 1)][k];
      v[k] = reduce(lane < nwaves ? x : (ops[k] == kSum ? 0.0 : -INFINITY), ops[k]);
    }
  }
  __syncthreads();
}

// Reduces v over the block and stores the block's partials at partials[k * kMaxPartials + bid].
template <int NV>
__device__ inline void store_partials(double (&v)[NV], const int (&ops)[NV], double* __restrict__ partials) {
  block_reduce<NV>(v, ops);
  if(threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) {
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for(int k = 0; k < NV; ++k) partials[(size_t)k * kMaxPartials + bid] = v[k];
  }
}

// Combines nb partials per value in a fixed order (one block); result valid in thread 0.
// Thread t folds partials t, t + blockDim.x, ... in that order. They are read BATCH at a time
// by unconditional loads at clamped indices and the out-of-range ones skipped when folded, so a
// batch costs one memory round trip (a load under the loop's bound test was waited for before
// the next iteration's loads were issued: eight round trips for 8192 partials). NT > 0 is the
// block's thread count (blockDim.x otherwise, which is read from the dispatch packet: one more
// round trip before the first partial can be requested).
template <int NV, int BATCH = (NV <= 4 ? 8 : 4), int NT = 0>
__device__ inline void fold_partials(const double* __restrict__ partials, unsigned nb, const int (&ops)[NV],
                                     double (&out)[NV]) {
#pragma unroll
  for(int k = 0; k < NV; ++k) out[k] = ops[k] == kSum ? 0.0 : -INFINITY;
  const unsigned nt = NT > 0 ? (unsigned)NT : blockDim.x;
  const unsigned last = nb > 0 ? nb - 1 : 0;
  for(unsigned base = threadIdx.x; base < nb; base += BATCH * nt) {
    double x[BATCH][NV];
#pragma unroll
    for(int u = 0; u < BATCH; ++u) {
      const unsigned b = min(base + u * nt, last);
#pragma unroll
      for(int k = 0; k < NV; ++k) x[u][k] = partials[(size_t)k * kMaxPartials + b];
    }
#pragma unroll
    for(int u = 0; u < BATCH; ++u) {
      const bool in = base + u * nt < nb;
#pragma unroll
      for(int k = 0; k < NV; ++k) {
        const double f = ops[k] == kSum ? out[k] + x[u][k] : fmax(out[k], x[u][k]);
        out[k] = in ? f : out[k];
      }
    }
  }
  block_reduce<NV>(out, ops);
}

// Finalize kernel: Epi provides NV, op(k), active() and operator()(res) (run by thread 0).
template <class Epi>
__global__ void __launch_bounds__(kFinalizeThreads) k_finalize(const double* __restrict__ partials, unsigned nb,
                                                               Epi epi) {
  if(!epi.active()) return;
  int ops[Epi::NV];
#pragma unroll
  for(int k = 0; k < Epi::NV; ++k) ops[k] = epi.op(k);
  double res[Epi::NV];
  fold_partials<Epi::NV, (Epi::NV <= 4 ? 8 : 4), kFinalizeThreads>(partials, nb, ops, res);
  if(threadIdx.x == 0) epi(res);
}

template <class Epi>
inline void launch_finalize(hipStream_t s, const double* partials, dim3 grid, const Epi& epi) {
  OFH_LAUNCH(k_finalize<Epi>, dim3(1), dim3(kFinalizeThreads), 0, s, partials, grid.x * grid.y * grid.z,
                     epi);
}

// Epilogue that stores the NV results at out[0..NV-1]; bit k of max_mask selects max for value k.
template <int N>
struct EpiStore {
  static constexpr int NV = N;
  double* out;
  unsigned max_mask;
  __device__ int op(int k) const { return (max_mask >> k) & 1u ? kMax : kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&r)[N]) const {
#pragma unroll
    for(int k = 0; k < N; ++k) out[k] = r[k];
  }
};

// ------------------------------------------------------------------ XCD-aware block order
// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each with a private
// L2. Remapping the linear block id makes each XCD own a contiguous run of tiles, so the
// halo planes shared by neighbouring tiles mostly hit the same L2.
__device__ inline unsigned xcd_remap(unsigned bid, unsigned nblocks) {
  constexpr unsigned kXcd = 8;
  const unsigned per = nblocks / kXcd;
  if(per == 0) return bid;
  const unsigned full = per * kXcd;
  if(bid >= full) return bid;
  return (bid % kXcd) * per + bid / kXcd;
}

// The block of a 3D grid after the XCD-aware remap of its linear id (x fastest): blocks of
// neighbouring rows and planes, which read each other's boundary rows, then mostly run on one XCD
// and share its L2.
__device__ inline uint3 remapped_block3() {
  unsigned r = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                         gridDim.x * gridDim.y * gridDim.z);
  uint3 b;
  b.x = r % gridDim.x;
  r /= gridDim.x;
  b.y = r % gridDim.y;
  b.z = r / gridDim.y;
  return b;
}

}  // namespace ofh
