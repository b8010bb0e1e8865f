// Conjugate-gradient Helmholtz predictor: (1 - a dxx)(1 - a dyy)(1 - a dzz) u* = rhs per velocity
// component, solved iteratively instead of by the exact line solves. This is the alternative the
// reference's Metal backend uses (metal/taylor_green_backend.metal:258-274 for the operator,
// metal/taylor_green_backend.mm:546-621 for the unpreconditioned CG): the initial guess is the
// right-hand side and the stop test is sqrt(r.r / r0.r0) <= tol. The Metal code also stops when
// r0.r0 or |d.q| falls to machine epsilon in absolute terms, which ends a solve of a small
// right-hand side before it reaches tol (r0 = b - A b is O(a/h^2) |b|); here the guards are
// scale-free: r0 = 0 is converged, and d.q <= 0 (impossible for the symmetric positive definite
// operator unless d = 0) stops the solve unconverged. It is defined for grids periodic on every
// axis, where each component's unknowns form a periodic nx x ny x nz grid (the duplicated normal
// face is not an unknown; it is copied from face 0 afterwards, as the line solves do).
//
// Device formulation: the three components iterate in lockstep inside the same launches (one
// CG per component, with its own scalars), on compact i-fastest arrays. The operator is applied
// in its factorised form, z then y then x, by threads that march k with the 3 x 3 columns of
// d around their own in registers. All scalars stay on the device; the host learns that every
// component has stopped from a mapped flag, one iteration behind.

#include "ofh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace ofh {

bool axis_periodic(const of_bc_set& b, int axis);

namespace {

constexpr int kHcgKZ = 8;
constexpr int kHcgSlots = 8;  // scalars per component
enum HS : int { H_RR = 0, H_RR0, H_ALPHA, H_BETA, H_DONE, H_CONV, H_ITER, H_REL };

struct Hcg {
  int nx, ny, nz;
  double bx, by, bz;  // a / h^2 per axis
  __host__ __device__ size_t n() const { return (size_t)nx * ny * nz; }
};

struct HcgTile {
  int i, j, k0, k1, comp;
  bool in;
};
__device__ inline HcgTile hcg_tile(const Hcg& H) {
  const int gx = (H.nx + 63) / 64, gy = (H.ny + 3) / 4;
  unsigned r = blockIdx.x;
  const int bx = (int)(r % (unsigned)gx);
  r /= (unsigned)gx;
  const int by = (int)(r % (unsigned)gy);
  const int bz = (int)(r / (unsigned)gy);
  HcgTile t;
  t.i = bx * 64 + threadIdx.x;
  t.j = by * 4 + threadIdx.y;
  t.k0 = bz * kHcgKZ;
  t.k1 = min(t.k0 + kHcgKZ, H.nz);
  t.comp = blockIdx.y;
  t.in = t.i < H.nx && t.j < H.ny;
  return t;
}

__device__ inline bool hcg_done(const double* hs, int comp) { return hs[comp * kHcgSlots + H_DONE] != 0.0; }

// Stores v (this block's component's value) at the component's partial slot and zero at the
// other two, so that a three-value fold yields one sum per component. Blocks of a component
// that has stopped still store (zeros): every slot of the fold is written by every launch.
__device__ inline void hcg_partials(double v, int comp, double* partials) {
  double vals[3] = {comp == 0 ? v : 0.0, comp == 1 ? v : 0.0, comp == 2 ? v : 0.0};
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(vals, ops, partials);
}

// Compact unknowns of each component from its ghosted storage (active cells 1..n per axis).
__global__ void k_hcg_gather(const double* __restrict__ u, const double* __restrict__ v,
                             const double* __restrict__ w, Lay Lu, Lay Lv, Lay Lw, Hcg H, double* __restrict__ x) {
  const HcgTile t = hcg_tile(H);
  if(!t.in) return;
  const double* f = t.comp == 0 ? u : t.comp == 1 ? v : w;
  const Lay& L = t.comp == 0 ? Lu : t.comp == 1 ? Lv : Lw;
  double* xc = x + (size_t)t.comp * H.n();
  for(int k = t.k0; k < t.k1; ++k)
    xc[(size_t)t.i + (size_t)H.nx * (t.j + (size_t)H.ny * k)] = f[L.at(1 + t.i, 1 + t.j, 1 + k)];
}

// Compact solution back into the ghosted storage; the normal component's duplicated face (index
// n + 1 along its own axis) copies face 0.
__global__ void k_hcg_scatter(const double* __restrict__ x, Hcg H, double* u, double* v, double* w, Lay Lu, Lay Lv,
                              Lay Lw) {
  const HcgTile t = hcg_tile(H);
  if(!t.in) return;
  double* f = t.comp == 0 ? u : t.comp == 1 ? v : w;
  const Lay& L = t.comp == 0 ? Lu : t.comp == 1 ? Lv : Lw;
  const double* xc = x + (size_t)t.comp * H.n();
  for(int k = t.k0; k < t.k1; ++k) {
    const double val = xc[(size_t)t.i + (size_t)H.nx * (t.j + (size_t)H.ny * k)];
    f[L.at(1 + t.i, 1 + t.j, 1 + k)] = val;
    if(t.comp == 0 && t.i == 0) f[L.at(1 + H.nx, 1 + t.j, 1 + k)] = val;
    if(t.comp == 1 && t.j == 0) f[L.at(1 + t.i, 1 + H.ny, 1 + k)] = val;
    if(t.comp == 2 && k == 0) f[L.at(1 + t.i, 1 + t.j, 1 + H.nz)] = val;
  }
}

// q = (1 - bx dxx)(1 - by dyy)(1 - bz dzz) d, periodic; DOT: sums of d.q per component.
template <bool DOT>
__global__ void __launch_bounds__(256) k_hcg_apply(const double* __restrict__ d, double* __restrict__ q, Hcg H,
                                                   double* partials, const double* hs) {
  const HcgTile t = hcg_tile(H);
  const bool skip = hcg_done(hs, t.comp);
  if(skip && !DOT) return;
  double acc = 0.0;
  if(t.in && !skip) {
    const size_t N = H.n(), plane = (size_t)H.nx * H.ny;
    const double* dc = d + (size_t)t.comp * N;
    double* qc = q + (size_t)t.comp * N;
    const int im = t.i > 0 ? t.i - 1 : H.nx - 1, ip = t.i + 1 < H.nx ? t.i + 1 : 0;
    const int jm = t.j > 0 ? t.j - 1 : H.ny - 1, jp = t.j + 1 < H.ny ? t.j + 1 : 0;
    const int ci[3] = {im, t.i, ip}, cj[3] = {jm, t.j, jp};
    size_t col[3][3];
#pragma unroll
    for(int a = 0; a < 3; ++a)
#pragma unroll
      for(int b = 0; b < 3; ++b) col[a][b] = (size_t)ci[a] + (size_t)H.nx * cj[b];
    auto kw = [&](int k) { return k < 0 ? k + H.nz : (k >= H.nz ? k - H.nz : k); };
    double w0[3][3], w1[3][3], w2[3][3];  // planes k-1, k, k+1 of the 3 x 3 columns
    const size_t pm = plane * (size_t)kw(t.k0 - 1), p0 = plane * (size_t)t.k0;
#pragma unroll
    for(int a = 0; a < 3; ++a)
#pragma unroll
      for(int b = 0; b < 3; ++b) {
        w0[a][b] = dc[col[a][b] + pm];
        w1[a][b] = dc[col[a][b] + p0];
      }
    for(int k = t.k0; k < t.k1; ++k) {
      const size_t pp = plane * (size_t)kw(k + 1);
#pragma unroll
      for(int a = 0; a < 3; ++a)
#pragma unroll
        for(int b = 0; b < 3; ++b) w2[a][b] = dc[col[a][b] + pp];
      double tz[3][3];
#pragma unroll
      for(int a = 0; a < 3; ++a)
#pragma unroll
        for(int b = 0; b < 3; ++b) tz[a][b] = w1[a][b] - H.bz * (w2[a][b] - 2.0 * w1[a][b] + w0[a][b]);
      double ty[3];
#pragma unroll
      for(int a = 0; a < 3; ++a) ty[a] = tz[a][1] - H.by * (tz[a][2] - 2.0 * tz[a][1] + tz[a][0]);
      const double qv = ty[1] - H.bx * (ty[2] - 2.0 * ty[1] + ty[0]);
      qc[col[1][1] + plane * (size_t)k] = qv;
      if(DOT) acc += w1[1][1] * qv;
#pragma unroll
      for(int a = 0; a < 3; ++a)
#pragma unroll
        for(int b = 0; b < 3; ++b) {
          w0[a][b] = w1[a][b];
          w1[a][b] = w2[a][b];
        }
    }
  }
  if(DOT) hcg_partials(acc, t.comp, partials);
}

// r = x - q (x holds the right-hand side, the initial guess), d = r; sums of r.r.
__global__ void __launch_bounds__(256) k_hcg_residual0(const double* __restrict__ x, const double* __restrict__ q,
                                                       double* __restrict__ r, double* __restrict__ d, Hcg H,
                                                       double* partials) {
  const HcgTile t = hcg_tile(H);
  double acc = 0.0;
  if(t.in) {
    const size_t N = H.n(), off = (size_t)t.comp * N;
    for(int k = t.k0; k < t.k1; ++k) {
      const size_t id = off + (size_t)t.i + (size_t)H.nx * (t.j + (size_t)H.ny * k);
      const double rv = x[id] - q[id];
      r[id] = rv;
      d[id] = rv;
      acc += rv * rv;
    }
  }
  hcg_partials(acc, t.comp, partials);
}

// x += alpha d, r -= alpha q; sums of r.r.
__global__ void __launch_bounds__(256) k_hcg_update(double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ d, const double* __restrict__ q,
                                                    Hcg H, double* partials, const double* hs) {
  const HcgTile t = hcg_tile(H);
  const bool skip = hcg_done(hs, t.comp);
  const double alpha = hs[t.comp * kHcgSlots + H_ALPHA];
  double acc = 0.0;
  if(t.in && !skip) {
    const size_t N = H.n(), off = (size_t)t.comp * N;
    for(int k = t.k0; k < t.k1; ++k) {
      const size_t id = off + (size_t)t.i + (size_t)H.nx * (t.j + (size_t)H.ny * k);
      x[id] = x[id] + alpha * d[id];
      const double rv = r[id] - alpha * q[id];
      r[id] = rv;
      acc += rv * rv;
    }
  }
  hcg_partials(acc, t.comp, partials);
}

// d = r + beta d.
__global__ void __launch_bounds__(256) k_hcg_direction(const double* __restrict__ r, double* __restrict__ d, Hcg H,
                                                       const double* hs) {
  const HcgTile t = hcg_tile(H);
  if(hcg_done(hs, t.comp)) return;
  const double beta = hs[t.comp * kHcgSlots + H_BETA];
  if(!t.in) return;
  const size_t N = H.n(), off = (size_t)t.comp * N;
  for(int k = t.k0; k < t.k1; ++k) {
    const size_t id = off + (size_t)t.i + (size_t)H.nx * (t.j + (size_t)H.ny * k);
    d[id] = r[id] + beta * d[id];
  }
}

// ---- epilogues (one-block finalize; per component)
struct EpiHcgInit {  // r0.r0; an exact initial guess (r0 = 0) needs no iteration
  static constexpr int NV = 3;
  double* hs;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&rr)[3]) const {
    for(int c = 0; c < 3; ++c) {
      double* h = hs + c * kHcgSlots;
      h[H_RR] = rr[c];
      h[H_RR0] = rr[c];
      h[H_ITER] = 0.0;
      h[H_REL] = 0.0;
      const bool tiny = !(rr[c] > 0.0);
      h[H_DONE] = tiny ? 1.0 : 0.0;
      h[H_CONV] = tiny ? 1.0 : 0.0;
      if(!tiny) h[H_REL] = 1.0;
    }
  }
};

struct EpiHcgAlpha {  // alpha = r.r / d.q, or stop unconverged when d.q <= 0 (or not finite)
  static constexpr int NV = 3;
  double* hs;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&dq)[3]) const {
    for(int c = 0; c < 3; ++c) {
      double* h = hs + c * kHcgSlots;
      if(h[H_DONE] != 0.0) continue;
      if(!(dq[c] > 0.0) || !isfinite(dq[c])) h[H_DONE] = 1.0;
      else h[H_ALPHA] = h[H_RR] / dq[c];
    }
  }
};

struct EpiHcgBeta {  // convergence test, beta; publishes "every component stopped" to the host
  static constexpr int NV = 3;
  double* hs;
  double* host_out;
  double tol;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&rr)[3]) const {
    double all = 1.0;
    for(int c = 0; c < 3; ++c) {
      double* h = hs + c * kHcgSlots;
      if(h[H_DONE] == 0.0) {
        h[H_ITER] += 1.0;
        const double rel = sqrt(rr[c] / h[H_RR0]);
        h[H_REL] = rel;
        if(rel <= tol) {
          h[H_CONV] = 1.0;
          h[H_DONE] = 1.0;
        } else {
          h[H_BETA] = rr[c] / h[H_RR];
          h[H_RR] = rr[c];
        }
      }
      if(h[H_DONE] == 0.0) all = 0.0;
    }
    __hip_atomic_store(host_out, all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
};

}  // namespace

// Work arrays of one grid (x, r, d, q for three components) and the scalar block; owned by a
// stepper or created per call.
struct HelmholtzCG {
  Context& c;
  Hcg H{};
  double *x = nullptr, *r = nullptr, *d = nullptr, *q = nullptr, *hs = nullptr;
  double* host_flag = nullptr;  // mapped pinned, two slots
  double* host_flag_dev = nullptr;
  hipEvent_t ev[2]{};
  HelmholtzCG(Context& ctx, const of_grid& g) : c(ctx) {
    H.nx = g.nx;
    H.ny = g.ny;
    H.nz = g.nz;
    const size_t bytes = 3 * H.n() * sizeof(double);
    OFH_CHECK(hipMalloc(&x, bytes));
    OFH_CHECK(hipMalloc(&r, bytes));
    OFH_CHECK(hipMalloc(&d, bytes));
    OFH_CHECK(hipMalloc(&q, bytes));
    OFH_CHECK(hipMalloc(&hs, 3 * kHcgSlots * sizeof(double)));
    OFH_CHECK(hipHostMalloc(&host_flag, 2 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    OFH_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&host_flag_dev), host_flag, 0));
    for(auto& e : ev) OFH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  ~HelmholtzCG() {
    OFH_RELEASE(hipStreamSynchronize(c.stream));
    for(double* p : {x, r, d, q, hs})
      if(p) OFH_RELEASE(hipFree(p));
    if(host_flag) OFH_RELEASE(hipHostFree(host_flag));
    for(auto& e : ev)
      if(e) OFH_RELEASE(hipEventDestroy(e));
  }
};

HelmholtzCG* helmholtz_cg_create(Context& c, const of_grid& g) { return new HelmholtzCG(c, g); }
void helmholtz_cg_destroy(HelmholtzCG* h) { delete h; }

// Predictor solve out = A^-1 rhs (all-periodic grids). Returns the largest iteration count
// over the three components; throws when a component did not converge within max_iterations.
int predictor_cg(HelmholtzCG& S, const of_grid& g, const of_bc_set& b, double alpha, const double* ru,
                 const double* rv, const double* rw, double* ou, double* ov, double* ow, double tol,
                 int max_iterations, double* worst_rel) {
  Context& c = S.c;
  for(int a = 0; a < 3; ++a)
    if(!axis_periodic(b, a)) throw InvalidArgument("the CG Helmholtz predictor supports periodic boundaries");
  if(g.nx != S.H.nx || g.ny != S.H.ny || g.nz != S.H.nz) throw InvalidArgument("CG Helmholtz work arrays: grid mismatch");
  if(alpha < 0.0) throw InvalidArgument("solve_predictor_cg expects a non-negative alpha");
  if(max_iterations <= 0) throw InvalidArgument("solve_predictor_cg expects max_iterations > 0");
  Hcg& H = S.H;
  H.bx = alpha / (g.dx * g.dx);
  H.by = alpha / (g.dy * g.dy);
  H.bz = alpha / (g.dz * g.dz);
  const double* rhs[3] = {ru, rv, rw};
  double* out[3] = {ou, ov, ow};
  const Lay Lu = make_lay(g, OF_FACE_X), Lv = make_lay(g, OF_FACE_Y), Lw = make_lay(g, OF_FACE_Z);
  // the right-hand side as the line solves see it: copied with its ghost fill, which averages
  // the duplicated normal faces
  for(int a = 0; a < 3; ++a) {
    const Lay L = make_lay(g, OF_FACE_X + a);
    if(out[a] != rhs[a])
      OFH_CHECK(hipMemcpyAsync(out[a], rhs[a], L.size() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  }
  launch_velocity_bc(c, g, b, ou, ov, ow);
  const dim3 block(64, 4, 1);
  const unsigned tiles = (unsigned)(((H.nx + 63) / 64) * ((H.ny + 3) / 4) * ((H.nz + kHcgKZ - 1) / kHcgKZ));
  if(tiles * 3u > (unsigned)kMaxPartials) throw InvalidArgument("grid too large for the reduction buffer");
  const dim3 grid(tiles, 3, 1);
  OFH_LAUNCH(k_hcg_gather, grid, block, 0, c.stream, ou, ov, ow, Lu, Lv, Lw, H, S.x);
  OFH_CHECK(hipMemsetAsync(S.hs, 0, 3 * kHcgSlots * sizeof(double), c.stream));
  S.host_flag[0] = S.host_flag[1] = 0.0;
  OFH_LAUNCH((k_hcg_apply<false>), grid, block, 0, c.stream, S.x, S.q, H, c.partials, S.hs);
  OFH_LAUNCH(k_hcg_residual0, grid, block, 0, c.stream, S.x, S.q, S.r, S.d, H, c.partials);
  launch_finalize(c.stream, c.partials, grid, EpiHcgInit{S.hs});
  for(int it = 0; it < max_iterations; ++it) {
    {
      ProfScope prof_(c, "k_hcg_apply");
      OFH_LAUNCH((k_hcg_apply<true>), grid, block, 0, c.stream, S.d, S.q, H, c.partials, S.hs);
    }
    launch_finalize(c.stream, c.partials, grid, EpiHcgAlpha{S.hs});
    OFH_LAUNCH(k_hcg_update, grid, block, 0, c.stream, S.x, S.r, S.d, S.q, H, c.partials, S.hs);
    launch_finalize(c.stream, c.partials, grid, EpiHcgBeta{S.hs, S.host_flag_dev + (it % 2), tol});
    OFH_LAUNCH(k_hcg_direction, grid, block, 0, c.stream, S.r, S.d, H, S.hs);
    OFH_CHECK(hipEventRecord(S.ev[it % 2], c.stream));
    if(it > 0) {  // one iteration of look-ahead: stop once the previous iteration stopped all three
      OFH_CHECK(hipEventSynchronize(S.ev[(it - 1) % 2]));
      if(reinterpret_cast<volatile double*>(S.host_flag)[(it - 1) % 2] != 0.0) break;
    }
  }
  OFH_LAUNCH(k_hcg_scatter, grid, block, 0, c.stream, S.x, H, ou, ov, ow, Lu, Lv, Lw);
  launch_velocity_bc(c, g, b, ou, ov, ow);
  OFH_CHECK(hipGetLastError());
  double h[3 * kHcgSlots];
  OFH_CHECK(hipMemcpyAsync(h, S.hs, sizeof h, hipMemcpyDeviceToHost, c.stream));
  OFH_CHECK(hipStreamSynchronize(c.stream));
  int worst_it = 0;
  double worst = 0.0;
  for(int a = 0; a < 3; ++a) {
    const double* ha = h + a * kHcgSlots;
    worst_it = std::max(worst_it, (int)ha[H_ITER]);
    worst = std::max(worst, ha[H_REL]);
    if(ha[H_CONV] == 0.0) {
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "CG Helmholtz solve of component %d did not converge (iterations=%d, relative_residual=%g, "
                    "tolerance=%g)",
                    a, (int)ha[H_ITER], ha[H_REL], tol);
      throw RuntimeError(buf);
    }
  }
  if(worst_rel) *worst_rel = worst;
  return worst_it;
}

}  // namespace ofh
