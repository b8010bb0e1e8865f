// Multigrid-preconditioned CG for the pressure Poisson equation on the device
// (linsolve/poisson_solver.cpp:195-809).
//
// Algorithm (identical to the reference): A = -lap_h with homogeneous BCs; preconditioner = one
// V-cycle of damped Jacobi (omega = 2/3, 2 pre + 2 post sweeps), restriction = average of the
// 2x2x2 (or 2x2x1) children, prolongation = piecewise-constant injection-add, rediscretised
// coarse operators, dense Gaussian elimination on the coarsest level with the last row replaced
// by ones when the problem has a null space; PCG converges on RMS(r_k)/RMS(r_0) <= tol.
//
// Device formulation:
//  * PCG vectors and multigrid levels are compact (nx*ny*nz, i fastest, no ghosts); the
//    reference's operator never reads ghosts (periodic neighbours wrap inside the active range).
//  * Every level kernel runs on 64x4 thread tiles that march along k. Neighbour indices and
//    Neumann/Dirichlet images are resolved per axis once per thread; the operator and the
//    Jacobi weight use precomputed reciprocals (1/h^2, omega/D per boundary class), so the inner
//    loop has no divisions and no integer division.
//  * Mean removal (null space) is deferred: a kernel that produces a vector also reduces its
//    sum, the last block stores the mean, and every consumer subtracts it when it reads the
//    vector. The values consumed are the values the reference stores after subtract_active_mean.
//  * Fusions: the first two pre-smoothing sweeps (the first starts from x = 0 and is pointwise);
//    residual + restriction (2x2x2 children combined with lane shuffles, LDS and the k-march);
//    prolongation + first post-smoothing sweep; the search-direction update + operator + dot
//    product; the residual update recomputes q = A d from the stored direction instead of
//    reading q back.
//  * Pressure accumulation: every search direction of a solve is kept (a ring of buffers) with
//    its alpha and mean, and one pass at the end adds them to p in iteration order. Each cell
//    sees exactly the operations of a per-iteration update, so p is bitwise the same, but it is
//    read and written once per solve instead of once per iteration.
//  * All scalars (alpha, beta, means, norms, convergence flags) stay on the device. Kernels of
//    iterations past convergence return immediately (flag check), so the host queues the next
//    iteration before it has read the previous one's residual.
//  * Precision: FP64, or MIXED = FP32 V-cycle (levels, smoothing, residual, transfers) inside
//    the FP64 CG (residual, dot products, pressure update). The FP32 V-cycle's two sweeps per
//    leg use Chebyshev weights (chebyshev_pair) instead of the reference's 2/3.
//  * Reductions are deterministic (fixed launch shape, fixed combination order).

#include "ofh.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace ofh {

namespace {

constexpr double kOmega = 2.0 / 3.0;  // linsolve/poisson_solver.hpp:17-23
// Smoother of the FP32 V-cycle in MIXED solves. The mixed mode is held to the outer tolerance,
// not to the reference's iteration counts, so its two Jacobi sweeps per leg use the weights of
// the degree-2 Chebyshev polynomial on [a, 2] (the roots of T2 mapped to that interval), which
// damps the modes with eigenvalues of D^-1 A in [a, 2] -- the modes the coarse levels cannot
// represent -- far better than one repeated weight: a = 0.4 on three-dimensional grids (the
// smallest such eigenvalue of the 7-point operator is 1/3) and 0.5 on two-dimensional ones
// (1/2 for the 5-point operator). On the 256^3 Taylor-Green step this takes 10 outer
// iterations per solve against 23 with the reference's 2/3 and 15 with the best single weight
// (0.87); on the Dirichlet, Neumann, mixed and periodic test families it never needs more
// iterations than the FP64 solve. The pair's polynomial (1 - w1 t)(1 - w2 t) lies in (-1, 1)
// for t in (0, 2], so the smoother converges and the V-cycle stays symmetric positive definite.
constexpr double kMixedCheb3D = 0.4, kMixedCheb2D = 0.5;
void chebyshev_pair(double a, double& w1, double& w2) {
  const double c = 0.5 * (2.0 + a), h = 0.5 * (2.0 - a) / std::sqrt(2.0);
  w1 = 1.0 / (c - h);
  w2 = 1.0 / (c + h);
}
constexpr double kSmallNumber = 1.0e-30;

// Poisson scalar block (device doubles)
enum PS : int {
  P_DONE = 0,      // 1 once converged or broken down
  P_CONV,          // converged flag
  P_ITER,          // iterations completed
  P_R0,            // initial residual RMS
  P_RMS,           // latest residual RMS
  P_REL,           // latest relative residual
  P_RZ,            // r.z of the current direction
  P_ALPHA,
  P_BETA,
  P_DQ,
  P_MR,            // mean of r
  P_MD,            // mean of the current direction d
  P_MQ,            // mean of q
  P_MP,            // mean of the warm start p
  P_MSYS,          // removed rhs mean
  P_PFINAL,        // final mean of p
  P_LEVEL = 32,    // per level: 8 slots (see LS)
  P_COUNT = 32 + 8 * 24
};
constexpr int kHostMirror = 8;  // host mirror slot stride; P_DONE is published after every residual update
constexpr int kHostSeq = 1;     // mirror slot of the sequence number (iteration + 1) of that update
enum LS : int { L_MB = 0, L_SBW, L_MX2, L_MX4, L_MX5, L_SUMB, L_SPARE1, L_SPARE2 };

struct PLevel {
  int nx, ny, nz;
  int n;
  double hx2, hy2, hz2;     // h*h per axis
  double ihx2, ihy2, ihz2;  // 1/(h*h)
  double inv[3];            // 1/pow(h,2) (operator diagonal, linsolve/poisson_solver.cpp:283)
  int per[3];               // periodic axes
  int lo_dir[3], hi_dir[3]; // Dirichlet faces
  // actual (inhomogeneous) face data, used only for the boundary source term
  double lo_val[3], hi_val[3], lo_grad[3], hi_grad[3], h[3];
  double winv[64];          // omega / D for the 4x4x4 boundary classes
  double winv_sum;          // sum over cells of omega / D
  double wr;                // second-sweep weight ratio: the second pre-smoothing sweep and the
                            // first post-smoothing sweep use wr omega / D (1 in FP64)
  int rx, ry, rz;           // refinement ratio to the next coarser level
};

// boundary class of a coordinate: 0 interior, 1 first, 2 last, 3 single
__device__ __host__ inline int bclass(int co, int n) { return (co == 0 ? 1 : 0) | (co == n - 1 ? 2 : 0); }

// Coarse coordinate of a fine one: the refinement ratio of every axis is 1 or 2, so the
// division is a shift (coordinates are non-negative).
__device__ inline int coarse_of(int co, int ratio) { return co >> (ratio >> 1); }

// n / d for 0 <= n < 2^22 and d >= 1: a float reciprocal estimate is within 0.5 of n / d,
// so one correction step gives the exact quotient (an integer division is ~30 instructions).
__device__ inline int small_div(int n, int d) {
  int q = (int)((float)n * (1.0f / (float)d));
  const int r = n - q * d;
  q += r >= d ? 1 : (r < 0 ? -1 : 0);
  return q;
}

// Neighbour coordinate and sign along one axis (homogeneous BCs): periodic wrap inside the
// active range, Neumann -> the cell itself, Dirichlet -> minus the cell itself
// (linsolve/poisson_solver.cpp:195-242 with value = gradient = 0).
struct Nb {
  int cm, cp;  // coordinates of the lower / upper neighbour
  int sm, sp;  // +1 or -1
};
__device__ inline Nb nb_axis(int co, int n, int per, int lod, int hid) {
  Nb r;
  if(co > 0) {
    r.cm = co - 1;
    r.sm = 1;
  } else if(per) {
    r.cm = n - 1;
    r.sm = 1;
  } else {
    r.cm = co;
    r.sm = lod ? -1 : 1;
  }
  if(co < n - 1) {
    r.cp = co + 1;
    r.sp = 1;
  } else if(per) {
    r.cp = 0;
    r.sp = 1;
  } else {
    r.cp = co;
    r.sp = hid ? -1 : 1;
  }
  return r;
}

template <typename T>
__device__ inline T sgn(int s, T v) {
  return s < 0 ? -v : v;
}

// -lap with reciprocal spacings: -(e - 2c + w)/hx^2 - (n - 2c + s)/hy^2 - (f - 2c + b)/hz^2
template <typename T>
__device__ inline T lap_neg(T c, T w, T e, T s, T n, T b, T f, T ihx2, T ihy2, T ihz2) {
  const T c2 = T(2) * c;
  return -((e - c2 + w) * ihx2) - ((n - c2 + s) * ihy2) - ((f - c2 + b) * ihz2);
}

__device__ inline bool stopped(const double* ps) { return ps[P_DONE] != 0.0; }

struct Tile {
  int i, j, k0, k1;
};
__device__ inline Tile tile(const PLevel& L, int kz) {
  Tile t;
  t.i = blockIdx.x * blockDim.x + threadIdx.x;
  t.j = blockIdx.y * blockDim.y + threadIdx.y;
  t.k0 = blockIdx.z * kz;
  t.k1 = min(t.k0 + kz, L.nz);
  return t;
}

struct Col {  // per-thread column data
  Nb x, y;
  int cx, cy;  // boundary classes of (i, j)
  bool in;
};
__device__ inline Col column(const PLevel& L, const Tile& t) {
  Col c;
  c.in = t.i < L.nx && t.j < L.ny;
  const int i = min(t.i, L.nx - 1), j = min(t.j, L.ny - 1);
  c.x = nb_axis(i, L.nx, L.per[0], L.lo_dir[0], L.hi_dir[0]);
  c.y = nb_axis(j, L.ny, L.per[1], L.lo_dir[1], L.hi_dir[1]);
  c.cx = bclass(i, L.nx);
  c.cy = bclass(j, L.ny);
  return c;
}

// A applied to the zero field with the actual BCs (linsolve/poisson_solver.cpp:635-642).
__device__ inline double boundary_source(const PLevel& L, int i, int j, int k) {
  double nb[3][2];
  const int co[3] = {i, j, k};
  const int nn[3] = {L.nx, L.ny, L.nz};
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    nb[a][0] = 0.0;
    nb[a][1] = 0.0;
    if(co[a] == 0 && !L.per[a]) nb[a][0] = L.lo_dir[a] ? 2.0 * L.lo_val[a] - 0.0 : 0.0 - L.lo_grad[a] * L.h[a];
    if(co[a] == nn[a] - 1 && !L.per[a]) nb[a][1] = L.hi_dir[a] ? 2.0 * L.hi_val[a] - 0.0 : 0.0 + L.hi_grad[a] * L.h[a];
  }
  return -(nb[0][1] - 2.0 * 0.0 + nb[0][0]) / L.hx2 - (nb[1][1] - 2.0 * 0.0 + nb[1][0]) / L.hy2 -
         (nb[2][1] - 2.0 * 0.0 + nb[2][0]) / L.hz2;
}

// ---------------------------------------------------------------- per-cell formulas
// Shared by the tiled level kernels and the one-workgroup tail. Neighbour data (Nb) per axis;
// the deferred means are subtracted on read.

// UNI: every axis periodic, so the Jacobi weight is the same for every cell (winv[0]) and
// every neighbour sign is +1 (the values are those of the general path). IX: index type (int
// where the level has fewer than 2^31 cells).

// Fused first two pre-smoothing sweeps at (i,j,k): x1 = (b - mb) w/D - m1 (pointwise, from
// x0 = 0), x2 = x1 + w/D (b - mb - A x1).
template <typename T, bool UNI = false, typename IX = size_t>
__device__ inline T smooth12_cell(const PLevel& L, const T* __restrict__ b, T mb, T m1, int i, int j, int k,
                                  const Nb& x, const Nb& y, const Nb& z) {
  const IX plane = (IX)L.nx * L.ny;
  auto W = [&](int cx, int cy, int cz) { return UNI ? T(L.winv[0]) : T(L.winv[cx + 4 * (cy + 4 * cz)]); };
  auto S = [&](int sg, T v) { return UNI ? v : sgn(sg, v); };
  auto B = [&](int ii, int jj, int kk) { return b[(IX)ii + (IX)L.nx * jj + plane * kk] - mb; };
  const int cx = bclass(i, L.nx), cy = bclass(j, L.ny), cz = bclass(k, L.nz);
  const T wc = W(cx, cy, cz);
  const T bt = B(i, j, k);
  const T c = bt * wc - m1;
  const T xw = S(x.sm, B(x.cm, j, k) * W(bclass(x.cm, L.nx), cy, cz) - m1);
  const T xe = S(x.sp, B(x.cp, j, k) * W(bclass(x.cp, L.nx), cy, cz) - m1);
  const T xs = S(y.sm, B(i, y.cm, k) * W(cx, bclass(y.cm, L.ny), cz) - m1);
  const T xn = S(y.sp, B(i, y.cp, k) * W(cx, bclass(y.cp, L.ny), cz) - m1);
  const T xb = S(z.sm, B(i, j, z.cm) * W(cx, cy, bclass(z.cm, L.nz)) - m1);
  const T xf = S(z.sp, B(i, j, z.cp) * W(cx, cy, bclass(z.cp, L.nz)) - m1);
  const T ax = lap_neg<T>(c, xw, xe, xs, xn, xb, xf, T(L.ihx2), T(L.ihy2), T(L.ihz2));
  return c + (wc * T(L.wr)) * (bt + T(-1) * ax);
}

// Fine residual (b - mb) - A (x2 - mx) at (i,j,k).
template <typename T, bool UNI = false, typename IX = size_t>
__device__ inline T residual_cell(const PLevel& L, const T* __restrict__ b, const T* __restrict__ x2, T mb, T mx,
                                  int i, int j, int k, const Nb& x, const Nb& y, const Nb& z) {
  const IX plane = (IX)L.nx * L.ny;
  auto S = [&](int sg, T v) { return UNI ? v : sgn(sg, v); };
  auto X = [&](int ii, int jj, int kk) { return x2[(IX)ii + (IX)L.nx * jj + plane * kk] - mx; };
  const T c = X(i, j, k);
  const T ax = lap_neg<T>(c, S(x.sm, X(x.cm, j, k)), S(x.sp, X(x.cp, j, k)), S(y.sm, X(i, y.cm, k)),
                          S(y.sp, X(i, y.cp, k)), S(z.sm, X(i, j, z.cm)), S(z.sp, X(i, j, z.cp)), T(L.ihx2),
                          T(L.ihy2), T(L.ihz2));
  return (b[(IX)i + (IX)L.nx * j + plane * k] - mb) + T(-1) * ax;
}

// Prolongation-add fused with the first post-smoothing sweep:
// x3 = (x2 - m2) + (xc - mc)[parent]; returns x4 = x3 + w/D (b - mb - A x3).
template <typename T, bool UNI = false, typename IX = size_t>
__device__ inline T prolong_smooth_cell(const PLevel& L, const PLevel& Cl, const T* __restrict__ b,
                                        const T* __restrict__ x2, const T* __restrict__ xc, T mb, T m2, T mc, int i,
                                        int j, int k, const Nb& x, const Nb& y, const Nb& z) {
  const IX plane = (IX)L.nx * L.ny;
  const IX cplane = (IX)Cl.nx * Cl.ny;
  auto S = [&](int sg, T v) { return UNI ? v : sgn(sg, v); };
  auto X3 = [&](int ii, int jj, int kk) {
    return (x2[(IX)ii + (IX)L.nx * jj + plane * kk] - m2) +
           (xc[(IX)coarse_of(ii, L.rx) + (IX)Cl.nx * coarse_of(jj, L.ry) + cplane * coarse_of(kk, L.rz)] - mc);
  };
  const T wc = UNI ? T(L.winv[0]) : T(L.winv[bclass(i, L.nx) + 4 * (bclass(j, L.ny) + 4 * bclass(k, L.nz))]);
  const T c = X3(i, j, k);
  const T ax = lap_neg<T>(c, S(x.sm, X3(x.cm, j, k)), S(x.sp, X3(x.cp, j, k)), S(y.sm, X3(i, y.cm, k)),
                          S(y.sp, X3(i, y.cp, k)), S(z.sm, X3(i, j, z.cm)), S(z.sp, X3(i, j, z.cp)), T(L.ihx2),
                          T(L.ihy2), T(L.ihz2));
  return c + (wc * T(L.wr)) * ((b[(IX)i + (IX)L.nx * j + plane * k] - mb) + T(-1) * ax);
}

// Second post-smoothing sweep: x4' = x4 - m4, returns x5 = x4' + w/D (bt - A x4'), bt = b - mb.
template <typename T, bool UNI = false, typename IX = size_t>
__device__ inline T smooth_last_cell(const PLevel& L, const T* __restrict__ b, const T* __restrict__ x4, T mb, T m4,
                                     int i, int j, int k, const Nb& x, const Nb& y, const Nb& z, T& bt) {
  const IX plane = (IX)L.nx * L.ny;
  auto S = [&](int sg, T v) { return UNI ? v : sgn(sg, v); };
  auto X = [&](int ii, int jj, int kk) { return x4[(IX)ii + (IX)L.nx * jj + plane * kk] - m4; };
  const T wc = UNI ? T(L.winv[0]) : T(L.winv[bclass(i, L.nx) + 4 * (bclass(j, L.ny) + 4 * bclass(k, L.nz))]);
  const T c = X(i, j, k);
  const T ax = lap_neg<T>(c, S(x.sm, X(x.cm, j, k)), S(x.sp, X(x.cp, j, k)), S(y.sm, X(i, y.cm, k)),
                          S(y.sp, X(i, y.cp, k)), S(z.sm, X(i, j, z.cm)), S(z.sp, X(i, j, z.cp)), T(L.ihx2),
                          T(L.ihy2), T(L.ihz2));
  bt = b[(IX)i + (IX)L.nx * j + plane * k] - mb;
  return c + wc * (bt + T(-1) * ax);
}

__device__ inline Nb nb_x(const PLevel& L, int i) { return nb_axis(i, L.nx, L.per[0], L.lo_dir[0], L.hi_dir[0]); }
__device__ inline Nb nb_y(const PLevel& L, int j) { return nb_axis(j, L.ny, L.per[1], L.lo_dir[1], L.hi_dir[1]); }
__device__ inline Nb nb_z(const PLevel& L, int k) { return nb_axis(k, L.nz, L.per[2], L.lo_dir[2], L.hi_dir[2]); }

// ---------------------------------------------------------------- reduction epilogues
// Run by the one-block finalize kernel after the producer. An epilogue is skipped once the
// solve has stopped (its producer returned early and stored no partials) unless noted.
struct EpiSetup1 {  // means of the system rhs and of the warm start
  static constexpr int NV = 2;
  double* ps;
  double n;
  int zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&r)[2]) const {
    ps[P_MSYS] = zm ? r[0] / n : 0.0;
    ps[P_MP] = zm ? r[1] / n : 0.0;
  }
};

struct EpiSetup2 {  // initial residual, early exit (linsolve/poisson_solver.cpp:703-725)
  static constexpr int NV = 3;
  double* ps;
  double* history;
  double n, tol;
  int zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&r)[3]) const {
    const double mr = zm ? r[0] / n : 0.0;
    ps[P_MR] = mr;
    ps[P_LEVEL + L_MB] = mr;
    ps[P_LEVEL + L_SBW] = r[2];
    const double rms = sqrt(fmax(r[1] - n * mr * mr, 0.0) / n);
    ps[P_R0] = rms;
    ps[P_RMS] = rms;
    ps[P_REL] = rms > 0.0 ? 1.0 : 0.0;
    ps[P_ITER] = 0.0;
    history[0] = rms;
    if(rms <= tol) {
      ps[P_DONE] = 1.0;
      ps[P_CONV] = 1.0;
      ps[P_REL] = 0.0;
    } else {
      ps[P_DONE] = 0.0;
      ps[P_CONV] = 0.0;
    }
  }
};

struct EpiMean {  // ps[idx] = mean of the produced vector
  static constexpr int NV = 1;
  double* ps;
  int idx;
  double n;
  int zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return !stopped(ps); }
  __device__ void operator()(const double (&r)[1]) const { ps[idx] = zm ? r[0] / n : 0.0; }
};

struct EpiRestrict {  // coarse rhs: mean, sum of b w/D, sum
  static constexpr int NV = 2;
  double* ps;
  int cslot;
  double nc;
  int zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return !stopped(ps); }
  __device__ void operator()(const double (&r)[2]) const {
    ps[cslot + L_MB] = zm ? r[0] / nc : 0.0;
    ps[cslot + L_SBW] = r[1];
    ps[cslot + L_SUMB] = r[0];
  }
};

struct EpiSmoothLast {  // mean of x5; on the finest level r.z and beta
  static constexpr int NV = 2;
  double* ps;
  int lslot;
  double n;
  int zm, finest;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return !stopped(ps); }
  __device__ void operator()(const double (&r)[2]) const {
    ps[lslot + L_MX5] = zm ? r[0] / n : 0.0;
    if(finest) {
      const double rz_old = ps[P_RZ];
      ps[P_BETA] = rz_old != 0.0 ? r[1] / rz_old : 0.0;
      ps[P_RZ] = r[1];
    }
  }
};

struct EpiDir {  // means of d and q, d.q, alpha or breakdown (linsolve/poisson_solver.cpp:752-762)
  static constexpr int NV = 3;
  double* ps;
  double n;
  int zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return !stopped(ps); }
  __device__ void operator()(const double (&r)[3]) const {
    const double mdn = zm ? r[0] / n : 0.0;
    const double mq = zm ? r[1] / n : 0.0;
    ps[P_MD] = mdn;
    ps[P_MQ] = mq;
    const double dq = r[2] - mdn * r[1];
    ps[P_DQ] = dq;
    if(fabs(dq) < kSmallNumber) {
      ps[P_DONE] = 1.0;  // breakdown: stop without convergence
    } else {
      ps[P_ALPHA] = ps[P_RZ] / dq;
    }
  }
};

// Residual RMS and convergence test (linsolve/poisson_solver.cpp:764-776). Always runs: it
// also publishes the stop flag to the mapped host mirror, from which the host learns that the
// solve has stopped. An update that runs records its iteration's alpha and direction mean for
// the pressure accumulation (coef[2 it], coef[2 it + 1]).
struct EpiUpdate {
  static constexpr int NV = 3;
  double* ps;
  double* history;
  double* coef;
  double* host_out;
  double n, tol;
  int it, zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&r)[3]) const {
    if(!stopped(ps)) {
      const double m = zm ? r[0] / n : 0.0;
      ps[P_MR] = m;
      ps[P_LEVEL + L_MB] = m;
      ps[P_LEVEL + L_SBW] = r[2];
      const double rms = sqrt(fmax(r[1] - n * m * m, 0.0) / n);
      ps[P_ITER] = (double)(it + 1);
      ps[P_RMS] = rms;
      coef[2 * it] = ps[P_ALPHA];
      coef[2 * it + 1] = ps[P_MD];
      const double rel = ps[P_R0] > 0.0 ? rms / ps[P_R0] : 0.0;
      ps[P_REL] = rel;
      history[it + 1] = rms;
      if(rel <= tol) {
        ps[P_CONV] = 1.0;
        ps[P_DONE] = 1.0;
      }
    }
    // the host only polls the stop flag and this update's sequence number (stored after the
    // flag, with release order); the other scalars are copied back once at the end
    __hip_atomic_store(host_out + P_DONE, ps[P_DONE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(host_out + kHostSeq, (double)(it + 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
};

struct EpiPfinal {  // final mean of p
  static constexpr int NV = 1;
  double* ps;
  double n;
  int zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&r)[1]) const {
    ps[P_PFINAL] = zm ? r[0] / n : 0.0;
  }
};

// ---------------------------------------------------------------- setup
// sys = -rhs - src (active cells), p_c = warm start; sums for their means.
__global__ void k_setup1(const double* __restrict__ rhs_g, const double* __restrict__ p_g, Lay Lc, PLevel L,
                         double* __restrict__ sys, double* __restrict__ pc, double* partials, int kz) {
  const Tile t = tile(L, kz);
  double v[2] = {0.0, 0.0};
  if(t.i < L.nx && t.j < L.ny) {
    for(int k = t.k0; k < t.k1; ++k) {
      const size_t id = (size_t)t.i + (size_t)L.nx * ((size_t)t.j + (size_t)L.ny * k);
      const size_t gi = Lc.at(t.i + 1, t.j + 1, k + 1);
      const double s = -rhs_g[gi] + -1.0 * boundary_source(L, t.i, t.j, k);
      const double p = p_g[gi];
      sys[id] = s;
      pc[id] = p;
      v[0] += s;
      v[1] += p;
    }
  }
  const int ops[2] = {kSum, kSum};
  store_partials<2>(v, ops, partials);
}

// k_setup1 for a zero warm start: p is neither read nor copied (the pressure accumulation
// starts from zero).
// Sum of the system rhs sys = -rhs - src for its mean (a zero warm start, whose mean is 0). sys
// is not stored: k_setup2_zero forms it again from rhs by the same expression (one read of rhs
// instead of a write of sys and its read back).
__global__ void k_setup1_zero(const double* __restrict__ rhs_g, Lay Lc, PLevel L, double* partials, int kz) {
  const Tile t = tile(L, kz);
  double v[2] = {0.0, 0.0};
  if(t.i < L.nx && t.j < L.ny) {
    for(int k = t.k0; k < t.k1; ++k) {
      const double s = -rhs_g[Lc.at(t.i + 1, t.j + 1, k + 1)] + -1.0 * boundary_source(L, t.i, t.j, k);
      v[0] += s;
    }
  }
  const int ops[2] = {kSum, kSum};
  store_partials<2>(v, ops, partials);
}

// k_setup2 for a zero warm start: A(p - m_p) = A 0 = 0 exactly, so r = sys - m_sys pointwise
// (the same values as the stencil path), sys formed again from rhs as k_setup1_zero forms it.
// lo (the mixed mode) also receives r rounded to T: the rhs of the finest FP32 level.
template <typename T>
__global__ void k_setup2_zero(const double* __restrict__ rhs_g, Lay Lc, double* __restrict__ r, T* __restrict__ lo,
                              PLevel L, double* partials, const double* ps, int kz) {
  const double msys = ps[P_MSYS];
  const Tile t = tile(L, kz);
  double v[3] = {0.0, 0.0, 0.0};
  if(t.i < L.nx && t.j < L.ny) {
    const int cxy = bclass(t.i, L.nx) + 4 * bclass(t.j, L.ny);
    for(int k = t.k0; k < t.k1; ++k) {
      const size_t id = (size_t)t.i + (size_t)L.nx * ((size_t)t.j + (size_t)L.ny * k);
      const double s = -rhs_g[Lc.at(t.i + 1, t.j + 1, k + 1)] + -1.0 * boundary_source(L, t.i, t.j, k);
      const double rv = s - msys;
      r[id] = rv;
      if(lo != nullptr) lo[id] = T(rv);
      v[0] += rv;
      v[1] += rv * rv;
      v[2] += rv * L.winv[cxy + 16 * bclass(k, L.nz)];
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

// r = (sys - m_sys) - A(p - m_p); sums of r, r^2, r w/D.
__global__ void k_setup2(double* __restrict__ r, const double* __restrict__ pc, PLevel L, double* partials,
                         const double* ps, int kz) {
  const double msys = ps[P_MSYS], mp = ps[P_MP];
  const Tile t = tile(L, kz);
  const Col C = column(L, t);
  double v[3] = {0.0, 0.0, 0.0};
  if(C.in) {
    const size_t plane = (size_t)L.nx * L.ny;
    for(int k = t.k0; k < t.k1; ++k) {
      const size_t id = (size_t)t.i + (size_t)L.nx * t.j + plane * k;
      const double rv = residual_cell<double>(L, r, pc, msys, mp, t.i, t.j, k, C.x, C.y, nb_z(L, k));
      r[id] = rv;
      v[0] += rv;
      v[1] += rv * rv;
      v[2] += rv * L.winv[C.cx + 4 * (C.cy + 4 * bclass(k, L.nz))];
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

// ---------------------------------------------------------------- tiled level kernels
// 64x4 thread tiles over (i, j); each thread marches KZ consecutive k (unrolled) and keeps the
// k-1 / k / k+1 values of the stencil operand in registers. The 1D grid is remapped so that
// each XCD owns a contiguous range of tiles (neighbouring tiles share halo rows in its L2).
// UNI: every axis periodic, so omega/D is one constant and no boundary image changes sign.
struct HotTile {
  int i, j, k0, k1;
  bool in;
};
template <int KZ>
__device__ inline HotTile hot_tile(const PLevel& L) {
  const int gx = (L.nx + 63) >> 6, gy = (L.ny + 3) >> 2;
  unsigned r = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = (int)(r % (unsigned)gx);
  r /= (unsigned)gx;
  const int by = (int)(r % (unsigned)gy);
  const int bz = (int)(r / (unsigned)gy);
  HotTile t;
  t.i = bx * 64 + threadIdx.x;
  t.j = by * 4 + threadIdx.y;
  t.k0 = bz * KZ;
  t.k1 = min(t.k0 + KZ, L.nz);
  t.in = t.i < L.nx && t.j < L.ny;
  return t;
}

// z-window bookkeeping for a column: index of the k+1 neighbour (k itself past a
// non-periodic top face, where the value is unused) and the boundary images.
template <bool UNI>
struct ZWin {
  int nz;
  bool per;
  int slo, shi;  // sign of the image below k = 0 / above k = nz - 1 (non-periodic)
  __device__ ZWin(const PLevel& L) : nz(L.nz), per(UNI || L.per[2] != 0), slo(L.lo_dir[2] ? -1 : 1), shi(L.hi_dir[2] ? -1 : 1) {}
  __device__ int kprev(int k0) const { return k0 > 0 ? k0 - 1 : (per ? nz - 1 : 0); }
  __device__ int knext(int k) const { return k + 1 < nz ? k + 1 : (per ? 0 : k); }
  template <typename T>
  __device__ T below(int k, T prev, T cur) const { return (!per && k == 0) ? sgn(slo, cur) : prev; }
  template <typename T>
  __device__ T above(int k, T next, T cur) const { return (!per && k == nz - 1) ? sgn(shi, cur) : next; }
};

template <typename T, bool UNI>
__device__ inline T wdiag(const PLevel& L, int cx, int cy, int cz) {
  return UNI ? T(L.winv[0]) : T(L.winv[cx + 4 * (cy + 4 * cz)]);
}

template <typename T, int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_smooth12(const T* __restrict__ b, T* __restrict__ x2, PLevel L,
                                                  double* partials, const double* ps, int lslot, int zm) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m1 = zm ? T((ps[lslot + L_SBW] - ps[lslot + L_MB] * L.winv_sum) / (double)L.n) : T(0);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const HotTile t = hot_tile<KZ>(L);
  double v[1] = {0.0};
  if(t.in) {
    const ZWin<UNI> zw(L);
    const size_t plane = (size_t)L.nx * L.ny;
    const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
    const Nb x = nb_x(L, t.i), y = nb_y(L, t.j);
    const int cx = bclass(t.i, L.nx), cy = bclass(t.j, L.ny);
    const int cxm = bclass(x.cm, L.nx), cxp = bclass(x.cp, L.nx), cym = bclass(y.cm, L.ny), cyp = bclass(y.cp, L.ny);
    const size_t oxm = (size_t)x.cm + (size_t)L.nx * t.j, oxp = (size_t)x.cp + (size_t)L.nx * t.j;
    const size_t oym = (size_t)t.i + (size_t)L.nx * y.cm, oyp = (size_t)t.i + (size_t)L.nx * y.cp;
    auto X1 = [&](size_t id, T w) { return (b[id] - mb) * w - m1; };
    const int kp0 = zw.kprev(t.k0);
    T prev = X1(col + plane * kp0, wdiag<T, UNI>(L, cx, cy, bclass(kp0, L.nz)));
    int cz = bclass(t.k0, L.nz);
    T wc = wdiag<T, UNI>(L, cx, cy, cz);
    T btc = b[col + plane * t.k0] - mb;
    T cur = btc * wc - m1;
#pragma unroll
    for(int kk = 0; kk < KZ; ++kk) {
      const int k = t.k0 + kk;
      if(k < t.k1) {
        const int kn = zw.knext(k);
        const int czn = bclass(kn, L.nz);
        const T wn = wdiag<T, UNI>(L, cx, cy, czn);
        const T btn = b[col + plane * kn] - mb;
        const T next = btn * wn - m1;
        const size_t pk = plane * k;
        const T xw = sgn(x.sm, X1(oxm + pk, wdiag<T, UNI>(L, cxm, cy, cz)));
        const T xe = sgn(x.sp, X1(oxp + pk, wdiag<T, UNI>(L, cxp, cy, cz)));
        const T xs = sgn(y.sm, X1(oym + pk, wdiag<T, UNI>(L, cx, cym, cz)));
        const T xn = sgn(y.sp, X1(oyp + pk, wdiag<T, UNI>(L, cx, cyp, cz)));
        const T ax = lap_neg<T>(cur, xw, xe, xs, xn, zw.below(k, prev, cur), zw.above(k, next, cur), ihx2, ihy2, ihz2);
        const T out = cur + (wc * T(L.wr)) * (btc + T(-1) * ax);
        x2[col + pk] = out;
        v[0] += (double)out;
        prev = cur;
        cur = next;
        btc = btn;
        wc = wn;
        cz = czn;
      }
    }
  }
  const int ops[1] = {kSum};
  store_partials<1>(v, ops, partials);
}

// Residual of x2 restricted to the coarse level: bc = avg_children(b - A x2). Each thread
// computes the fine residual of its own cell; x-pairs are combined across lanes, y-pairs through
// LDS, z-pairs along the thread's k-march (KZ is even).
template <typename T, int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_residual_restrict(const T* __restrict__ b, const T* __restrict__ x2, PLevel L,
                                                           PLevel Cl, T* __restrict__ bc, double* partials,
                                                           const double* ps, int lslot, int zm) {
  if(stopped(ps)) return;
  __shared__ T ypair[4][64];
  const T mb = T(ps[lslot + L_MB]);
  const T mx = T(ps[lslot + L_MX2]);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const T inv_children = T(1.0 / (double)(L.rx * L.ry * L.rz));
  const HotTile t = hot_tile<KZ>(L);
  const ZWin<UNI> zw(L);
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t cplane = (size_t)Cl.nx * Cl.ny;
  const int ii = min(t.i, L.nx - 1), jj = min(t.j, L.ny - 1);
  const size_t col = (size_t)ii + (size_t)L.nx * jj;
  const Nb x = nb_x(L, ii), y = nb_y(L, jj);
  const size_t oxm = (size_t)x.cm + (size_t)L.nx * jj, oxp = (size_t)x.cp + (size_t)L.nx * jj;
  const size_t oym = (size_t)ii + (size_t)L.nx * y.cm, oyp = (size_t)ii + (size_t)L.nx * y.cp;
  auto X = [&](size_t id) { return x2[id] - mx; };
  T prev = X(col + plane * zw.kprev(t.k0));
  T cur = X(col + plane * t.k0);
  double v[2] = {0.0, 0.0};
  T acc = T(0);
#pragma unroll
  for(int kk = 0; kk < KZ; ++kk) {
    const int k = t.k0 + kk;
    if(k < t.k1) {
      const T next = X(col + plane * zw.knext(k));
      const size_t pk = plane * k;
      const T ax = lap_neg<T>(cur, sgn(x.sm, X(oxm + pk)), sgn(x.sp, X(oxp + pk)), sgn(y.sm, X(oym + pk)),
                              sgn(y.sp, X(oyp + pk)), zw.below(k, prev, cur), zw.above(k, next, cur), ihx2, ihy2, ihz2);
      T rr = t.in ? (b[col + pk] - mb) + T(-1) * ax : T(0);
      prev = cur;
      cur = next;
      if(L.rx == 2) rr += __shfl_xor(rr, 1, 64);  // x pair (lanes i, i+1)
      if(L.ry == 2) {                             // y pair (rows j, j+1 of the tile)
        ypair[threadIdx.y][threadIdx.x] = rr;
        __syncthreads();
        rr = rr + ypair[threadIdx.y ^ 1][threadIdx.x];
        __syncthreads();
      }
      const bool second = (kk & 1) != 0;
      acc = (L.rz == 2 && second) ? acc + rr : rr;
      const bool complete = L.rz == 1 || second;
      if(complete && t.in && (L.rx == 1 || (t.i & 1) == 0) && (L.ry == 1 || (t.j & 1) == 0)) {
        const int ic = coarse_of(t.i, L.rx), jc = coarse_of(t.j, L.ry), kc = coarse_of(k, L.rz);
        const T val = inv_children * acc;
        bc[(size_t)ic + (size_t)Cl.nx * jc + cplane * kc] = val;
        v[0] += (double)val;
        v[1] += (double)val *
                (UNI ? Cl.winv[0] : Cl.winv[bclass(ic, Cl.nx) + 4 * (bclass(jc, Cl.ny) + 4 * bclass(kc, Cl.nz))]);
      }
    }
  }
  const int ops[2] = {kSum, kSum};
  store_partials<2>(v, ops, partials);
}

// Prolongation-add fused with the first post-smoothing sweep:
// x3 = (x2 - m2) + (xc - mc)[parent]; x4 = x3 + w/D (b - mb - A x3).
template <typename T, int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_prolong_smooth(const T* __restrict__ b, const T* __restrict__ x2,
                                                        const T* __restrict__ xc, PLevel L, PLevel Cl,
                                                        T* __restrict__ x4, double* partials, const double* ps,
                                                        int lslot, int cslot) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m2 = T(ps[lslot + L_MX2]);
  const T mc = T(ps[cslot + L_MX5]);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const HotTile t = hot_tile<KZ>(L);
  double v[1] = {0.0};
  if(t.in) {
    const ZWin<UNI> zw(L);
    const size_t plane = (size_t)L.nx * L.ny;
    const size_t cplane = (size_t)Cl.nx * Cl.ny;
    const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
    const Nb x = nb_x(L, t.i), y = nb_y(L, t.j);
    const int cx = bclass(t.i, L.nx), cy = bclass(t.j, L.ny);
    // parent columns of the cell and of its x / y neighbours
    const size_t pc = (size_t)coarse_of(t.i, L.rx) + (size_t)Cl.nx * coarse_of(t.j, L.ry);
    const size_t pxm = (size_t)coarse_of(x.cm, L.rx) + (size_t)Cl.nx * coarse_of(t.j, L.ry);
    const size_t pxp = (size_t)coarse_of(x.cp, L.rx) + (size_t)Cl.nx * coarse_of(t.j, L.ry);
    const size_t pym = (size_t)coarse_of(t.i, L.rx) + (size_t)Cl.nx * coarse_of(y.cm, L.ry);
    const size_t pyp = (size_t)coarse_of(t.i, L.rx) + (size_t)Cl.nx * coarse_of(y.cp, L.ry);
    const size_t oxm = (size_t)x.cm + (size_t)L.nx * t.j, oxp = (size_t)x.cp + (size_t)L.nx * t.j;
    const size_t oym = (size_t)t.i + (size_t)L.nx * y.cm, oyp = (size_t)t.i + (size_t)L.nx * y.cp;
    auto X3 = [&](size_t fine, size_t coarse) { return (x2[fine] - m2) + (xc[coarse] - mc); };
    const int kp0 = zw.kprev(t.k0);
    T prev = X3(col + plane * kp0, pc + cplane * coarse_of(kp0, L.rz));
    T cur = X3(col + plane * t.k0, pc + cplane * coarse_of(t.k0, L.rz));
#pragma unroll
    for(int kk = 0; kk < KZ; ++kk) {
      const int k = t.k0 + kk;
      if(k < t.k1) {
        const int kn = zw.knext(k);
        const T next = X3(col + plane * kn, pc + cplane * coarse_of(kn, L.rz));
        const size_t pk = plane * k, cpk = cplane * coarse_of(k, L.rz);
        const T ax = lap_neg<T>(cur, sgn(x.sm, X3(oxm + pk, pxm + cpk)), sgn(x.sp, X3(oxp + pk, pxp + cpk)),
                                sgn(y.sm, X3(oym + pk, pym + cpk)), sgn(y.sp, X3(oyp + pk, pyp + cpk)),
                                zw.below(k, prev, cur), zw.above(k, next, cur), ihx2, ihy2, ihz2);
        const T wc = wdiag<T, UNI>(L, cx, cy, bclass(k, L.nz));
        const T out = cur + (wc * T(L.wr)) * ((b[col + pk] - mb) + T(-1) * ax);
        x4[col + pk] = out;
        v[0] += (double)out;
        prev = cur;
        cur = next;
      }
    }
  }
  const int ops[1] = {kSum};
  store_partials<1>(v, ops, partials);
}

// Second post-smoothing sweep; also sums of x5 and (b - mb) x5 (r.z on the finest level,
// where b is the CG residual and z = x5 - m5; the sum of r is zero up to rounding).
template <typename T, int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_smooth_last(const T* __restrict__ b, const T* __restrict__ x4, PLevel L,
                                                     T* __restrict__ x5, double* partials, const double* ps,
                                                     int lslot) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m4 = T(ps[lslot + L_MX4]);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const HotTile t = hot_tile<KZ>(L);
  double v[2] = {0.0, 0.0};
  if(t.in) {
    const ZWin<UNI> zw(L);
    const size_t plane = (size_t)L.nx * L.ny;
    const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
    const Nb x = nb_x(L, t.i), y = nb_y(L, t.j);
    const int cx = bclass(t.i, L.nx), cy = bclass(t.j, L.ny);
    const size_t oxm = (size_t)x.cm + (size_t)L.nx * t.j, oxp = (size_t)x.cp + (size_t)L.nx * t.j;
    const size_t oym = (size_t)t.i + (size_t)L.nx * y.cm, oyp = (size_t)t.i + (size_t)L.nx * y.cp;
    auto X = [&](size_t id) { return x4[id] - m4; };
    T prev = X(col + plane * zw.kprev(t.k0));
    T cur = X(col + plane * t.k0);
#pragma unroll
    for(int kk = 0; kk < KZ; ++kk) {
      const int k = t.k0 + kk;
      if(k < t.k1) {
        const T next = X(col + plane * zw.knext(k));
        const size_t pk = plane * k;
        const T ax = lap_neg<T>(cur, sgn(x.sm, X(oxm + pk)), sgn(x.sp, X(oxp + pk)), sgn(y.sm, X(oym + pk)),
                                sgn(y.sp, X(oyp + pk)), zw.below(k, prev, cur), zw.above(k, next, cur), ihx2, ihy2,
                                ihz2);
        const T wc = wdiag<T, UNI>(L, cx, cy, bclass(k, L.nz));
        const T bt = b[col + pk] - mb;
        const T out = cur + wc * (bt + T(-1) * ax);
        x5[col + pk] = out;
        v[0] += (double)out;
        v[1] += (double)bt * (double)out;
        prev = cur;
        cur = next;
      }
    }
  }
  const int ops[2] = {kSum, kSum};
  store_partials<2>(v, ops, partials);
}

// ---------------------------------------------------------------- folded reductions
// A consumer kernel can fold the partials of the kernel before it instead of a one-block
// finalize launch between them (for producers of at most a few thousand blocks): every block
// combines the partials in the same fixed order, so every block sees the same values, and
// block 0 stores the epilogue's scalars for the later kernels. The producer and the consumer
// use different partial buffers.
// NT: the calling block's thread count. Producers folded here have at most a few blocks per
// thread, so two partials per value are requested at a time.
// Folds a producer's nb partials per value into sums valid in every thread of the block (NT
// threads, whole waves, blocks of 64 x ROWS / 2 threads indexed by threadIdx.x + 64 threadIdx.y):
// thread t adds partials t, t + NT, ... (read two at a time at clamped indices), each wave sums
// its threads by DPP moves, and every thread adds the wave sums in wave order after one barrier
// (a block reduction and a broadcast took three).
template <int N, int NT>
__device__ inline void fold_all(const double* __restrict__ src, unsigned nb, double (&out)[N]) {
  static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
  constexpr int NW = NT / 64;
  __shared__ double fred[NW][N];
  const unsigned tid = threadIdx.x + blockDim.x * threadIdx.y;
  const unsigned last = nb > 0 ? nb - 1 : 0;
  double acc[N];
#pragma unroll
  for(int k = 0; k < N; ++k) acc[k] = 0.0;
  for(unsigned base = tid; base < nb; base += 2 * NT) {
    double x[2][N];
#pragma unroll
    for(int u = 0; u < 2; ++u) {
      const unsigned b = min(base + u * NT, last);
#pragma unroll
      for(int k = 0; k < N; ++k) x[u][k] = src[(size_t)k * kMaxPartials + b];
    }
#pragma unroll
    for(int u = 0; u < 2; ++u) {
      const bool in = base + u * NT < nb;
#pragma unroll
      for(int k = 0; k < N; ++k) acc[k] = in ? acc[k] + x[u][k] : acc[k];
    }
  }
#pragma unroll
  for(int k = 0; k < N; ++k) acc[k] = wave_sum_all(acc[k]);
  if((tid & 63) == 0) {
#pragma unroll
    for(int k = 0; k < N; ++k) fred[tid >> 6][k] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for(int k = 0; k < N; ++k) {
    double v = fred[0][k];
#pragma unroll
    for(int w = 1; w < NW; ++w) v = v + fred[w][k];
    out[k] = v;
  }
}

// Folded EpiDown of the finer level pl into level l (the coarse one): returns mb and the sum of
// b w/D of level l; block 0 stores what EpiDown stores.
struct DownFold {
  const double* src;  // partials of the finer level's k_down_fused (null: read ps)
  unsigned nb;
  int pslot;  // slot of the finer level
  double pn;  // cells of the finer level
};
template <int NT>
__device__ inline void down_fold(const DownFold& f, double* ps, int lslot, double n, int zm, double& mb,
                                 double& sbw) {
  if(f.src == nullptr) {
    mb = ps[lslot + L_MB];
    sbw = ps[lslot + L_SBW];
    return;
  }
  double r[3];
  fold_all<3, NT>(f.src, f.nb, r);
  mb = zm ? r[1] / n : 0.0;
  sbw = r[2];
  if(blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    ps[f.pslot + L_MX2] = zm ? r[0] / f.pn : 0.0;
    ps[lslot + L_MB] = mb;
    ps[lslot + L_SBW] = sbw;
    ps[lslot + L_SUMB] = r[1];
  }
}

// ---------------------------------------------------------------- fused all-periodic levels
// On a level that is periodic on every axis and halves every axis, the down leg (two
// pre-smoothing sweeps, residual, restriction) and the up leg (prolongation, two post-smoothing
// sweeps) each run as one kernel by temporal blocking in LDS: a 64 x ROWS tile (ROWS = 16 or 8,
// two rows per thread) marches KZ planes along k through three-plane rings (two for b - mb,
// which is read one step after it is written), producing each intermediate on the tile plus the halo its
// consumer needs (x1 on halo 2, x2 on halo 1, residual on the tile; x3 on halo 2, x4 on halo 1,
// x5 on the tile). Every value is computed by the same per-cell expression as the unfused
// kernels. The one difference: the periodic operator annihilates constants exactly, so the
// residual reads x2 and the last sweep reads x4 without their means subtracted (the means of
// intermediates are not known inside the kernel); those values differ from the unfused path by
// rounding only. The down kernel reads b once and writes x2 and the coarse rhs (17 B per cell in
// FP64, against 33 B for the two unfused kernels); the up kernel reads x2, b and the coarse
// correction and writes x5 (25 B against 49 B).
constexpr int kFX = 64;  // tile width; a tile has ROWS rows and 32 ROWS threads (two rows per thread)
constexpr int kMaxFoldPartials = 4096;  // producers with more blocks keep their finalize launch

__device__ inline int wrap_index(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ inline int ring3(int k) { return (k + 6) % 3; }  // k >= -6

struct FusedTile {
  int i0, j0, k0, k1;
};
template <int ROWS>
__device__ inline FusedTile fused_tile(const PLevel& L, int kz) {
  const int gx = (L.nx + kFX - 1) / kFX, gy = (L.ny + ROWS - 1) / ROWS;
  unsigned r = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = (int)(r % (unsigned)gx);
  r /= (unsigned)gx;
  const int by = (int)(r % (unsigned)gy);
  const int bz = (int)(r / (unsigned)gy);
  FusedTile t;
  t.i0 = bx * kFX;
  t.j0 = by * ROWS;
  t.k0 = bz * kz;
  t.k1 = min(t.k0 + kz, L.nz);
  return t;
}

// Global loads two planes ahead are issued while the current plane is computed (three register
// buffers rotate, so the loop is unrolled by three); each thread also carries the
// right-hand side of its own two tile cells in a three-plane register queue, so the LDS rings
// need only three planes. Two barriers per plane: after the LDS stores, and between the stage
// on the halo and the stage on the tile.
template <typename T, int ROWS>
__global__ void __launch_bounds__(32 * ROWS) k_down_fused(const T* __restrict__ b, T* __restrict__ x2g, PLevel L,
                                                          PLevel Cl, T* __restrict__ bc, double* partials,
                                                          double* ps, int lslot, int zm, int kz, DownFold fold) {
  if(stopped(ps)) return;
  constexpr int kThreads = 32 * ROWS;
  constexpr int BW = kFX + 4, BH = ROWS + 4, BP = BW * BH;  // x1, halo 2
  constexpr int XW = kFX + 2, XH = ROWS + 2, XP = XW * XH;  // b - mb and x2, halo 1
  constexpr int NB = (BP + kThreads - 1) / kThreads;
  constexpr int NH = (XP + kThreads - 1) / kThreads;
  __shared__ T sb[3][BP];   // x1 = (b - mb) w - m1, computed once per point
  __shared__ T sbt[2][XP];  // b - mb (read one step after it is written)
  __shared__ T sx[3][XP];
  const int tid = threadIdx.x;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  const size_t cplane = (size_t)Cl.nx * Cl.ny;
  const FusedTile t = fused_tile<ROWS>(L, kz);
  const int tx = tid & 63, ty = tid >> 6;
  const int gi = t.i0 + tx, gj0 = t.j0 + 2 * ty;
  const bool own0 = gi < nx && gj0 < ny, own1 = gi < nx && gj0 + 1 < ny;
  // per-thread element offsets of the halo-2 plane (and the matching halo-1 index, or -1)
  size_t off[NB];
  bool okb[NB];
  int e1[NB];
#pragma unroll
  for(int m = 0; m < NB; ++m) {
    const int e = tid + m * kThreads;
    const int ex = e % BW, ey = e / BW;
    const int gx = t.i0 + ex - 2, gy = t.j0 + ey - 2;
    okb[m] = e < BP && gx < nx + 2 && gy < ny + 2;
    off[m] = okb[m] ? (size_t)wrap_index(gx, nx) + (size_t)nx * wrap_index(gy, ny) : 0;
    e1[m] = e < BP && ex >= 1 && ex <= XW && ey >= 1 && ey <= XH ? (ex - 1) + XW * (ey - 1) : -1;
  }
  // per-thread points of the halo-1 stage: LDS index in the halo-2 plane, and the global
  // offset when the point is a tile cell inside the level (else -1)
  int hb[NH];
  long long hg[NH];
#pragma unroll
  for(int m = 0; m < NH; ++m) {
    const int e = tid + m * kThreads;
    const int xi = e % XW, xj = e / XW;
    hb[m] = (xi + 1) + BW * (xj + 1);
    const int ci = t.i0 + xi - 1, cj = t.j0 + xj - 1;
    hg[m] = e < XP && xi >= 1 && xi <= kFX && xj >= 1 && xj <= ROWS && ci < nx && cj < ny
                ? (long long)ci + (long long)nx * cj
                : -1;
  }
  const size_t own_off0 = own0 ? (size_t)gi + (size_t)nx * gj0 : 0, own_off1 = own1 ? own_off0 + nx : 0;
  // Loads are unconditional (masked-out elements read offset 0) and masked afterwards: a load
  // under a lane-divergent condition is compiled with a wait for its completion before the
  // branches rejoin, which would serialise every load of the plane.
  // The raw values stay in registers until the next plane's LDS store, so the loads are only
  // waited for after the current plane's compute.
  auto fetch = [&](int kp, T (&pre)[NB], T (&tb)[2]) {
    const size_t kz = plane * (size_t)wrap_index(kp, nz);
#pragma unroll
    for(int m = 0; m < NB; ++m) pre[m] = b[off[m] + kz];
    tb[0] = b[own_off0 + kz];
    tb[1] = b[own_off1 + kz];
  };
  // three register buffers rotate: a step stores plane kp and fetches plane kp + 2. The first
  // two planes are requested before the finer level's partials are folded, so the two round
  // trips overlap.
  T pa[NB], pb[NB], pc[NB], ta[2], tb[2], tc[2];
  fetch(t.k0 - 2, pa, ta);
  fetch(t.k0 - 1, pb, tb);
  double mbd, sbw;
  down_fold<kThreads>(fold, ps, lslot, (double)L.n, zm, mbd, sbw);
  const T mb = T(mbd);
  const T m1 = zm ? T((sbw - mbd * L.winv_sum) / (double)L.n) : T(0);
  const T w = T(L.winv[0]);
  const T w2 = w * T(L.wr);  // second sweep
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const T inv_children = T(1.0 / 8.0);
  double v[3] = {0.0, 0.0, 0.0};
  T acc = T(0);
  T q1[2] = {T(0), T(0)}, q2[2] = {T(0), T(0)};  // tile rhs of planes kp - 1, kp - 2
  const int last = (t.k1 - t.k0) + 1;
  // ALWAYS: the step is known to fetch (steady state), so the fetch is not under a branch. A
  // fetch under a branch makes the wait counts at the join assume the path without it, which
  // waits for every prefetched plane one step after it was requested.
  auto step = [&](int s, T (&cur)[NB], T (&ctb)[2], T (&nxt)[NB], T (&ntb)[2], auto always) {
    const int kp = t.k0 + s;
    {
      T* d1 = sb[ring3(kp)];
      T* dt = sbt[kp & 1];
#pragma unroll
      for(int m = 0; m < NB; ++m) {
        const int e = tid + m * kThreads;
        const T bt = okb[m] ? cur[m] - mb : T(0);
        if(e < BP) d1[e] = bt * w - m1;
        if(e1[m] >= 0) dt[e1[m]] = bt;
      }
    }
    const T tb0 = own0 ? ctb[0] - mb : T(0), tb1 = own1 ? ctb[1] - mb : T(0);
    __syncthreads();
    if(decltype(always)::value || s + 2 <= last) fetch(kp + 2, nxt, ntb);
    if(s >= 0) {  // x2 on plane kq = kp - 1, tile + halo 1
      const int kq = kp - 1;
      const T* x1m = sb[ring3(kq - 1)];
      const T* x1c = sb[ring3(kq)];
      const T* x1p = sb[ring3(kq + 1)];
      const T* btc = sbt[kq & 1];
      T* dst = sx[ring3(kq)];
      const bool own = kq >= t.k0 && kq < t.k1;
      const size_t pq = plane * (size_t)kq;
#pragma unroll
      for(int m = 0; m < NH; ++m) {
        const int e = tid + m * kThreads;
        if(e < XP) {
          const int be = hb[m];
          const T c = x1c[be];
          const T ax = lap_neg<T>(c, x1c[be - 1], x1c[be + 1], x1c[be - BW], x1c[be + BW], x1m[be], x1p[be], ihx2,
                                  ihy2, ihz2);
          const T out = c + w2 * (btc[e] + T(-1) * ax);
          dst[e] = out;
          if(own && hg[m] >= 0) {
            x2g[(size_t)hg[m] + pq] = out;
            v[0] += (double)out;
          }
        }
      }
    }
    __syncthreads();
    if(s >= 2) {  // residual on plane kr = kp - 2 (tile), restricted 2x2x2
      const int kr = kp - 2;
      const int sm = ring3(kr - 1), sc = ring3(kr), sp = ring3(kr + 1);
      const bool second = ((kr - t.k0) & 1) != 0;
      T xs[2];
#pragma unroll
      for(int rr = 0; rr < 2; ++rr) {
        const int xe = (tx + 1) + XW * (2 * ty + rr + 1);
        const T c = sx[sc][xe];
        const T ax = lap_neg<T>(c, sx[sc][xe - 1], sx[sc][xe + 1], sx[sc][xe - XW], sx[sc][xe + XW], sx[sm][xe],
                                sx[sp][xe], ihx2, ihy2, ihz2);
        T r = q2[rr] + T(-1) * ax;
        r += lane_xor1(r);  // x pair
        xs[rr] = r;
      }
      const T ysum = xs[0] + xs[1];  // y pair
      const T tot = second ? acc + ysum : ysum;
      acc = tot;
      if(second && (tx & 1) == 0 && own0) {
        const int ic = gi >> 1, jc = gj0 >> 1, kc = kr >> 1;
        const T val = inv_children * tot;
        bc[(size_t)ic + (size_t)Cl.nx * jc + cplane * kc] = val;
        v[1] += (double)val;
        v[2] += (double)val * Cl.winv[0];
      }
    }
    q2[0] = q1[0];
    q2[1] = q1[1];
    q1[0] = tb0;
    q1[1] = tb1;
  };
  int s = -2;
  for(; s + 4 <= last; s += 3) {  // every step of the group fetches
    step(s, pa, ta, pc, tc, std::true_type{});
    step(s + 1, pb, tb, pa, ta, std::true_type{});
    step(s + 2, pc, tc, pb, tb, std::true_type{});
  }
  // at most four steps remain; whether they fetch depends on the position
  if(s <= last) step(s, pa, ta, pc, tc, std::false_type{});
  if(s + 1 <= last) step(s + 1, pb, tb, pa, ta, std::false_type{});
  if(s + 2 <= last) step(s + 2, pc, tc, pb, tb, std::false_type{});
  if(s + 3 <= last) step(s + 3, pa, ta, pc, tc, std::false_type{});
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

// NBUF = 3: loads two planes ahead through three register buffers. NBUF = 2: one plane ahead
// through two, which frees registers for more resident blocks (the launch bound asks for 4
// waves per SIMD in FP64, two 512-thread blocks per compute unit, and 6 in FP32); the other
// blocks' work covers the shorter lead.
template <typename T, int ROWS, int NBUF = 3>
__global__ void __launch_bounds__(32 * ROWS, NBUF == 2 ? (sizeof(T) == 4 ? 6 : 4) : 1)
    k_up_fused(const T* __restrict__ b, const T* __restrict__ x2,
                                                        const T* __restrict__ xc, PLevel L, PLevel Cl,
                                                        T* __restrict__ x5g, double* partials, double* ps,
                                                        int lslot, int cslot, int kz, const double* fold_src,
                                                        unsigned fold_nb, int zm) {
  if(stopped(ps)) return;
  constexpr int kThreads = 32 * ROWS;
  constexpr int AW = kFX + 4, AH = ROWS + 4, AP = AW * AH;  // x3, halo 2
  constexpr int XW = kFX + 2, XH = ROWS + 2, XP = XW * XH;  // x4 and b - mb, halo 1
  constexpr int NA = (AP + kThreads - 1) / kThreads, NX = (XP + kThreads - 1) / kThreads;
  __shared__ T s3[3][AP];
  __shared__ T s4[3][XP];
  __shared__ T sbt[2][XP];
  const int tid = threadIdx.x;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  const size_t cplane = (size_t)Cl.nx * Cl.ny;
  const FusedTile t = fused_tile<ROWS>(L, kz);
  const int tx = tid & 63, ty = tid >> 6;
  const int gi = t.i0 + tx, gj0 = t.j0 + 2 * ty;
  const bool own0 = gi < nx && gj0 < ny, own1 = gi < nx && gj0 + 1 < ny;
  // in-plane offsets fit 32 bits (a level plane holds < 2^32 cells); 32-bit keeps the
  // loop-invariant offsets within the register budget of four waves per SIMD
  unsigned offa[NA], offc[NA], offx[NX];
  bool oka[NA], okx[NX];
  int ha[NX];  // halo-1 stage points: index in the halo-2 plane
#pragma unroll
  for(int m = 0; m < NA; ++m) {
    const int e = tid + m * kThreads;
    const int gx = t.i0 + (e % AW) - 2, gy = t.j0 + (e / AW) - 2;
    oka[m] = e < AP && gx < nx + 2 && gy < ny + 2;
    const int ii = oka[m] ? wrap_index(gx, nx) : 0, jj = oka[m] ? wrap_index(gy, ny) : 0;
    offa[m] = (unsigned)ii + (unsigned)nx * (unsigned)jj;
    offc[m] = (unsigned)(ii >> 1) + (unsigned)Cl.nx * (unsigned)(jj >> 1);
  }
#pragma unroll
  for(int m = 0; m < NX; ++m) {
    const int e = tid + m * kThreads;
    const int gx = t.i0 + (e % XW) - 1, gy = t.j0 + (e / XW) - 1;
    okx[m] = e < XP && gx < nx + 1 && gy < ny + 1;
    offx[m] = okx[m] ? (unsigned)wrap_index(gx, nx) + (unsigned)nx * (unsigned)wrap_index(gy, ny) : 0u;
    ha[m] = (e % XW + 1) + AW * (e / XW + 1);
  }
  const size_t own_off0 = own0 ? (size_t)gi + (size_t)nx * gj0 : 0, own_off1 = own1 ? own_off0 + nx : 0;
  struct Pre {
    T a[NA];  // x2
    T c[NA];  // parent of x2 in the coarse correction
    T x[NX];  // b
    T tb[2];  // b of the thread's two tile cells
  };
  auto fetch = [&](int kp, Pre& p) {  // unconditional raw loads (see k_down_fused)
    const int kw = wrap_index(kp, nz);
    const size_t kz = plane * (size_t)kw, cz = cplane * (size_t)(kw >> 1);
#pragma unroll
    for(int m = 0; m < NA; ++m) {
      p.a[m] = x2[offa[m] + kz];
      p.c[m] = xc[offc[m] + cz];
    }
#pragma unroll
    for(int m = 0; m < NX; ++m) p.x[m] = b[offx[m] + kz];
    p.tb[0] = b[own_off0 + kz];
    p.tb[1] = b[own_off1 + kz];
  };
  // the first planes are requested before the means are read and the coarser level's partials
  // are folded, so those round trips overlap
  Pre pa, pb, pc;
  fetch(t.k0 - 2, pa);
  if constexpr(NBUF == 3) fetch(t.k0 - 1, pb);
  const T mb = T(ps[lslot + L_MB]);
  const T m2 = T(ps[lslot + L_MX2]);
  // mean of the coarse correction: EpiSmoothLast of the coarser level, folded here when its
  // partials are passed (fold_src), else stored in ps by its finalize or by the tail
  double mcd;
  if(fold_src != nullptr) {
    double r[2];
    fold_all<2, kThreads>(fold_src, fold_nb, r);
    mcd = zm ? r[0] / (double)Cl.n : 0.0;
    if(blockIdx.x == 0 && tid == 0) ps[cslot + L_MX5] = mcd;
  } else {
    mcd = ps[cslot + L_MX5];
  }
  const T mc = T(mcd);
  const T w = T(L.winv[0]);
  const T w2 = w * T(L.wr);  // first post-smoothing sweep
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  double v[2] = {0.0, 0.0};
  T q1[2] = {T(0), T(0)}, q2[2] = {T(0), T(0)};  // tile rhs of planes kp - 1, kp - 2
  const int last = (t.k1 - t.k0) + 1;
  constexpr int D = NBUF - 1;  // planes between a fetch and its LDS store
  auto step = [&](int s, Pre& cur, Pre& nxt, auto always) {  // ALWAYS: see k_down_fused
    const int kp = t.k0 + s;
    {
      T* d3 = s3[ring3(kp)];
#pragma unroll
      for(int m = 0; m < NA; ++m) {  // x3 = (x2 - m2) + (xc - mc)[parent]
        const int e = tid + m * kThreads;
        if(e < AP) d3[e] = oka[m] ? (cur.a[m] - m2) + (cur.c[m] - mc) : T(0);
      }
      T* db = sbt[kp & 1];
#pragma unroll
      for(int m = 0; m < NX; ++m) {
        const int e = tid + m * kThreads;
        if(e < XP) db[e] = okx[m] ? cur.x[m] - mb : T(0);
      }
    }
    const T tb0 = own0 ? cur.tb[0] - mb : T(0), tb1 = own1 ? cur.tb[1] - mb : T(0);
    __syncthreads();
    if(decltype(always)::value || s + D <= last) fetch(kp + D, nxt);
    if(s >= 0) {  // x4 on plane kq = kp - 1, tile + halo 1
      const int kq = kp - 1;
      const int sm = ring3(kq - 1), sc = ring3(kq), sp = ring3(kq + 1);
      T* dst = s4[ring3(kq)];
      const T* bq = sbt[kq & 1];
      const T* x3m = s3[sm];
      const T* x3c = s3[sc];
      const T* x3p = s3[sp];
#pragma unroll
      for(int m = 0; m < NX; ++m) {
        const int e = tid + m * kThreads;
        if(e < XP) {
          const int ae = ha[m];
          const T c = x3c[ae];
          const T ax =
              lap_neg<T>(c, x3c[ae - 1], x3c[ae + 1], x3c[ae - AW], x3c[ae + AW], x3m[ae], x3p[ae], ihx2, ihy2, ihz2);
          dst[e] = c + w2 * (bq[e] + T(-1) * ax);
        }
      }
    }
    __syncthreads();
    if(s >= 2) {  // x5 on plane kr = kp - 2 (tile)
      const int kr = kp - 2;
      const int sm = ring3(kr - 1), sc = ring3(kr), sp = ring3(kr + 1);
#pragma unroll
      for(int rr = 0; rr < 2; ++rr) {
        const int xe = (tx + 1) + XW * (2 * ty + rr + 1);
        const T c = s4[sc][xe];
        const T ax = lap_neg<T>(c, s4[sc][xe - 1], s4[sc][xe + 1], s4[sc][xe - XW], s4[sc][xe + XW], s4[sm][xe],
                                s4[sp][xe], ihx2, ihy2, ihz2);
        const T bt = q2[rr];
        const T out = c + w * (bt + T(-1) * ax);
        if(rr == 0 ? own0 : own1) {
          x5g[(rr == 0 ? own_off0 : own_off1) + plane * kr] = out;
          v[0] += (double)out;
          v[1] += (double)bt * (double)out;
        }
      }
    }
    q2[0] = q1[0];
    q2[1] = q1[1];
    q1[0] = tb0;
    q1[1] = tb1;
  };
  if constexpr(NBUF == 3) {  // pa, pb, pc rotate as in k_down_fused
    int s = -2;
    for(; s + 4 <= last; s += 3) {  // every step of the group fetches
      step(s, pa, pc, std::true_type{});
      step(s + 1, pb, pa, std::true_type{});
      step(s + 2, pc, pb, std::true_type{});
    }
    if(s <= last) step(s, pa, pc, std::false_type{});
    if(s + 1 <= last) step(s + 1, pb, pa, std::false_type{});
    if(s + 2 <= last) step(s + 2, pc, pb, std::false_type{});
    if(s + 3 <= last) step(s + 3, pa, pc, std::false_type{});
  } else {
    int s = -2;
    for(; s + 2 <= last; s += 2) {  // every step of the pair fetches
      step(s, pa, pb, std::true_type{});
      step(s + 1, pb, pa, std::true_type{});
    }
    if(s <= last) step(s, pa, pb, std::false_type{});
    if(s + 1 <= last) step(s + 1, pb, pa, std::false_type{});
  }
  const int ops[2] = {kSum, kSum};
  store_partials<2>(v, ops, partials);
}

// ---------------------------------------------------------------- up leg on per-point z windows
// The up leg of a fused level (prolongation and the two post-smoothing sweeps) with all three
// stages on one thread-to-point map: thread t owns the points e = t, t + 512, t + 1024 of the
// tile's halo-2 plane (68 x 20) and keeps in registers, per point, x3 of planes k-2..k, x4 of
// planes k-3..k-1 and b - mb of planes k-2 and k-1. A stage then takes its z neighbours and its
// right-hand side from its own registers and reads only the four in-plane neighbours from LDS,
// from the plane of x3 or x4 stored one step earlier (two planes each, by parity), so a plane
// needs one barrier and 4 LDS reads per stencil (k_up_fused: two barriers, three-plane rings and
// 8 reads). Each value is the same per-point expression as in k_up_fused; only the order in which
// a block accumulates its partial sums differs.
template <typename T, int NBUF>
__global__ void __launch_bounds__(512, 4) k_up_zw(const T* __restrict__ b, const T* __restrict__ x2,
                                               const T* __restrict__ xc, PLevel L, PLevel Cl, T* __restrict__ x5g,
                                               double* partials, double* ps, int lslot, int cslot, int kz,
                                               const double* fold_src, unsigned fold_nb, int zm) {
  if(stopped(ps)) return;
  constexpr int ROWS = 16, kThreads = 32 * ROWS;
  constexpr int AW = kFX + 4, AH = ROWS + 4, AP = AW * AH;  // halo-2 plane
  constexpr int NA = (AP + kThreads - 1) / kThreads;
  __shared__ T s3[2][AP];
  __shared__ T s4[2][AP];
  const int tid = threadIdx.x;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  const size_t cplane = (size_t)Cl.nx * Cl.ny;
  const FusedTile t = fused_tile<ROWS>(L, kz);
  // per point: in-plane offset (and of its coarse parent), whether x3 is taken from the level
  // (oka; else 0), whether x4 is formed (h1: halo 1) and b enters (okx), whether x5 is stored
  unsigned offa[NA], offc[NA];
  bool oka[NA], okx[NA], h1[NA], own[NA];
#pragma unroll
  for(int m = 0; m < NA; ++m) {
    const int e = tid + m * kThreads;
    const int ex = e % AW, ey = e / AW;
    const int gx = t.i0 + ex - 2, gy = t.j0 + ey - 2;
    const bool in = e < AP;
    oka[m] = in && gx < nx + 2 && gy < ny + 2;
    h1[m] = in && ex >= 1 && ex <= AW - 2 && ey >= 1 && ey <= AH - 2;
    okx[m] = h1[m] && gx < nx + 1 && gy < ny + 1;
    own[m] = in && ex >= 2 && ex < AW - 2 && ey >= 2 && ey < AH - 2 && gx < nx && gy < ny;
    const int ii = oka[m] ? wrap_index(gx, nx) : 0, jj = oka[m] ? wrap_index(gy, ny) : 0;
    offa[m] = (unsigned)ii + (unsigned)nx * (unsigned)jj;
    offc[m] = (unsigned)(ii >> 1) + (unsigned)Cl.nx * (unsigned)(jj >> 1);
  }
  struct Pre {
    T a[NA];  // x2 of plane kp
    T c[NA];  // the coarse correction at the parent, plane kp
    T x[NA];  // b of plane kp - 1 (the plane of the step's x4)
  };
  auto fetch = [&](int kp, Pre& p) {  // unconditional loads at in-bounds offsets
    const int kw = wrap_index(kp, nz);
    const size_t kzo = plane * (size_t)kw, cz = cplane * (size_t)(kw >> 1);
    const size_t kbo = plane * (size_t)wrap_index(kp - 1, nz);
#pragma unroll
    for(int m = 0; m < NA; ++m) {
      p.a[m] = x2[offa[m] + kzo];
      p.c[m] = xc[offc[m] + cz];
      p.x[m] = b[offa[m] + kbo];
    }
  };
  // the first planes are requested before the means are read and the coarser level's partials
  // are folded, so those round trips overlap
  Pre pa, pb, pc;
  fetch(t.k0 - 2, pa);
  if constexpr(NBUF == 3) fetch(t.k0 - 1, pb);
  const T mb = T(ps[lslot + L_MB]);
  const T m2 = T(ps[lslot + L_MX2]);
  double mcd;
  if(fold_src != nullptr) {
    double r[2];
    fold_all<2, kThreads>(fold_src, fold_nb, r);
    mcd = zm ? r[0] / (double)Cl.n : 0.0;
    if(blockIdx.x == 0 && tid == 0) ps[cslot + L_MX5] = mcd;
  } else {
    mcd = ps[cslot + L_MX5];
  }
  const T mc = T(mcd);
  const T w = T(L.winv[0]);
  const T w2 = w * T(L.wr);  // first post-smoothing sweep
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  T w3[NA][3], w4[NA][3], bold[NA];  // x3 and x4 of three planes (oldest first), b - mb of kp - 2
#pragma unroll
  for(int m = 0; m < NA; ++m) {
    bold[m] = T(0);
#pragma unroll
    for(int q = 0; q < 3; ++q) w3[m][q] = w4[m][q] = T(0);
  }
  double v[2] = {0.0, 0.0};
  const int last = (t.k1 - t.k0) + 1;
  constexpr int D = NBUF - 1;  // planes between a fetch and its use
  // ALWAYS: the step fetches (not under a branch); STEADY: s >= 2, so all three stages run with
  // no test, in one basic block the scheduler can interleave
  auto step = [&](int s, Pre& cur, Pre& nxt, auto always, auto steady) {
    const int kp = t.k0 + s;
    // x3 on plane kp at every point of the halo-2 plane; b - mb of plane kp - 1
    T* d3 = s3[kp & 1];
    T bnew[NA];
#pragma unroll
    for(int m = 0; m < NA; ++m) {
      const T x3 = oka[m] ? (cur.a[m] - m2) + (cur.c[m] - mc) : T(0);
      if(tid + m * kThreads < AP) d3[tid + m * kThreads] = x3;
      w3[m][0] = w3[m][1];
      w3[m][1] = w3[m][2];
      w3[m][2] = x3;
      bnew[m] = okx[m] ? cur.x[m] - mb : T(0);
    }
    if(decltype(always)::value || s + D <= last) fetch(kp + D, nxt);
    if(decltype(steady)::value || s >= 0) {  // x4 on plane kq = kp - 1 (halo 1): x3 stored last step
      const int kq = kp - 1;
      const T* a3 = s3[kq & 1];
      T* d4 = s4[kq & 1];
#pragma unroll
      for(int m = 0; m < NA; ++m) {
        T x4 = T(0);
        if(h1[m]) {
          const int e = tid + m * kThreads;
          const T c = w3[m][1];
          const T ax = lap_neg<T>(c, a3[e - 1], a3[e + 1], a3[e - AW], a3[e + AW], w3[m][0], w3[m][2], ihx2, ihy2,
                                  ihz2);
          x4 = c + w2 * (bnew[m] + T(-1) * ax);
          d4[e] = x4;
        }
        w4[m][0] = w4[m][1];
        w4[m][1] = w4[m][2];
        w4[m][2] = x4;
      }
    }
    if(decltype(steady)::value || s >= 2) {  // x5 on plane kr = kp - 2 (tile): x4 stored last step
      const int kr = kp - 2;
      const T* a4 = s4[kr & 1];
#pragma unroll
      for(int m = 0; m < NA; ++m) {
        if(own[m]) {
          const int e = tid + m * kThreads;
          const T c = w4[m][1];
          const T ax = lap_neg<T>(c, a4[e - 1], a4[e + 1], a4[e - AW], a4[e + AW], w4[m][0], w4[m][2], ihx2, ihy2,
                                  ihz2);
          const T bt = bold[m];
          const T out = c + w * (bt + T(-1) * ax);
          x5g[offa[m] + plane * (size_t)kr] = out;
          v[0] += (double)out;
          v[1] += (double)bt * (double)out;
        }
      }
    }
#pragma unroll
    for(int m = 0; m < NA; ++m) bold[m] = bnew[m];
    __syncthreads();
  };
  // the four ramp steps (s = -2 .. 1) fetch (last >= 3), then the steady steps
  const std::true_type yes{};
  const std::false_type no{};
  if constexpr(NBUF == 3) {  // pa, pb, pc rotate: a step commits one and refills another
    step(-2, pa, pc, yes, no);
    step(-1, pb, pa, yes, no);
    step(0, pc, pb, yes, no);
    step(1, pa, pc, yes, no);
    int s = 2;
    for(; s + 4 <= last; s += 3) {
      step(s, pb, pa, yes, yes);
      step(s + 1, pc, pb, yes, yes);
      step(s + 2, pa, pc, yes, yes);
    }
    if(s <= last) step(s, pb, pa, no, yes);
    if(s + 1 <= last) step(s + 1, pc, pb, no, yes);
    if(s + 2 <= last) step(s + 2, pa, pc, no, yes);
    if(s + 3 <= last) step(s + 3, pb, pa, no, yes);
  } else {
    step(-2, pa, pb, yes, no);
    step(-1, pb, pa, yes, no);
    step(0, pa, pb, yes, no);
    step(1, pb, pa, yes, no);
    int s = 2;
    for(; s + 2 <= last; s += 2) {
      step(s, pa, pb, yes, yes);
      step(s + 1, pb, pa, yes, yes);
    }
    if(s <= last) step(s, pa, pb, no, yes);
    if(s + 1 <= last) step(s + 1, pb, pa, no, yes);
  }
  const int ops[2] = {kSum, kSum};
  store_partials<2>(v, ops, partials);
}

// ---------------------------------------------------------------- small fused levels in one shot
// On the small all-periodic levels (64^3 and 32^3 at 256^3) each block of the fused kernels
// owns two planes and its march is six dependent plane steps (four of them halo planes) with
// two barriers each, so those launches cost about 13 us for very little work. These variants
// compute the same stages with the same per-cell expressions, in the same order, without the
// march: every load of the six planes is issued at once, and the stages run one after another
// over all their planes (three barrier-separated phases). LDS holds all six planes (141 KB in
// FP64), so a block owns a compute unit; these levels have at most a few hundred blocks.
template <typename T>
__global__ void __launch_bounds__(512) k_down_small(const T* __restrict__ b, T* __restrict__ x2g, PLevel L, PLevel Cl,
                                                    T* __restrict__ bc, double* partials, double* ps, int lslot,
                                                    int zm, DownFold fold) {
  if(stopped(ps)) return;
  constexpr int ROWS = 16, kThreads = 32 * ROWS;
  constexpr int BW = kFX + 4, BH = ROWS + 4, BP = BW * BH;  // x1, halo 2
  constexpr int XW = kFX + 2, XH = ROWS + 2, XP = XW * XH;  // b - mb and x2, halo 1
  constexpr int NB = (BP + kThreads - 1) / kThreads;
  constexpr int NH = (XP + kThreads - 1) / kThreads;
  constexpr int P1 = 6, P2 = 4;  // x1 on planes k0-2 .. k0+3; x2 and b - mb on k0-1 .. k0+2
  __shared__ T sb[P1][BP];
  __shared__ T sbt[P2][XP];
  __shared__ T sx[P2][XP];
  const int tid = threadIdx.x;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  const size_t cplane = (size_t)Cl.nx * Cl.ny;
  const FusedTile t = fused_tile<ROWS>(L, 2);
  const int tx = tid & 63, ty = tid >> 6;
  const int gi = t.i0 + tx, gj0 = t.j0 + 2 * ty;
  const bool own0 = gi < nx && gj0 < ny, own1 = gi < nx && gj0 + 1 < ny;
  size_t off[NB];
  bool okb[NB];
  int e1[NB];
#pragma unroll
  for(int m = 0; m < NB; ++m) {
    const int e = tid + m * kThreads;
    const int ex = e % BW, ey = e / BW;
    const int gx = t.i0 + ex - 2, gy = t.j0 + ey - 2;
    okb[m] = e < BP && gx < nx + 2 && gy < ny + 2;
    off[m] = okb[m] ? (size_t)wrap_index(gx, nx) + (size_t)nx * wrap_index(gy, ny) : 0;
    e1[m] = e < BP && ex >= 1 && ex <= XW && ey >= 1 && ey <= XH ? (ex - 1) + XW * (ey - 1) : -1;
  }
  int hb[NH];
  long long hg[NH];
#pragma unroll
  for(int m = 0; m < NH; ++m) {
    const int e = tid + m * kThreads;
    const int xi = e % XW, xj = e / XW;
    hb[m] = (xi + 1) + BW * (xj + 1);
    const int ci = t.i0 + xi - 1, cj = t.j0 + xj - 1;
    hg[m] = e < XP && xi >= 1 && xi <= kFX && xj >= 1 && xj <= ROWS && ci < nx && cj < ny
                ? (long long)ci + (long long)nx * cj
                : -1;
  }
  const size_t own_off0 = own0 ? (size_t)gi + (size_t)nx * gj0 : 0, own_off1 = own1 ? own_off0 + nx : 0;
  // phase A: every load of the six planes, issued before the finer level's partials are folded
  // so the two round trips overlap; then x1 = (b - mb) w - m1 and b - mb into LDS
  T raw[P1][NB], tb[2][2];
#pragma unroll
  for(int p = 0; p < P1; ++p) {
    const size_t kz = plane * (size_t)wrap_index(t.k0 - 2 + p, nz);
#pragma unroll
    for(int m = 0; m < NB; ++m) raw[p][m] = b[off[m] + kz];
  }
#pragma unroll
  for(int r = 0; r < 2; ++r) {
    const size_t kz = plane * (size_t)(t.k0 + r);
    tb[r][0] = b[own_off0 + kz];
    tb[r][1] = b[own_off1 + kz];
  }
  double mbd, sbw;
  down_fold<kThreads>(fold, ps, lslot, (double)L.n, zm, mbd, sbw);
  const T mb = T(mbd);
  const T m1 = zm ? T((sbw - mbd * L.winv_sum) / (double)L.n) : T(0);
  const T w = T(L.winv[0]);
  const T w2 = w * T(L.wr);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const T inv_children = T(1.0 / 8.0);
#pragma unroll
  for(int p = 0; p < P1; ++p) {
#pragma unroll
    for(int m = 0; m < NB; ++m) {
      const int e = tid + m * kThreads;
      const T bt = okb[m] ? raw[p][m] - mb : T(0);
      if(e < BP) sb[p][e] = bt * w - m1;
      if(p >= 1 && p <= P2 && e1[m] >= 0) sbt[p - 1][e1[m]] = bt;
    }
  }
  __syncthreads();
  // phase B: x2 on planes k0-1 .. k0+2, tile + halo 1 (stored for the block's own planes)
  double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for(int q = 0; q < P2; ++q) {
    const int kq = t.k0 - 1 + q;
    const T* x1m = sb[q];
    const T* x1c = sb[q + 1];
    const T* x1p = sb[q + 2];
    const T* btc = sbt[q];
    T* dst = sx[q];
    const bool own = q == 1 || q == 2;
    const size_t pq = plane * (size_t)(own ? kq : 0);
#pragma unroll
    for(int m = 0; m < NH; ++m) {
      const int e = tid + m * kThreads;
      if(e < XP) {
        const int be = hb[m];
        const T c = x1c[be];
        const T ax = lap_neg<T>(c, x1c[be - 1], x1c[be + 1], x1c[be - BW], x1c[be + BW], x1m[be], x1p[be], ihx2, ihy2,
                                ihz2);
        const T out = c + w2 * (btc[e] + T(-1) * ax);
        dst[e] = out;
        if(own && hg[m] >= 0) {
          x2g[(size_t)hg[m] + pq] = out;
          v[0] += (double)out;
        }
      }
    }
  }
  __syncthreads();
  // phase C: residual on planes k0 and k0+1 (tile), restricted 2x2x2
  T acc = T(0);
#pragma unroll
  for(int r = 0; r < 2; ++r) {
    const T* sm = sx[r];
    const T* sc = sx[r + 1];
    const T* sp = sx[r + 2];
    T xs[2];
#pragma unroll
    for(int rr = 0; rr < 2; ++rr) {
      const int xe = (tx + 1) + XW * (2 * ty + rr + 1);
      const T c = sc[xe];
      const T ax = lap_neg<T>(c, sc[xe - 1], sc[xe + 1], sc[xe - XW], sc[xe + XW], sm[xe], sp[xe], ihx2, ihy2, ihz2);
      const T bq = (rr == 0 ? own0 : own1) ? tb[r][rr] - mb : T(0);
      T res = bq + T(-1) * ax;
      res += lane_xor1(res);  // x pair
      xs[rr] = res;
    }
    const T ysum = xs[0] + xs[1];  // y pair
    const T tot = r == 1 ? acc + ysum : ysum;
    acc = tot;
    if(r == 1 && (tx & 1) == 0 && own0) {
      const int ic = gi >> 1, jc = gj0 >> 1, kc = t.k0 >> 1;
      const T val = inv_children * tot;
      bc[(size_t)ic + (size_t)Cl.nx * jc + cplane * kc] = val;
      v[1] += (double)val;
      v[2] += (double)val * Cl.winv[0];
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

template <typename T>
__global__ void __launch_bounds__(512) k_up_small(const T* __restrict__ b, const T* __restrict__ x2,
                                                  const T* __restrict__ xc, PLevel L, PLevel Cl,
                                                  T* __restrict__ x5g, double* partials, double* ps, int lslot,
                                                  int cslot, const double* fold_src, unsigned fold_nb, int zm) {
  if(stopped(ps)) return;
  constexpr int ROWS = 16, kThreads = 32 * ROWS;
  constexpr int AW = kFX + 4, AH = ROWS + 4, AP = AW * AH;  // x3, halo 2
  constexpr int XW = kFX + 2, XH = ROWS + 2, XP = XW * XH;  // x4 and b - mb, halo 1
  constexpr int NA = (AP + kThreads - 1) / kThreads, NX = (XP + kThreads - 1) / kThreads;
  constexpr int P1 = 6, P2 = 4;  // x3 on planes k0-2 .. k0+3; x4 and b - mb on k0-1 .. k0+2
  __shared__ T s3[P1][AP];
  __shared__ T s4[P2][XP];
  __shared__ T sbt[P2][XP];
  const int tid = threadIdx.x;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  const size_t cplane = (size_t)Cl.nx * Cl.ny;
  const FusedTile t = fused_tile<ROWS>(L, 2);
  const int tx = tid & 63, ty = tid >> 6;
  const int gi = t.i0 + tx, gj0 = t.j0 + 2 * ty;
  const bool own0 = gi < nx && gj0 < ny, own1 = gi < nx && gj0 + 1 < ny;
  unsigned offa[NA], offc[NA], offx[NX];
  bool oka[NA], okx[NX];
  int ha[NX];
#pragma unroll
  for(int m = 0; m < NA; ++m) {
    const int e = tid + m * kThreads;
    const int gx = t.i0 + (e % AW) - 2, gy = t.j0 + (e / AW) - 2;
    oka[m] = e < AP && gx < nx + 2 && gy < ny + 2;
    const int ii = oka[m] ? wrap_index(gx, nx) : 0, jj = oka[m] ? wrap_index(gy, ny) : 0;
    offa[m] = (unsigned)ii + (unsigned)nx * (unsigned)jj;
    offc[m] = (unsigned)(ii >> 1) + (unsigned)Cl.nx * (unsigned)(jj >> 1);
  }
#pragma unroll
  for(int m = 0; m < NX; ++m) {
    const int e = tid + m * kThreads;
    const int gx = t.i0 + (e % XW) - 1, gy = t.j0 + (e / XW) - 1;
    okx[m] = e < XP && gx < nx + 1 && gy < ny + 1;
    offx[m] = okx[m] ? (unsigned)wrap_index(gx, nx) + (unsigned)nx * (unsigned)wrap_index(gy, ny) : 0u;
    ha[m] = (e % XW + 1) + AW * (e / XW + 1);
  }
  const size_t own_off0 = own0 ? (size_t)gi + (size_t)nx * gj0 : 0, own_off1 = own1 ? own_off0 + nx : 0;
  // phase A: every load, issued before the means are read and the coarser level's partials are
  // folded so the round trips overlap, then x3 = (x2 - m2) + (xc - mc)[parent] and b - mb into
  // LDS
  T ra[P1][NA], rc[P1][NA], rx[P2][NX], tb[2][2];
#pragma unroll
  for(int p = 0; p < P1; ++p) {
    const int kw = wrap_index(t.k0 - 2 + p, nz);
    const size_t kz = plane * (size_t)kw, cz = cplane * (size_t)(kw >> 1);
#pragma unroll
    for(int m = 0; m < NA; ++m) {
      ra[p][m] = x2[offa[m] + kz];
      rc[p][m] = xc[offc[m] + cz];
    }
  }
#pragma unroll
  for(int q = 0; q < P2; ++q) {
    const size_t kz = plane * (size_t)wrap_index(t.k0 - 1 + q, nz);
#pragma unroll
    for(int m = 0; m < NX; ++m) rx[q][m] = b[offx[m] + kz];
  }
#pragma unroll
  for(int r = 0; r < 2; ++r) {
    const size_t kz = plane * (size_t)(t.k0 + r);
    tb[r][0] = b[own_off0 + kz];
    tb[r][1] = b[own_off1 + kz];
  }
  const T mb = T(ps[lslot + L_MB]);
  const T m2 = T(ps[lslot + L_MX2]);
  double mcd;
  if(fold_src != nullptr) {
    double r[2];
    fold_all<2, kThreads>(fold_src, fold_nb, r);
    mcd = zm ? r[0] / (double)Cl.n : 0.0;
    if(blockIdx.x == 0 && tid == 0) ps[cslot + L_MX5] = mcd;
  } else {
    mcd = ps[cslot + L_MX5];
  }
  const T mc = T(mcd);
  const T w = T(L.winv[0]);
  const T w2 = w * T(L.wr);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
#pragma unroll
  for(int p = 0; p < P1; ++p) {
#pragma unroll
    for(int m = 0; m < NA; ++m) {
      const int e = tid + m * kThreads;
      if(e < AP) s3[p][e] = oka[m] ? (ra[p][m] - m2) + (rc[p][m] - mc) : T(0);
    }
  }
#pragma unroll
  for(int q = 0; q < P2; ++q) {
#pragma unroll
    for(int m = 0; m < NX; ++m) {
      const int e = tid + m * kThreads;
      if(e < XP) sbt[q][e] = okx[m] ? rx[q][m] - mb : T(0);
    }
  }
  __syncthreads();
  // phase B: x4 on planes k0-1 .. k0+2, tile + halo 1
#pragma unroll
  for(int q = 0; q < P2; ++q) {
    const T* x3m = s3[q];
    const T* x3c = s3[q + 1];
    const T* x3p = s3[q + 2];
    const T* bq = sbt[q];
    T* dst = s4[q];
#pragma unroll
    for(int m = 0; m < NX; ++m) {
      const int e = tid + m * kThreads;
      if(e < XP) {
        const int ae = ha[m];
        const T c = x3c[ae];
        const T ax =
            lap_neg<T>(c, x3c[ae - 1], x3c[ae + 1], x3c[ae - AW], x3c[ae + AW], x3m[ae], x3p[ae], ihx2, ihy2, ihz2);
        dst[e] = c + w2 * (bq[e] + T(-1) * ax);
      }
    }
  }
  __syncthreads();
  // phase C: x5 on planes k0 and k0+1 (tile)
  double v[2] = {0.0, 0.0};
#pragma unroll
  for(int r = 0; r < 2; ++r) {
    const int kr = t.k0 + r;
    const T* sm = s4[r];
    const T* sc = s4[r + 1];
    const T* sp = s4[r + 2];
#pragma unroll
    for(int rr = 0; rr < 2; ++rr) {
      const int xe = (tx + 1) + XW * (2 * ty + rr + 1);
      const T c = sc[xe];
      const T ax = lap_neg<T>(c, sc[xe - 1], sc[xe + 1], sc[xe - XW], sc[xe + XW], sm[xe], sp[xe], ihx2, ihy2, ihz2);
      const bool own = rr == 0 ? own0 : own1;
      const T bt = own ? tb[r][rr] - mb : T(0);
      const T out = c + w * (bt + T(-1) * ax);
      if(own) {
        x5g[(rr == 0 ? own_off0 : own_off1) + plane * kr] = out;
        v[0] += (double)out;
        v[1] += (double)bt * (double)out;
      }
    }
  }
  const int ops[2] = {kSum, kSum};
  store_partials<2>(v, ops, partials);
}

struct EpiDown {  // mean of x2; coarse rhs mean, sum of b w/D, sum
  static constexpr int NV = 3;
  double* ps;
  int lslot, cslot;
  double n, nc;
  int zm;
  __device__ int op(int) const { return kSum; }
  __device__ bool active() const { return !stopped(ps); }
  __device__ void operator()(const double (&r)[3]) const {
    ps[lslot + L_MX2] = zm ? r[0] / n : 0.0;
    ps[cslot + L_MB] = zm ? r[1] / nc : 0.0;
    ps[cslot + L_SBW] = r[2];
    ps[cslot + L_SUMB] = r[1];
  }
};

// ---------------------------------------------------------------- one-workgroup tail
// Levels l0..nl-1 of the V-cycle (the small ones) in a single workgroup: the same per-cell
// formulas, block-wide sums for the deferred means, and the coarsest-level elimination with the
// pre-factored row operations of the reference's partial-pivoting Gaussian elimination
// (linsolve/poisson_solver.cpp:482-567). One launch replaces ~8 launches per level.
constexpr int kMaxLevels = 24;
constexpr int kTailThreads = 1024;
constexpr int kTailCells = 4096;  // levels of at most this many cells run in the tail

template <typename T>
struct TailBufs {
  T* b[kMaxLevels];  // global arrays
  T* xa[kMaxLevels];
  T* xb[kMaxLevels];
  int ob[kMaxLevels];  // LDS offsets (elements of T) when the tail levels are LDS-resident
  int oxa[kMaxLevels];
  int oxb[kMaxLevels];
  int owork;  // LDS offset of the coarse work vector (doubles)
  int n0;     // cells of the first tail level
  int ncs;    // unknowns of the coarsest level
  const T* b0;  // b of the first tail level (global)
};
constexpr int kTailLdsBytes = 152 * 1024;
constexpr int kCoarseStage = 16;  // coarsest systems of up to this many unknowns are read from LDS
constexpr int kTailPre = 4;       // level-l0 values per thread read ahead of the fold (4096 cells)

// Thread groups of the tail. Levels of more than kWaveCells cells (and the column-mapped ones of
// more than 64) run on the whole workgroup; the smaller ones and the coarsest solve run on its
// first wave alone, which needs no workgroup barrier: a wave-level fence orders its LDS and
// global accesses. sums() leaves the group's sums in every thread and orders the group's earlier
// writes before its later reads.
constexpr int kWaveCells = 512;
constexpr int kTailWaves = kTailThreads / 64;
struct TailBlock {
  int tid, nt;
  double (*red)[kTailWaves][4];  // two LDS buffers of wave sums, used in turn
  mutable int flip;
  __device__ void sync() const { __syncthreads(); }
  // One barrier: each wave's sum is stored, then every thread adds the wave sums in wave order.
  // A buffer is rewritten two calls later, after every thread has passed the barrier of the call
  // between, so its reads are done.
  template <int N>
  __device__ void sums(double (&v)[N]) const {
    static_assert(N <= 4, "at most four sums at once");
#pragma unroll
    for(int k = 0; k < N; ++k) v[k] = wave_sum_all(v[k]);
    double (*r)[4] = red[flip];
    flip ^= 1;
    if((tid & 63) == 0) {
#pragma unroll
      for(int k = 0; k < N; ++k) r[tid >> 6][k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < N; ++k) {
      double a = r[0][k];
#pragma unroll
      for(int w = 1; w < kTailWaves; ++w) a = a + r[w][k];
      v[k] = a;
    }
  }
};
struct TailWave {
  int tid, nt;
  __device__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
  template <int N>
  __device__ void sums(double (&v)[N]) const {
#pragma unroll
    for(int k = 0; k < N; ++k) v[k] = wave_sum_all(v[k]);
    sync();
  }
};

// Column map of a tail level that is periodic on every axis, has power-of-two sizes and halves
// every axis: thread t of a group of nt threads owns the cells (i, j, kq km + m), m < km, with
// i = t mod nx, j = (t / nx) mod ny, kq = t / (nx ny) and km = cells / nt. Its neighbour and
// parent offsets come from shifts and masks once per thread (the generic loops divide the cell
// id and test every boundary per cell: about 130 instructions a cell on the tail's one CU,
// against about 50 here). The eight children of a coarse cell sit in lanes t and t ^ 1 (x),
// t ^ nx (y) and in cells m, m + 1 of a thread (z; lanes t ^ (nx ny) when km = 1), so the
// restriction sums them by lane exchanges in the generic loops' order: x pairs, then y pairs,
// then z pairs. Every per-cell value is the generic formula's with all signs +1 and the uniform
// weight.
__device__ inline bool pow2i(int v) { return v > 0 && (v & (v - 1)) == 0; }
__device__ inline int log2i(int v) { return 31 - __clz(v); }

// Cells per thread: cells / nt on a level of at least nt cells, else 2 (on cells / 2 threads).
// Every count here is a power of two (nt is 64 or the workgroup size, a level's sizes are checked
// first), so quotients are shifts: a division by a run-time value is some forty dependent
// instructions, and the small levels run on one wave, where nothing hides them.
__device__ inline int col_km(const PLevel& L, int nt) { return L.n >= nt ? L.n >> log2i(nt) : 2; }
__device__ inline bool col_ok(const PLevel& L, const PLevel& Cl, int nt) {
  if(!(pow2i(L.nx) && pow2i(L.ny) && pow2i(L.nz))) return false;
  if(L.nx < 2 || L.ny < 2 || L.nz < 2 || L.nx > 32) return false;
  if(L.rx != 2 || L.ry != 2 || L.rz != 2) return false;
  if(2 * Cl.nx != L.nx || 2 * Cl.ny != L.ny || 2 * Cl.nz != L.nz) return false;
  const int km = col_km(L, nt), na = L.n >> log2i(km), plane = L.nx * L.ny;
  if((na & 63) != 0 || na > nt || km > L.nz) return false;  // whole waves own cells
  return km >= 2 ? (km & 1) == 0 : 2 * plane <= 64;
}

struct ColMap {
  int km, k0, o0;          // cells per thread, first k, offset of the first cell
  int plane, nz;
  int dxm, dxp, dym, dyp;  // in-plane neighbour offsets (periodic wrap)
  int i, j;
  bool act;  // the thread owns cells (whole waves do)
  __device__ int dzm(int k) const { return k == 0 ? (nz - 1) * plane : -plane; }
  __device__ int dzp(int k) const { return k == nz - 1 ? (1 - nz) * plane : plane; }
};
__device__ inline ColMap col_map(const PLevel& L, int t, int nt) {
  ColMap c;
  const int lx = log2i(L.nx), ly = log2i(L.ny);
  c.i = t & (L.nx - 1);
  c.j = (t >> lx) & (L.ny - 1);
  const int kq = t >> (lx + ly);
  c.km = col_km(L, nt);
  c.act = t < (L.n >> log2i(c.km));
  c.plane = L.nx * L.ny;
  c.nz = L.nz;
  c.k0 = kq * c.km;
  c.o0 = c.i + (c.j << lx) + c.k0 * c.plane;
  c.dxm = c.i == 0 ? L.nx - 1 : -1;
  c.dxp = c.i == L.nx - 1 ? 1 - L.nx : 1;
  c.dym = c.j == 0 ? (L.ny - 1) * L.nx : -L.nx;
  c.dyp = c.j == L.ny - 1 ? (1 - L.ny) * L.nx : L.nx;
  return c;
}

// LDS: every tail level lives in one LDS buffer (b of level l0 is copied in, x of level l0
// copied out); otherwise the levels are read and written in global memory.
template <typename T, bool LDS, bool UNI>
__global__ void __launch_bounds__(kTailThreads)
    k_vtail(const PLevel* __restrict__ lvs, int l0, int nl, TailBufs<T> B, const int* __restrict__ perm,
            const double* __restrict__ factor, const double* __restrict__ U, double* __restrict__ gwork, double* ps,
            int zm, DownFold fold) {
  __shared__ double sbuf[kTailLdsBytes / sizeof(double)];
  T* sT = reinterpret_cast<T*>(sbuf);
  auto Bp = [&](int l) -> T* { return LDS ? sT + B.ob[l] : B.b[l]; };
  auto XA = [&](int l) -> T* { return LDS ? sT + B.oxa[l] : B.xa[l]; };
  auto XB = [&](int l) -> T* { return LDS ? sT + B.oxb[l] : B.xb[l]; };
  double* work = LDS ? sbuf + B.owork : gwork;
  // The start's global reads are issued together, ahead of the stop test: the stop flag, level
  // l0's b when the levels are LDS-resident, a small coarsest system's solution operator, the
  // partials the fold combines and the tail levels' parameter structs (the last two only to bring
  // them into the caches: the fold's own loads and each level's scalar loads of its fields then
  // hit there). The empty asm statements take the loaded values, so the loads complete before
  // the test and are not moved past it: one memory round trip instead of three (flag, level
  // sizes, data). A coarsest system of at most kCoarseStage unknowns is solved by its solution
  // operator C (in `factor`, built on the host from the recorded elimination): one matrix-vector
  // product on the first wave instead of the serial elimination and back substitution.
  __shared__ double cop[kCoarseStage * kCoarseStage];
  static_assert(kCoarseStage * kCoarseStage <= kTailThreads, "one staged entry per thread");
  const int tid = threadIdx.x;
  const double done = ps[P_DONE];
  const int n0 = B.n0;
  T bpre[kTailPre];
  if(LDS) {
#pragma unroll
    for(int u = 0; u < kTailPre; ++u) bpre[u] = B.b0[min(tid + u * kTailThreads, n0 - 1)];
  }
  const int ncs = B.ncs;
  const bool cstage = ncs <= kCoarseStage;
  const double cload = factor[min(tid, ncs * ncs - 1)];  // in the operator or the multipliers
  const bool fsrc = fold.src != nullptr && fold.nb > 0;
  const unsigned fb = fsrc ? min((unsigned)tid, fold.nb - 1) : 0u;
  const double* f0 = fsrc ? fold.src + fb : ps;  // ps[0]: an in-bounds stand-in
  const double* f1 = fsrc ? fold.src + kMaxPartials + fb : ps;
  const double* f2 = fsrc ? fold.src + 2 * (size_t)kMaxPartials + fb : ps;
  const double fwarm = *f0 + *f1 + *f2;
  const int nw = (int)((nl - l0) * sizeof(PLevel) / sizeof(float));
  const float lwarm = reinterpret_cast<const float*>(lvs + l0)[min(tid, nw - 1)];
  asm volatile("" ::"v"(cload), "v"(fwarm), "v"(lwarm));
  if(LDS) {
#pragma unroll
    for(int u = 0; u < kTailPre; ++u) asm volatile("" ::"v"(bpre[u]));
  }
  if(done != 0.0) return;
  const double cpre = cstage && tid < ncs * ncs ? cload : 0.0;
  __shared__ double lm[kMaxLevels][4];  // per level: mb, m2, m4, m5
  __shared__ double lsbw[kMaxLevels];   // sum of raw b w/D
  __shared__ double rz_sh;
  const int s0 = P_LEVEL + 8 * l0;
  double mb0, sbw0;
  // The fold of the finer level's partials (its EpiDown) on wave 0 alone: the few dozen partials
  // of a small level need no workgroup reduction and its barriers, and the barrier after the
  // copy-in publishes the result. Partials are read four a lane at a time at clamped indices.
  if(tid < 64) {
    if(fold.src == nullptr) {
      mb0 = ps[s0 + L_MB];
      sbw0 = ps[s0 + L_SBW];
    } else {
      double r[3] = {0.0, 0.0, 0.0};
      const unsigned last = fold.nb > 0 ? fold.nb - 1 : 0;
      for(unsigned base = (unsigned)tid; base < fold.nb; base += 256) {
        double x[4][3];
#pragma unroll
        for(int u = 0; u < 4; ++u) {
          const unsigned b = min(base + 64u * u, last);
#pragma unroll
          for(int k = 0; k < 3; ++k) x[u][k] = fold.src[(size_t)k * kMaxPartials + b];
        }
#pragma unroll
        for(int u = 0; u < 4; ++u) {
          const bool in = base + 64u * u < fold.nb;
#pragma unroll
          for(int k = 0; k < 3; ++k) r[k] = in ? r[k] + x[u][k] : r[k];
        }
      }
#pragma unroll
      for(int k = 0; k < 3; ++k) r[k] = wave_sum_all(r[k]);
      mb0 = zm ? r[1] / (double)n0 : 0.0;
      sbw0 = r[2];
      if(tid == 0) {
        ps[fold.pslot + L_MX2] = zm ? r[0] / fold.pn : 0.0;
        ps[s0 + L_MB] = mb0;
        ps[s0 + L_SBW] = sbw0;
        ps[s0 + L_SUMB] = r[1];
      }
    }
  }
  if(LDS) {
    T* dst = Bp(l0);
#pragma unroll
    for(int u = 0; u < kTailPre; ++u) {
      const int id = tid + u * kTailThreads;
      if(id < n0) dst[id] = bpre[u];
    }
    for(int id = tid + kTailPre * kTailThreads; id < n0; id += kTailThreads) dst[id] = B.b[l0][id];
  }
  if(cstage && tid < ncs * ncs) cop[tid] = cpre;
  if(tid == 0) {
    lm[l0][0] = mb0;
    lsbw[l0] = sbw0;
  }
  __syncthreads();
  // down leg of level l: two sweeps (x2 and its mean), residual restricted to level l + 1. On
  // entry mbl and sbwl hold level l's rhs mean and sum of raw b w/D in every thread of G; on
  // return, level l + 1's (also kept in lm and lsbw for the up leg).
  auto down = [&](int l, const auto& G, double& mbl, double& sbwl) {
    const PLevel& L = lvs[l];
    const PLevel& Cl = lvs[l + 1];
    const T* b = Bp(l);
    T* x2 = XA(l);
    const T mb = T(mbl);
    const T m1 = zm ? T((sbwl - mbl * L.winv_sum) / (double)L.n) : T(0);
    double sc[2] = {0.0, 0.0};  // coarse rhs: sum, sum of raw b w/D
    if(UNI && col_ok(L, Cl, G.nt)) {  // smooth12_cell and residual_cell on the column map
      const ColMap cm = col_map(L, G.tid, G.nt);
      const T wc = T(L.winv[0]), wcr = wc * T(L.wr);
      const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
      auto X1 = [&](int o) { return (b[o] - mb) * wc - m1; };
      double acc[1] = {0.0};
      for(int m = 0; cm.act && m < cm.km; ++m) {
        const int k = cm.k0 + m, o = cm.o0 + m * cm.plane;
        const T bt = b[o] - mb;
        const T c = bt * wc - m1;
        const T ax = lap_neg<T>(c, X1(o + cm.dxm), X1(o + cm.dxp), X1(o + cm.dym), X1(o + cm.dyp), X1(o + cm.dzm(k)),
                                X1(o + cm.dzp(k)), ihx2, ihy2, ihz2);
        const T xv = c + wcr * (bt + T(-1) * ax);
        x2[o] = xv;
        acc[0] += (double)xv;
      }
      G.sums(acc);
      const double mxd = zm ? acc[0] / (double)L.n : 0.0;
      if(G.tid == 0) lm[l][1] = mxd;
      const T mx = T(mxd);
      auto X = [&](int o) { return x2[o] - mx; };
      auto res = [&](int m) {  // fine residual, summed over the x pair and then the y pair
        const int k = cm.k0 + m, o = cm.o0 + m * cm.plane;
        const T c = X(o);
        const T ax = lap_neg<T>(c, X(o + cm.dxm), X(o + cm.dxp), X(o + cm.dym), X(o + cm.dyp), X(o + cm.dzm(k)),
                                X(o + cm.dzp(k)), ihx2, ihy2, ihz2);
        T r = (b[o] - mb) + T(-1) * ax;
        r = r + lane_xor1(r);
        return r + __shfl_xor(r, L.nx, 64);
      };
      const T inv_children = T(1.0 / 8.0);
      const bool owner = ((cm.i | cm.j) & 1) == 0;
      T* bc = Bp(l + 1);
      const double cw = Cl.winv[0];
      const int cbase = (cm.i >> 1) + Cl.nx * (cm.j >> 1);
      const int cplane = Cl.nx * Cl.ny;
      auto put = [&](int k, T a) {  // k: the even fine plane of the pair
        const T val = inv_children * a;
        bc[cbase + cplane * (k >> 1)] = val;
        sc[0] += (double)val;
        sc[1] += (double)val * cw;
      };
      if(!cm.act) {
      } else if(cm.km == 1) {
        const T p = res(0);
        const T a = p + __shfl_xor(p, cm.plane, 64);
        if(owner && (cm.k0 & 1) == 0) put(cm.k0, a);
      } else {
        for(int m = 0; m < cm.km; m += 2) {
          const T p0 = res(m), p1 = res(m + 1);
          if(owner) put(cm.k0 + m, p0 + p1);
        }
      }
    } else {
      const int pxy = L.nx * L.ny;
      double acc[1] = {0.0};
      for(int id = G.tid; id < L.n; id += G.nt) {
        const int k = small_div(id, pxy), rem = id - k * pxy, j = small_div(rem, L.nx), i = rem - j * L.nx;
        const T xv = smooth12_cell<T, UNI, int>(L, b, mb, m1, i, j, k, nb_x(L, i), nb_y(L, j), nb_z(L, k));
        x2[id] = xv;
        acc[0] += (double)xv;
      }
      G.sums(acc);
      const double mxd = zm ? acc[0] / (double)L.n : 0.0;
      if(G.tid == 0) lm[l][1] = mxd;
      const T mx = T(mxd);
      const T inv_children = T(1.0 / (double)(L.rx * L.ry * L.rz));
      T* bc = Bp(l + 1);
      const int cxy = Cl.nx * Cl.ny;
      for(int id = G.tid; id < Cl.n; id += G.nt) {
        const int kc = small_div(id, cxy), rem = id - kc * cxy, jc = small_div(rem, Cl.nx), ic = rem - jc * Cl.nx;
        const int i0 = ic * L.rx, j0 = jc * L.ry, k0 = kc * L.rz;
        auto R = [&](int i, int j, int k) {
          return residual_cell<T, UNI, int>(L, b, x2, mb, mx, i, j, k, nb_x(L, i), nb_y(L, j), nb_z(L, k));
        };
        // same combination order as the tiled kernel: x pairs, then y pairs, then z pairs
        auto row = [&](int j, int k) {
          T r = R(i0, j, k);
          if(L.rx == 2) r = r + R(i0 + 1, j, k);
          return r;
        };
        auto pl = [&](int k) {
          T r = row(j0, k);
          if(L.ry == 2) r = r + row(j0 + 1, k);
          return r;
        };
        T a = pl(k0);
        if(L.rz == 2) a = a + pl(k0 + 1);
        const T val = inv_children * a;
        bc[id] = val;
        sc[0] += (double)val;
        sc[1] += (double)val * Cl.winv[bclass(ic, Cl.nx) + 4 * (bclass(jc, Cl.ny) + 4 * bclass(kc, Cl.nz))];
      }
    }
    G.sums(sc);
    mbl = zm ? sc[0] / (double)Cl.n : 0.0;
    sbwl = sc[1];
    if(G.tid == 0) {
      lm[l + 1][0] = mbl;
      lsbw[l + 1] = sbwl;
    }
  };
  // up leg of level l: prolongation + sweep (x4 and its mean), last sweep (x5, its mean, r.z). On
  // entry mcl holds the mean of level l + 1's x in every thread of G; on return, level l's.
  auto up = [&](int l, const auto& G, double& mcl) {
    const PLevel& L = lvs[l];
    const PLevel& Cl = lvs[l + 1];
    const T* b = Bp(l);
    const T mb = T(lm[l][0]), m2 = T(lm[l][1]), mc = T(mcl);
    const bool col = UNI && col_ok(L, Cl, G.nt);
    const int pxy = L.nx * L.ny;
    double acc[1] = {0.0};
    if(col) {  // prolong_smooth_cell on the column map
      const ColMap cm = col_map(L, G.tid, G.nt);
      const T wcr = T(L.winv[0]) * T(L.wr);
      const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
      const T* x2 = XA(l);
      const T* xc = XA(l + 1);
      T* x4 = XB(l);
      const int cnx = Cl.nx, cplane = Cl.nx * Cl.ny;
      const int px = cm.i >> 1, pxm = ((cm.i - 1) & (L.nx - 1)) >> 1, pxp = ((cm.i + 1) & (L.nx - 1)) >> 1;
      const int py = (cm.j >> 1) * cnx, pym = (((cm.j - 1) & (L.ny - 1)) >> 1) * cnx,
                pyp = (((cm.j + 1) & (L.ny - 1)) >> 1) * cnx;
      for(int m = 0; cm.act && m < cm.km; ++m) {
        const int k = cm.k0 + m, o = cm.o0 + m * cm.plane;
        const int pz = (k >> 1) * cplane, pzm = (((k - 1) & (L.nz - 1)) >> 1) * cplane,
                  pzp = (((k + 1) & (L.nz - 1)) >> 1) * cplane;
        auto X3 = [&](int of, int oc) { return (x2[of] - m2) + (xc[oc] - mc); };
        const T c = X3(o, px + py + pz);
        const T ax = lap_neg<T>(c, X3(o + cm.dxm, pxm + py + pz), X3(o + cm.dxp, pxp + py + pz),
                                X3(o + cm.dym, px + pym + pz), X3(o + cm.dyp, px + pyp + pz),
                                X3(o + cm.dzm(k), px + py + pzm), X3(o + cm.dzp(k), px + py + pzp), ihx2, ihy2, ihz2);
        const T xv = c + wcr * ((b[o] - mb) + T(-1) * ax);
        x4[o] = xv;
        acc[0] += (double)xv;
      }
    } else {
      for(int id = G.tid; id < L.n; id += G.nt) {
        const int k = small_div(id, pxy), rem = id - k * pxy, j = small_div(rem, L.nx), i = rem - j * L.nx;
        const T xv = prolong_smooth_cell<T, UNI, int>(L, Cl, b, XA(l), XA(l + 1), mb, m2, mc, i, j, k, nb_x(L, i),
                                                      nb_y(L, j), nb_z(L, k));
        XB(l)[id] = xv;
        acc[0] += (double)xv;
      }
    }
    G.sums(acc);
    const double m4d = zm ? acc[0] / (double)L.n : 0.0;
    if(G.tid == 0) lm[l][2] = m4d;
    const T m4 = T(m4d);
    double s5[2] = {0.0, 0.0};  // sum of x5, r.z
    if(col) {  // smooth_last_cell on the column map
      const ColMap cm = col_map(L, G.tid, G.nt);
      const T wc = T(L.winv[0]);
      const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
      const T* x4 = XB(l);
      T* x5 = XA(l);
      auto X = [&](int o) { return x4[o] - m4; };
      for(int m = 0; cm.act && m < cm.km; ++m) {
        const int k = cm.k0 + m, o = cm.o0 + m * cm.plane;
        const T c = X(o);
        const T ax = lap_neg<T>(c, X(o + cm.dxm), X(o + cm.dxp), X(o + cm.dym), X(o + cm.dyp), X(o + cm.dzm(k)),
                                X(o + cm.dzp(k)), ihx2, ihy2, ihz2);
        const T bt = b[o] - mb;
        const T xv = c + wc * (bt + T(-1) * ax);
        x5[o] = xv;
        s5[0] += (double)xv;
        s5[1] += (double)bt * (double)xv;
      }
    } else {
      for(int id = G.tid; id < L.n; id += G.nt) {
        const int k = small_div(id, pxy), rem = id - k * pxy, j = small_div(rem, L.nx), i = rem - j * L.nx;
        T bt;
        const T xv =
            smooth_last_cell<T, UNI, int>(L, b, XB(l), mb, m4, i, j, k, nb_x(L, i), nb_y(L, j), nb_z(L, k), bt);
        XA(l)[id] = xv;
        s5[0] += (double)xv;
        s5[1] += (double)bt * (double)xv;
      }
    }
    G.sums(s5);
    mcl = zm ? s5[0] / (double)L.n : 0.0;
    if(G.tid == 0) {
      lm[l][3] = mcl;
      if(l == 0) rz_sh = s5[1];
    }
  };
  // coarsest level, rhs mean mbc: elimination with the recorded row swaps and multipliers
  auto coarsest = [&](const auto& G, double mbc) {
    const int lc = nl - 1;
    const int n = lvs[lc].n;
    const T* b = Bp(lc);
    T* x = XA(lc);
    const T mb = T(mbc);
    for(int i = G.tid; i < n; i += G.nt) work[i] = (double)(b[i] - mb);
    G.sync();
    if(cstage) {  // x = C (b - mb), row i on lane i; unrolled, so the reads go out together
      double xi = 0.0;
      if(G.tid < n) {
#pragma unroll
        for(int j = 0; j < kCoarseStage; ++j)
          if(j < n) xi += cop[G.tid * n + j] * work[j];
      }
      G.sync();
      if(G.tid < n) work[G.tid] = xi;
      if(G.tid == 0) lm[lc][3] = 0.0;
      G.sync();
    } else {  // the recorded elimination (linsolve/poisson_solver.cpp:482-567)
      const int* pv = perm;
      const double* fv = factor;
      const double* uv = U;
      if(zm && G.tid == 0) work[n - 1] = 0.0;
      G.sync();
      for(int p = 0; p < n; ++p) {
        if(G.tid == 0) {
          const int best = pv[p];
          if(best != p) {
            const double t = work[p];
            work[p] = work[best];
            work[best] = t;
          }
        }
        G.sync();
        const double wp = work[p];
        for(int row = p + 1 + G.tid; row < n; row += G.nt) {
          const double f = fv[(size_t)row * n + p];
          if(f != 0.0) work[row] -= f * wp;
        }
        G.sync();
      }
      if(G.tid == 0) {
        for(int row = n - 1; row >= 0; --row) {
          double s = work[row];
          for(int c = row + 1; c < n; ++c) s -= uv[(size_t)row * n + c] * work[c];
          work[row] = s / uv[(size_t)row * n + row];
        }
        if(zm) {
          double s = 0.0;
          for(int i = 0; i < n; ++i) s += work[i];
          const double m = s / (double)n;
          for(int i = 0; i < n; ++i) work[i] -= m;
        }
        lm[lc][3] = 0.0;
      }
      G.sync();
    }
    double rz[1] = {0.0};
    for(int i = G.tid; i < n; i += G.nt) {
      const T xv = T(work[i]);
      x[i] = xv;
      rz[0] += (double)(b[i] - mb) * (double)xv;
    }
    if(lc == 0) {  // the finest level is the coarsest: r.z of the direct solve
      G.sums(rz);
      if(G.tid == 0) rz_sh = rz[0];
    }
    G.sync();
  };
  __shared__ double tred[2][kTailWaves][4];
  const TailBlock GB{tid, (int)blockDim.x, tred, 0};
  // levels l0 .. lw - 1 on the workgroup, lw .. nl - 1 on the first wave
  auto on_block = [&](int l) {
    return lvs[l].n > kWaveCells || (UNI && lvs[l].n > 64 && col_ok(lvs[l], lvs[l + 1], kTailThreads));
  };
  int lw = l0;
  while(lw < nl - 1 && on_block(lw)) ++lw;
  const bool wave_coarsest = lvs[nl - 1].n <= kWaveCells;
  double mbl = lm[l0][0], sbwl = lsbw[l0];
  for(int l = l0; l < lw; ++l) down(l, GB, mbl, sbwl);
  if(tid < 64) {
    const TailWave GW{tid, 64};
    for(int l = lw; l < nl - 1; ++l) down(l, GW, mbl, sbwl);
    if(wave_coarsest) coarsest(GW, mbl);
  }
  __syncthreads();
  if(!wave_coarsest) coarsest(GB, mbl);  // every level above it ran on the workgroup
  double mcl = 0.0;  // mean of the coarsest level's x
  if(tid < 64) {
    const TailWave GW{tid, 64};
    for(int l = nl - 2; l >= lw; --l) up(l, GW, mcl);
  }
  __syncthreads();
  mcl = lm[min(lw, nl - 1)][3];
  for(int l = min(lw, nl - 1) - 1; l >= l0; --l) up(l, GB, mcl);
  if(LDS) {
    const int n0 = lvs[l0].n;
    const T* src = XA(l0);
    T* dst = B.xa[l0];
    for(int id = threadIdx.x; id < n0; id += blockDim.x) dst[id] = src[id];
  }
  if(tid == 0) {
    ps[s0 + L_MX5] = lm[l0][3];
    if(l0 == 0) {
      const double rz_old = ps[P_RZ];
      ps[P_BETA] = rz_old != 0.0 ? rz_sh / rz_old : 0.0;
      ps[P_RZ] = rz_sh;
    }
  }
}

// ---------------------------------------------------------------- PCG vector kernels
// Search direction (pointwise): d = (z - mz) + beta (d_old - md) (first: d = z - mz).
constexpr int kPointIlp = 4;
template <typename T>
__global__ void __launch_bounds__(256) k_direction(const T* __restrict__ z, const double* __restrict__ d_old,
                                                   double* __restrict__ d_new, int n, const double* ps, int first) {
  if(stopped(ps)) return;
  const double mz = ps[P_LEVEL + L_MX5];
  const double md = ps[P_MD];
  const double beta = first ? 0.0 : ps[P_BETA];
  const int stride = gridDim.x * 256 * kPointIlp;
  for(int base = blockIdx.x * 256 * kPointIlp + threadIdx.x; base < n; base += stride) {
    double zt[kPointIlp], dv[kPointIlp];
    if(base + (kPointIlp - 1) * 256 < n) {
      // whole chunk in range: unconditional loads (d_old is valid memory even when `first`
      // leaves it unused), so they are issued together
#pragma unroll
      for(int e = 0; e < kPointIlp; ++e) {
        const int id = base + e * 256;
        zt[e] = (double)z[id];
        dv[e] = d_old[id];
      }
#pragma unroll
      for(int e = 0; e < kPointIlp; ++e) {
        const double zc = zt[e] - mz;
        d_new[base + e * 256] = first ? zc : zc + beta * (dv[e] - md);
      }
      continue;
    }
#pragma unroll
    for(int e = 0; e < kPointIlp; ++e) {
      const int id = base + e * 256;
      if(id < n) {
        const double zc = (double)z[id] - mz;
        d_new[id] = first ? zc : zc + beta * (d_old[id] - md);
      }
    }
  }
}

// The k march of k_apply_dot. FULL: the block owns KZ planes, so there is no per-plane bound
// test (loads under such a test are waited for plane by plane). The pointers are restrict
// parameters of this function, so the next planes' loads may be scheduled above the current
// store. SHFL (the wave covers 64 active cells of a row): the x neighbours come from the
// adjacent lanes, and only the first and last lane load the neighbour beyond the tile (the
// others load their own cell, in the cache line they already read), so a plane costs four
// loads instead of five.
template <int KZ, bool UNI, bool FULL, bool SHFL>
__device__ __forceinline__ void apply_march(const double* __restrict__ d, double* __restrict__ q, const PLevel& L,
                                            const HotTile& t, double (&v)[3]) {
  const ZWin<UNI> zw(L);
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
  const Nb x = nb_x(L, t.i), y = nb_y(L, t.j);
  const size_t oxm = (size_t)x.cm + (size_t)L.nx * t.j, oxp = (size_t)x.cp + (size_t)L.nx * t.j;
  const size_t oym = (size_t)t.i + (size_t)L.nx * y.cm, oyp = (size_t)t.i + (size_t)L.nx * y.cp;
  const int tx = threadIdx.x;
  const size_t oxe = tx == 0 ? oxm : tx == 63 ? oxp : col;  // SHFL: the edge lanes' outside neighbour
  double prev = d[col + plane * zw.kprev(t.k0)];
  double cur = d[col + plane * t.k0];
  // software pipeline: the lateral neighbours of plane k + 1 and the centre of plane k + 2 are
  // loaded before plane k is computed
  double lw = d[(SHFL ? oxe : oxm) + plane * t.k0], le = SHFL ? 0.0 : d[oxp + plane * t.k0],
         ls = d[oym + plane * t.k0], ln = d[oyp + plane * t.k0];
  double next = d[col + plane * zw.knext(t.k0)];
#pragma unroll
  for(int kk = 0; kk < KZ; ++kk) {
    const int k = t.k0 + kk;
    if(FULL || k < t.k1) {
      double nlw = 0.0, nle = 0.0, nls = 0.0, nln = 0.0, nnext = 0.0;
      if(kk + 1 < KZ && (FULL || k + 1 < t.k1)) {
        const size_t pk1 = plane * (size_t)(k + 1);
        nlw = d[(SHFL ? oxe : oxm) + pk1];
        if(!SHFL) nle = d[oxp + pk1];
        nls = d[oym + pk1];
        nln = d[oyp + pk1];
        nnext = d[col + plane * zw.knext(k + 1)];
      }
      const size_t pk = plane * k;
      double wv, ev;
      if(SHFL) {
        const double up = __shfl_up(cur, 1, 64), dn = __shfl_down(cur, 1, 64);
        wv = tx == 0 ? sgn(x.sm, lw) : up;
        ev = tx == 63 ? sgn(x.sp, lw) : dn;
      } else {
        wv = sgn(x.sm, lw);
        ev = sgn(x.sp, le);
      }
      const double qv = lap_neg<double>(cur, wv, ev, sgn(y.sm, ls), sgn(y.sp, ln), zw.below(k, prev, cur),
                                        zw.above(k, next, cur), L.ihx2, L.ihy2, L.ihz2);
      q[col + pk] = qv;
      v[0] += cur;
      v[1] += qv;
      v[2] += cur * qv;
      prev = cur;
      cur = next;
      next = nnext;
      lw = nlw;
      le = nle;
      ls = nls;
      ln = nln;
    }
  }
}

// q = A d; sums of d, q, d.q.
template <int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_apply_dot(const double* __restrict__ d, double* __restrict__ q, PLevel L,
                                                   double* partials, const double* ps) {
  if(stopped(ps)) return;
  const HotTile t = hot_tile<KZ>(L);
  double v[3] = {0.0, 0.0, 0.0};
  if(t.in) {
    const bool rowfull = t.i - (int)threadIdx.x + 64 <= L.nx;  // uniform over the wave
    if(t.k1 - t.k0 == KZ) {
      if(rowfull) apply_march<KZ, UNI, true, true>(d, q, L, t, v);
      else apply_march<KZ, UNI, true, false>(d, q, L, t, v);
    } else {
      if(rowfull) apply_march<KZ, UNI, false, true>(d, q, L, t, v);
      else apply_march<KZ, UNI, false, false>(d, q, L, t, v);
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

// Search direction, operator and dot products in one pass over the grid (k_direction followed
// by k_apply_dot without the round trip of d through HBM): d = (z - mz) + beta (d_old - md)
// (first: z - mz) is formed wherever the stencil needs it (the cell, the next plane and the
// lateral neighbours, the latter from z and d_old read through the cache) and stored once;
// q = A d is formed for the sums of d, q and d.q and not stored (the residual update
// recomputes it from the stored d). Every d is the expression of k_direction, so d and the sums
// are the values of the two separate kernels.
struct DirCoef {
  double mz, md, beta;
  int first;
  template <typename T>
  __device__ double d(T zv, double dold) const {
    const double zc = (double)zv - mz;
    return first ? zc : zc + beta * (dold - md);
  }
};

template <typename T, int KZ, bool UNI, bool FULL, bool SHFL>
__device__ __forceinline__ void dir_apply_march(const T* __restrict__ z, const double* __restrict__ d_old,
                                                double* __restrict__ d_new, const PLevel& L, const HotTile& t,
                                                const DirCoef& cf, double (&v)[3]) {
  const ZWin<UNI> zw(L);
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
  const Nb x = nb_x(L, t.i), y = nb_y(L, t.j);
  const size_t oxm = (size_t)x.cm + (size_t)L.nx * t.j, oxp = (size_t)x.cp + (size_t)L.nx * t.j;
  const size_t oym = (size_t)t.i + (size_t)L.nx * y.cm, oyp = (size_t)t.i + (size_t)L.nx * y.cp;
  const int tx = threadIdx.x;
  const size_t oxe = tx == 0 ? oxm : tx == 63 ? oxp : col;  // SHFL: the edge lanes' outside neighbour
  // lateral inputs of one plane: z and d_old at the x neighbour(s) and the two y neighbours
  struct Lat {
    T zw, ze, zs, zn;
    double dw, de, ds, dn;
  };
  auto lat = [&](size_t pk, Lat& o) {
    o.zw = z[(SHFL ? oxe : oxm) + pk];
    o.dw = d_old[(SHFL ? oxe : oxm) + pk];
    if(!SHFL) {
      o.ze = z[oxp + pk];
      o.de = d_old[oxp + pk];
    }
    o.zs = z[oym + pk];
    o.ds = d_old[oym + pk];
    o.zn = z[oyp + pk];
    o.dn = d_old[oyp + pk];
  };
  const size_t kp0 = plane * (size_t)zw.kprev(t.k0);
  double prev = cf.d(z[col + kp0], d_old[col + kp0]);
  T zc = z[col + plane * t.k0];
  double dc = d_old[col + plane * t.k0];
  Lat la;
  lat(plane * t.k0, la);
  const size_t kn0 = plane * (size_t)zw.knext(t.k0);
  T zn = z[col + kn0];
  double dn = d_old[col + kn0];
#pragma unroll
  for(int kk = 0; kk < KZ; ++kk) {
    const int k = t.k0 + kk;
    if(FULL || k < t.k1) {
      // software pipeline: the lateral inputs of plane k + 1 and the centre of plane k + 2
      Lat nla{};
      T nzn = T(0);
      double ndn = 0.0;
      if(kk + 1 < KZ && (FULL || k + 1 < t.k1)) {
        lat(plane * (size_t)(k + 1), nla);
        const size_t kn = plane * (size_t)zw.knext(k + 1);
        nzn = z[col + kn];
        ndn = d_old[col + kn];
      }
      const double cur = cf.d(zc, dc);
      const double next = cf.d(zn, dn);
      double wv, ev;
      if(SHFL) {
        const double up = __shfl_up(cur, 1, 64), dnl = __shfl_down(cur, 1, 64);
        const double edge = cf.d(la.zw, la.dw);
        wv = tx == 0 ? sgn(x.sm, edge) : up;
        ev = tx == 63 ? sgn(x.sp, edge) : dnl;
      } else {
        wv = sgn(x.sm, cf.d(la.zw, la.dw));
        ev = sgn(x.sp, cf.d(la.ze, la.de));
      }
      const double sv = sgn(y.sm, cf.d(la.zs, la.ds)), nv = sgn(y.sp, cf.d(la.zn, la.dn));
      const double qv = lap_neg<double>(cur, wv, ev, sv, nv, zw.below(k, prev, cur), zw.above(k, next, cur), L.ihx2,
                                        L.ihy2, L.ihz2);
      d_new[col + plane * k] = cur;
      v[0] += cur;
      v[1] += qv;
      v[2] += cur * qv;
      prev = cur;
      zc = zn;
      dc = dn;
      zn = nzn;
      dn = ndn;
      la = nla;
    }
  }
}

template <typename T, int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_dir_apply(const T* __restrict__ z, const double* __restrict__ d_old,
                                                   double* __restrict__ d_new, PLevel L, double* partials,
                                                   const double* ps, int first) {
  if(stopped(ps)) return;
  DirCoef cf;
  cf.mz = ps[P_LEVEL + L_MX5];
  cf.md = ps[P_MD];
  cf.beta = first ? 0.0 : ps[P_BETA];
  cf.first = first;
  const HotTile t = hot_tile<KZ>(L);
  double v[3] = {0.0, 0.0, 0.0};
  if(t.in) {
    const bool rowfull = t.i - (int)threadIdx.x + 64 <= L.nx;  // uniform over the wave
    if(t.k1 - t.k0 == KZ) {
      if(rowfull) dir_apply_march<T, KZ, UNI, true, true>(z, d_old, d_new, L, t, cf, v);
      else dir_apply_march<T, KZ, UNI, true, false>(z, d_old, d_new, L, t, cf, v);
    } else {
      if(rowfull) dir_apply_march<T, KZ, UNI, false, true>(z, d_old, d_new, L, t, cf, v);
      else dir_apply_march<T, KZ, UNI, false, false>(z, d_old, d_new, L, t, cf, v);
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

// r = (r - mr) - alpha (q - mq); sums of r, r^2, r w/D (q stored by k_apply_dot; the path
// with separate direction and operator kernels). The matching p update is left to the pressure
// accumulation.
template <typename T, int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_update(double* __restrict__ r, const double* __restrict__ q,
                                                T* __restrict__ b_lo, PLevel L, double* partials, const double* ps,
                                                int write_lo) {
  if(stopped(ps)) return;
  const double alpha = ps[P_ALPHA], mr = ps[P_MR], mq = ps[P_MQ];
  const HotTile t = hot_tile<KZ>(L);
  double v[3] = {0.0, 0.0, 0.0};
  if(t.in) {
    const int cxy = bclass(t.i, L.nx) + 4 * bclass(t.j, L.ny);
    const size_t plane = (size_t)L.nx * L.ny;
    const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
    auto march = [&](auto full) {  // branch-free planes for a full k range (see k_apply_dot)
      constexpr bool FULL = decltype(full)::value;
#pragma unroll
      for(int kk = 0; kk < KZ; ++kk) {
        const int k = t.k0 + kk;
        if(FULL || k < t.k1) {
          const size_t id = col + plane * k;
          const double rn = (r[id] - mr) + -alpha * (q[id] - mq);
          r[id] = rn;
          if(write_lo) b_lo[id] = T(rn);
          v[0] += rn;
          v[1] += rn * rn;
          v[2] += rn * (UNI ? L.winv[0] : L.winv[cxy + 16 * bclass(k, L.nz)]);
        }
      }
    };
    if(t.k1 - t.k0 == KZ) march(std::true_type{});
    else march(std::false_type{});
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

// The residual update of the fused path: q = A d is recomputed from the stored direction by
// the march of k_apply_dot (the same neighbour values in the same expression, so q is bitwise
// the q whose sums k_dir_apply formed), then r = (r - mr) - alpha (q - mq), its FP32 copy
// (mixed) and the sums of r, r^2, r w/D. Reading d with its stencil instead of q saves the
// store and the load of q.
template <typename T, int KZ, bool UNI, bool FULL, bool SHFL>
__device__ __forceinline__ void update_apply_march(const double* __restrict__ d, double* __restrict__ r,
                                                   T* __restrict__ b_lo, const PLevel& L, const HotTile& t,
                                                   double alpha, double mr, double mq, int write_lo, double (&v)[3]) {
  const ZWin<UNI> zw(L);
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
  const Nb x = nb_x(L, t.i), y = nb_y(L, t.j);
  const size_t oxm = (size_t)x.cm + (size_t)L.nx * t.j, oxp = (size_t)x.cp + (size_t)L.nx * t.j;
  const size_t oym = (size_t)t.i + (size_t)L.nx * y.cm, oyp = (size_t)t.i + (size_t)L.nx * y.cp;
  const int tx = threadIdx.x;
  const size_t oxe = tx == 0 ? oxm : tx == 63 ? oxp : col;  // SHFL: the edge lanes' outside neighbour
  const int cxy = bclass(t.i, L.nx) + 4 * bclass(t.j, L.ny);
  double prev = d[col + plane * zw.kprev(t.k0)];
  double cur = d[col + plane * t.k0];
  double rc = r[col + plane * t.k0];
  double lw = d[(SHFL ? oxe : oxm) + plane * t.k0], le = SHFL ? 0.0 : d[oxp + plane * t.k0],
         ls = d[oym + plane * t.k0], ln = d[oyp + plane * t.k0];
  double next = d[col + plane * zw.knext(t.k0)];
#pragma unroll
  for(int kk = 0; kk < KZ; ++kk) {
    const int k = t.k0 + kk;
    if(FULL || k < t.k1) {
      // software pipeline: plane k + 1's lateral neighbours and residual, plane k + 2's centre
      double nlw = 0.0, nle = 0.0, nls = 0.0, nln = 0.0, nnext = 0.0, nr = 0.0;
      if(kk + 1 < KZ && (FULL || k + 1 < t.k1)) {
        const size_t pk1 = plane * (size_t)(k + 1);
        nlw = d[(SHFL ? oxe : oxm) + pk1];
        if(!SHFL) nle = d[oxp + pk1];
        nls = d[oym + pk1];
        nln = d[oyp + pk1];
        nnext = d[col + plane * zw.knext(k + 1)];
        nr = r[col + pk1];
      }
      double wv, ev;
      if(SHFL) {
        const double up = __shfl_up(cur, 1, 64), dn = __shfl_down(cur, 1, 64);
        wv = tx == 0 ? sgn(x.sm, lw) : up;
        ev = tx == 63 ? sgn(x.sp, lw) : dn;
      } else {
        wv = sgn(x.sm, lw);
        ev = sgn(x.sp, le);
      }
      const double qv = lap_neg<double>(cur, wv, ev, sgn(y.sm, ls), sgn(y.sp, ln), zw.below(k, prev, cur),
                                        zw.above(k, next, cur), L.ihx2, L.ihy2, L.ihz2);
      const size_t id = col + plane * k;
      const double rn = (rc - mr) + -alpha * (qv - mq);
      r[id] = rn;
      if(write_lo) b_lo[id] = T(rn);
      v[0] += rn;
      v[1] += rn * rn;
      v[2] += rn * (UNI ? L.winv[0] : L.winv[cxy + 16 * bclass(k, L.nz)]);
      prev = cur;
      cur = next;
      next = nnext;
      lw = nlw;
      le = nle;
      ls = nls;
      ln = nln;
      rc = nr;
    }
  }
}

template <typename T, int KZ, bool UNI>
__global__ void __launch_bounds__(256) k_update_apply(const double* __restrict__ d, double* __restrict__ r,
                                                      T* __restrict__ b_lo, PLevel L, double* partials,
                                                      const double* ps, int write_lo) {
  if(stopped(ps)) return;
  const double alpha = ps[P_ALPHA], mr = ps[P_MR], mq = ps[P_MQ];
  const HotTile t = hot_tile<KZ>(L);
  double v[3] = {0.0, 0.0, 0.0};
  if(t.in) {
    const bool rowfull = t.i - (int)threadIdx.x + 64 <= L.nx;  // uniform over the wave
    if(t.k1 - t.k0 == KZ) {
      if(rowfull) update_apply_march<T, KZ, UNI, true, true>(d, r, b_lo, L, t, alpha, mr, mq, write_lo, v);
      else update_apply_march<T, KZ, UNI, true, false>(d, r, b_lo, L, t, alpha, mr, mq, write_lo, v);
    } else {
      if(rowfull) update_apply_march<T, KZ, UNI, false, true>(d, r, b_lo, L, t, alpha, mr, mq, write_lo, v);
      else update_apply_march<T, KZ, UNI, false, false>(d, r, b_lo, L, t, alpha, mr, mq, write_lo, v);
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

// ---- the two CG stencil passes with the lateral neighbours shared through LDS
// On a grid periodic on every axis whose 64 x 4 tiles and 8-plane ranges are all full, each
// block stages one plane of the stencil operand at a time: its 4 rows plus the one-cell ring
// around them (rows j0-1 and j0+4 loaded by the first and last wave, the two halo columns by
// the second), formed once per cell and read by the neighbours from LDS instead of through the
// cache. The plane buffers alternate by parity, so a plane needs one barrier: a buffer is
// rewritten two planes later, after every thread has passed the barrier of the plane between.
// Tile mapping, per-thread order, values and partial sums are those of k_dir_apply and
// k_update_apply (the periodic full-row march), bitwise.
constexpr int kLdsKz = 8;
struct RingHalo {
  int hr, hc;  // LDS row/column of this thread's ring cell (hr < 0: none)
  size_t off;  // its column offset within a plane
  __device__ RingHalo(const PLevel& L, int i0, int j0, int tx, int ty) : hr(-1), hc(-1), off(0) {
    const int nx = L.nx, ny = L.ny;
    if(ty == 0) {
      hr = 0;
      hc = tx + 1;
      off = (size_t)(i0 + tx) + (size_t)nx * (size_t)(j0 == 0 ? ny - 1 : j0 - 1);
    } else if(ty == 3) {
      hr = 5;
      hc = tx + 1;
      off = (size_t)(i0 + tx) + (size_t)nx * (size_t)(j0 + 4 == ny ? 0 : j0 + 4);
    } else if(ty == 1 && tx < 12) {
      const int side = tx / 6;
      hr = tx - 6 * side;
      hc = side ? 65 : 0;
      const int gi = side ? (i0 + 64 == nx ? 0 : i0 + 64) : (i0 == 0 ? nx - 1 : i0 - 1);
      const int gj = j0 - 1 + hr;
      off = (size_t)gi + (size_t)nx * (size_t)(gj < 0 ? gj + ny : (gj >= ny ? gj - ny : gj));
    }
  }
};
__device__ inline int wrap_plane(int k, int nz) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); }

template <typename T>
__global__ void __launch_bounds__(256) k_dir_apply_lds(const T* __restrict__ z, const double* __restrict__ d_old,
                                                       double* __restrict__ d_new, PLevel L, double* partials,
                                                       const double* ps, int first) {
  if(stopped(ps)) return;
  constexpr int KZ = kLdsKz;
  __shared__ double sd[2][6][68];
  DirCoef cf;
  cf.mz = ps[P_LEVEL + L_MX5];
  cf.md = ps[P_MD];
  cf.beta = first ? 0.0 : ps[P_BETA];
  cf.first = first;
  const HotTile t = hot_tile<KZ>(L);
  const int nz = L.nz, tx = threadIdx.x, ty = threadIdx.y;
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
  const RingHalo h(L, t.i - tx, t.j - ty, tx, ty);
  const bool has_h = h.hr >= 0;
  // A rolled march with two planes of look-ahead in named registers: the centre values of
  // plane k + 3 and the ring values of plane k + 2 are requested while plane k is computed.
  // Every load is unconditional (h.off is 0 without a ring cell), so no branch join waits for
  // one; a fully unrolled march let the compiler hoist every plane's loads to the start (94
  // VGPRs, five waves per SIMD instead of eight).
  auto at_c = [&](int k) { return col + plane * (size_t)wrap_plane(k, nz); };
  auto at_h = [&](int k) { return h.off + plane * (size_t)k; };
  const T z0 = z[at_c(t.k0 - 1)], z1 = z[at_c(t.k0)];
  const double d0 = d_old[at_c(t.k0 - 1)], d1 = d_old[at_c(t.k0)];
  T zn1 = z[at_c(t.k0 + 1)], zn2 = z[at_c(t.k0 + 2)], zh1 = z[at_h(t.k0)], zh2 = z[at_h(t.k0 + 1)];
  double dn1 = d_old[at_c(t.k0 + 1)], dn2 = d_old[at_c(t.k0 + 2)], dh1 = d_old[at_h(t.k0)],
         dh2 = d_old[at_h(t.k0 + 1)];
  double prev = cf.d(z0, d0);
  double cur = cf.d(z1, d1);
  double v[3] = {0.0, 0.0, 0.0};
  auto plane_step = [&](int kk, double next, T zhv, double dhv) {
    double(&buf)[6][68] = sd[kk & 1];
    buf[ty + 1][tx + 1] = cur;
    if(has_h) buf[h.hr][h.hc] = cf.d(zhv, dhv);
    __syncthreads();
    const double qv = lap_neg<double>(cur, buf[ty + 1][tx], buf[ty + 1][tx + 2], buf[ty][tx + 1], buf[ty + 2][tx + 1],
                                      prev, next, L.ihx2, L.ihy2, L.ihz2);
    d_new[col + plane * (size_t)(t.k0 + kk)] = cur;
    v[0] += cur;
    v[1] += qv;
    v[2] += cur * qv;
    prev = cur;
    cur = next;
  };
#pragma unroll 1
  for(int kk = 0; kk < KZ - 2; ++kk) {
    const int k = t.k0 + kk;
    const T zl = z[at_c(k + 3)], zhl = z[at_h(k + 2)];
    const double dl = d_old[at_c(k + 3)], dhl = d_old[at_h(k + 2)];
    plane_step(kk, cf.d(zn1, dn1), zh1, dh1);
    zn1 = zn2;
    dn1 = dn2;
    zn2 = zl;
    dn2 = dl;
    zh1 = zh2;
    dh1 = dh2;
    zh2 = zhl;
    dh2 = dhl;
  }
  plane_step(KZ - 2, cf.d(zn1, dn1), zh1, dh1);
  plane_step(KZ - 1, cf.d(zn2, dn2), zh2, dh2);
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

template <typename T>
__global__ void __launch_bounds__(256) k_update_apply_lds(const double* __restrict__ d, double* __restrict__ r,
                                                          T* __restrict__ b_lo, PLevel L, double* partials,
                                                          const double* ps, int write_lo) {
  if(stopped(ps)) return;
  constexpr int KZ = kLdsKz;
  __shared__ double sd[2][6][68];
  const double alpha = ps[P_ALPHA], mr = ps[P_MR], mq = ps[P_MQ];
  const HotTile t = hot_tile<KZ>(L);
  const int nz = L.nz, tx = threadIdx.x, ty = threadIdx.y;
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t col = (size_t)t.i + (size_t)L.nx * t.j;
  const RingHalo h(L, t.i - tx, t.j - ty, tx, ty);
  const bool has_h = h.hr >= 0;
  double dc[KZ + 2], rc[KZ], dh[KZ];
  auto load_c = [&](int m) { dc[m] = d[col + plane * (size_t)wrap_plane(t.k0 - 1 + m, nz)]; };
  auto load_p = [&](int m) {
    const size_t kp = plane * (size_t)(t.k0 + m);
    rc[m] = r[col + kp];
    dh[m] = has_h ? d[h.off + kp] : 0.0;
  };
  load_c(0);
  load_c(1);
  load_c(2);
  load_p(0);
  double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for(int kk = 0; kk < KZ; ++kk) {
    if(kk + 3 < KZ + 2) load_c(kk + 3);
    if(kk + 1 < KZ) load_p(kk + 1);
    const double cur = dc[kk + 1];
    double(&buf)[6][68] = sd[kk & 1];
    buf[ty + 1][tx + 1] = cur;
    if(has_h) buf[h.hr][h.hc] = dh[kk];
    __syncthreads();
    const double qv = lap_neg<double>(cur, buf[ty + 1][tx], buf[ty + 1][tx + 2], buf[ty][tx + 1], buf[ty + 2][tx + 1],
                                      dc[kk], dc[kk + 2], L.ihx2, L.ihy2, L.ihz2);
    const size_t id = col + plane * (size_t)(t.k0 + kk);
    const double rn = (rc[kk] - mr) + -alpha * (qv - mq);
    r[id] = rn;
    if(write_lo) b_lo[id] = T(rn);
    v[0] += rn;
    v[1] += rn * rn;
    v[2] += rn * L.winv[0];
  }
  const int ops[3] = {kSum, kSum, kSum};
  store_partials<3>(v, ops, partials);
}

template <typename T>
__global__ void k_cast(const double* __restrict__ in, T* __restrict__ out, int n, const double* ps) {
  if(stopped(ps)) return;
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) out[id] = T(in[id]);
}

// ---------------------------------------------------------------- pressure accumulation
// The search directions of a solve live in a ring of R buffers: iteration k writes slot k % R.
// The pressure p = p0 + sum_k alpha_k (d_k - md_k) is formed per cell in iteration order,
// p = p + alpha_k (d_k - md_k), the update the reference applies after every iteration, so the
// result is bitwise that of a per-iteration update. Only iterations whose residual update ran
// count (the first ps[P_ITER]; a breakdown stops before its update). The slot pointers travel
// in the kernel arguments, alpha_k and md_k in coef (written by EpiUpdate).
constexpr int kMaxRing = 32;
struct DirRing {
  const double* d[kMaxRing];
  const double* coef;
  int R;
};

__device__ inline double accumulate(double p, const DirRing& ring, int id, int k0, int k1) {
#pragma unroll 4
  for(int k = k0; k < k1; ++k) p = p + ring.coef[2 * k] * (ring.d[k % ring.R][id] - ring.coef[2 * k + 1]);
  return p;
}

// Folds directions [k0, min(k1, iterations)) into the compact p (from zero when from_zero) before
// their slots are reused.
__global__ void k_pflush(double* __restrict__ pc, DirRing ring, int k0, int k1, int n, int from_zero,
                         const double* ps) {
  const int kend = min(k1, (int)ps[P_ITER]);
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x)
    pc[id] = accumulate(from_zero ? 0.0 : pc[id], ring, id, k0, kend);
}

// The last directions [k0, iterations) added to the compact p (or to zero), written into the
// active range of the ghosted result, with the sum of p for its mean.
__global__ void k_pfinal(double* __restrict__ p_g, const double* __restrict__ pc, DirRing ring, int k0,
                         int from_zero, PLevel L, Lay Lc, double* partials, const double* ps) {
  const int kend = (int)ps[P_ITER];
  double v[1] = {0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    const double p = accumulate(from_zero ? 0.0 : pc[id], ring, id, k0, kend);
    const int row = id / L.nx;
    p_g[Lc.at(id - row * L.nx + 1, row % L.ny + 1, row / L.ny + 1)] = p;
    v[0] += p;
  }
  const int ops[1] = {kSum};
  store_partials<1>(v, ops, partials);
}

// Removes the final mean from the active range of p (problems without a Dirichlet face).
__global__ void k_psub(double* __restrict__ p_g, PLevel L, Lay Lc, const double* ps, int kz) {
  const double m = ps[P_PFINAL];
  const Tile t = tile(L, kz);
  if(t.i >= L.nx || t.j >= L.ny) return;
  for(int k = t.k0; k < t.k1; ++k) {
    const size_t gi = Lc.at(t.i + 1, t.j + 1, k + 1);
    p_g[gi] = p_g[gi] - m;
  }
}

// Poisson operator on ghosted storage with the actual BCs (linsolve/poisson_solver.cpp:244-269).
__global__ void k_poisson_apply_g(const double* __restrict__ in, double* __restrict__ out, Lay Lc, PLevel L,
                                  int homogeneous, int kz) {
  const Tile t = tile(L, kz);
  if(t.i >= L.nx || t.j >= L.ny) return;
  const int i = t.i, j = t.j;
  for(int k = t.k0; k < t.k1; ++k) {
    const double c = in[Lc.at(i + 1, j + 1, k + 1)];
    auto X = [&](int a, int b, int e) { return in[Lc.at(a + 1, b + 1, e + 1)]; };
    double nb[3][2];
    const int co[3] = {i, j, k};
    const int nn[3] = {L.nx, L.ny, L.nz};
    for(int a = 0; a < 3; ++a) {
      const int di = a == 0, dj = a == 1, dk = a == 2;
      const double lv = homogeneous ? 0.0 : L.lo_val[a], hv = homogeneous ? 0.0 : L.hi_val[a];
      const double lg = homogeneous ? 0.0 : L.lo_grad[a], hg = homogeneous ? 0.0 : L.hi_grad[a];
      if(co[a] > 0) nb[a][0] = X(i - di, j - dj, k - dk);
      else if(L.per[a]) nb[a][0] = X(a == 0 ? nn[0] - 1 : i, a == 1 ? nn[1] - 1 : j, a == 2 ? nn[2] - 1 : k);
      else nb[a][0] = L.lo_dir[a] ? 2.0 * lv - c : c - lg * L.h[a];
      if(co[a] < nn[a] - 1) nb[a][1] = X(i + di, j + dj, k + dk);
      else if(L.per[a]) nb[a][1] = X(a == 0 ? 0 : i, a == 1 ? 0 : j, a == 2 ? 0 : k);
      else nb[a][1] = L.hi_dir[a] ? 2.0 * hv - c : c + hg * L.h[a];
    }
    out[Lc.at(i + 1, j + 1, k + 1)] = -(nb[0][1] - 2.0 * c + nb[0][0]) / L.hx2 -
                                      (nb[1][1] - 2.0 * c + nb[1][0]) / L.hy2 -
                                      (nb[2][1] - 2.0 * c + nb[2][0]) / L.hz2;
  }
}

inline bool axis_coarsens(int n) { return n > 1 && n % 2 == 0 && n > 2; }  // :316-345

// Launch shape of a level kernel: 64x4 tiles, each thread marching kz planes (kz even).
struct LevelLaunch {
  dim3 grid, block;
  int kz;
};
LevelLaunch level_launch(const PLevel& L) {
  LevelLaunch l;
  l.block = dim3(64, 4, 1);
  const unsigned gx = (unsigned)((L.nx + 63) / 64), gy = (unsigned)((L.ny + 3) / 4);
  const size_t plane = (size_t)gx * gy;
  int kz = 2;
  while(plane * (size_t)((L.nz + kz - 1) / kz) > 4096 && kz < L.nz) kz *= 2;
  l.kz = kz;
  l.grid = dim3(gx, gy, (unsigned)((L.nz + kz - 1) / kz));
  return l;
}

// Launch shape of the hot level kernels: 1D grid of 64x4 tiles, KZ = 8 planes per thread when
// that still gives >= 1024 tiles, else 2.
bool lds_cg_disabled() {
  const char* e = std::getenv("SOLVER_AMD_CG_LDS");
  return e != nullptr && e[0] == '0';
}
struct HotLaunch {
  dim3 grid, block;
  int kz;
  bool uni;
  bool lds;  // the LDS-staged CG passes apply: periodic, all tiles and plane ranges full
};
HotLaunch hot_launch(const PLevel& L) {
  HotLaunch h;
  h.block = dim3(64, 4, 1);
  const size_t plane = (size_t)((L.nx + 63) / 64) * (size_t)((L.ny + 3) / 4);
  h.kz = plane * (size_t)((L.nz + 7) / 8) >= 1024 ? 8 : 2;
  h.grid = dim3((unsigned)(plane * (size_t)((L.nz + h.kz - 1) / h.kz)));
  h.uni = L.per[0] && L.per[1] && L.per[2];
  h.lds = h.uni && h.kz == kLdsKz && L.nx % 64 == 0 && L.ny % 4 == 0 && L.nz % kLdsKz == 0 && !lds_cg_disabled();
  if(h.grid.x > (unsigned)kMaxPartials) throw InvalidArgument("grid too large for the reduction buffer");
  return h;
}

// Launch shape of the fused all-periodic level kernels: 64 x rows tiles, kz planes per block (even,
// so that z pairs of the restriction stay inside a block). Each block re-reads the 4 planes
// around its range and pays a pipeline fill of 4 plane steps, so long ranges are cheaper; the
// grid is sized to one round of resident blocks (occupancy x compute units), which also leaves
// no partial last round: kz is the shortest even range that needs at most `resident` blocks
// (at least 2 planes).
struct FusedLaunch {
  dim3 grid;
  int kz;
};
FusedLaunch fused_launch(const PLevel& L, int resident, int rows) {
  FusedLaunch f;
  const size_t tiles = (size_t)((L.nx + kFX - 1) / kFX) * (size_t)((L.ny + rows - 1) / rows);
  const size_t chunks = std::max<size_t>(1, (size_t)std::max(resident, 1) / tiles);
  int kz = (int)(((size_t)L.nz + chunks - 1) / chunks);
  kz = std::max(2, kz + (kz & 1));
  f.kz = kz;
  const size_t n = tiles * (size_t)((L.nz + f.kz - 1) / f.kz);
  if(n > (size_t)kMaxPartials) throw InvalidArgument("grid too large for the reduction buffer");
  f.grid = dim3((unsigned)n);
  return f;
}

// The one-shot small-level kernels (k_down_small, k_up_small) own 16-row tiles of two planes
// whatever row count the marching kernels use: their grid is every such tile of the level.
FusedLaunch small_launch(const PLevel& L) {
  FusedLaunch f;
  f.kz = 2;
  const size_t tiles = (size_t)((L.nx + kFX - 1) / kFX) * (size_t)((L.ny + 15) / 16);
  const size_t n = tiles * (size_t)((L.nz + 1) / 2);
  if(n > (size_t)kMaxPartials) throw InvalidArgument("grid too large for the reduction buffer");
  f.grid = dim3((unsigned)n);
  return f;
}

// Blocks of `kernel` (launched with `threads` threads) that the device holds at once.
template <typename K>
int resident_blocks(K kernel, int threads, int device) {
  int per_cu = 0, cus = 0;
  OFH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  OFH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  return std::max(1, per_cu) * std::max(1, cus);
}

// Instantiates KERN<pre..., KZ, UNI> for the launch shape h.
#define OFH_HOT_LAUNCH(KERN, h, stream, ...)                                                  \
  do {                                                                                        \
    if((h).kz == 8) {                                                                         \
      if((h).uni)                                                                             \
        OFH_LAUNCH((KERN, 8, true>), (h).grid, (h).block, 0, stream, __VA_ARGS__);    \
      else                                                                                    \
        OFH_LAUNCH((KERN, 8, false>), (h).grid, (h).block, 0, stream, __VA_ARGS__);   \
    } else {                                                                                  \
      if((h).uni)                                                                             \
        OFH_LAUNCH((KERN, 2, true>), (h).grid, (h).block, 0, stream, __VA_ARGS__);    \
      else                                                                                    \
        OFH_LAUNCH((KERN, 2, false>), (h).grid, (h).block, 0, stream, __VA_ARGS__);   \
    }                                                                                         \
  } while(0)

}  // namespace

// ------------------------------------------------------------------ solver object
class PoissonSolver {
 public:
  Context& c;
  of_grid g;
  of_pressure_bc_set bcs;
  int precision;
  bool zm = true;
  std::vector<PLevel> lv;
  // finest-level CG vectors (compact FP64): the pressure, the residual and the ring of search
  // directions (q only for the path with separate direction and operator kernels)
  double *pc = nullptr, *r = nullptr, *q = nullptr;
  std::vector<double*> ring;
  double* coef = nullptr;  // alpha_k and md_k per iteration (capacity hist_cap - 1 iterations)
  // level arrays: b (rhs) and two x buffers, element type T
  std::vector<void*> lb, lxa, lxb;
  double* ps = nullptr;       // scalar block
  double* history = nullptr;  // residual history (device)
  double* host_ps = nullptr;  // mapped pinned mirror of the leading scalars (2 slots)
  double* host_ps_dev = nullptr;
  PLevel* dlv = nullptr;      // level descriptors for the tail kernel
  int l0 = 0;                 // first level handled by the one-workgroup tail
  std::vector<char> fused;    // per level: the fused all-periodic kernels run it
  int resident_down = 0, resident_up = 0;  // blocks of the fused kernels the device holds at once
  int rows_down = 8, rows_up = 8;          // tile rows of the fused kernels (8: 256 threads, 16: 512)
  int up_bufs = 3;                         // prefetch buffers of the 16-row fused up kernel (2 or 3)
  bool up_zw = true;                       // 16-row up leg on per-point z windows (k_up_zw)
  double* partials2 = nullptr;  // second partial buffer: a folding consumer writes beside its input
  // coarse solve data
  int nc = 0;
  int* perm = nullptr;
  double *factor = nullptr, *U = nullptr, *work = nullptr;
  int hist_cap = 0;
  double omega = kOmega;  // Jacobi weight of every level's first pre-smoothing sweep
  // Device memory source: null -> hipMalloc per array (freed by the destructor); otherwise every
  // array is carved from the arena and nothing is allocated or freed. With an arena the host
  // reads the stop flag from device memory after each iteration (no mapped pinned mirror), so
  // the solve allocates nothing at all.
  DeviceArena* arena = nullptr;
  void dev_alloc(void** out, size_t bytes) {
    if(arena) *out = arena->take(bytes);
    else OFH_CHECK(hipMalloc(out, bytes));
  }

  PoissonSolver(Context& ctx, const of_grid& grid, const of_pressure_bc_set& b, int prec, DeviceArena* a = nullptr)
      : c(ctx), g(grid), bcs(b), precision(prec), arena(a) {}
  ~PoissonSolver();
  void build();
  template <typename T>
  void vcycle();
  template <typename T>
  PoissonResult solve_t(const double* rhs, double* p, int maxit, double tol, bool want_history, bool zero_guess,
                        const double** deferred_mean);
  // the array holding a level's smoothed correction after the up leg: the fused up kernel
  // writes it beside x2 (its halo is still being read by neighbouring tiles)
  void* x_final(int l) const { return l < l0 && fused[l] ? lxb[l] : lxa[l]; }
  // Waits until the update of iteration `it` has published its stop flag in the mapped mirror
  // and returns the flag. The host spins on the coherent mapping; after about a second without
  // the sequence number it falls back to waiting for the whole stream (the update has then run).
  double wait_mirror(int it) const {
    const volatile double* hm = host_ps + (it % 2) * kHostMirror;
    for(long spin = 0; hm[kHostSeq] < (double)(it + 1); ++spin) {
      if(spin == (1L << 26)) {
        OFH_CHECK(hipStreamSynchronize(c.stream));
        if(hm[kHostSeq] < (double)(it + 1)) throw HipError("the residual update's stop flag did not arrive");
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return hm[P_DONE];
  }
};

namespace {

// SOLVER_AMD_MG_UNFUSED=1 selects the unfused level kernels everywhere (they differ from the
// fused ones only in rounding); read when a solver is built.
bool fused_disabled() {
  const char* e = std::getenv("SOLVER_AMD_MG_UNFUSED");
  return e != nullptr && e[0] == '1';
}

// SOLVER_AMD_MG_SMALL=0 keeps the marching fused kernels on the small levels (two planes per
// block) instead of the one-shot variants; read when a V-cycle is queued.
bool small_levels() {
  const char* e = std::getenv("SOLVER_AMD_MG_SMALL");
  return e == nullptr || e[0] != '0';
}

// SOLVER_AMD_CG_UNFUSED=1 runs the search direction and the operator as two kernels
// (k_direction, k_apply_dot) instead of k_dir_apply; read when a solve starts.
bool cg_fused_disabled() {
  const char* e = std::getenv("SOLVER_AMD_CG_UNFUSED");
  return e != nullptr && e[0] == '1';
}

// Slots of the search-direction ring: directions are kept until the pressure pass adds them
// to p, so a solve of at most R iterations reads and writes p once; a longer one folds R
// directions into p every R iterations. About 4 GiB of directions (32 slots up to 256^3 cells,
// 4 at 512^3, at least 2); 4 slots in a no-allocation arena. SOLVER_AMD_CG_RING (2 to 32),
// read when a solver is built, overrides the count.
int ring_slots(size_t cells, bool arena) {
  const char* e = std::getenv("SOLVER_AMD_CG_RING");
  if(e != nullptr && e[0] != 0) {
    const int v = std::atoi(e);
    if(v < 2 || v > kMaxRing) throw InvalidArgument("SOLVER_AMD_CG_RING must be between 2 and 32");
    return v;
  }
  if(arena) return 4;
  const size_t budget = (size_t)4 << 30;
  return (int)std::min<size_t>(std::max<size_t>(budget / (cells * sizeof(double)), 2), kMaxRing);
}

// Jacobi weights of the FP32 V-cycle of a MIXED solve on a grid of dimension `dims`.
// SOLVER_AMD_MIXED_OMEGA, read when a solver is built, overrides them: one weight w for both
// sweeps, or a pair "w1,w2" (first and second pre-smoothing sweep; the post-smoothing sweeps
// run them in reverse order) whose polynomial (1 - w1 t)(1 - w2 t) stays within (-1, 1) on
// t in (0, 2].
void mixed_omega(int dims, double& w1, double& w2) {
  chebyshev_pair(dims >= 3 ? kMixedCheb3D : kMixedCheb2D, w1, w2);
  const char* e = std::getenv("SOLVER_AMD_MIXED_OMEGA");
  if(e == nullptr || e[0] == 0) return;
  const char* comma = std::strchr(e, ',');
  w1 = std::atof(e);
  w2 = comma != nullptr ? std::atof(comma + 1) : w1;
  const double p2 = (1.0 - 2.0 * w1) * (1.0 - 2.0 * w2);  // at t = 2
  const double tm = (w1 + w2) / (2.0 * w1 * w2);           // where the polynomial is smallest
  const double pm = 1.0 - (w1 + w2) * (w1 + w2) / (4.0 * w1 * w2);
  if(!(w1 > 0.0 && w2 > 0.0 && std::fabs(p2) < 1.0 && (tm > 2.0 || pm > -1.0)))
    throw InvalidArgument("SOLVER_AMD_MIXED_OMEGA must give a convergent smoother");
}

// SOLVER_AMD_FP64_CHEBYSHEV=1, read when a solver is built, gives an FP64 solve the MIXED mode's
// Chebyshev sweep weights (or SOLVER_AMD_MIXED_OMEGA's). Not the reference's smoother, so FP64
// iteration counts then differ from the reference's: it exists to measure how much of the
// mixed mode's speed comes from the weights and how much from FP32 arithmetic.
bool fp64_chebyshev() {
  const char* e = std::getenv("SOLVER_AMD_FP64_CHEBYSHEV");
  return e != nullptr && e[0] == '1';
}

// Prefetch buffers of the 16-row fused up kernel per precision, SOLVER_AMD_UP_BUFFERS =
// "fp64,mixed" (each 2 or 3), read when a solver is built. Defaults: 2 in FP64 (128 VGPRs, two
// blocks per compute unit instead of one: the 256^3 step ran 17.34 -> 17.10 ms), 3 in FP32
// (its two-buffer variant needs scratch spills to reach three blocks and ran slower).
int up_buffers(int precision) {
  int r[2] = {2, 3};
  if(const char* e = std::getenv("SOLVER_AMD_UP_BUFFERS")) {
    if(std::sscanf(e, "%d,%d", &r[0], &r[1]) != 2 || r[0] < 2 || r[0] > 3 || r[1] < 2 || r[1] > 3)
      throw InvalidArgument("SOLVER_AMD_UP_BUFFERS must hold two buffer counts (2 or 3)");
  }
  return precision == OF_PRECISION_MIXED ? r[1] : r[0];
}

// SOLVER_AMD_UP_ZW=0 keeps the two-barrier 16-row up kernel (k_up_fused) instead of the one on
// per-point z windows (k_up_zw); read when a solver is built.
bool up_windows() {
  const char* e = std::getenv("SOLVER_AMD_UP_ZW");
  return e == nullptr || e[0] != '0';
}

// Tile rows of the fused level kernels per precision, overridden by SOLVER_AMD_FUSED_ROWS =
// "down64,up64,down32,up32" (each 8 or 16) when a solver is built. 16 rows (512 threads) is the
// default: the halo overhead falls from 1.59x to 1.33x of the tile and the FP64 down kernel
// reaches four waves per SIMD (80 KB of LDS per block); the 256^3 steps ran 2.2% (FP64) and 1.6%
// (mixed) faster than with 8 rows.
void fused_rows(int precision, int& down, int& up) {
  int r[4] = {16, 16, 16, 16};
  if(const char* e = std::getenv("SOLVER_AMD_FUSED_ROWS")) {
    if(std::sscanf(e, "%d,%d,%d,%d", &r[0], &r[1], &r[2], &r[3]) != 4)
      throw InvalidArgument("SOLVER_AMD_FUSED_ROWS must hold four row counts");
    for(int v : r)
      if(v != 8 && v != 16) throw InvalidArgument("SOLVER_AMD_FUSED_ROWS: rows must be 8 or 16");
  }
  const bool fp64 = precision != OF_PRECISION_MIXED;
  down = fp64 ? r[0] : r[2];
  up = fp64 ? r[1] : r[3];
}

bool p_axis_periodic(const of_pressure_bc_set& b, int axis) {
  const bool lo = b.faces[2 * axis].type == OF_PRESSURE_PERIODIC, hi = b.faces[2 * axis + 1].type == OF_PRESSURE_PERIODIC;
  if(lo != hi) throw InvalidArgument("periodic pressure boundaries must be paired on both sides");
  return lo;
}

PLevel make_level(int nx, int ny, int nz, double dx, double dy, double dz, const of_pressure_bc_set& b,
                  double omega = kOmega) {
  PLevel L{};
  L.nx = nx;
  L.ny = ny;
  L.nz = nz;
  L.n = nx * ny * nz;
  const double h[3] = {dx, dy, dz};
  L.hx2 = dx * dx;
  L.hy2 = dy * dy;
  L.hz2 = dz * dz;
  L.ihx2 = 1.0 / L.hx2;
  L.ihy2 = 1.0 / L.hy2;
  L.ihz2 = 1.0 / L.hz2;
  for(int a = 0; a < 3; ++a) {
    L.inv[a] = 1.0 / std::pow(h[a], 2.0);
    L.per[a] = p_axis_periodic(b, a);
    L.lo_dir[a] = b.faces[2 * a].type == OF_PRESSURE_DIRICHLET;
    L.hi_dir[a] = b.faces[2 * a + 1].type == OF_PRESSURE_DIRICHLET;
    L.lo_val[a] = b.faces[2 * a].value;
    L.hi_val[a] = b.faces[2 * a + 1].value;
    L.lo_grad[a] = b.faces[2 * a].gradient;
    L.hi_grad[a] = b.faces[2 * a + 1].gradient;
    L.h[a] = h[a];
  }
  L.rx = L.ry = L.rz = 1;
  // diagonal per boundary class (linsolve/poisson_solver.cpp:271-314)
  auto axis_coef = [&](int a, int cls) {
    double coef = 2.0 * L.inv[a];
    if(L.per[a]) return coef;
    if(cls & 1) coef += L.lo_dir[a] ? L.inv[a] : -L.inv[a];
    if(cls & 2) coef += L.hi_dir[a] ? L.inv[a] : -L.inv[a];
    return coef;
  };
  for(int cz = 0; cz < 4; ++cz)
    for(int cy = 0; cy < 4; ++cy)
      for(int cx = 0; cx < 4; ++cx) {
        const double dval = 0.0 + axis_coef(0, cx) + axis_coef(1, cy) + axis_coef(2, cz);
        L.winv[cx + 4 * (cy + 4 * cz)] = dval != 0.0 ? omega / dval : 0.0;
      }
  // sum of omega/D over the level
  auto count = [](int n, int cls) -> double {
    if(n == 1) return cls == 3 ? 1.0 : 0.0;
    if(cls == 0) return (double)(n - 2);
    return (cls == 1 || cls == 2) ? 1.0 : 0.0;
  };
  double s = 0.0;
  for(int cz = 0; cz < 4; ++cz)
    for(int cy = 0; cy < 4; ++cy)
      for(int cx = 0; cx < 4; ++cx)
        s += count(nx, cx) * count(ny, cy) * count(nz, cz) * L.winv[cx + 4 * (cy + 4 * cz)];
  L.winv_sum = s;
  L.wr = 1.0;
  return L;
}

// Dense operator of the coarsest level from basis images, then the reference's elimination
// (linsolve/poisson_solver.cpp:482-567) recorded as row swaps, multipliers and U.
void factor_coarse(const PLevel& L, bool zm, std::vector<int>& perm, std::vector<double>& factor,
                   std::vector<double>& U) {
  const int n = L.n;
  std::vector<double> m((size_t)n * n, 0.0);
  auto apply = [&](const std::vector<double>& x, std::vector<double>& y) {
    const int nn[3] = {L.nx, L.ny, L.nz};
    for(int k = 0; k < L.nz; ++k)
      for(int j = 0; j < L.ny; ++j)
        for(int i = 0; i < L.nx; ++i) {
          const int co[3] = {i, j, k};
          auto X = [&](int a, int b, int e) { return x[a + L.nx * (b + L.ny * e)]; };
          const double cval = X(i, j, k);
          double nb[3][2];
          for(int a = 0; a < 3; ++a) {
            const int di = a == 0, dj = a == 1, dk = a == 2;
            if(co[a] > 0) nb[a][0] = X(i - di, j - dj, k - dk);
            else if(L.per[a]) nb[a][0] = X(a == 0 ? nn[0] - 1 : i, a == 1 ? nn[1] - 1 : j, a == 2 ? nn[2] - 1 : k);
            else nb[a][0] = L.lo_dir[a] ? 2.0 * 0.0 - cval : cval - 0.0 * L.h[a];
            if(co[a] < nn[a] - 1) nb[a][1] = X(i + di, j + dj, k + dk);
            else if(L.per[a]) nb[a][1] = X(a == 0 ? 0 : i, a == 1 ? 0 : j, a == 2 ? 0 : k);
            else nb[a][1] = L.hi_dir[a] ? 2.0 * 0.0 - cval : cval + 0.0 * L.h[a];
          }
          y[i + L.nx * (j + L.ny * k)] = -(nb[0][1] - 2.0 * cval + nb[0][0]) / L.hx2 -
                                         (nb[1][1] - 2.0 * cval + nb[1][0]) / L.hy2 -
                                         (nb[2][1] - 2.0 * cval + nb[2][0]) / L.hz2;
        }
  };
  std::vector<double> e(n), img(n);
  for(int col = 0; col < n; ++col) {
    std::fill(e.begin(), e.end(), 0.0);
    e[col] = 1.0;
    apply(e, img);
    for(int row = 0; row < n; ++row) m[(size_t)row * n + col] = img[row];
  }
  if(zm)
    for(int col = 0; col < n; ++col) m[(size_t)(n - 1) * n + col] = 1.0;
  perm.assign(n, 0);
  factor.assign((size_t)n * n, 0.0);
  for(int p = 0; p < n; ++p) {
    int best = p;
    double best_abs = std::abs(m[(size_t)p * n + p]);
    for(int row = p + 1; row < n; ++row) {
      const double cand = std::abs(m[(size_t)row * n + p]);
      if(cand > best_abs) {
        best = row;
        best_abs = cand;
      }
    }
    if(best_abs < kSmallNumber) throw RuntimeError("direct coarse solve encountered a singular matrix");
    perm[p] = best;
    if(best != p)
      for(int col = 0; col < n; ++col) std::swap(m[(size_t)p * n + col], m[(size_t)best * n + col]);
    const double ip = 1.0 / m[(size_t)p * n + p];
    for(int row = p + 1; row < n; ++row) {
      const double f = m[(size_t)row * n + p] * ip;
      if(std::abs(f) < kSmallNumber) continue;
      factor[(size_t)row * n + p] = f;
      for(int col = p; col < n; ++col) m[(size_t)row * n + col] -= f * m[(size_t)p * n + col];
    }
  }
  U = m;
}

// Solution operator of a small coarsest system: column j is the recorded elimination applied
// to the unit vector e_j exactly as the device applies it to a right-hand side (the last entry
// zeroed when the last row was replaced by ones, the mean removed afterwards), evaluated in long
// double and rounded once. x = C (b - mb) is then the elimination's result to rounding.
std::vector<double> coarse_operator(int n, bool zm, const std::vector<int>& perm, const std::vector<double>& factor,
                                    const std::vector<double>& U) {
  std::vector<double> C((size_t)n * n, 0.0);
  std::vector<long double> w(n);
  for(int col = 0; col < n; ++col) {
    std::fill(w.begin(), w.end(), 0.0L);
    w[col] = 1.0L;
    if(zm) w[n - 1] = 0.0L;
    for(int p = 0; p < n; ++p) {
      if(perm[p] != p) std::swap(w[p], w[perm[p]]);
      for(int row = p + 1; row < n; ++row) w[row] -= (long double)factor[(size_t)row * n + p] * w[p];
    }
    for(int row = n - 1; row >= 0; --row) {
      long double s = w[row];
      for(int c = row + 1; c < n; ++c) s -= (long double)U[(size_t)row * n + c] * w[c];
      w[row] = s / (long double)U[(size_t)row * n + row];
    }
    if(zm) {
      long double m = 0.0L;
      for(int i = 0; i < n; ++i) m += w[i];
      m /= (long double)n;
      for(int i = 0; i < n; ++i) w[i] -= m;
    }
    for(int row = 0; row < n; ++row) C[(size_t)row * n + col] = (double)w[row];
  }
  return C;
}

}  // namespace

PoissonSolver::~PoissonSolver() {
  if(arena) return;  // the arena's owner keeps its memory
  auto fr = [](void* p) {
    if(p) OFH_RELEASE(hipFree(p));
  };
  fr(pc);
  fr(r);
  for(double* p : ring) fr(p);
  fr(q);
  fr(coef);
  for(void* p : lb) fr(p);
  for(void* p : lxa) fr(p);
  for(void* p : lxb) fr(p);
  fr(ps);
  fr(partials2);
  fr(history);
  fr(perm);
  fr(factor);
  fr(U);
  fr(work);
  fr(dlv);
  if(host_ps) OFH_RELEASE(hipHostFree(host_ps));
}

void PoissonSolver::build() {
  double w2 = omega;
  if(precision == OF_PRECISION_MIXED || fp64_chebyshev())
    mixed_omega((g.nx > 1) + (g.ny > 1) + (g.nz > 1), omega, w2);
  for(int f = 0; f < 6; ++f)
    if(bcs.faces[f].type == OF_PRESSURE_DIRICHLET) zm = false;
  // hierarchy (linsolve/poisson_solver.cpp:316-384)
  int nx = g.nx, ny = g.ny, nz = g.nz;
  double dx = g.dx, dy = g.dy, dz = g.dz;
  lv.push_back(make_level(nx, ny, nz, dx, dy, dz, bcs, omega));
  while(axis_coarsens(nx) || axis_coarsens(ny) || axis_coarsens(nz)) {
    const int cx = axis_coarsens(nx) ? nx / 2 : nx, cy = axis_coarsens(ny) ? ny / 2 : ny,
              cz = axis_coarsens(nz) ? nz / 2 : nz;
    const double hx = axis_coarsens(nx) ? 2.0 * dx : dx, hy = axis_coarsens(ny) ? 2.0 * dy : dy,
                 hz = axis_coarsens(nz) ? 2.0 * dz : dz;
    if(cx == nx && cy == ny && cz == nz) break;
    lv.back().rx = nx / cx;
    lv.back().ry = ny / cy;
    lv.back().rz = nz / cz;
    nx = cx;
    ny = cy;
    nz = cz;
    dx = hx;
    dy = hy;
    dz = hz;
    lv.push_back(make_level(nx, ny, nz, dx, dy, dz, bcs, omega));
  }
  if((int)lv.size() > kMaxLevels) throw InvalidArgument("multigrid hierarchy deeper than supported");
  for(PLevel& L : lv) L.wr = w2 / omega;  // exactly 1 with one weight
  l0 = (int)lv.size() - 1;
  for(int l = 0; l < (int)lv.size(); ++l)
    if(lv[l].n <= kTailCells) {
      l0 = l;
      break;
    }
  fused.assign(lv.size(), 0);
  for(int l = 0; l < l0; ++l) {
    const PLevel& P = lv[l];
    const bool uni = P.per[0] && P.per[1] && P.per[2];
    fused[l] = uni && P.rx == 2 && P.ry == 2 && P.rz == 2 && P.nx >= 4 && P.ny >= 4 && P.nz >= 4 && !fused_disabled();
  }
  fused_rows(precision, rows_down, rows_up);
  up_bufs = up_buffers(precision);
  up_zw = up_windows() && rows_up == 16;
  if(precision == OF_PRECISION_MIXED) {
    resident_down = rows_down == 16 ? resident_blocks(k_down_fused<float, 16>, 512, c.device)
                                    : resident_blocks(k_down_fused<float, 8>, 256, c.device);
    resident_up = rows_up == 16 ? (up_bufs == 2 ? resident_blocks(k_up_fused<float, 16, 2>, 512, c.device)
                                                : resident_blocks(k_up_fused<float, 16>, 512, c.device))
                                : resident_blocks(k_up_fused<float, 8>, 256, c.device);
  } else {
    resident_down = rows_down == 16 ? resident_blocks(k_down_fused<double, 16>, 512, c.device)
                                    : resident_blocks(k_down_fused<double, 8>, 256, c.device);
    resident_up = rows_up == 16 ? (up_bufs == 2 ? resident_blocks(k_up_fused<double, 16, 2>, 512, c.device)
                                                : resident_blocks(k_up_fused<double, 16>, 512, c.device))
                                : resident_blocks(k_up_fused<double, 8>, 256, c.device);
  }
  if(up_zw) {
    if(precision == OF_PRECISION_MIXED)
      resident_up = up_bufs == 2 ? resident_blocks(k_up_zw<float, 2>, 512, c.device)
                                 : resident_blocks(k_up_zw<float, 3>, 512, c.device);
    else
      resident_up = up_bufs == 2 ? resident_blocks(k_up_zw<double, 2>, 512, c.device)
                                 : resident_blocks(k_up_zw<double, 3>, 512, c.device);
  }
  const size_t n0 = (size_t)lv[0].n;
  dev_alloc(reinterpret_cast<void**>(&pc), n0 * sizeof(double));
  dev_alloc(reinterpret_cast<void**>(&r), n0 * sizeof(double));
  ring.assign(ring_slots(n0, arena != nullptr), nullptr);
  for(double*& dk : ring) dev_alloc(reinterpret_cast<void**>(&dk), n0 * sizeof(double));
  const size_t es = precision == OF_PRECISION_MIXED ? sizeof(float) : sizeof(double);
  for(size_t l = 0; l < lv.size(); ++l) {
    void *b = nullptr, *xa = nullptr, *xb = nullptr;
    const size_t n = (size_t)lv[l].n;
    // the finest FP64 level reads the CG residual directly as its rhs
    if(!(l == 0 && precision == OF_PRECISION_FP64)) dev_alloc(&b, n * es);
    dev_alloc(&xa, n * es);
    dev_alloc(&xb, n * es);
    lb.push_back(b);
    lxa.push_back(xa);
    lxb.push_back(xb);
  }
  dev_alloc(reinterpret_cast<void**>(&ps), P_COUNT * sizeof(double));
  OFH_CHECK(hipMemsetAsync(ps, 0, P_COUNT * sizeof(double), c.stream));
  dev_alloc(reinterpret_cast<void**>(&partials2), (size_t)kMaxRed * kMaxPartials * sizeof(double));
  if(arena) {  // the stop flag stays in device memory; the host copies it back after each iteration
    host_ps_dev = static_cast<double*>(arena->take(2 * kHostMirror * sizeof(double)));
    OFH_CHECK(hipMemsetAsync(host_ps_dev, 0, 2 * kHostMirror * sizeof(double), c.stream));
  } else {
    OFH_CHECK(hipHostMalloc(&host_ps, 2 * kHostMirror * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    OFH_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&host_ps_dev), host_ps, 0));
    for(int q = 0; q < 2 * kHostMirror; ++q) host_ps[q] = 0.0;
  }
  dev_alloc(reinterpret_cast<void**>(&dlv), lv.size() * sizeof(PLevel));
  OFH_CHECK(hipMemcpyAsync(dlv, lv.data(), lv.size() * sizeof(PLevel), hipMemcpyHostToDevice, c.stream));
  std::vector<int> hperm;
  std::vector<double> hf, hU;
  factor_coarse(lv.back(), zm, hperm, hf, hU);
  nc = lv.back().n;
  if(nc <= kCoarseStage) hf = coarse_operator(nc, zm, hperm, hf, hU);  // the tail reads C from `factor`
  dev_alloc(reinterpret_cast<void**>(&perm), nc * sizeof(int));
  dev_alloc(reinterpret_cast<void**>(&factor), (size_t)nc * nc * sizeof(double));
  dev_alloc(reinterpret_cast<void**>(&U), (size_t)nc * nc * sizeof(double));
  dev_alloc(reinterpret_cast<void**>(&work), (size_t)nc * sizeof(double));
  // copies on the solver's stream, waited for here: the host vectors go out of scope
  OFH_CHECK(hipMemcpyAsync(perm, hperm.data(), nc * sizeof(int), hipMemcpyHostToDevice, c.stream));
  OFH_CHECK(hipMemcpyAsync(factor, hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  OFH_CHECK(hipMemcpyAsync(U, hU.data(), hU.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  OFH_CHECK(hipStreamSynchronize(c.stream));
}

template <typename T>
void PoissonSolver::vcycle() {
  const int nl = (int)lv.size();
  const int zmi = zm ? 1 : 0;
  auto slot = [](int l) { return P_LEVEL + 8 * l; };
  // finest rhs: CG residual (FP64) or its FP32 copy (mixed)
  auto bptr = [&](int l) -> T* {
    if(l == 0 && precision == OF_PRECISION_FP64) return reinterpret_cast<T*>(r);
    return static_cast<T*>(lb[l]);
  };
  // A fused level whose consumer is the next fused level or the tail leaves its partials for the
  // consumer to fold (no finalize launch); the two partial buffers alternate along the chain.
  double* pbuf[2] = {c.partials, partials2};
  int which = 0;
  DownFold pend{nullptr, 0u, 0, 0.0};
  auto foldable = [](const dim3& g) { return g.x <= (unsigned)kMaxFoldPartials; };
  for(int l = 0; l < l0; ++l) {
    const PLevel& P = lv[l];
    const PLevel& C = lv[l + 1];
    if(fused[l]) {
      FusedLaunch f = fused_launch(P, resident_down, rows_down);
      const bool small = f.kz == 2 && small_levels();
      if(small) f = small_launch(P);  // its own tiles, whatever rows_down is
      double* out = pbuf[which];
      {
        ProfScope prof_(c, "k_down_fused", l == 0);
        if(small)
          OFH_LAUNCH((k_down_small<T>), f.grid, dim3(512), 0, c.stream, bptr(l), static_cast<T*>(lxa[l]), P, C,
                     static_cast<T*>(lb[l + 1]), out, ps, slot(l), zmi, pend);
        else if(rows_down == 16)
          OFH_LAUNCH((k_down_fused<T, 16>), f.grid, dim3(512), 0, c.stream, bptr(l), static_cast<T*>(lxa[l]),
                             P, C, static_cast<T*>(lb[l + 1]), out, ps, slot(l), zmi, f.kz, pend);
        else
          OFH_LAUNCH((k_down_fused<T, 8>), f.grid, dim3(256), 0, c.stream, bptr(l), static_cast<T*>(lxa[l]),
                             P, C, static_cast<T*>(lb[l + 1]), out, ps, slot(l), zmi, f.kz, pend);
      }
      const bool consumer_folds = (l + 1 == l0 || fused[l + 1]) && foldable(f.grid);
      if(consumer_folds) {
        pend = DownFold{out, f.grid.x, slot(l), (double)P.n};
        which ^= 1;
      } else {
        launch_finalize(c.stream, out, f.grid, EpiDown{ps, slot(l), slot(l + 1), (double)P.n, (double)C.n, zmi});
        pend = DownFold{nullptr, 0u, 0, 0.0};
      }
      continue;
    }
    const HotLaunch h = hot_launch(P);
    {
      ProfScope prof_(c, "k_smooth12", l == 0);
      OFH_HOT_LAUNCH(k_smooth12<T, h, c.stream, bptr(l), static_cast<T*>(lxa[l]), P, c.partials, ps, slot(l), zmi);
    }
    launch_finalize(c.stream, c.partials, h.grid, EpiMean{ps, slot(l) + L_MX2, (double)P.n, zmi});
    {
      ProfScope prof_(c, "k_residual_restrict", l == 0);
      OFH_HOT_LAUNCH(k_residual_restrict<T, h, c.stream, bptr(l), static_cast<const T*>(lxa[l]), P, C,
                     static_cast<T*>(lb[l + 1]), c.partials, ps, slot(l), zmi);
    }
    launch_finalize(c.stream, c.partials, h.grid, EpiRestrict{ps, slot(l + 1), (double)C.n, zmi});
  }
  {
    TailBufs<T> tb{};
    size_t tot = 0;  // LDS layout in elements of T
    for(int l = l0; l < nl; ++l) {
      tb.b[l] = bptr(l);
      tb.xa[l] = static_cast<T*>(lxa[l]);
      tb.xb[l] = static_cast<T*>(lxb[l]);
      tb.ob[l] = (int)tot;
      tb.oxa[l] = (int)(tot + lv[l].n);
      tb.oxb[l] = (int)(tot + 2 * (size_t)lv[l].n);
      tot += 3 * (size_t)lv[l].n;
    }
    const size_t work_off = (tot * sizeof(T) + sizeof(double) - 1) / sizeof(double);
    tb.owork = (int)work_off;
    tb.n0 = lv[l0].n;
    tb.ncs = lv[nl - 1].n;
    tb.b0 = bptr(l0);
    const bool in_lds = (work_off + (size_t)nc) * sizeof(double) <= (size_t)kTailLdsBytes;
    ProfScope prof_(c, "k_vtail");
    // all tail levels share the boundary types of the finest level
    const bool tuni = lv[l0].per[0] && lv[l0].per[1] && lv[l0].per[2];
    if(in_lds) {
      if(tuni)
        OFH_LAUNCH((k_vtail<T, true, true>), dim3(1), dim3(kTailThreads), 0, c.stream, dlv, l0, nl, tb, perm,
                           factor, U, work, ps, zmi, pend);
      else
        OFH_LAUNCH((k_vtail<T, true, false>), dim3(1), dim3(kTailThreads), 0, c.stream, dlv, l0, nl, tb, perm,
                           factor, U, work, ps, zmi, pend);
    } else {
      if(tuni)
        OFH_LAUNCH((k_vtail<T, false, true>), dim3(1), dim3(kTailThreads), 0, c.stream, dlv, l0, nl, tb, perm,
                           factor, U, work, ps, zmi, pend);
      else
        OFH_LAUNCH((k_vtail<T, false, false>), dim3(1), dim3(kTailThreads), 0, c.stream, dlv, l0, nl, tb,
                           perm, factor, U, work, ps, zmi, pend);
    }
  }
  const double* up_src = nullptr;  // partials of the coarser fused up kernel, to be folded
  unsigned up_nb = 0;
  for(int l = l0 - 1; l >= 0; --l) {
    const PLevel& P = lv[l];
    const PLevel& C = lv[l + 1];
    const T* xc = static_cast<const T*>(x_final(l + 1));
    if(fused[l]) {
      FusedLaunch f = fused_launch(P, resident_up, rows_up);
      const bool small = f.kz == 2 && small_levels();
      if(small) f = small_launch(P);
      double* out = pbuf[which];
      {
        ProfScope prof_(c, "k_up_fused", l == 0);
        if(small)
          OFH_LAUNCH((k_up_small<T>), f.grid, dim3(512), 0, c.stream, bptr(l), static_cast<const T*>(lxa[l]), xc, P,
                     C, static_cast<T*>(lxb[l]), out, ps, slot(l), slot(l + 1), up_src, up_nb, zmi);
        else if(up_zw && up_bufs == 2)
          OFH_LAUNCH((k_up_zw<T, 2>), f.grid, dim3(512), 0, c.stream, bptr(l), static_cast<const T*>(lxa[l]), xc, P,
                     C, static_cast<T*>(lxb[l]), out, ps, slot(l), slot(l + 1), f.kz, up_src, up_nb, zmi);
        else if(up_zw)
          OFH_LAUNCH((k_up_zw<T, 3>), f.grid, dim3(512), 0, c.stream, bptr(l), static_cast<const T*>(lxa[l]), xc, P,
                     C, static_cast<T*>(lxb[l]), out, ps, slot(l), slot(l + 1), f.kz, up_src, up_nb, zmi);
        else if(rows_up == 16 && up_bufs == 2)
          OFH_LAUNCH((k_up_fused<T, 16, 2>), f.grid, dim3(512), 0, c.stream, bptr(l),
                             static_cast<const T*>(lxa[l]), xc, P, C, static_cast<T*>(lxb[l]), out, ps, slot(l),
                             slot(l + 1), f.kz, up_src, up_nb, zmi);
        else if(rows_up == 16)
          OFH_LAUNCH((k_up_fused<T, 16>), f.grid, dim3(512), 0, c.stream, bptr(l),
                             static_cast<const T*>(lxa[l]), xc, P, C, static_cast<T*>(lxb[l]), out, ps, slot(l),
                             slot(l + 1), f.kz, up_src, up_nb, zmi);
        else
          OFH_LAUNCH((k_up_fused<T, 8>), f.grid, dim3(256), 0, c.stream, bptr(l),
                             static_cast<const T*>(lxa[l]), xc, P, C, static_cast<T*>(lxb[l]), out, ps, slot(l),
                             slot(l + 1), f.kz, up_src, up_nb, zmi);
      }
      // the finer fused level folds this level's mean of x5; the finest level's finalize also
      // forms r.z and beta
      if(l > 0 && fused[l - 1] && foldable(f.grid)) {
        up_src = out;
        up_nb = f.grid.x;
        which ^= 1;
      } else {
        launch_finalize(c.stream, out, f.grid, EpiSmoothLast{ps, slot(l), (double)P.n, zmi, l == 0 ? 1 : 0});
        up_src = nullptr;
        up_nb = 0;
      }
      continue;
    }
    const HotLaunch h = hot_launch(P);
    {
      ProfScope prof_(c, "k_prolong_smooth", l == 0);
      OFH_HOT_LAUNCH(k_prolong_smooth<T, h, c.stream, bptr(l), static_cast<const T*>(lxa[l]), xc, P, C,
                     static_cast<T*>(lxb[l]), c.partials, ps, slot(l), slot(l + 1));
    }
    launch_finalize(c.stream, c.partials, h.grid, EpiMean{ps, slot(l) + L_MX4, (double)P.n, zmi});
    {
      ProfScope prof_(c, "k_smooth_last", l == 0);
      OFH_HOT_LAUNCH(k_smooth_last<T, h, c.stream, bptr(l), static_cast<const T*>(lxb[l]), P,
                     static_cast<T*>(lxa[l]), c.partials, ps, slot(l));
    }
    launch_finalize(c.stream, c.partials, h.grid, EpiSmoothLast{ps, slot(l), (double)P.n, zmi, l == 0 ? 1 : 0});
  }
}

template <typename T>
PoissonResult PoissonSolver::solve_t(const double* rhs, double* p, int maxit, double tol, bool want_history,
                                    bool zero_guess, const double** deferred_mean) {
  if(hist_cap < maxit + 1) {
    if(history && !arena) OFH_CHECK(hipFree(history));
    if(coef && !arena) OFH_CHECK(hipFree(coef));
    history = coef = nullptr;
    dev_alloc(reinterpret_cast<void**>(&history), (size_t)(maxit + 1) * sizeof(double));
    dev_alloc(reinterpret_cast<void**>(&coef), (size_t)2 * maxit * sizeof(double));
    hist_cap = maxit + 1;
  }
  const PLevel& F = lv[0];
  const Lay Lc = make_lay(g, OF_CELL);
  const int zmi = zm ? 1 : 0;
  const LevelLaunch ll = level_launch(F);
  const bool mixed = precision == OF_PRECISION_MIXED;
  const double n0 = (double)F.n;
  const HotLaunch hf = hot_launch(F);
  const bool cg_fused = !cg_fused_disabled();
  if(!cg_fused && q == nullptr) dev_alloc(reinterpret_cast<void**>(&q), (size_t)F.n * sizeof(double));
  const int R = (int)ring.size();
  DirRing dr{};
  for(int k = 0; k < R; ++k) dr.d[k] = ring[k];
  dr.coef = coef;
  dr.R = R;
  const unsigned pgrid = std::min<unsigned>(div_up((size_t)F.n, 256 * kPointIlp), 8192u);
  OFH_CHECK(hipMemsetAsync(ps, 0, P_COUNT * sizeof(double), c.stream));
  if(host_ps) {  // the previous solve has finished (its end synchronised the stream)
    volatile double* hm = host_ps;
    for(int q = 0; q < 2 * kHostMirror; ++q) hm[q] = 0.0;
  }
  if(zero_guess)
    OFH_LAUNCH(k_setup1_zero, ll.grid, ll.block, 0, c.stream, rhs, Lc, F, c.partials, ll.kz);
  else
    OFH_LAUNCH(k_setup1, ll.grid, ll.block, 0, c.stream, rhs, p, Lc, F, r, pc, c.partials, ll.kz);
  launch_finalize(c.stream, c.partials, ll.grid, EpiSetup1{ps, n0, zmi});
  // the FP32 copy of the initial residual (the rhs of the finest FP32 level) comes with it for a
  // zero warm start, else from k_cast
  T* lo0 = mixed ? static_cast<T*>(lb[0]) : nullptr;
  if(zero_guess)
    OFH_LAUNCH(k_setup2_zero<T>, ll.grid, ll.block, 0, c.stream, rhs, Lc, r, lo0, F, c.partials, ps, ll.kz);
  else
    OFH_LAUNCH(k_setup2, ll.grid, ll.block, 0, c.stream, r, pc, F, c.partials, ps, ll.kz);
  launch_finalize(c.stream, c.partials, ll.grid, EpiSetup2{ps, history, n0, tol, zmi});
  if(mixed && !zero_guess) OFH_LAUNCH(k_cast<T>, 2048, 256, 0, c.stream, r, lo0, F.n, ps);
  vcycle<T>();
  const dim3 pg(2048);  // launch shape of the pressure passes (fixes the order of p's sum)
  int folded = 0;       // directions [0, folded) have been added to pc
  int it = 0;
  for(; it < maxit; ++it) {
    if(it > 0 && it % R == 0) {  // slot it % R is about to be reused: fold the last R directions
      OFH_LAUNCH(k_pflush, pg, dim3(256), 0, c.stream, pc, dr, it - R, it, F.n, folded == 0 && zero_guess ? 1 : 0,
                 ps);
      folded = it;
    }
    double* dnew = ring[it % R];
    const double* dold = ring[(it + R - 1) % R];  // unused (not read into any result) when it == 0
    if(cg_fused && hf.lds) {
      ProfScope prof_(c, "k_dir_apply");
      OFH_LAUNCH(k_dir_apply_lds<T>, hf.grid, hf.block, 0, c.stream, static_cast<const T*>(x_final(0)), dold, dnew, F,
                 c.partials, ps, it == 0 ? 1 : 0);
    } else if(cg_fused) {
      ProfScope prof_(c, "k_dir_apply");
      OFH_HOT_LAUNCH(k_dir_apply<T, hf, c.stream, static_cast<const T*>(x_final(0)), dold, dnew, F, c.partials, ps,
                     it == 0 ? 1 : 0);
    } else {
      {
        ProfScope prof_(c, "k_direction");
        OFH_LAUNCH(k_direction<T>, dim3(pgrid), dim3(256), 0, c.stream, static_cast<const T*>(x_final(0)), dold,
                   dnew, F.n, ps, it == 0 ? 1 : 0);
      }
      ProfScope prof_(c, "k_apply_dot");
      if(hf.kz == 8) {
        if(hf.uni)
          OFH_LAUNCH((k_apply_dot<8, true>), hf.grid, hf.block, 0, c.stream, dnew, q, F, c.partials, ps);
        else
          OFH_LAUNCH((k_apply_dot<8, false>), hf.grid, hf.block, 0, c.stream, dnew, q, F, c.partials, ps);
      } else {
        if(hf.uni)
          OFH_LAUNCH((k_apply_dot<2, true>), hf.grid, hf.block, 0, c.stream, dnew, q, F, c.partials, ps);
        else
          OFH_LAUNCH((k_apply_dot<2, false>), hf.grid, hf.block, 0, c.stream, dnew, q, F, c.partials, ps);
      }
    }
    launch_finalize(c.stream, c.partials, hf.grid, EpiDir{ps, n0, zmi});
    {
      ProfScope prof_(c, "k_update");
      T* blo = mixed ? static_cast<T*>(lb[0]) : nullptr;
      if(cg_fused && hf.lds)
        OFH_LAUNCH(k_update_apply_lds<T>, hf.grid, hf.block, 0, c.stream, dnew, r, blo, F, c.partials, ps,
                   mixed ? 1 : 0);
      else if(cg_fused)
        OFH_HOT_LAUNCH(k_update_apply<T, hf, c.stream, dnew, r, blo, F, c.partials, ps, mixed ? 1 : 0);
      else
        OFH_HOT_LAUNCH(k_update<T, hf, c.stream, r, q, blo, F, c.partials, ps, mixed ? 1 : 0);
    }
    launch_finalize(c.stream, c.partials, hf.grid,
                    EpiUpdate{ps, history, coef, host_ps_dev + (it % 2) * kHostMirror, n0, tol, it, zmi});
    if(arena) {
      double done = 0.0;
      OFH_CHECK(hipMemcpyAsync(&done, host_ps_dev + (it % 2) * kHostMirror + P_DONE, sizeof(double),
                               hipMemcpyDeviceToHost, c.stream));
      OFH_CHECK(hipStreamSynchronize(c.stream));
      if(done != 0.0) break;
    } else if(it > 0) {
      // one iteration of look-ahead: the previous iteration's update has published its stop flag
      // once its sequence number shows in the mapped mirror (no event in the queue: an event
      // record left the device idle for about 5 us per iteration)
      if(wait_mirror(it - 1) != 0.0) break;
    }
    vcycle<T>();
  }
  // The host has seen the stop flag of iteration it - 1 (or it), so at most one queued
  // iteration follows the last one that ran: every direction the pressure needs is in the ring,
  // and the folds already queued covered [0, folded) clipped at the iteration count.
  OFH_LAUNCH(k_pfinal, pg, dim3(256), 0, c.stream, p, pc, dr, folded, folded == 0 && zero_guess ? 1 : 0, F, Lc,
             c.partials, ps);
  launch_finalize(c.stream, c.partials, pg, EpiPfinal{ps, n0, zmi});
  if(deferred_mean != nullptr) *deferred_mean = zm ? ps + P_PFINAL : nullptr;
  if(zm && deferred_mean == nullptr) OFH_LAUNCH(k_psub, ll.grid, ll.block, 0, c.stream, p, F, Lc, ps, ll.kz);
  OFH_CHECK(hipGetLastError());
  std::vector<double> hps(P_COUNT);
  OFH_CHECK(hipMemcpyAsync(hps.data(), ps, P_COUNT * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  OFH_CHECK(hipStreamSynchronize(c.stream));
  PoissonResult res{};
  of_poisson_diagnostics& dg = res.diag;
  dg.iterations = (int)hps[P_ITER];
  dg.initial_residual_l2 = hps[P_R0];
  dg.final_residual_l2 = hps[P_RMS];
  dg.relative_residual = hps[P_REL];
  dg.removed_rhs_mean = hps[P_MSYS];
  dg.multigrid_levels = (int)lv.size();
  dg.coarse_unknowns = lv.back().n;
  dg.pre_smoothing_steps = 2;
  dg.post_smoothing_steps = 2;
  dg.converged = hps[P_CONV] != 0.0;
  dg.zero_mean_enforced = zm;
  dg.history_length = dg.iterations + 1;
  if(want_history) {
    res.history.resize(dg.history_length);
    OFH_CHECK(hipMemcpy(res.history.data(), history, dg.history_length * sizeof(double), hipMemcpyDeviceToHost));
  }
  return res;
}

PoissonSolver* poisson_create(Context& c, const of_grid& g, const of_pressure_bc_set& b, int precision) {
  if(precision != OF_PRECISION_FP64 && precision != OF_PRECISION_MIXED) throw InvalidArgument("unknown precision");
  auto* s = new PoissonSolver(c, g, b, precision);
  try {
    s->build();
  } catch(...) {
    delete s;
    throw;
  }
  return s;
}

PoissonSolver* poisson_create_in(Context& c, const of_grid& g, const of_pressure_bc_set& b, int precision,
                                 DeviceArena& arena) {
  if(precision != OF_PRECISION_FP64 && precision != OF_PRECISION_MIXED) throw InvalidArgument("unknown precision");
  auto* s = new PoissonSolver(c, g, b, precision, &arena);
  try {
    s->build();
  } catch(...) {
    delete s;
    throw;
  }
  return s;
}

void poisson_destroy(PoissonSolver* s) { delete s; }

PoissonResult poisson_solve(PoissonSolver* s, const double* rhs, double* p, int maxit, double tol, bool want_history,
                            bool zero_guess, const double** deferred_mean) {
  if(maxit <= 0) throw InvalidArgument("solve_pressure_poisson expects iterations > 0");
  if(!(tol > 0.0)) throw InvalidArgument("solve_pressure_poisson expects tolerance > 0");
  if(s->precision == OF_PRECISION_MIXED)
    return s->solve_t<float>(rhs, p, maxit, tol, want_history, zero_guess, deferred_mean);
  return s->solve_t<double>(rhs, p, maxit, tol, want_history, zero_guess, deferred_mean);
}

void launch_poisson_apply(Context& c, const of_grid& g, const of_pressure_bc_set& b, bool homogeneous,
                          const double* in, double* out) {
  const PLevel L = make_level(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, b);
  const LevelLaunch ll = level_launch(L);
  OFH_LAUNCH(k_poisson_apply_g, ll.grid, ll.block, 0, c.stream, in, out, make_lay(g, OF_CELL), L,
                     homogeneous ? 1 : 0, ll.kz);
  OFH_CHECK(hipGetLastError());
}

}  // namespace ofh
