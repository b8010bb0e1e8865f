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
//    product (+ the deferred pressure update of the previous iteration).
//  * All scalars (alpha, beta, means, norms, convergence flags) stay on the device. Kernels of
//    iterations past convergence return immediately (flag check), so the host queues the next
//    iteration before it has read the previous one's residual.
//  * Precision: FP64, or MIXED = FP32 V-cycle (levels, smoothing, residual, transfers) inside
//    the FP64 CG (residual, dot products, pressure update).
//  * Reductions are deterministic (fixed launch shape, fixed combination order).

#include "ofh.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace ofh {

namespace {

constexpr double kOmega = 2.0 / 3.0;  // linsolve/poisson_solver.hpp:17-23
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
  P_PEND,          // 1 while p += alpha d of the last residual update is still to be applied
  P_DSEL,          // which direction buffer holds the current d
  P_LEVEL = 32,    // per level: 8 slots (see LS)
  P_COUNT = 32 + 8 * 24
};
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
  int rx, ry, rz;           // refinement ratio to the next coarser level
};

// boundary class of a coordinate: 0 interior, 1 first, 2 last, 3 single
__device__ __host__ inline int bclass(int co, int n) { return (co == 0 ? 1 : 0) | (co == n - 1 ? 2 : 0); }

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

// ---------------------------------------------------------------- setup
// sys = -rhs - src (active cells), p_c = warm start; sums for their means.
__global__ void k_setup1(const double* __restrict__ rhs_g, const double* __restrict__ p_g, Lay Lc, PLevel L,
                         double* __restrict__ sys, double* __restrict__ pc, double* partials, unsigned* counter,
                         double* ps, int zm, int kz) {
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
  double res[2];
  if(reduce_partials<2>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0) {
    ps[P_MSYS] = zm ? res[0] / (double)L.n : 0.0;
    ps[P_MP] = zm ? res[1] / (double)L.n : 0.0;
  }
}

// r = (sys - m_sys) - A(p - m_p); sums of r, r^2, r w/D; initial residual and early exit.
__global__ void k_setup2(double* __restrict__ r, const double* __restrict__ pc, PLevel L, double* partials,
                         unsigned* counter, double* ps, double* history, double tol, int zm, int kz) {
  const double msys = ps[P_MSYS], mp = ps[P_MP];
  const Tile t = tile(L, kz);
  const Col C = column(L, t);
  double v[3] = {0.0, 0.0, 0.0};
  if(C.in) {
    const size_t plane = (size_t)L.nx * L.ny;
    auto X = [&](int ii, int jj, int kk) { return pc[(size_t)ii + (size_t)L.nx * jj + plane * kk] - mp; };
    for(int k = t.k0; k < t.k1; ++k) {
      const Nb z = nb_axis(k, L.nz, L.per[2], L.lo_dir[2], L.hi_dir[2]);
      const double c = X(t.i, t.j, k);
      const double q = lap_neg<double>(c, sgn(C.x.sm, X(C.x.cm, t.j, k)), sgn(C.x.sp, X(C.x.cp, t.j, k)),
                                       sgn(C.y.sm, X(t.i, C.y.cm, k)), sgn(C.y.sp, X(t.i, C.y.cp, k)),
                                       sgn(z.sm, X(t.i, t.j, z.cm)), sgn(z.sp, X(t.i, t.j, z.cp)), L.ihx2, L.ihy2,
                                       L.ihz2);
      const size_t id = (size_t)t.i + (size_t)L.nx * t.j + plane * k;
      const double rv = (r[id] - msys) + -1.0 * q;
      r[id] = rv;
      v[0] += rv;
      v[1] += rv * rv;
      v[2] += rv * L.winv[C.cx + 4 * (C.cy + 4 * bclass(k, L.nz))];
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  double res[3];
  if(reduce_partials<3>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0) {
    const double n = (double)L.n;
    const double mr = zm ? res[0] / n : 0.0;
    ps[P_MR] = mr;
    ps[P_LEVEL + L_MB] = mr;
    ps[P_LEVEL + L_SBW] = res[2];
    const double rms = sqrt(fmax(res[1] - n * mr * mr, 0.0) / n);
    ps[P_R0] = rms;
    ps[P_RMS] = rms;
    ps[P_REL] = rms > 0.0 ? 1.0 : 0.0;
    ps[P_ITER] = 0.0;
    ps[P_PEND] = 0.0;
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
}

// ---------------------------------------------------------------- V-cycle kernels
// Level rhs b of type T with deferred mean mb; iterates of type T.

// S12: x2 = x1 + w/D (b - A x1), x1 = (b - mb) w/D - m1 (fused first two pre-smoothing sweeps;
// m1 = mean of x1 from the producer's sum of b w/D).
template <typename T>
__global__ void k_smooth12(const T* __restrict__ b, T* __restrict__ x2, PLevel L, double* partials, unsigned* counter,
                           double* ps, int lslot, int zm, int kz) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m1 = zm ? T((ps[lslot + L_SBW] - ps[lslot + L_MB] * L.winv_sum) / (double)L.n) : T(0);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const Tile t = tile(L, kz);
  const Col C = column(L, t);
  double v[1] = {0.0};
  if(C.in) {
    const size_t plane = (size_t)L.nx * L.ny;
    auto W = [&](int cx, int cy, int cz) { return T(L.winv[cx + 4 * (cy + 4 * cz)]); };
    const int cxm = bclass(C.x.cm, L.nx), cxp = bclass(C.x.cp, L.nx);
    const int cym = bclass(C.y.cm, L.ny), cyp = bclass(C.y.cp, L.ny);
    auto B = [&](int ii, int jj, int kk) { return b[(size_t)ii + (size_t)L.nx * jj + plane * kk] - mb; };
    for(int k = t.k0; k < t.k1; ++k) {
      const Nb z = nb_axis(k, L.nz, L.per[2], L.lo_dir[2], L.hi_dir[2]);
      const int cz = bclass(k, L.nz), czm = bclass(z.cm, L.nz), czp = bclass(z.cp, L.nz);
      const T wc = W(C.cx, C.cy, cz);
      const T bt = B(t.i, t.j, k);
      const T c = bt * wc - m1;
      const T xw = sgn(C.x.sm, B(C.x.cm, t.j, k) * W(cxm, C.cy, cz) - m1);
      const T xe = sgn(C.x.sp, B(C.x.cp, t.j, k) * W(cxp, C.cy, cz) - m1);
      const T xs = sgn(C.y.sm, B(t.i, C.y.cm, k) * W(C.cx, cym, cz) - m1);
      const T xn = sgn(C.y.sp, B(t.i, C.y.cp, k) * W(C.cx, cyp, cz) - m1);
      const T xb = sgn(z.sm, B(t.i, t.j, z.cm) * W(C.cx, C.cy, czm) - m1);
      const T xf = sgn(z.sp, B(t.i, t.j, z.cp) * W(C.cx, C.cy, czp) - m1);
      const T ax = lap_neg<T>(c, xw, xe, xs, xn, xb, xf, ihx2, ihy2, ihz2);
      const T xn2 = c + wc * (bt + T(-1) * ax);
      x2[(size_t)t.i + (size_t)L.nx * t.j + plane * k] = xn2;
      v[0] += (double)xn2;
    }
  }
  const int ops[1] = {kSum};
  double res[1];
  if(reduce_partials<1>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0)
    ps[lslot + L_MX2] = zm ? res[0] / (double)L.n : 0.0;
}

// Residual of x2 restricted to the coarse level: bc = avg_children(b - A x2). Each thread
// computes the fine residual of its own cell; x-pairs are combined across lanes, y-pairs through
// LDS, z-pairs along the thread's k-march (kz is even).
template <typename T>
__global__ void k_residual_restrict(const T* __restrict__ b, const T* __restrict__ x2, PLevel L, PLevel Cl,
                                    T* __restrict__ bc, double* partials, unsigned* counter, double* ps, int lslot,
                                    int cslot, int zm, int kz) {
  if(stopped(ps)) return;
  __shared__ T ypair[4][64];
  const T mb = T(ps[lslot + L_MB]);
  const T mx = T(ps[lslot + L_MX2]);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const T inv_children = T(1.0 / (double)(L.rx * L.ry * L.rz));
  const Tile t = tile(L, kz);
  const Col C = column(L, t);
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t cplane = (size_t)Cl.nx * Cl.ny;
  auto X = [&](int ii, int jj, int kk) { return x2[(size_t)ii + (size_t)L.nx * jj + plane * kk] - mx; };
  double v[2] = {0.0, 0.0};
  T acc = T(0);
  for(int k = t.k0; k < t.k0 + kz; ++k) {
    T rr = T(0);
    if(C.in && k < L.nz) {
      const Nb z = nb_axis(k, L.nz, L.per[2], L.lo_dir[2], L.hi_dir[2]);
      const T c = X(t.i, t.j, k);
      const T ax = lap_neg<T>(c, sgn(C.x.sm, X(C.x.cm, t.j, k)), sgn(C.x.sp, X(C.x.cp, t.j, k)),
                              sgn(C.y.sm, X(t.i, C.y.cm, k)), sgn(C.y.sp, X(t.i, C.y.cp, k)),
                              sgn(z.sm, X(t.i, t.j, z.cm)), sgn(z.sp, X(t.i, t.j, z.cp)), ihx2, ihy2, ihz2);
      rr = (b[(size_t)t.i + (size_t)L.nx * t.j + plane * k] - mb) + T(-1) * ax;
    }
    if(L.rx == 2) rr += __shfl_xor(rr, 1, 64);  // x pair (lanes i, i+1)
    if(L.ry == 2) {                             // y pair (rows j, j+1 of the tile)
      ypair[threadIdx.y][threadIdx.x] = rr;
      __syncthreads();
      rr = rr + ypair[threadIdx.y ^ 1][threadIdx.x];
      __syncthreads();
    }
    const bool second = ((k - t.k0) & 1) != 0;
    acc = (L.rz == 2 && second) ? acc + rr : rr;
    const bool complete = L.rz == 1 || second;
    if(complete && C.in && k < L.nz && (L.rx == 1 || (t.i & 1) == 0) && (L.ry == 1 || (t.j & 1) == 0)) {
      const int ic = t.i / L.rx, jc = t.j / L.ry, kc = k / L.rz;
      const T val = inv_children * acc;
      bc[(size_t)ic + (size_t)Cl.nx * jc + cplane * kc] = val;
      v[0] += (double)val;
      v[1] += (double)val * Cl.winv[bclass(ic, Cl.nx) + 4 * (bclass(jc, Cl.ny) + 4 * bclass(kc, Cl.nz))];
    }
  }
  const int ops[2] = {kSum, kSum};
  double res[2];
  if(reduce_partials<2>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0) {
    ps[cslot + L_MB] = zm ? res[0] / (double)Cl.n : 0.0;
    ps[cslot + L_SBW] = res[1];
    ps[cslot + L_SUMB] = res[0];
  }
}

// Prolongation-add fused with the first post-smoothing sweep:
// x3 = (x2 - m2) + (xc - mc)[parent];  x4 = x3 + w/D (b - A x3).
template <typename T>
__global__ void k_prolong_smooth(const T* __restrict__ b, const T* __restrict__ x2, const T* __restrict__ xc,
                                 PLevel L, PLevel Cl, T* __restrict__ x4, double* partials, unsigned* counter,
                                 double* ps, int lslot, int cslot, int zm, int kz) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m2 = T(ps[lslot + L_MX2]);
  const T mc = T(ps[cslot + L_MX5]);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const Tile t = tile(L, kz);
  const Col C = column(L, t);
  double v[1] = {0.0};
  if(C.in) {
    const size_t plane = (size_t)L.nx * L.ny;
    const size_t cplane = (size_t)Cl.nx * Cl.ny;
    auto X3 = [&](int ii, int jj, int kk) {
      return (x2[(size_t)ii + (size_t)L.nx * jj + plane * kk] - m2) +
             (xc[(size_t)(ii / L.rx) + (size_t)Cl.nx * (jj / L.ry) + cplane * (kk / L.rz)] - mc);
    };
    for(int k = t.k0; k < t.k1; ++k) {
      const Nb z = nb_axis(k, L.nz, L.per[2], L.lo_dir[2], L.hi_dir[2]);
      const T wc = T(L.winv[C.cx + 4 * (C.cy + 4 * bclass(k, L.nz))]);
      const T c = X3(t.i, t.j, k);
      const T ax = lap_neg<T>(c, sgn(C.x.sm, X3(C.x.cm, t.j, k)), sgn(C.x.sp, X3(C.x.cp, t.j, k)),
                              sgn(C.y.sm, X3(t.i, C.y.cm, k)), sgn(C.y.sp, X3(t.i, C.y.cp, k)),
                              sgn(z.sm, X3(t.i, t.j, z.cm)), sgn(z.sp, X3(t.i, t.j, z.cp)), ihx2, ihy2, ihz2);
      const size_t id = (size_t)t.i + (size_t)L.nx * t.j + plane * k;
      const T xn = c + wc * ((b[id] - mb) + T(-1) * ax);
      x4[id] = xn;
      v[0] += (double)xn;
    }
  }
  const int ops[1] = {kSum};
  double res[1];
  if(reduce_partials<1>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0)
    ps[lslot + L_MX4] = zm ? res[0] / (double)L.n : 0.0;
}

// Second post-smoothing sweep: x5 = x4' + w/D (b - A x4'), x4' = x4 - m4. On the finest level
// also r.z and beta (r = b - mb, z = x5 - m5; the sum of r is zero up to rounding).
template <typename T>
__global__ void k_smooth_last(const T* __restrict__ b, const T* __restrict__ x4, PLevel L, T* __restrict__ x5,
                              double* partials, unsigned* counter, double* ps, int lslot, int zm, int finest, int kz) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m4 = T(ps[lslot + L_MX4]);
  const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
  const Tile t = tile(L, kz);
  const Col C = column(L, t);
  double v[2] = {0.0, 0.0};
  if(C.in) {
    const size_t plane = (size_t)L.nx * L.ny;
    auto X = [&](int ii, int jj, int kk) { return x4[(size_t)ii + (size_t)L.nx * jj + plane * kk] - m4; };
    for(int k = t.k0; k < t.k1; ++k) {
      const Nb z = nb_axis(k, L.nz, L.per[2], L.lo_dir[2], L.hi_dir[2]);
      const T wc = T(L.winv[C.cx + 4 * (C.cy + 4 * bclass(k, L.nz))]);
      const T c = X(t.i, t.j, k);
      const T ax = lap_neg<T>(c, sgn(C.x.sm, X(C.x.cm, t.j, k)), sgn(C.x.sp, X(C.x.cp, t.j, k)),
                              sgn(C.y.sm, X(t.i, C.y.cm, k)), sgn(C.y.sp, X(t.i, C.y.cp, k)),
                              sgn(z.sm, X(t.i, t.j, z.cm)), sgn(z.sp, X(t.i, t.j, z.cp)), ihx2, ihy2, ihz2);
      const size_t id = (size_t)t.i + (size_t)L.nx * t.j + plane * k;
      const T bt = b[id] - mb;
      const T xn = c + wc * (bt + T(-1) * ax);
      x5[id] = xn;
      v[0] += (double)xn;
      v[1] += (double)bt * (double)xn;
    }
  }
  const int ops[2] = {kSum, kSum};
  double res[2];
  if(reduce_partials<2>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0) {
    const double m5 = zm ? res[0] / (double)L.n : 0.0;
    ps[lslot + L_MX5] = m5;
    if(finest) {
      const double rz_new = res[1];
      const double rz_old = ps[P_RZ];
      ps[P_BETA] = rz_old != 0.0 ? rz_new / rz_old : 0.0;
      ps[P_RZ] = rz_new;
    }
  }
}

// Coarsest level: b - mb, Gaussian elimination with the pre-factored row operations of the
// reference's partial-pivoting elimination (linsolve/poisson_solver.cpp:482-567), mean removal.
template <typename T>
__global__ void k_coarse_solve(const T* __restrict__ b, T* __restrict__ x, int n, const int* __restrict__ perm,
                               const double* __restrict__ factor, const double* __restrict__ U, double* work,
                               double* ps, int lslot, int zm) {
  if(stopped(ps)) return;
  if(threadIdx.x != 0) return;
  const double mb = (double)T(ps[lslot + L_MB]);
  for(int i = 0; i < n; ++i) work[i] = (double)(b[i] - T(mb));
  if(zm) work[n - 1] = 0.0;
  for(int p = 0; p < n; ++p) {
    const int best = perm[p];
    if(best != p) {
      const double t = work[p];
      work[p] = work[best];
      work[best] = t;
    }
    for(int row = p + 1; row < n; ++row) {
      const double f = factor[(size_t)row * n + p];
      if(f != 0.0) work[row] -= f * work[p];
    }
  }
  for(int row = n - 1; row >= 0; --row) {
    double s = work[row];
    for(int c = row + 1; c < n; ++c) s -= U[(size_t)row * n + c] * work[c];
    work[row] = s / U[(size_t)row * n + row];
  }
  if(zm) {
    double s = 0.0;
    for(int i = 0; i < n; ++i) s += work[i];
    const double m = s / (double)n;
    for(int i = 0; i < n; ++i) work[i] -= m;
  }
  for(int i = 0; i < n; ++i) x[i] = T(work[i]);
  ps[lslot + L_MX5] = 0.0;
}

// ---------------------------------------------------------------- PCG vector kernels
// Fused: deferred p += alpha_prev (d_old - md_old); d = (z - mz) + beta (d_old - md_old)
// (first: d = z - mz); q = A d; sums of d, q, d.q; alpha = rz / dq or breakdown.
template <typename T>
__global__ void k_dir_apply(const T* __restrict__ z, const double* __restrict__ d_old, double* __restrict__ d_new,
                            double* __restrict__ q, double* __restrict__ p, PLevel L, double* partials,
                            unsigned* counter, double* ps, int first, int zm, int dsel, int kz) {
  if(stopped(ps)) return;
  const double mz = ps[P_LEVEL + L_MX5];
  const double md = ps[P_MD];
  const double beta = first ? 0.0 : ps[P_BETA];
  const double alpha_prev = ps[P_ALPHA];
  const int pend = !first && ps[P_PEND] != 0.0;
  const Tile t = tile(L, kz);
  const Col C = column(L, t);
  double v[3] = {0.0, 0.0, 0.0};
  if(C.in) {
    const size_t plane = (size_t)L.nx * L.ny;
    auto D = [&](int ii, int jj, int kk) {
      const size_t id = (size_t)ii + (size_t)L.nx * jj + plane * kk;
      const double zt = (double)z[id] - mz;
      return first ? zt : zt + beta * (d_old[id] - md);
    };
    for(int k = t.k0; k < t.k1; ++k) {
      const Nb zz = nb_axis(k, L.nz, L.per[2], L.lo_dir[2], L.hi_dir[2]);
      const size_t id = (size_t)t.i + (size_t)L.nx * t.j + plane * k;
      const double c = D(t.i, t.j, k);
      const double qv = lap_neg<double>(c, sgn(C.x.sm, D(C.x.cm, t.j, k)), sgn(C.x.sp, D(C.x.cp, t.j, k)),
                                        sgn(C.y.sm, D(t.i, C.y.cm, k)), sgn(C.y.sp, D(t.i, C.y.cp, k)),
                                        sgn(zz.sm, D(t.i, t.j, zz.cm)), sgn(zz.sp, D(t.i, t.j, zz.cp)), L.ihx2, L.ihy2,
                                        L.ihz2);
      if(pend) p[id] = p[id] + alpha_prev * (d_old[id] - md);
      d_new[id] = c;
      q[id] = qv;
      v[0] += c;
      v[1] += qv;
      v[2] += c * qv;
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  double res[3];
  if(reduce_partials<3>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0) {
    const double n = (double)L.n;
    const double mdn = zm ? res[0] / n : 0.0;
    const double mq = zm ? res[1] / n : 0.0;
    ps[P_MD] = mdn;
    ps[P_MQ] = mq;
    ps[P_PEND] = 0.0;
    ps[P_DSEL] = (double)dsel;
    const double dq = res[2] - mdn * res[1];
    ps[P_DQ] = dq;
    if(fabs(dq) < kSmallNumber) {
      ps[P_DONE] = 1.0;  // breakdown: stop without convergence (linsolve/poisson_solver.cpp:756-759)
    } else {
      ps[P_ALPHA] = ps[P_RZ] / dq;
    }
  }
}

// r = (r - mr) - alpha (q - mq); residual RMS, convergence test; the matching p update is
// deferred to the next direction kernel (or the final update).
template <typename T>
__global__ void k_update(double* __restrict__ r, const double* __restrict__ q, T* __restrict__ b_lo, PLevel L,
                         double* partials, unsigned* counter, double* ps, double* history, double tol, int it, int zm,
                         int write_lo, int kz) {
  if(stopped(ps)) return;
  const double alpha = ps[P_ALPHA], mr = ps[P_MR], mq = ps[P_MQ];
  const Tile t = tile(L, kz);
  double v[3] = {0.0, 0.0, 0.0};
  if(t.i < L.nx && t.j < L.ny) {
    const int cxy = bclass(t.i, L.nx) + 4 * bclass(t.j, L.ny);
    const size_t plane = (size_t)L.nx * L.ny;
    for(int k = t.k0; k < t.k1; ++k) {
      const size_t id = (size_t)t.i + (size_t)L.nx * t.j + plane * k;
      const double rn = (r[id] - mr) + -alpha * (q[id] - mq);
      r[id] = rn;
      if(write_lo) b_lo[id] = T(rn);
      v[0] += rn;
      v[1] += rn * rn;
      v[2] += rn * L.winv[cxy + 16 * bclass(k, L.nz)];
    }
  }
  const int ops[3] = {kSum, kSum, kSum};
  double res[3];
  if(reduce_partials<3>(v, ops, partials, counter, res) && threadIdx.x == 0 && threadIdx.y == 0) {
    const double n = (double)L.n;
    const double m = zm ? res[0] / n : 0.0;
    ps[P_MR] = m;
    ps[P_LEVEL + L_MB] = m;
    ps[P_LEVEL + L_SBW] = res[2];
    const double rms = sqrt(fmax(res[1] - n * m * m, 0.0) / n);
    ps[P_ITER] = (double)(it + 1);
    ps[P_RMS] = rms;
    ps[P_PEND] = 1.0;
    const double rel = ps[P_R0] > 0.0 ? rms / ps[P_R0] : 0.0;
    ps[P_REL] = rel;
    history[it + 1] = rms;
    if(rel <= tol) {
      ps[P_CONV] = 1.0;
      ps[P_DONE] = 1.0;
    }
  }
}

template <typename T>
__global__ void k_cast(const double* __restrict__ in, T* __restrict__ out, int n, const double* ps) {
  if(stopped(ps)) return;
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) out[id] = T(in[id]);
}

// Pending pressure update of the last iteration, then the mean of p.
__global__ void k_pfinal(double* __restrict__ pc, const double* __restrict__ d0, const double* __restrict__ d1, int n,
                         double* partials, unsigned* counter, double* ps, int zm) {
  const int pend = ps[P_PEND] != 0.0;
  const double* d = ps[P_DSEL] != 0.0 ? d1 : d0;
  const double alpha = ps[P_ALPHA], md = ps[P_MD];
  double v[1] = {0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) {
    double p = pc[id];
    if(pend) {
      p = p + alpha * (d[id] - md);
      pc[id] = p;
    }
    v[0] += p;
  }
  const int ops[1] = {kSum};
  double res[1];
  if(reduce_partials<1>(v, ops, partials, counter, res) && threadIdx.x == 0) {
    ps[P_PFINAL] = zm ? res[0] / (double)n : 0.0;
    ps[P_PEND] = 0.0;
  }
}

__global__ void k_pout(const double* __restrict__ pc, PLevel L, Lay Lc, double* __restrict__ p_g, const double* ps,
                       int kz) {
  const double m = ps[P_PFINAL];
  const Tile t = tile(L, kz);
  if(t.i >= L.nx || t.j >= L.ny) return;
  for(int k = t.k0; k < t.k1; ++k)
    p_g[Lc.at(t.i + 1, t.j + 1, k + 1)] = pc[(size_t)t.i + (size_t)L.nx * (t.j + (size_t)L.ny * k)] - m;
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
  // finest-level CG vectors (compact FP64); two direction buffers alternate
  double *pc = nullptr, *r = nullptr, *d[2] = {nullptr, nullptr}, *q = nullptr;
  // level arrays: b (rhs) and two x buffers, element type T
  std::vector<void*> lb, lxa, lxb;
  double* ps = nullptr;       // scalar block
  double* history = nullptr;  // residual history (device)
  double* host_ps = nullptr;  // pinned copies (2 slots)
  hipEvent_t ev[2]{};
  // coarse solve data
  int nc = 0;
  int* perm = nullptr;
  double *factor = nullptr, *U = nullptr, *work = nullptr;
  int hist_cap = 0;

  PoissonSolver(Context& ctx, const of_grid& grid, const of_pressure_bc_set& b, int prec)
      : c(ctx), g(grid), bcs(b), precision(prec) {}
  ~PoissonSolver();
  void build();
  template <typename T>
  void vcycle();
  template <typename T>
  PoissonResult solve_t(const double* rhs, double* p, int maxit, double tol, bool want_history);
};

namespace {

bool p_axis_periodic(const of_pressure_bc_set& b, int axis) {
  const bool lo = b.faces[2 * axis].type == OF_PRESSURE_PERIODIC, hi = b.faces[2 * axis + 1].type == OF_PRESSURE_PERIODIC;
  if(lo != hi) throw InvalidArgument("periodic pressure boundaries must be paired on both sides");
  return lo;
}

PLevel make_level(int nx, int ny, int nz, double dx, double dy, double dz, const of_pressure_bc_set& b) {
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
        L.winv[cx + 4 * (cy + 4 * cz)] = dval != 0.0 ? kOmega / dval : 0.0;
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

}  // namespace

PoissonSolver::~PoissonSolver() {
  auto fr = [](void* p) {
    if(p) (void)hipFree(p);
  };
  fr(pc);
  fr(r);
  fr(d[0]);
  fr(d[1]);
  fr(q);
  for(void* p : lb) fr(p);
  for(void* p : lxa) fr(p);
  for(void* p : lxb) fr(p);
  fr(ps);
  fr(history);
  fr(perm);
  fr(factor);
  fr(U);
  fr(work);
  if(host_ps) (void)hipHostFree(host_ps);
  for(auto& e : ev)
    if(e) (void)hipEventDestroy(e);
}

void PoissonSolver::build() {
  for(int f = 0; f < 6; ++f)
    if(bcs.faces[f].type == OF_PRESSURE_DIRICHLET) zm = false;
  // hierarchy (linsolve/poisson_solver.cpp:316-384)
  int nx = g.nx, ny = g.ny, nz = g.nz;
  double dx = g.dx, dy = g.dy, dz = g.dz;
  lv.push_back(make_level(nx, ny, nz, dx, dy, dz, bcs));
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
    lv.push_back(make_level(nx, ny, nz, dx, dy, dz, bcs));
  }
  if((int)lv.size() > 24) throw InvalidArgument("multigrid hierarchy deeper than supported");
  const size_t n0 = (size_t)lv[0].n;
  OFH_CHECK(hipMalloc(&pc, n0 * sizeof(double)));
  OFH_CHECK(hipMalloc(&r, n0 * sizeof(double)));
  OFH_CHECK(hipMalloc(&d[0], n0 * sizeof(double)));
  OFH_CHECK(hipMalloc(&d[1], n0 * sizeof(double)));
  OFH_CHECK(hipMalloc(&q, n0 * sizeof(double)));
  const size_t es = precision == OF_PRECISION_MIXED ? sizeof(float) : sizeof(double);
  for(size_t l = 0; l < lv.size(); ++l) {
    void *b = nullptr, *xa = nullptr, *xb = nullptr;
    const size_t n = (size_t)lv[l].n;
    // the finest FP64 level reads the CG residual directly as its rhs
    if(!(l == 0 && precision == OF_PRECISION_FP64)) OFH_CHECK(hipMalloc(&b, n * es));
    OFH_CHECK(hipMalloc(&xa, n * es));
    OFH_CHECK(hipMalloc(&xb, n * es));
    lb.push_back(b);
    lxa.push_back(xa);
    lxb.push_back(xb);
  }
  OFH_CHECK(hipMalloc(&ps, P_COUNT * sizeof(double)));
  OFH_CHECK(hipMemsetAsync(ps, 0, P_COUNT * sizeof(double), c.stream));
  OFH_CHECK(hipHostMalloc(&host_ps, 2 * P_COUNT * sizeof(double), hipHostMallocDefault));
  for(auto& e : ev) OFH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  std::vector<int> hperm;
  std::vector<double> hf, hU;
  factor_coarse(lv.back(), zm, hperm, hf, hU);
  nc = lv.back().n;
  OFH_CHECK(hipMalloc(&perm, nc * sizeof(int)));
  OFH_CHECK(hipMalloc(&factor, (size_t)nc * nc * sizeof(double)));
  OFH_CHECK(hipMalloc(&U, (size_t)nc * nc * sizeof(double)));
  OFH_CHECK(hipMalloc(&work, (size_t)nc * sizeof(double)));
  OFH_CHECK(hipMemcpy(perm, hperm.data(), nc * sizeof(int), hipMemcpyHostToDevice));
  OFH_CHECK(hipMemcpy(factor, hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice));
  OFH_CHECK(hipMemcpy(U, hU.data(), hU.size() * sizeof(double), hipMemcpyHostToDevice));
}

template <typename T>
void PoissonSolver::vcycle() {
  const int L = (int)lv.size();
  const int zmi = zm ? 1 : 0;
  unsigned* cnt = c.counters + 8;
  auto slot = [](int l) { return P_LEVEL + 8 * l; };
  // finest rhs: CG residual (FP64) or its FP32 copy (mixed)
  auto bptr = [&](int l) -> const T* {
    if(l == 0 && precision == OF_PRECISION_FP64) return reinterpret_cast<const T*>(r);
    return static_cast<const T*>(lb[l]);
  };
  for(int l = 0; l < L - 1; ++l) {
    const PLevel& P = lv[l];
    const PLevel& C = lv[l + 1];
    const LevelLaunch ll = level_launch(P);
    {
      ProfScope prof_(c, "k_smooth12", l == 0);
      hipLaunchKernelGGL(k_smooth12<T>, ll.grid, ll.block, 0, c.stream, bptr(l), static_cast<T*>(lxa[l]), P,
                         c.partials, cnt, ps, slot(l), zmi, ll.kz);
    }
    {
      ProfScope prof_(c, "k_residual_restrict", l == 0);
      hipLaunchKernelGGL(k_residual_restrict<T>, ll.grid, ll.block, 0, c.stream, bptr(l),
                         static_cast<const T*>(lxa[l]), P, C, static_cast<T*>(lb[l + 1]), c.partials, cnt, ps,
                         slot(l), slot(l + 1), zmi, ll.kz);
    }
  }
  hipLaunchKernelGGL(k_coarse_solve<T>, 1, 64, 0, c.stream, static_cast<const T*>(lb[L - 1]),
                     static_cast<T*>(lxa[L - 1]), nc, perm, factor, U, work, ps, slot(L - 1), zmi);
  for(int l = L - 2; l >= 0; --l) {
    const PLevel& P = lv[l];
    const PLevel& C = lv[l + 1];
    const LevelLaunch ll = level_launch(P);
    {
      ProfScope prof_(c, "k_prolong_smooth", l == 0);
      hipLaunchKernelGGL(k_prolong_smooth<T>, ll.grid, ll.block, 0, c.stream, bptr(l),
                         static_cast<const T*>(lxa[l]), static_cast<const T*>(lxa[l + 1]), P, C,
                         static_cast<T*>(lxb[l]), c.partials, cnt, ps, slot(l), slot(l + 1), zmi, ll.kz);
    }
    {
      ProfScope prof_(c, "k_smooth_last", l == 0);
      hipLaunchKernelGGL(k_smooth_last<T>, ll.grid, ll.block, 0, c.stream, bptr(l), static_cast<const T*>(lxb[l]),
                         P, static_cast<T*>(lxa[l]), c.partials, cnt, ps, slot(l), zmi, l == 0 ? 1 : 0, ll.kz);
    }
  }
}

template <typename T>
PoissonResult PoissonSolver::solve_t(const double* rhs, double* p, int maxit, double tol, bool want_history) {
  if(hist_cap < maxit + 1) {
    if(history) OFH_CHECK(hipFree(history));
    OFH_CHECK(hipMalloc(&history, (size_t)(maxit + 1) * sizeof(double)));
    hist_cap = maxit + 1;
  }
  const PLevel& F = lv[0];
  const Lay Lc = make_lay(g, OF_CELL);
  const int zmi = zm ? 1 : 0;
  unsigned* cnt = c.counters + 8;
  const LevelLaunch ll = level_launch(F);
  const bool mixed = precision == OF_PRECISION_MIXED;
  OFH_CHECK(hipMemsetAsync(ps, 0, P_COUNT * sizeof(double), c.stream));
  hipLaunchKernelGGL(k_setup1, ll.grid, ll.block, 0, c.stream, rhs, p, Lc, F, r, pc, c.partials, cnt, ps, zmi, ll.kz);
  hipLaunchKernelGGL(k_setup2, ll.grid, ll.block, 0, c.stream, r, pc, F, c.partials, cnt, ps, history, tol, zmi,
                     ll.kz);
  if(mixed)  // FP32 copy of the initial residual: the rhs of the finest FP32 level
    hipLaunchKernelGGL(k_cast<T>, 2048, 256, 0, c.stream, r, static_cast<T*>(lb[0]), F.n, ps);
  vcycle<T>();
  int it = 0;
  for(; it < maxit; ++it) {
    {
      ProfScope prof_(c, "k_dir_apply");
      hipLaunchKernelGGL(k_dir_apply<T>, ll.grid, ll.block, 0, c.stream, static_cast<const T*>(lxa[0]),
                         d[(it + 1) % 2], d[it % 2], q, pc, F, c.partials, cnt, ps, it == 0 ? 1 : 0, zmi, it % 2,
                         ll.kz);
    }
    {
      ProfScope prof_(c, "k_update");
      hipLaunchKernelGGL(k_update<T>, ll.grid, ll.block, 0, c.stream, r, q, mixed ? static_cast<T*>(lb[0]) : nullptr,
                         F, c.partials, cnt, ps, history, tol, it, zmi, mixed ? 1 : 0, ll.kz);
    }
    OFH_CHECK(hipMemcpyAsync(host_ps + (it % 2) * P_COUNT, ps, 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    OFH_CHECK(hipEventRecord(ev[it % 2], c.stream));
    if(it > 0) {
      OFH_CHECK(hipEventSynchronize(ev[(it - 1) % 2]));
      if(host_ps[((it - 1) % 2) * P_COUNT + P_DONE] != 0.0) break;
    }
    vcycle<T>();
  }
  hipLaunchKernelGGL(k_pfinal, 2048, 256, 0, c.stream, pc, d[0], d[1], F.n, c.partials, cnt, ps, zmi);
  hipLaunchKernelGGL(k_pout, ll.grid, ll.block, 0, c.stream, pc, F, Lc, p, ps, ll.kz);
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

void poisson_destroy(PoissonSolver* s) { delete s; }

PoissonResult poisson_solve(PoissonSolver* s, const double* rhs, double* p, int maxit, double tol, bool want_history) {
  if(maxit <= 0) throw InvalidArgument("solve_pressure_poisson expects iterations > 0");
  if(!(tol > 0.0)) throw InvalidArgument("solve_pressure_poisson expects tolerance > 0");
  if(s->precision == OF_PRECISION_MIXED) return s->solve_t<float>(rhs, p, maxit, tol, want_history);
  return s->solve_t<double>(rhs, p, maxit, tol, want_history);
}

void launch_poisson_apply(Context& c, const of_grid& g, const of_pressure_bc_set& b, bool homogeneous,
                          const double* in, double* out) {
  const PLevel L = make_level(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, b);
  const LevelLaunch ll = level_launch(L);
  hipLaunchKernelGGL(k_poisson_apply_g, ll.grid, ll.block, 0, c.stream, in, out, make_lay(g, OF_CELL), L,
                     homogeneous ? 1 : 0, ll.kz);
  OFH_CHECK(hipGetLastError());
}

}  // namespace ofh
