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
//  * Mean removal (null space) is deferred: a kernel that produces a vector also reduces its
//    sum, the last block stores the mean, and every consumer subtracts it when it reads the
//    vector. The values consumed are the values the reference stores after subtract_active_mean.
//  * The first pre-smoothing sweep starts from x = 0, so it is pointwise (x1 = w b / D); it is
//    fused into the second sweep, whose stencil evaluates x1 at the neighbours on the fly.
//    The residual is fused with the restriction and the prolongation with the first
//    post-smoothing sweep.
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
constexpr int kPT = 256;  // threads per block for the compact-array kernels

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
  P_MD,            // mean of d
  P_MQ,            // mean of q
  P_MP,            // mean of the warm start p
  P_MSYS,          // removed rhs mean
  P_PFINAL,        // final mean of p
  P_LEVEL = 32,    // per level: 8 slots (see LS)
  P_COUNT = 32 + 8 * 24
};
enum LS : int { L_MB = 0, L_SBD, L_MX2, L_MX4, L_MX5, L_SUMB, L_SPARE1, L_SPARE2 };

struct PLevel {
  int nx, ny, nz;
  int n;
  double hx2, hy2, hz2;    // h*h per axis
  double inv[3];           // 1/pow(h,2) (operator diagonal)
  int per[3];              // periodic axes
  int lo_dir[3], hi_dir[3];// Dirichlet faces
  // actual (inhomogeneous) face data, used only for the boundary source term
  double lo_val[3], hi_val[3], lo_grad[3], hi_grad[3], h[3];
  double inv_d_sum;        // sum over cells of 1/D
  int rx, ry, rz;          // refinement ratio to the next coarser level
};

__device__ inline void unflat(const PLevel& L, int id, int& i, int& j, int& k) {
  i = id % L.nx;
  const int t = id / L.nx;
  j = t % L.ny;
  k = t / L.ny;
}
__device__ inline int flat(const PLevel& L, int i, int j, int k) { return i + L.nx * (j + L.ny * k); }

// linsolve/poisson_solver.cpp:271-314
__device__ inline double diag_at(const PLevel& L, int i, int j, int k) {
  const int co[3] = {i, j, k};
  const int nn[3] = {L.nx, L.ny, L.nz};
  double v = 0.0;
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    double coef = 2.0 * L.inv[a];
    if(L.per[a]) {
      v += coef;
      continue;
    }
    if(co[a] == 0) coef += L.lo_dir[a] ? L.inv[a] : -L.inv[a];
    if(co[a] == nn[a] - 1) coef += L.hi_dir[a] ? L.inv[a] : -L.inv[a];
    v += coef;
  }
  return v;
}

// linsolve/poisson_solver.cpp:195-269 with homogeneous BCs; X(i,j,k) returns the (mean-removed)
// value at a compact index; the centre value is c.
template <typename T, typename F>
__device__ inline T apply_hom(const PLevel& L, int i, int j, int k, T c, F&& X) {
  T nb[3][2];
  const int co[3] = {i, j, k};
  const int nn[3] = {L.nx, L.ny, L.nz};
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    const int di = a == 0, dj = a == 1, dk = a == 2;
    if(co[a] > 0) nb[a][0] = X(i - di, j - dj, k - dk);
    else if(L.per[a]) nb[a][0] = X(a == 0 ? nn[0] - 1 : i, a == 1 ? nn[1] - 1 : j, a == 2 ? nn[2] - 1 : k);
    else nb[a][0] = L.lo_dir[a] ? T(2.0) * T(0.0) - c : c;
    if(co[a] < nn[a] - 1) nb[a][1] = X(i + di, j + dj, k + dk);
    else if(L.per[a]) nb[a][1] = X(a == 0 ? 0 : i, a == 1 ? 0 : j, a == 2 ? 0 : k);
    else nb[a][1] = L.hi_dir[a] ? T(2.0) * T(0.0) - c : c;
  }
  return -(nb[0][1] - T(2.0) * c + nb[0][0]) / T(L.hx2) - (nb[1][1] - T(2.0) * c + nb[1][0]) / T(L.hy2) -
         (nb[2][1] - T(2.0) * c + nb[2][0]) / T(L.hz2);
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

__device__ inline bool stopped(const double* ps) { return ps[P_DONE] != 0.0; }

// ---------------------------------------------------------------- setup
// sys = -rhs - src (active cells), p_c = warm start; sums for their means.
__global__ void k_setup1(const double* __restrict__ rhs_g, const double* __restrict__ p_g, Lay Lc, PLevel L,
                         double* __restrict__ sys, double* __restrict__ pc, double* partials, unsigned* counter,
                         double* ps, int zm) {
  double v[2] = {0.0, 0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    const double s = -rhs_g[Lc.at(i + 1, j + 1, k + 1)] + -1.0 * boundary_source(L, i, j, k);
    const double p = p_g[Lc.at(i + 1, j + 1, k + 1)];
    sys[id] = s;
    pc[id] = p;
    v[0] += s;
    v[1] += p;
  }
  const int ops[2] = {kSum, kSum};
  double res[2];
  if(reduce_partials<2>(v, ops, partials, counter, res) && threadIdx.x == 0) {
    ps[P_MSYS] = zm ? res[0] / (double)L.n : 0.0;
    ps[P_MP] = zm ? res[1] / (double)L.n : 0.0;
  }
}

// r = (sys - m_sys) - A(p - m_p); sums of r, r^2, r/D; initial residual and early exit.
__global__ void k_setup2(double* __restrict__ r, const double* __restrict__ pc, PLevel L, double* partials,
                         unsigned* counter, double* ps, double* history, double tol, int zm) {
  const double msys = ps[P_MSYS], mp = ps[P_MP];
  double v[3] = {0.0, 0.0, 0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    const double c = pc[id] - mp;
    const double q = apply_hom<double>(L, i, j, k, c, [&](int a, int b, int e) { return pc[flat(L, a, b, e)] - mp; });
    const double rv = (r[id] - msys) + -1.0 * q;
    r[id] = rv;
    v[0] += rv;
    v[1] += rv * rv;
    v[2] += rv / diag_at(L, i, j, k);
  }
  const int ops[3] = {kSum, kSum, kSum};
  double res[3];
  if(reduce_partials<3>(v, ops, partials, counter, res) && threadIdx.x == 0) {
    const double n = (double)L.n;
    const double mr = zm ? res[0] / n : 0.0;
    ps[P_MR] = mr;
    ps[P_LEVEL + L_MB] = mr;
    ps[P_LEVEL + L_SBD] = res[2];
    const double rms = sqrt(fmax(res[1] - n * mr * mr, 0.0) / n);
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
}

// ---------------------------------------------------------------- V-cycle kernels
// Level b (rhs) of type T with deferred mean mb; x arrays of type T.
// S12: x2 = x1 + w (b - A x1)/D with x1 = w b/D - m1 (fused first two pre-smoothing sweeps).
template <typename T>
__global__ void k_smooth12(const T* __restrict__ b, T* __restrict__ x2, PLevel L, double* partials, unsigned* counter,
                           double* ps, int lslot, int zm) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  // mean of x1 = w (b - mb) / D  from the producer's sums: w (S(b/D) - mb S(1/D)) / n
  const T m1 = zm ? T(kOmega * (ps[lslot + L_SBD] - ps[lslot + L_MB] * L.inv_d_sum) / (double)L.n) : T(0);
  const T om = T(kOmega);
  double v[1] = {0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    auto X1 = [&](int a, int bb, int e) {
      return (om * (b[flat(L, a, bb, e)] - mb)) / T(diag_at(L, a, bb, e)) - m1;
    };
    const T d = T(diag_at(L, i, j, k));
    const T bt = b[id] - mb;
    const T c = (om * bt) / d - m1;
    const T ax = apply_hom<T>(L, i, j, k, c, X1);
    const T rr = bt + T(-1.0) * ax;
    const T xn = c + (om * rr) / d;
    x2[id] = xn;
    v[0] += (double)xn;
  }
  const int ops[1] = {kSum};
  double res[1];
  if(reduce_partials<1>(v, ops, partials, counter, res) && threadIdx.x == 0)
    ps[lslot + L_MX2] = zm ? res[0] / (double)L.n : 0.0;
}

// Residual of x2 restricted to the coarse level: bc = avg_children(b - A x2).
template <typename T>
__global__ void k_residual_restrict(const T* __restrict__ b, const T* __restrict__ x2, PLevel L, PLevel C,
                                    T* __restrict__ bc, double* partials, unsigned* counter, double* ps, int lslot,
                                    int cslot, int zm) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T mx = T(ps[lslot + L_MX2]);
  const T inv_children = T(1.0 / (double)(L.rx * L.ry * L.rz));
  double v[2] = {0.0, 0.0};
  for(int cid = blockIdx.x * blockDim.x + threadIdx.x; cid < C.n; cid += gridDim.x * blockDim.x) {
    int ic, jc, kc;
    unflat(C, cid, ic, jc, kc);
    T sum = T(0);
    for(int dz = 0; dz < L.rz; ++dz)
      for(int dy = 0; dy < L.ry; ++dy)
        for(int dx = 0; dx < L.rx; ++dx) {
          const int i = ic * L.rx + dx, j = jc * L.ry + dy, k = kc * L.rz + dz;
          const int id = flat(L, i, j, k);
          const T c = x2[id] - mx;
          const T ax = apply_hom<T>(L, i, j, k, c, [&](int a, int bb, int e) { return x2[flat(L, a, bb, e)] - mx; });
          const T rr = (b[id] - mb) + T(-1.0) * ax;
          sum += rr;
        }
    const T val = inv_children * sum;
    bc[cid] = val;
    v[0] += (double)val;
    v[1] += (double)val / diag_at(C, ic, jc, kc);
  }
  const int ops[2] = {kSum, kSum};
  double res[2];
  if(reduce_partials<2>(v, ops, partials, counter, res) && threadIdx.x == 0) {
    ps[cslot + L_MB] = zm ? res[0] / (double)C.n : 0.0;
    ps[cslot + L_SBD] = res[1];
    ps[cslot + L_SUMB] = res[0];
  }
}

// Prolongation-add fused with the first post-smoothing sweep:
// x3 = (x2 - m2) + (xc - mc)[parent];  x4 = x3 + w (b - A x3)/D.
template <typename T>
__global__ void k_prolong_smooth(const T* __restrict__ b, const T* __restrict__ x2, const T* __restrict__ xc,
                                 PLevel L, PLevel C, T* __restrict__ x4, double* partials, unsigned* counter,
                                 double* ps, int lslot, int cslot, int zm) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m2 = T(ps[lslot + L_MX2]);
  const T mc = T(ps[cslot + L_MX5]);
  const T om = T(kOmega);
  double v[1] = {0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    auto X3 = [&](int a, int bb, int e) {
      return (x2[flat(L, a, bb, e)] - m2) + (xc[flat(C, a / L.rx, bb / L.ry, e / L.rz)] - mc);
    };
    const T c = X3(i, j, k);
    const T d = T(diag_at(L, i, j, k));
    const T ax = apply_hom<T>(L, i, j, k, c, X3);
    const T rr = (b[id] - mb) + T(-1.0) * ax;
    const T xn = c + (om * rr) / d;
    x4[id] = xn;
    v[0] += (double)xn;
  }
  const int ops[1] = {kSum};
  double res[1];
  if(reduce_partials<1>(v, ops, partials, counter, res) && threadIdx.x == 0)
    ps[lslot + L_MX4] = zm ? res[0] / (double)L.n : 0.0;
}

// Second post-smoothing sweep: x5 = x4' + w (b - A x4')/D, x4' = x4 - m4. On the finest level
// also r.z and beta (with r = b - mb, z = x5 - m5; the sum of r is zero up to rounding).
template <typename T>
__global__ void k_smooth_last(const T* __restrict__ b, const T* __restrict__ x4, PLevel L, T* __restrict__ x5,
                              double* partials, unsigned* counter, double* ps, int lslot, int zm, int finest) {
  if(stopped(ps)) return;
  const T mb = T(ps[lslot + L_MB]);
  const T m4 = T(ps[lslot + L_MX4]);
  const T om = T(kOmega);
  double v[2] = {0.0, 0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    const T c = x4[id] - m4;
    const T d = T(diag_at(L, i, j, k));
    const T ax = apply_hom<T>(L, i, j, k, c, [&](int a, int bb, int e) { return x4[flat(L, a, bb, e)] - m4; });
    const T bt = b[id] - mb;
    const T rr = bt + T(-1.0) * ax;
    const T xn = c + (om * rr) / d;
    x5[id] = xn;
    v[0] += (double)xn;
    v[1] += (double)bt * (double)xn;
  }
  const int ops[2] = {kSum, kSum};
  double res[2];
  if(reduce_partials<2>(v, ops, partials, counter, res) && threadIdx.x == 0) {
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
// d = (z - mz) + beta (d - md)  [first call: beta = 0 and d := z - mz]
template <typename T>
__global__ void k_direction(const T* __restrict__ z, double* __restrict__ d, PLevel L, double* partials,
                            unsigned* counter, double* ps, int first, int zm) {
  if(stopped(ps)) return;
  const double mz = ps[P_LEVEL + L_MX5];
  const double md = ps[P_MD];
  const double beta = ps[P_BETA];
  double v[1] = {0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    const double zt = (double)z[id] - mz;
    const double dn = first ? zt : zt + beta * (d[id] - md);
    d[id] = dn;
    v[0] += dn;
  }
  const int ops[1] = {kSum};
  double res[1];
  if(reduce_partials<1>(v, ops, partials, counter, res) && threadIdx.x == 0)
    ps[P_MD] = zm ? res[0] / (double)L.n : 0.0;
}

// q = A (d - md); dq = (d - md).q; alpha = rz / dq (or breakdown).
__global__ void k_apply_dot(const double* __restrict__ d, double* __restrict__ q, PLevel L, double* partials,
                            unsigned* counter, double* ps, int zm) {
  if(stopped(ps)) return;
  const double md = ps[P_MD];
  double v[2] = {0.0, 0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    const double c = d[id] - md;
    const double qv = apply_hom<double>(L, i, j, k, c, [&](int a, int b, int e) { return d[flat(L, a, b, e)] - md; });
    q[id] = qv;
    v[0] += qv;
    v[1] += c * qv;
  }
  const int ops[2] = {kSum, kSum};
  double res[2];
  if(reduce_partials<2>(v, ops, partials, counter, res) && threadIdx.x == 0) {
    const double mq = zm ? res[0] / (double)L.n : 0.0;
    ps[P_MQ] = mq;
    const double dq = res[1];
    ps[P_DQ] = dq;
    if(fabs(dq) < kSmallNumber) {
      ps[P_DONE] = 1.0;  // breakdown: stop without convergence (linsolve/poisson_solver.cpp:756-759)
    } else {
      ps[P_ALPHA] = ps[P_RZ] / dq;
    }
  }
}

// p += alpha (d - md); r = (r - mr) - alpha (q - mq); residual RMS, convergence test.
template <typename T>
__global__ void k_update(double* __restrict__ p, const double* __restrict__ d, double* __restrict__ r,
                         const double* __restrict__ q, T* __restrict__ b_lo, PLevel L, double* partials,
                         unsigned* counter, double* ps, double* history, double tol, int it, int zm, int write_lo) {
  if(stopped(ps)) return;
  const double alpha = ps[P_ALPHA], md = ps[P_MD], mr = ps[P_MR], mq = ps[P_MQ];
  double v[3] = {0.0, 0.0, 0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    p[id] = p[id] + alpha * (d[id] - md);
    const double rn = (r[id] - mr) + -alpha * (q[id] - mq);
    r[id] = rn;
    if(write_lo) b_lo[id] = T(rn);
    v[0] += rn;
    v[1] += rn * rn;
    v[2] += rn / diag_at(L, i, j, k);
  }
  const int ops[3] = {kSum, kSum, kSum};
  double res[3];
  if(reduce_partials<3>(v, ops, partials, counter, res) && threadIdx.x == 0) {
    const double n = (double)L.n;
    const double m = zm ? res[0] / n : 0.0;
    ps[P_MR] = m;
    ps[P_LEVEL + L_MB] = m;
    ps[P_LEVEL + L_SBD] = res[2];
    const double rms = sqrt(fmax(res[1] - n * m * m, 0.0) / n);
    ps[P_ITER] = (double)(it + 1);
    ps[P_RMS] = rms;
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

// Final pressure: mean of p_c, then p_out (active range of the ghosted field) = p_c - mean.
__global__ void k_psum(const double* __restrict__ pc, PLevel L, double* partials, unsigned* counter, double* ps,
                       int zm) {
  double v[1] = {0.0};
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) v[0] += pc[id];
  const int ops[1] = {kSum};
  double res[1];
  if(reduce_partials<1>(v, ops, partials, counter, res) && threadIdx.x == 0)
    ps[P_PFINAL] = zm ? res[0] / (double)L.n : 0.0;
}

__global__ void k_pout(const double* __restrict__ pc, PLevel L, Lay Lc, double* __restrict__ p_g, const double* ps) {
  const double m = ps[P_PFINAL];
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    p_g[Lc.at(i + 1, j + 1, k + 1)] = pc[id] - m;
  }
}

// Poisson operator on ghosted storage with actual or homogeneous BCs (for the API).
__global__ void k_poisson_apply_g(const double* __restrict__ in, double* __restrict__ out, Lay Lc, PLevel L,
                                  int homogeneous) {
  for(int id = blockIdx.x * blockDim.x + threadIdx.x; id < L.n; id += gridDim.x * blockDim.x) {
    int i, j, k;
    unflat(L, id, i, j, k);
    const double c = in[Lc.at(i + 1, j + 1, k + 1)];
    auto X = [&](int a, int b, int e) { return in[Lc.at(a + 1, b + 1, e + 1)]; };
    double r = apply_hom<double>(L, i, j, k, c, X);
    if(!homogeneous) {
      // add the inhomogeneous boundary part: A(x; bc) = A(x; 0) + A(0; bc)
      // evaluated in the reference's single expression order below instead
      double nb[3][2];
      const int co[3] = {i, j, k};
      const int nn[3] = {L.nx, L.ny, L.nz};
      for(int a = 0; a < 3; ++a) {
        const int di = a == 0, dj = a == 1, dk = a == 2;
        if(co[a] > 0) nb[a][0] = X(i - di, j - dj, k - dk);
        else if(L.per[a]) nb[a][0] = X(a == 0 ? nn[0] - 1 : i, a == 1 ? nn[1] - 1 : j, a == 2 ? nn[2] - 1 : k);
        else nb[a][0] = L.lo_dir[a] ? 2.0 * L.lo_val[a] - c : c - L.lo_grad[a] * L.h[a];
        if(co[a] < nn[a] - 1) nb[a][1] = X(i + di, j + dj, k + dk);
        else if(L.per[a]) nb[a][1] = X(a == 0 ? 0 : i, a == 1 ? 0 : j, a == 2 ? 0 : k);
        else nb[a][1] = L.hi_dir[a] ? 2.0 * L.hi_val[a] - c : c + L.hi_grad[a] * L.h[a];
      }
      r = -(nb[0][1] - 2.0 * c + nb[0][0]) / L.hx2 - (nb[1][1] - 2.0 * c + nb[1][0]) / L.hy2 -
          (nb[2][1] - 2.0 * c + nb[2][0]) / L.hz2;
    }
    out[Lc.at(i + 1, j + 1, k + 1)] = r;
  }
}

unsigned blocks_for(int n) { return (unsigned)std::max(1, std::min((n + kPT - 1) / kPT, 2048)); }

inline bool axis_coarsens(int n) { return n > 1 && n % 2 == 0 && n > 2; }  // :316-345

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
  // finest-level CG vectors (compact FP64)
  double *pc = nullptr, *r = nullptr, *d = nullptr, *q = nullptr;
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
  // sum of 1/D over the level (host, same diagonal formula)
  double s = 0.0;
  const int nn[3] = {nx, ny, nz};
  for(int k = 0; k < nz; ++k)
    for(int j = 0; j < ny; ++j)
      for(int i = 0; i < nx; ++i) {
        const int co[3] = {i, j, k};
        double v = 0.0;
        for(int a = 0; a < 3; ++a) {
          double coef = 2.0 * L.inv[a];
          if(L.per[a]) {
            v += coef;
            continue;
          }
          if(co[a] == 0) coef += L.lo_dir[a] ? L.inv[a] : -L.inv[a];
          if(co[a] == nn[a] - 1) coef += L.hi_dir[a] ? L.inv[a] : -L.inv[a];
          v += coef;
        }
        s += 1.0 / v;
      }
  L.inv_d_sum = s;
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
  fr(d);
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
  OFH_CHECK(hipMalloc(&d, n0 * sizeof(double)));
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
    {
      ProfScope prof_(c, "k_smooth12", l == 0);
      hipLaunchKernelGGL(k_smooth12<T>, blocks_for(P.n), kPT, 0, c.stream, bptr(l), static_cast<T*>(lxa[l]), P,
                         c.partials, cnt, ps, slot(l), zmi);
    }
    {
      ProfScope prof_(c, "k_residual_restrict", l == 0);
      hipLaunchKernelGGL(k_residual_restrict<T>, blocks_for(C.n), kPT, 0, c.stream, bptr(l),
                         static_cast<const T*>(lxa[l]), P, C, static_cast<T*>(lb[l + 1]), c.partials, cnt, ps, slot(l),
                         slot(l + 1), zmi);
    }
  }
  hipLaunchKernelGGL(k_coarse_solve<T>, 1, 64, 0, c.stream, static_cast<const T*>(lb[L - 1]),
                     static_cast<T*>(lxa[L - 1]), nc, perm, factor, U, work, ps, slot(L - 1), zmi);
  for(int l = L - 2; l >= 0; --l) {
    const PLevel& P = lv[l];
    const PLevel& C = lv[l + 1];
    {
      ProfScope prof_(c, "k_prolong_smooth", l == 0);
      hipLaunchKernelGGL(k_prolong_smooth<T>, blocks_for(P.n), kPT, 0, c.stream, bptr(l), static_cast<const T*>(lxa[l]),
                         static_cast<const T*>(lxa[l + 1]), P, C, static_cast<T*>(lxb[l]), c.partials, cnt, ps,
                         slot(l), slot(l + 1), zmi);
    }
    {
      ProfScope prof_(c, "k_smooth_last", l == 0);
      hipLaunchKernelGGL(k_smooth_last<T>, blocks_for(P.n), kPT, 0, c.stream, bptr(l), static_cast<const T*>(lxb[l]),
                         P, static_cast<T*>(lxa[l]), c.partials, cnt, ps, slot(l), zmi, l == 0 ? 1 : 0);
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
  const unsigned nb = blocks_for(F.n);
  const bool mixed = precision == OF_PRECISION_MIXED;
  OFH_CHECK(hipMemsetAsync(ps, 0, P_COUNT * sizeof(double), c.stream));
  hipLaunchKernelGGL(k_setup1, nb, kPT, 0, c.stream, rhs, p, Lc, F, r, pc, c.partials, cnt, ps, zmi);
  hipLaunchKernelGGL(k_setup2, nb, kPT, 0, c.stream, r, pc, F, c.partials, cnt, ps, history, tol, zmi);
  if(mixed) {
    // FP32 copy of the initial residual: the rhs of the finest FP32 level
    hipLaunchKernelGGL(k_cast<T>, nb, kPT, 0, c.stream, r, static_cast<T*>(lb[0]), F.n, ps);
  }
  vcycle<T>();
  {
    ProfScope prof_(c, "k_direction", true);
    hipLaunchKernelGGL(k_direction<T>, nb, kPT, 0, c.stream, static_cast<const T*>(lxa[0]), d, F, c.partials, cnt, ps,
                       1, zmi);
  }
  int it = 0;
  for(; it < maxit; ++it) {
    {
      ProfScope prof_(c, "k_apply_dot", true);
      hipLaunchKernelGGL(k_apply_dot, nb, kPT, 0, c.stream, d, q, F, c.partials, cnt, ps, zmi);
    }
    {
      ProfScope prof_(c, "k_update", true);
      hipLaunchKernelGGL(k_update<T>, nb, kPT, 0, c.stream, pc, d, r, q, mixed ? static_cast<T*>(lb[0]) : nullptr, F,
                         c.partials, cnt, ps, history, tol, it, zmi, mixed ? 1 : 0);
    }
    OFH_CHECK(hipMemcpyAsync(host_ps + (it % 2) * P_COUNT, ps, 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    OFH_CHECK(hipEventRecord(ev[it % 2], c.stream));
    if(it > 0) {
      OFH_CHECK(hipEventSynchronize(ev[(it - 1) % 2]));
      if(host_ps[((it - 1) % 2) * P_COUNT + P_DONE] != 0.0) break;
    }
    vcycle<T>();
    {
      ProfScope prof_(c, "k_direction", true);
      hipLaunchKernelGGL(k_direction<T>, nb, kPT, 0, c.stream, static_cast<const T*>(lxa[0]), d, F, c.partials, cnt, ps,
                         0, zmi);
    }
  }
  hipLaunchKernelGGL(k_psum, nb, kPT, 0, c.stream, pc, F, c.partials, cnt, ps, zmi);
  hipLaunchKernelGGL(k_pout, nb, kPT, 0, c.stream, pc, F, Lc, p, ps);
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
  hipLaunchKernelGGL(k_poisson_apply_g, blocks_for(L.n), kPT, 0, c.stream, in, out, make_lay(g, OF_CELL), L,
                     homogeneous ? 1 : 0);
  OFH_CHECK(hipGetLastError());
}

}  // namespace ofh
