// Field kernels: boundary-condition fills, MAC operators, TVD advection, the factorisation
// correction and the fused predictor right-hand side, all on the reference's ghosted storage.
//
// Floating-point expressions keep the reference's association order (this file is built with
// -ffp-contract=off), so per-point results match the reference to the last bit where the
// inputs do.

#include "ofh.hpp"

#include <algorithm>
#include <cmath>

namespace ofh {

ColumnLaunch column_launch(int ex, int ey, int ez, int min_kz) {
  ColumnLaunch l;
  l.block = dim3(kBX, kBY, 1);
  const unsigned gx = (unsigned)((ex + kBX - 1) / kBX), gy = (unsigned)((ey + kBY - 1) / kBY);
  const size_t plane = (size_t)gx * gy;
  int kz = std::max(1, min_kz);
  // keep the block count below the partial-slot capacity and near 8 blocks per CU x 256 CUs
  while(plane * (size_t)((ez + kz - 1) / kz) > (size_t)std::min(kMaxPartials, 16384) && kz < ez) kz *= 2;
  kz = std::min(kz, std::max(ez, 1));
  l.kz = kz;
  l.grid = dim3(gx, gy, (unsigned)((ez + kz - 1) / kz));
  return l;
}

namespace {

// ------------------------------------------------------------------ boundary conditions
struct Slab {
  int axis, o0, o1;
};
__device__ inline Slab slab_of(int axis) {
  Slab s;
  s.axis = axis;
  s.o0 = axis == 0 ? 1 : 0;
  s.o1 = axis == 2 ? 1 : 2;
  return s;
}
__device__ inline size_t slab_at(const Lay& L, const Slab& s, int along, int t0, int t1) {
  int c[3];
  c[s.axis] = along;
  c[s.o0] = 1 + t0;
  c[s.o1] = 1 + t1;
  return L.at(c[0], c[1], c[2]);
}

// solver/projection.cpp:233-300 (periodic) and :384-449 (walls) for one component along one axis
__device__ inline void component_axis_bc(double* f, const Lay& L, int comp, int axis, int periodic, const of_bc& lo,
                                         const of_bc& hi, int t0, int t1) {
  const Slab s = slab_of(axis);
  if(t0 >= L.a(s.o0) || t1 >= L.a(s.o1)) return;
  const int A = L.a(axis);
  auto F = [&](int along) -> double& { return f[slab_at(L, s, along, t0, t1)]; };
  if(periodic) {
    if(comp == axis) {
      const double shared = 0.5 * (F(1) + F(A));
      F(1) = shared;
      F(A) = shared;
      F(0) = F(A - 1);
      F(A + 1) = F(2);
    } else {
      F(0) = F(A);
      F(A + 1) = F(1);
    }
    return;
  }
  for(int side = 0; side < 2; ++side) {
    const of_bc& c = side == 0 ? lo : hi;
    const bool lower = side == 0;
    if(c.type == OF_PERIODIC) continue;
    const int b = lower ? 1 : A;
    const int gh = lower ? 0 : A + 1;
    if(comp == axis) {
      const int ii = lower ? b + 1 : b - 1;
      if(c.type == OF_FIXED_PRESSURE) {
        F(b) = F(ii);
        F(gh) = F(b);
        continue;
      }
      const double bv = c.type == OF_SYMMETRY ? 0.0 : c.velocity[comp];
      F(b) = bv;
      F(gh) = 2.0 * bv - F(ii);
      continue;
    }
    if(c.type == OF_NO_SLIP_WALL || c.type == OF_PRESCRIBED_VELOCITY) {
      F(gh) = 2.0 * c.velocity[comp] - F(b);
    } else if(c.type == OF_SYMMETRY || c.type == OF_FIXED_PRESSURE) {
      F(gh) = F(b);
    }
  }
}

__global__ void k_velocity_bc_axis(double* u, double* v, double* w, Lay Lu, Lay Lv, Lay Lw, int axis, int periodic,
                                   of_bc lo, of_bc hi, int comp_mask) {
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  if(comp_mask & 1) component_axis_bc(u, Lu, 0, axis, periodic, lo, hi, t0, t1);
  if(comp_mask & 2) component_axis_bc(v, Lv, 1, axis, periodic, lo, hi, t0, t1);
  if(comp_mask & 4) component_axis_bc(w, Lw, 2, axis, periodic, lo, hi, t0, t1);
}

// solver/projection.cpp:207-231,302-337,973-984
__global__ void k_pressure_bc_axis(double* p, Lay L, int axis, int periodic, of_pressure_bc lo, of_pressure_bc hi,
                                   double h) {
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  const Slab s = slab_of(axis);
  if(t0 >= L.a(s.o0) || t1 >= L.a(s.o1)) return;
  const int A = L.a(axis);
  auto F = [&](int along) -> double& { return p[slab_at(L, s, along, t0, t1)]; };
  if(periodic) {
    F(0) = F(A);
    F(A + 1) = F(1);
    return;
  }
  for(int side = 0; side < 2; ++side) {
    const of_pressure_bc& c = side == 0 ? lo : hi;
    const bool lower = side == 0;
    const double av = F(lower ? 1 : A);
    double& g = F(lower ? 0 : A + 1);
    if(c.type == OF_PRESSURE_NEUMANN) g = lower ? av - c.gradient * h : av + c.gradient * h;
    else if(c.type == OF_PRESSURE_DIRICHLET) g = 2.0 * c.value - av;
  }
}

// solver/projection.cpp:339-382,986-1026
__global__ void k_total_pressure_bc_axis(double* p, Lay L, const double* src, Lay Ls, int axis, int periodic,
                                         of_bc lo, of_bc hi, double h) {
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  const Slab s = slab_of(axis);
  if(t0 >= L.a(s.o0) || t1 >= L.a(s.o1)) return;
  const int A = L.a(axis);
  auto F = [&](int along) -> double& { return p[slab_at(L, s, along, t0, t1)]; };
  if(periodic) {
    F(0) = F(A);
    F(A + 1) = F(1);
    return;
  }
  for(int side = 0; side < 2; ++side) {
    const of_bc& c = side == 0 ? lo : hi;
    const bool lower = side == 0;
    const double av = F(lower ? 1 : A);
    double& g = F(lower ? 0 : A + 1);
    if(c.type == OF_NO_SLIP_WALL || c.type == OF_PRESCRIBED_VELOCITY || c.type == OF_SYMMETRY) {
      const double grad = src[slab_at(Ls, s, lower ? 1 : Ls.a(axis), t0, t1)];
      g = lower ? av - grad * h : av + grad * h;
    } else if(c.type == OF_FIXED_PRESSURE) {
      g = 2.0 * c.pressure - av;
    }
  }
}

// ------------------------------------------------------------------ MAC operators
// operators/discrete_operators.cpp:77-109 for one component
__global__ void k_gradient(const double* __restrict__ p, Lay Lp, double* __restrict__ g, Lay Lg, int axis, double h,
                           int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > Lg.ax || j > Lg.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + Lg.az);
  const int di = axis == 0, dj = axis == 1, dk = axis == 2;
  for(int k = kb; k < ke; ++k) g[Lg.at(i, j, k)] = (p[Lp.at(i, j, k)] - p[Lp.at(i - di, j - dj, k - dk)]) / h;
}

// operators/discrete_operators.cpp:111-126 (times `scale`, as build_pressure_rhs does when
// scale != 1: solver/projection.cpp:1062-1063)
__global__ void k_divergence(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                             Lay Lu, Lay Lv, Lay Lw, Lay Lc, of_grid g, double scale, int apply_scale,
                             double* __restrict__ d, int kz) {
  const uint3 bk = remapped_block3();
  const int i = 1 + (int)bk.x * blockDim.x + threadIdx.x;
  const int j = 1 + (int)bk.y * blockDim.y + threadIdx.y;
  if(i > Lc.ax || j > Lc.ay) return;
  const int kb = 1 + (int)bk.z * kz, ke = min(kb + kz, 1 + Lc.az);
  for(int k = kb; k < ke; ++k) {
    const double a = (u[Lu.at(i + 1, j, k)] - u[Lu.at(i, j, k)]) / g.dx;
    const double b = (v[Lv.at(i, j + 1, k)] - v[Lv.at(i, j, k)]) / g.dy;
    const double c = (w[Lw.at(i, j, k + 1)] - w[Lw.at(i, j, k)]) / g.dz;
    const double dv = a + b + c;
    d[Lc.at(i, j, k)] = apply_scale ? dv * scale : dv;
  }
}

// operators/discrete_operators.cpp:37-73 (+ the viscosity scaling of compute_diffusion_term,
// solver/momentum_terms.cpp:362-371, when apply_scale)
__device__ inline double lap7(const double* f, const Lay& L, int i, int j, int k, double ix, double iy, double iz) {
  const double c = f[L.at(i, j, k)];
  const double a = (f[L.at(i + 1, j, k)] - 2.0 * c + f[L.at(i - 1, j, k)]) * ix;
  const double b = (f[L.at(i, j + 1, k)] - 2.0 * c + f[L.at(i, j - 1, k)]) * iy;
  const double e = (f[L.at(i, j, k + 1)] - 2.0 * c + f[L.at(i, j, k - 1)]) * iz;
  return a + b + e;
}

__global__ void k_laplacian(const double* __restrict__ in, double* __restrict__ out, Lay L, double ix, double iy,
                            double iz, double scale, int apply_scale, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  for(int k = kb; k < ke; ++k) {
    const double l = lap7(in, L, i, j, k, ix, iy, iz);
    out[L.at(i, j, k)] = apply_scale ? l * scale : l;
  }
}

// ------------------------------------------------------------------ TVD advection
constexpr double kSmall = 1.0e-14;  // solver/momentum_terms.cpp:62

__device__ inline double van_leer(double r) { return (r + fabs(r)) / (1.0 + fabs(r)); }
__device__ inline double safe_ratio(double n, double d) { return fabs(d) < kSmall ? 0.0 : n / d; }

// solver/momentum_terms.cpp:88-197. `f` points at the value at storage index `r` along the
// axis; `stride` is the storage stride of that axis; `max_index` its last storage index.
__device__ inline double reconstruct(const double* f, ptrdiff_t stride, int r, int max_index, double vel,
                                     int scheme) {
  const double left = f[-stride], right = f[0];
  if(scheme == 2) return 0.5 * (left + right);
  if(scheme == 1 || fabs(vel) < kSmall) return vel >= 0.0 ? left : right;
  if(vel >= 0.0) {
    if(r - 2 < 0) return left;
    const double dd = right - left;
    const double ratio = safe_ratio(left - f[-2 * stride], dd);
    return left + 0.5 * van_leer(ratio) * dd;
  }
  if(r + 1 > max_index) return right;
  const double d = right - left;
  const double ratio = safe_ratio(f[stride] - right, d);
  return right - 0.5 * van_leer(ratio) * d;
}

struct VelPtrs {
  const double* u;
  const double* v;
  const double* w;
  Lay Lu, Lv, Lw;
};

// solver/momentum_terms.cpp:199-312: conservative flux divergence of component `c` at
// storage point (i,j,k) of that component's layout.
__device__ inline double advect_point(const VelPtrs& V, int c, int i, int j, int k, const of_grid& g, int scheme) {
  double wv, ev, sv, nv, bv, fv;
  const double* u = V.u;
  const double* v = V.v;
  const double* w = V.w;
  const Lay &Lu = V.Lu, &Lv = V.Lv, &Lw = V.Lw;
  if(c == 0) {
    wv = 0.5 * (u[Lu.at(i - 1, j, k)] + u[Lu.at(i, j, k)]);
    ev = 0.5 * (u[Lu.at(i, j, k)] + u[Lu.at(i + 1, j, k)]);
    sv = 0.5 * (v[Lv.at(i - 1, j, k)] + v[Lv.at(i, j, k)]);
    nv = 0.5 * (v[Lv.at(i - 1, j + 1, k)] + v[Lv.at(i, j + 1, k)]);
    bv = 0.5 * (w[Lw.at(i - 1, j, k)] + w[Lw.at(i, j, k)]);
    fv = 0.5 * (w[Lw.at(i - 1, j, k + 1)] + w[Lw.at(i, j, k + 1)]);
  } else if(c == 1) {
    wv = 0.5 * (u[Lu.at(i, j - 1, k)] + u[Lu.at(i, j, k)]);
    ev = 0.5 * (u[Lu.at(i + 1, j - 1, k)] + u[Lu.at(i + 1, j, k)]);
    sv = 0.5 * (v[Lv.at(i, j - 1, k)] + v[Lv.at(i, j, k)]);
    nv = 0.5 * (v[Lv.at(i, j, k)] + v[Lv.at(i, j + 1, k)]);
    bv = 0.5 * (w[Lw.at(i, j - 1, k)] + w[Lw.at(i, j, k)]);
    fv = 0.5 * (w[Lw.at(i, j - 1, k + 1)] + w[Lw.at(i, j, k + 1)]);
  } else {
    wv = 0.5 * (u[Lu.at(i, j, k - 1)] + u[Lu.at(i, j, k)]);
    ev = 0.5 * (u[Lu.at(i + 1, j, k - 1)] + u[Lu.at(i + 1, j, k)]);
    sv = 0.5 * (v[Lv.at(i, j, k - 1)] + v[Lv.at(i, j, k)]);
    nv = 0.5 * (v[Lv.at(i, j + 1, k - 1)] + v[Lv.at(i, j + 1, k)]);
    bv = 0.5 * (w[Lw.at(i, j, k - 1)] + w[Lw.at(i, j, k)]);
    fv = 0.5 * (w[Lw.at(i, j, k)] + w[Lw.at(i, j, k + 1)]);
  }
  const double* q = c == 0 ? u : c == 1 ? v : w;
  const Lay& L = c == 0 ? Lu : c == 1 ? Lv : Lw;
  const ptrdiff_t sj = L.sx, sk = (ptrdiff_t)L.sx * L.sy;
  const double* p = q + L.at(i, j, k);
  const double wf = wv * reconstruct(p, 1, i, L.sx - 1, wv, scheme);
  const double ef = ev * reconstruct(p + 1, 1, i + 1, L.sx - 1, ev, scheme);
  const double sf = sv * reconstruct(p, sj, j, L.sy - 1, sv, scheme);
  const double nf = nv * reconstruct(p + sj, sj, j + 1, L.sy - 1, nv, scheme);
  const double bf = bv * reconstruct(p, sk, k, L.sz - 1, bv, scheme);
  const double ff = fv * reconstruct(p + sk, sk, k + 1, L.sz - 1, fv, scheme);
  return (ef - wf) / g.dx + (nf - sf) / g.dy + (ff - bf) / g.dz;
}

__global__ void k_advection(VelPtrs V, int c, of_grid g, int scheme, double* __restrict__ out, int kz) {
  const Lay& L = c == 0 ? V.Lu : c == 1 ? V.Lv : V.Lw;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  for(int k = kb; k < ke; ++k) out[L.at(i, j, k)] = advect_point(V, c, i, j, k, g, scheme);
}

// ------------------------------------------------------------------ factorisation correction
// solver/taylor_green.cpp:132-225 (3D/2D) at storage point (i,j,k)
__device__ inline double sd_axis(const double* f, const Lay& L, int i, int j, int k, int axis, double inv) {
  const int di = axis == 0, dj = axis == 1, dk = axis == 2;
  return (f[L.at(i + di, j + dj, k + dk)] - 2.0 * f[L.at(i, j, k)] + f[L.at(i - di, j - dj, k - dk)]) * inv;
}
__device__ inline double mixed_sd(const double* f, const Lay& L, int i, int j, int k, int outer, int inner,
                                  const double* inv) {
  const int di = outer == 0, dj = outer == 1, dk = outer == 2;
  return (sd_axis(f, L, i + di, j + dj, k + dk, inner, inv[inner]) - 2.0 * sd_axis(f, L, i, j, k, inner, inv[inner]) +
          sd_axis(f, L, i - di, j - dj, k - dk, inner, inv[inner])) *
         inv[outer];
}
__device__ inline double fc_point_tg(const double* f, const Lay& L, int i, int j, int k, const double* inv,
                                     bool three_d, double alpha) {
  double corr = mixed_sd(f, L, i, j, k, 0, 1, inv);
  if(three_d) {
    corr += mixed_sd(f, L, i, j, k, 0, 2, inv);
    corr += mixed_sd(f, L, i, j, k, 1, 2, inv);
    const double t = (mixed_sd(f, L, i + 1, j, k, 1, 2, inv) - 2.0 * mixed_sd(f, L, i, j, k, 1, 2, inv) +
                      mixed_sd(f, L, i - 1, j, k, 1, 2, inv)) *
                     inv[0];
    corr -= alpha * t;
  }
  return corr;
}
// solver/lid_driven_cavity.cpp:219-252
__device__ inline double fc_point_cavity(const double* f, const Lay& L, int i, int j, int k, double dy2,
                                         double idx2) {
  const double a = (f[L.at(i - 1, j + 1, k)] - 2.0 * f[L.at(i - 1, j, k)] + f[L.at(i - 1, j - 1, k)]) / dy2;
  const double b = (f[L.at(i, j + 1, k)] - 2.0 * f[L.at(i, j, k)] + f[L.at(i, j - 1, k)]) / dy2;
  const double c = (f[L.at(i + 1, j + 1, k)] - 2.0 * f[L.at(i + 1, j, k)] + f[L.at(i + 1, j - 1, k)]) / dy2;
  return (c - 2.0 * b + a) * idx2;
}

struct FcParams {
  double inv[3];
  double alpha, dy2, idx2;
  int variant, three_d;
};

__device__ inline double fc_point(const double* f, const Lay& L, int i, int j, int k, const FcParams& P) {
  return P.variant == 0 ? fc_point_tg(f, L, i, j, k, P.inv, P.three_d != 0, P.alpha)
                        : fc_point_cavity(f, L, i, j, k, P.dy2, P.idx2);
}

__global__ void k_fc(const double* __restrict__ f, Lay L, FcParams P, double* __restrict__ out, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  for(int k = kb; k < ke; ++k) out[L.at(i, j, k)] = fc_point(f, L, i, j, k, P);
}

// ------------------------------------------------------------------ fused predictor RHS
// solver/taylor_green.cpp:227-256 / solver/lid_driven_cavity.cpp:489-518 for component c:
//   rhs = q - dt*grad p + (dt/2)*nu*lap q - dt*(3/2 A - 1/2 A_prev) + alpha^2 * FC(q)
// evaluated as the reference's chain of axpy updates; A is written to `adv` after A_prev has
// been read from `prev` (they may alias).
struct RhsParams {
  of_grid g;
  int c, scheme, has_prev;
  double ix, iy, iz, nu;
  double idx, idy, idz;  // 1/dx, 1/dy, 1/dz (the tiled predictor multiplies by them)
  double s_grad, s_dif, s_adv1, s_prev, s_adv0, s_fc;
  FcParams fc;
};

__global__ void k_predictor_rhs(VelPtrs V, const double* __restrict__ p, Lay Lp, const double* prev, double* adv,
                                double* __restrict__ rhs, RhsParams P, int kz) {
  const int c = P.c;
  const Lay& L = c == 0 ? V.Lu : c == 1 ? V.Lv : V.Lw;
  const double* q = c == 0 ? V.u : c == 1 ? V.v : V.w;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  const int di = c == 0, dj = c == 1, dk = c == 2;
  const double h = c == 0 ? P.g.dx : c == 1 ? P.g.dy : P.g.dz;
  for(int k = kb; k < ke; ++k) {
    const size_t at = L.at(i, j, k);
    const double a = advect_point(V, c, i, j, k, P.g, P.scheme);
    const double dif = lap7(q, L, i, j, k, P.ix, P.iy, P.iz) * P.nu;
    const double grad = (p[Lp.at(i, j, k)] - p[Lp.at(i - di, j - dj, k - dk)]) / h;
    const double fc = fc_point(q, L, i, j, k, P.fc);
    double r = q[at];
    r += P.s_grad * grad;
    r += P.s_dif * dif;
    if(P.has_prev) {
      const double pa = prev[at];
      r += P.s_adv1 * a;
      r += P.s_prev * pa;
    } else {
      r += P.s_adv0 * a;
    }
    r += P.s_fc * fc;
    adv[at] = a;
    rhs[at] = r;
  }
}

// ------------------------------------------------------------------ tiled predictor RHS
// The same per-point formulas as k_predictor_rhs for the Taylor-Green variant, evaluated on
// 64x8 tiles that march 8 planes along k: the component q is staged in an LDS ring of five
// planes (k-2..k+2, x/y halo 2), and the second differences shared by neighbouring points of
// the factorisation correction (delta_z q, delta_y q, delta_y delta_z q) are computed once
// per plane into LDS. Every value is produced by the same expression as the per-point path
// except two that trade divisions for cheaper forms (the kernel is VALU-bound): the limiter in
// one-division form (reconstruct4_1div) and multiplications by 1/dx, 1/dy, 1/dz; both agree
// with the per-point path to a few ulps. Loads per point drop from ~100 to ~12.
constexpr int kRX = 64, kRY = 8, kRKZ = 8;
constexpr int kRHX = kRX + 4, kRHY = kRY + 4;

// reconstruct() on explicit values: ll, l, r, rr are q at indices r_index-2 .. r_index+1
// (ll / rr enter only when those indices lie in storage). Branch-free: both upwind
// directions share one ratio and one limiter evaluation, selected by the sign of vel, so
// waves whose face velocities change sign do not run both paths; each selected value is the
// same expression as in reconstruct().
__device__ inline double reconstruct4(double ll, double left, double right, double rr, int r_index, int max_index,
                                      double vel, int scheme) {
  if(scheme == 2) return 0.5 * (left + right);
  const bool up = vel >= 0.0;
  const double upwind = up ? left : right;
  if(scheme == 1) return upwind;
  const double d = right - left;
  const double ratio = safe_ratio(up ? left - ll : rr - right, d);
  const double lim = 0.5 * van_leer(ratio);
  const double tvd = up ? left + lim * d : right - lim * d;
  const bool plain = fabs(vel) < kSmall || (up ? r_index - 2 < 0 : r_index + 1 > max_index);
  return plain ? upwind : tvd;
}

// reconstruct4 with the van Leer limiter in one-division form, for the tiled predictor: with
// r = n / d the limited slope is psi(r) d / 2 = r d / (1 + r) = n d / (n + d) for r > 0, and 0
// for r <= 0 or |d| < kSmall (safe_ratio's zero). One division per face instead of two; the
// values equal the two-division form up to a few ulps.
__device__ inline double reconstruct4_1div(double ll, double left, double right, double rr, int r_index,
                                           int max_index, double vel, int scheme) {
  if(scheme == 2) return 0.5 * (left + right);
  const bool up = vel >= 0.0;
  const double upwind = up ? left : right;
  if(scheme == 1) return upwind;
  const double d = right - left;
  const double n = up ? left - ll : rr - right;
  const bool pos = fabs(d) >= kSmall && n * d > 0.0;
  const double half_psi_d = pos ? n * d / (n + d) : 0.0;
  const double tvd = up ? left + half_psi_d : right - half_psi_d;
  const bool plain = fabs(vel) < kSmall || (up ? r_index - 2 < 0 : r_index + 1 > max_index);
  return plain ? upwind : tvd;
}

// The van Leer faces of one point in two halves around one shared reciprocal. lim_prep forms the
// operands of reconstruct4_1div's limited slope n d / (n + d); the reciprocals of the faces'
// denominators come from one division of their product (each 1/q_f = the product of the others
// times 1/prod: Montgomery's batch inversion), and lim_finish selects the face value exactly as
// reconstruct4_1div does. Five divisions become one and twelve multiplications; the slope
// differs from the divided form by a few ulps. A product outside [1e-280, 1e280] (denominators
// are at least kSmall where the limiter is active, 1 elsewhere) falls back to one division per
// face.
struct LimFace {
  double upwind, nd, den;
  bool up, pos, plain;
};
__device__ inline LimFace lim_prep(double ll, double left, double right, double rr, int r_index, int max_index,
                                   double vel) {
  LimFace f;
  f.up = vel >= 0.0;
  f.upwind = f.up ? left : right;
  const double d = right - left;
  const double n = f.up ? left - ll : rr - right;
  f.pos = fabs(d) >= kSmall && n * d > 0.0;
  f.nd = n * d;
  f.den = f.pos ? n + d : 1.0;
  f.plain = fabs(vel) < kSmall || (f.up ? r_index - 2 < 0 : r_index + 1 > max_index);
  return f;
}
__device__ inline double lim_finish(const LimFace& f, double rinv) {
  const double half_psi_d = f.pos ? f.nd * rinv : 0.0;
  const double tvd = f.up ? f.upwind + half_psi_d : f.upwind - half_psi_d;
  return f.plain ? f.upwind : tvd;
}
// reciprocals of five denominators
__device__ inline void batch_reciprocal5(const double (&q)[5], double (&r)[5]) {
  const double p1 = q[0] * q[1], p2 = p1 * q[2], p3 = p2 * q[3], p4 = p3 * q[4];
  if(fabs(p4) > 1e-280 && fabs(p4) < 1e280) {
    double inv = 1.0 / p4;
    r[4] = inv * p3;
    inv = inv * q[4];
    r[3] = inv * p2;
    inv = inv * q[3];
    r[2] = inv * p1;
    inv = inv * q[2];
    r[1] = inv * q[0];
    r[0] = inv * q[1];
  } else {
#pragma unroll
    for(int f = 0; f < 5; ++f) r[f] = 1.0 / q[f];
  }
}

// Loads are issued well before their use: the q plane k+3 is fetched into registers while
// plane k is computed (and committed to the LDS ring at the top of the next iteration), and
// the point values of plane k (the other two components, p, A_prev) are requested before the
// plane's LDS phases. Every load is unconditional at clamped indices (threads outside the
// active range read valid storage and store nothing); out-of-storage halo values are masked
// to zero afterwards, so no load waits inside a divergent branch.
__global__ void __launch_bounds__(kRX * kRY) k_predictor_rhs_tiled(VelPtrs V, const double* __restrict__ p, Lay Lp,
                                                                   const double* prev, double* adv,
                                                                   double* __restrict__ rhs, RhsParams P) {
  // q ring of six planes (k - 2 .. k + 2 in use while plane k + 3 is committed) and the second
  // differences double-buffered by plane parity: with these, a plane needs two barriers (after
  // its commit, after its second differences) and none at its end
  __shared__ double sq[6][kRHY][kRHX];
  __shared__ double ssz_[2][kRY + 2][kRX + 2];
  __shared__ double ssy_[2][kRY][kRX + 2];
  __shared__ double sm12_[2][kRY][kRX + 2];
  const int c = P.c;
  const Lay& L = c == 0 ? V.Lu : c == 1 ? V.Lv : V.Lw;
  const double* q = c == 0 ? V.u : c == 1 ? V.v : V.w;
  const int gx = (L.ax + kRX - 1) / kRX, gy = (L.ay + kRY - 1) / kRY;
  unsigned rb = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = (int)(rb % (unsigned)gx);
  rb /= (unsigned)gx;
  const int by = (int)(rb % (unsigned)gy);
  const int bz = (int)(rb / (unsigned)gy);
  const int i0 = 1 + bx * kRX, j0 = 1 + by * kRY, kb = 1 + bz * kRKZ, ke = min(kb + kRKZ, 1 + L.az);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + kRX * ty;
  const int i = i0 + tx, j = j0 + ty;
  const bool in = i <= L.ax && j <= L.ay;
  const int ic = min(i, L.ax), jc = min(j, L.ay);  // load coordinates
  const double inv0 = P.fc.inv[0], inv1 = P.fc.inv[1], inv2 = P.fc.inv[2];
  const bool three_d = P.fc.three_d != 0;
  auto slot = [](int kk) { return (kk + 12) % 6; };
  // staged plane elements of this thread: e = tid + m * threads
  constexpr int kPlaneElems = kRHY * kRHX;
  constexpr int kPE = (kPlaneElems + kRX * kRY - 1) / (kRX * kRY);
  const size_t plane = (size_t)L.sx * L.sy;
  size_t qoff[kPE];
  bool okxy[kPE];
#pragma unroll
  for(int m = 0; m < kPE; ++m) {
    const int e = min(tid + m * kRX * kRY, kPlaneElems - 1);
    const int hy = e / kRHX, hx = e - hy * kRHX;
    const int gi = i0 - 2 + hx, gj = j0 - 2 + hy;
    okxy[m] = gi >= 0 && gi < L.sx && gj >= 0 && gj < L.sy;
    qoff[m] = L.at(min(max(gi, 0), L.sx - 1), min(max(gj, 0), L.sy - 1), 0);
  }
  double qraw[kPE];
  bool qok[kPE];
  auto fetch = [&](int kk) {
    const bool okz = kk >= 0 && kk < L.sz;
    const size_t kofs = plane * (size_t)min(max(kk, 0), L.sz - 1);
#pragma unroll
    for(int m = 0; m < kPE; ++m) {
      qraw[m] = q[qoff[m] + kofs];
      qok[m] = okz && okxy[m];
    }
  };
  auto commit = [&](int kk) {
    double* dst = &sq[slot(kk)][0][0];
#pragma unroll
    for(int m = 0; m < kPE; ++m) {
      const int e = tid + m * kRX * kRY;
      if(e < kPlaneElems) dst[e] = qok[m] ? qraw[m] : 0.0;
    }
  };
  const int di = c == 0, dj = c == 1, dk = c == 2;
  const double ih = c == 0 ? P.idx : c == 1 ? P.idy : P.idz;  // reciprocal spacings
  const double* u = V.u;
  const double* v = V.v;
  const double* w = V.w;
  const Lay &Lu = V.Lu, &Lv = V.Lv, &Lw = V.Lw;
  // point values of plane k: the two face-velocity pairs of the other components (a0..a3,
  // b0..b3, in the order of advect_point), p at the two cells of the gradient, A_prev. Each is
  // read at a per-thread offset within its plane plus (k + dz) planes, a uniform product (the
  // 64-bit index products of Lay::at per point were a large share of the kernel's VALU work).
  const double* A = c == 0 ? v : u;  // a0..a3
  const double* B = c == 2 ? v : w;  // b0..b3
  const Lay& LA = c == 0 ? Lv : Lu;
  const Lay& LB = c == 2 ? Lv : Lw;
  const size_t plA = (size_t)LA.sx * LA.sy, plB = (size_t)LB.sx * LB.sy, plP = (size_t)Lp.sx * Lp.sy;
  size_t oa[4], ob[4];
  int dza[4] = {0, 0, 0, 0}, dzb[4] = {0, 0, 1, 1};
  if(c == 0) {
    oa[0] = LA.at(ic - 1, jc, 0), oa[1] = LA.at(ic, jc, 0), oa[2] = LA.at(ic - 1, jc + 1, 0), oa[3] = LA.at(ic, jc + 1, 0);
    ob[0] = LB.at(ic - 1, jc, 0), ob[1] = LB.at(ic, jc, 0), ob[2] = LB.at(ic - 1, jc, 0), ob[3] = LB.at(ic, jc, 0);
  } else if(c == 1) {
    oa[0] = LA.at(ic, jc - 1, 0), oa[1] = LA.at(ic, jc, 0), oa[2] = LA.at(ic + 1, jc - 1, 0), oa[3] = LA.at(ic + 1, jc, 0);
    ob[0] = LB.at(ic, jc - 1, 0), ob[1] = LB.at(ic, jc, 0), ob[2] = LB.at(ic, jc - 1, 0), ob[3] = LB.at(ic, jc, 0);
  } else {
    oa[0] = LA.at(ic, jc, 0), oa[1] = LA.at(ic, jc, 0), oa[2] = LA.at(ic + 1, jc, 0), oa[3] = LA.at(ic + 1, jc, 0);
    ob[0] = LB.at(ic, jc, 0), ob[1] = LB.at(ic, jc, 0), ob[2] = LB.at(ic, jc + 1, 0), ob[3] = LB.at(ic, jc + 1, 0);
    dza[0] = dza[2] = dzb[0] = dzb[2] = -1;
    dzb[1] = dzb[3] = 0;
  }
  const size_t op0 = Lp.at(ic, jc, 0), op1 = Lp.at(ic - di, jc - dj, 0), oq = L.at(ic, jc, 0);
  const size_t ost = L.at(min(i, L.sx - 1), min(j, L.sy - 1), 0);  // store offset (threads in range)
  double a0, a1, a2, a3, b0, b1, b2, b3, p0, p1, pa;
  auto point_loads = [&](int k) {
    a0 = A[oa[0] + plA * (size_t)(k + dza[0])];
    a1 = A[oa[1] + plA * (size_t)(k + dza[1])];
    a2 = A[oa[2] + plA * (size_t)(k + dza[2])];
    a3 = A[oa[3] + plA * (size_t)(k + dza[3])];
    b0 = B[ob[0] + plB * (size_t)(k + dzb[0])];
    b1 = B[ob[1] + plB * (size_t)(k + dzb[1])];
    b2 = B[ob[2] + plB * (size_t)(k + dzb[2])];
    b3 = B[ob[3] + plB * (size_t)(k + dzb[3])];
    p0 = p[op0 + plP * (size_t)k];
    p1 = p[op1 + plP * (size_t)(k - dk)];
    pa = P.has_prev ? prev[oq + plane * (size_t)k] : 0.0;
  };
  // the per-plane LDS phases: each thread's elements of the (kRY + 2) x (kRX + 2) and
  // kRY x (kRX + 2) patches (at most two each), decomposed once
  constexpr int kE1 = (kRY + 2) * (kRX + 2), kE2 = kRY * (kRX + 2), kThreads = kRX * kRY;
  static_assert(kE1 <= 2 * kThreads && kE2 <= 2 * kThreads, "two patch elements per thread");
  const int e1b = tid + kThreads, e2b = tid + kThreads;
  const int y1a = tid / (kRX + 2), x1a = tid - y1a * (kRX + 2), y1b = e1b / (kRX + 2), x1b = e1b - y1b * (kRX + 2);
  const int y2a = y1a, x2a = x1a, y2b = y1b, x2b = x1b;  // the same decomposition (row length kRX + 2)
  const bool has1b = e1b < kE1, has2a = tid < kE2, has2b = e2b < kE2;
  for(int kk = kb - 2; kk < kb + 2; ++kk) {
    fetch(kk);
    commit(kk);
  }
  fetch(kb + 2);
  double ff_prev = 0.0;  // upper-face flux of the previous plane
  for(int k = kb; k < ke; ++k) {
    commit(k + 2);
    point_loads(k);
    if(k + 1 < ke) fetch(k + 3);
    __syncthreads();
    const int s0 = slot(k), sp = slot(k + 1), sm = slot(k - 1);
    auto (&ssz)[kRY + 2][kRX + 2] = ssz_[k & 1];
    auto (&ssy)[kRY][kRX + 2] = ssy_[k & 1];
    auto (&sm12)[kRY][kRX + 2] = sm12_[k & 1];
    // delta_z q over x/y halo 1, delta_y q over x halo 1 (sd_axis), and delta_y delta_z q over
    // x halo 1 (mixed_sd with outer = y, inner = z) from delta_z q values formed here by the
    // same expression as ssz's, so no barrier separates it from them
    auto dz = [&](int hy, int hx) {
      return (sq[sp][hy + 1][hx + 1] - 2.0 * sq[s0][hy + 1][hx + 1] + sq[sm][hy + 1][hx + 1]) * inv2;
    };
    auto dy_at = [&](int hy, int hx) {
      ssy[hy][hx] = (sq[s0][hy + 3][hx + 1] - 2.0 * sq[s0][hy + 2][hx + 1] + sq[s0][hy + 1][hx + 1]) * inv1;
    };
    auto dyz_at = [&](int hy, int hx) {
      sm12[hy][hx] = (dz(hy + 2, hx) - 2.0 * dz(hy + 1, hx) + dz(hy, hx)) * inv1;
    };
    ssz[y1a][x1a] = dz(y1a, x1a);
    if(has1b) ssz[y1b][x1b] = dz(y1b, x1b);
    if(has2a) {
      dy_at(y2a, x2a);
      dyz_at(y2a, x2a);
    }
    if(has2b) {
      dy_at(y2b, x2b);
      dyz_at(y2b, x2b);
    }
    __syncthreads();
    if(in) {
      const int cy = ty + 2, cx = tx + 2;
      auto Q = [&](int s, int dy, int dx) { return sq[s][cy + dy][cx + dx]; };
      // factorisation correction (fc_point_tg)
      double corr = (ssy[ty][tx + 2] - 2.0 * ssy[ty][tx + 1] + ssy[ty][tx]) * inv0;
      if(three_d) {
        corr += (ssz[ty + 1][tx + 2] - 2.0 * ssz[ty + 1][tx + 1] + ssz[ty + 1][tx]) * inv0;
        corr += sm12[ty][tx + 1];
        const double t = (sm12[ty][tx + 2] - 2.0 * sm12[ty][tx + 1] + sm12[ty][tx]) * inv0;
        corr -= P.fc.alpha * t;
      }
      // diffusion (lap7)
      const double qc = Q(s0, 0, 0);
      const double la = (Q(s0, 0, 1) - 2.0 * qc + Q(s0, 0, -1)) * P.ix;
      const double lb = (Q(s0, 1, 0) - 2.0 * qc + Q(s0, -1, 0)) * P.iy;
      const double le = (Q(sp, 0, 0) - 2.0 * qc + Q(sm, 0, 0)) * P.iz;
      const double dif = (la + lb + le) * P.nu;
      // advection (advect_point): face velocities
      double wv, ev, sv, nv, bv, fv;
      if(c == 0) {
        wv = 0.5 * (Q(s0, 0, -1) + qc);
        ev = 0.5 * (qc + Q(s0, 0, 1));
        sv = 0.5 * (a0 + a1);
        nv = 0.5 * (a2 + a3);
        bv = 0.5 * (b0 + b1);
        fv = 0.5 * (b2 + b3);
      } else if(c == 1) {
        wv = 0.5 * (a0 + a1);
        ev = 0.5 * (a2 + a3);
        sv = 0.5 * (Q(s0, -1, 0) + qc);
        nv = 0.5 * (qc + Q(s0, 1, 0));
        bv = 0.5 * (b0 + b1);
        fv = 0.5 * (b2 + b3);
      } else {
        wv = 0.5 * (a0 + a1);
        ev = 0.5 * (a2 + a3);
        sv = 0.5 * (b0 + b1);
        nv = 0.5 * (b2 + b3);
        bv = 0.5 * (Q(sm, 0, 0) + qc);
        fv = 0.5 * (qc + Q(sp, 0, 0));
      }
      const int s2m = slot(k - 2), s2p = slot(k + 2);
      const int mx = L.sx - 1, my = L.sy - 1, mz = L.sz - 1;
      double wf, ef, sf, nf, ff;
      if(P.scheme == 0) {  // van Leer: the five faces share one division
        const LimFace fw = lim_prep(Q(s0, 0, -2), Q(s0, 0, -1), qc, Q(s0, 0, 1), i, mx, wv);
        const LimFace fe = lim_prep(Q(s0, 0, -1), qc, Q(s0, 0, 1), Q(s0, 0, 2), i + 1, mx, ev);
        const LimFace fs = lim_prep(Q(s0, -2, 0), Q(s0, -1, 0), qc, Q(s0, 1, 0), j, my, sv);
        const LimFace fn = lim_prep(Q(s0, -1, 0), qc, Q(s0, 1, 0), Q(s0, 2, 0), j + 1, my, nv);
        const LimFace ft = lim_prep(Q(sm, 0, 0), qc, Q(sp, 0, 0), Q(s2p, 0, 0), k + 1, mz, fv);
        const double q[5] = {fw.den, fe.den, fs.den, fn.den, ft.den};
        double r[5];
        batch_reciprocal5(q, r);
        wf = wv * lim_finish(fw, r[0]);
        ef = ev * lim_finish(fe, r[1]);
        sf = sv * lim_finish(fs, r[2]);
        nf = nv * lim_finish(fn, r[3]);
        ff = fv * lim_finish(ft, r[4]);
      } else {
        wf = wv * reconstruct4_1div(Q(s0, 0, -2), Q(s0, 0, -1), qc, Q(s0, 0, 1), i, mx, wv, P.scheme);
        ef = ev * reconstruct4_1div(Q(s0, 0, -1), qc, Q(s0, 0, 1), Q(s0, 0, 2), i + 1, mx, ev, P.scheme);
        sf = sv * reconstruct4_1div(Q(s0, -2, 0), Q(s0, -1, 0), qc, Q(s0, 1, 0), j, my, sv, P.scheme);
        nf = nv * reconstruct4_1div(Q(s0, -1, 0), qc, Q(s0, 1, 0), Q(s0, 2, 0), j + 1, my, nv, P.scheme);
        ff = fv * reconstruct4_1div(Q(sm, 0, 0), qc, Q(sp, 0, 0), Q(s2p, 0, 0), k + 1, mz, fv, P.scheme);
      }
      // the lower face of plane k is the upper face of plane k - 1 with the same operands (values,
      // face velocity, index), so its flux is carried over from the previous plane
      const double bf =
          k > kb ? ff_prev : bv * reconstruct4_1div(Q(s2m, 0, 0), Q(sm, 0, 0), qc, Q(sp, 0, 0), k, mz, bv, P.scheme);
      ff_prev = ff;
      const double a = (ef - wf) * P.idx + (nf - sf) * P.idy + (ff - bf) * P.idz;
      const double grad = (p0 - p1) * ih;
      const size_t at = ost + plane * (size_t)k;
      double r = qc;
      r += P.s_grad * grad;
      r += P.s_dif * dif;
      if(P.has_prev) {
        r += P.s_adv1 * a;
        r += P.s_prev * pa;
      } else {
        r += P.s_adv0 * a;
      }
      r += P.s_fc * corr;
      adv[at] = a;
      rhs[at] = r;
    }
  }
}

// ------------------------------------------------------------------ CFL diagnostic
// solver/momentum_terms.cpp:373-401
__global__ void k_cfl(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w, Lay Lu,
                      Lay Lv, Lay Lw, of_grid g, double dt, double* partials, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  double vals[4] = {0.0, 0.0, 0.0, 0.0};
  if(i <= g.nx && j <= g.ny) {
    const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + g.nz);
    for(int k = kb; k < ke; ++k) {
      const double uc = 0.5 * (u[Lu.at(i, j, k)] + u[Lu.at(i + 1, j, k)]);
      const double vc = 0.5 * (v[Lv.at(i, j, k)] + v[Lv.at(i, j + 1, k)]);
      const double wc = 0.5 * (w[Lw.at(i, j, k)] + w[Lw.at(i, j, k + 1)]);
      vals[1] = fmax(vals[1], fabs(uc));
      vals[2] = fmax(vals[2], fabs(vc));
      vals[3] = fmax(vals[3], fabs(wc));
      const double local = dt * (fabs(uc) / g.dx + fabs(vc) / g.dy + fabs(wc) / g.dz);
      vals[0] = fmax(vals[0], local);
    }
  }
  const int ops[4] = {kMax, kMax, kMax, kMax};
  store_partials<4>(vals, ops, partials);
}

// out = {dt, max cfl, max |uc|, max |vc|, max |wc|}
struct EpiCfl {
  static constexpr int NV = 4;
  double* out;
  double dt;
  __device__ int op(int) const { return kMax; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&r)[4]) const {
    out[0] = dt;
    for(int k = 0; k < 4; ++k) out[1 + k] = r[k];
  }
};

// sum of squares and max |x| over the active box
__global__ void k_norms(const double* __restrict__ x, Lay L, double* partials, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  double vals[2] = {0.0, 0.0};
  if(i <= L.ax && j <= L.ay) {
    const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
    for(int k = kb; k < ke; ++k) {
      const double a = x[L.at(i, j, k)];
      vals[0] += a * a;
      vals[1] = fmax(vals[1], fabs(a));
    }
  }
  const int ops[2] = {kSum, kMax};
  store_partials<2>(vals, ops, partials);
}

__global__ void k_scale_active(const double* in, double* out, Lay L, double s, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  for(int k = kb; k < ke; ++k) out[L.at(i, j, k)] = in[L.at(i, j, k)] * s;
}

__global__ void k_axpy_active(const double* __restrict__ x, double* __restrict__ y, Lay L, double a, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  for(int k = kb; k < ke; ++k) y[L.at(i, j, k)] = y[L.at(i, j, k)] + a * x[L.at(i, j, k)];
}

// solver/projection.cpp:1082-1093: out = pred + scale * (phi(i) - phi(i-1))/h on every
// active face (scale = -dt/density); phi carries ghosts.
__global__ void k_correct(const double* __restrict__ pred, Lay L, const double* __restrict__ phi, Lay Lp, int axis,
                          double h, double scale, double* __restrict__ out, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  const int di = axis == 0, dj = axis == 1, dk = axis == 2;
  for(int k = kb; k < ke; ++k) {
    const double gr = (phi[Lp.at(i, j, k)] - phi[Lp.at(i - di, j - dj, k - dk)]) / h;
    out[L.at(i, j, k)] = pred[L.at(i, j, k)] + scale * gr;
  }
}

dim3 slab_grid(int e0, int e1, dim3 block) {
  return dim3((unsigned)((e0 + block.x - 1) / block.x), (unsigned)((e1 + block.y - 1) / block.y), 1);
}

void slab_extents(const of_grid& g, int axis, int& e0, int& e1) {
  // largest other-axis extents over the three face layouts (each is at most n+1)
  const int n[3] = {g.nx, g.ny, g.nz};
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  e0 = n[o0] + 1;
  e1 = n[o1] + 1;
}

VelPtrs vel_ptrs(const of_grid& g, const double* u, const double* v, const double* w) {
  VelPtrs V;
  V.u = u;
  V.v = v;
  V.w = w;
  V.Lu = make_lay(g, OF_FACE_X);
  V.Lv = make_lay(g, OF_FACE_Y);
  V.Lw = make_lay(g, OF_FACE_Z);
  return V;
}

FcParams fc_params(const of_grid& g, int variant, double alpha) {
  FcParams P;
  P.inv[0] = 1.0 / (g.dx * g.dx);
  P.inv[1] = 1.0 / (g.dy * g.dy);
  P.inv[2] = 1.0 / (g.dz * g.dz);
  P.alpha = alpha;
  P.dy2 = g.dy * g.dy;
  P.idx2 = 1.0 / (g.dx * g.dx);
  P.variant = variant;
  P.three_d = g.nz != 1;
  return P;
}

}  // namespace

bool axis_periodic(const of_bc_set& b, int axis) {
  const bool lo = b.faces[2 * axis].type == OF_PERIODIC, hi = b.faces[2 * axis + 1].type == OF_PERIODIC;
  if(lo != hi) throw InvalidArgument("periodic velocity boundaries must be paired on both sides");
  return lo;
}

bool pressure_axis_periodic(const of_pressure_bc_set& b, int axis) {
  const bool lo = b.faces[2 * axis].type == OF_PRESSURE_PERIODIC, hi = b.faces[2 * axis + 1].type == OF_PRESSURE_PERIODIC;
  if(lo != hi) throw InvalidArgument("periodic pressure boundaries must be paired on both sides");
  return lo;
}

void launch_component_bc_mask(Context& c, const of_grid& g, const of_bc_set& b, double* u, double* v, double* w,
                              int mask) {
  const Lay Lu = make_lay(g, OF_FACE_X), Lv = make_lay(g, OF_FACE_Y), Lw = make_lay(g, OF_FACE_Z);
  const dim3 block(64, 4, 1);
  for(int axis = 0; axis < 3; ++axis) {
    int e0, e1;
    slab_extents(g, axis, e0, e1);
    const int per = axis_periodic(b, axis) ? 1 : 0;
    OFH_LAUNCH(k_velocity_bc_axis, slab_grid(e0, e1, block), block, 0, c.stream, u, v, w, Lu, Lv, Lw, axis,
                       per, b.faces[2 * axis], b.faces[2 * axis + 1], mask);
  }
  OFH_CHECK(hipGetLastError());
}

void launch_velocity_bc(Context& c, const of_grid& g, const of_bc_set& b, double* u, double* v, double* w) {
  launch_component_bc_mask(c, g, b, u, v, w, 7);
}

void launch_component_bc(Context& c, const of_grid& g, const of_bc_set& b, int comp, double* f) {
  launch_component_bc_mask(c, g, b, comp == 0 ? f : nullptr, comp == 1 ? f : nullptr, comp == 2 ? f : nullptr,
                           1 << comp);
}

void launch_pressure_bc(Context& c, const of_grid& g, const of_pressure_bc_set& b, double* p) {
  const Lay L = make_lay(g, OF_CELL);
  const dim3 block(64, 4, 1);
  for(int axis = 0; axis < 3; ++axis) {
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    const int per = pressure_axis_periodic(b, axis) ? 1 : 0;
    OFH_LAUNCH(k_pressure_bc_axis, slab_grid(L.a(o0), L.a(o1), block), block, 0, c.stream, p, L, axis, per,
                       b.faces[2 * axis], b.faces[2 * axis + 1], spacing(g, axis));
  }
  OFH_CHECK(hipGetLastError());
}

void launch_total_pressure_bc(Context& c, const of_grid& g, const of_bc_set& b, const double* su, const double* sv,
                              const double* sw, double* p) {
  const Lay L = make_lay(g, OF_CELL);
  const dim3 block(64, 4, 1);
  for(int axis = 0; axis < 3; ++axis) {
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    const int per = axis_periodic(b, axis) ? 1 : 0;
    const double* src = axis == 0 ? su : axis == 1 ? sv : sw;
    OFH_LAUNCH(k_total_pressure_bc_axis, slab_grid(L.a(o0), L.a(o1), block), block, 0, c.stream, p, L, src,
                       make_lay(g, OF_FACE_X + axis), axis, per, b.faces[2 * axis], b.faces[2 * axis + 1],
                       spacing(g, axis));
  }
  OFH_CHECK(hipGetLastError());
}

void launch_gradient(Context& c, const of_grid& g, const double* p, double* gx, double* gy, double* gz) {
  const Lay Lp = make_lay(g, OF_CELL);
  double* outs[3] = {gx, gy, gz};
  for(int a = 0; a < 3; ++a) {
    const Lay L = make_lay(g, OF_FACE_X + a);
    const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
    OFH_LAUNCH(k_gradient, cl.grid, cl.block, 0, c.stream, p, Lp, outs[a], L, a, spacing(g, a), cl.kz);
  }
  OFH_CHECK(hipGetLastError());
}

void launch_divergence(Context& c, const of_grid& g, const double* u, const double* v, const double* w, double* d) {
  const Lay Lc = make_lay(g, OF_CELL);
  const ColumnLaunch cl = column_launch(Lc.ax, Lc.ay, Lc.az);
  {
    ProfScope prof_(c, "k_divergence", true);
    OFH_LAUNCH(k_divergence, cl.grid, cl.block, 0, c.stream, u, v, w, make_lay(g, OF_FACE_X),
                       make_lay(g, OF_FACE_Y), make_lay(g, OF_FACE_Z), Lc, g, 1.0, 0, d, cl.kz);
  }
  OFH_CHECK(hipGetLastError());
}

void launch_divergence_scaled(Context& c, const of_grid& g, const double* u, const double* v, const double* w,
                              double scale, double* d) {
  const Lay Lc = make_lay(g, OF_CELL);
  const ColumnLaunch cl = column_launch(Lc.ax, Lc.ay, Lc.az);
  {
    ProfScope prof_(c, "k_divergence", true);
    OFH_LAUNCH(k_divergence, cl.grid, cl.block, 0, c.stream, u, v, w, make_lay(g, OF_FACE_X),
                       make_lay(g, OF_FACE_Y), make_lay(g, OF_FACE_Z), Lc, g, scale, 1, d, cl.kz);
  }
  OFH_CHECK(hipGetLastError());
}

void launch_laplacian(Context& c, const of_grid& g, int loc, double scale, const double* in, double* out) {
  const Lay L = make_lay(g, loc);
  const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
  OFH_LAUNCH(k_laplacian, cl.grid, cl.block, 0, c.stream, in, out, L, 1.0 / (g.dx * g.dx),
                     1.0 / (g.dy * g.dy), 1.0 / (g.dz * g.dz), scale, scale != 1.0 ? 1 : 0, cl.kz);
  OFH_CHECK(hipGetLastError());
}

void launch_advection(Context& c, const of_grid& g, int scheme, const double* u, const double* v, const double* w,
                      double* au, double* av, double* aw) {
  const VelPtrs V = vel_ptrs(g, u, v, w);
  double* outs[3] = {au, av, aw};
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
    OFH_LAUNCH(k_advection, cl.grid, cl.block, 0, c.stream, V, comp, g, scheme, outs[comp], cl.kz);
  }
  OFH_CHECK(hipGetLastError());
}

void launch_cfl(Context& c, const of_grid& g, const double* u, const double* v, const double* w, double dt,
                double* out5) {
  const ColumnLaunch cl = column_launch(g.nx, g.ny, g.nz);
  OFH_LAUNCH(k_cfl, cl.grid, cl.block, 0, c.stream, u, v, w, make_lay(g, OF_FACE_X), make_lay(g, OF_FACE_Y),
                     make_lay(g, OF_FACE_Z), g, dt, c.partials, cl.kz);
  launch_finalize(c.stream, c.partials, cl.grid, EpiCfl{out5, dt});
  OFH_CHECK(hipGetLastError());
}

void launch_norms(Context& c, const Lay& L, const double* x, int slot) {
  const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
  OFH_LAUNCH(k_norms, cl.grid, cl.block, 0, c.stream, x, L, c.partials, cl.kz);
  launch_finalize(c.stream, c.partials, cl.grid, EpiStore<2>{c.scal + slot, 2u});
  OFH_CHECK(hipGetLastError());
}

void launch_scale_active(Context& c, const Lay& L, double scale, const double* in, double* out) {
  const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
  OFH_LAUNCH(k_scale_active, cl.grid, cl.block, 0, c.stream, in, out, L, scale, cl.kz);
  OFH_CHECK(hipGetLastError());
}

void launch_axpy_active(Context& c, const Lay& L, double a, const double* x, double* y) {
  const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
  OFH_LAUNCH(k_axpy_active, cl.grid, cl.block, 0, c.stream, x, y, L, a, cl.kz);
  OFH_CHECK(hipGetLastError());
}

void launch_factorized_correction(Context& c, const of_grid& g, int variant, double alpha, const double* u,
                                  const double* v, const double* w, double* fu, double* fv, double* fw) {
  const double* ins[3] = {u, v, w};
  double* outs[3] = {fu, fv, fw};
  const FcParams P = fc_params(g, variant, alpha);
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    OFH_CHECK(hipMemsetAsync(outs[comp], 0, L.size() * sizeof(double), c.stream));
    const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
    OFH_LAUNCH(k_fc, cl.grid, cl.block, 0, c.stream, ins[comp], L, P, outs[comp], cl.kz);
  }
  OFH_CHECK(hipGetLastError());
}

void launch_predictor_rhs(Context& c, const of_grid& g, int variant, int scheme, double nu, double dt,
                          const double* u, const double* v, const double* w, const double* p, const double* pu,
                          const double* pv, const double* pw, double* au, double* av, double* aw, double* ru,
                          double* rv, double* rw) {
  const VelPtrs V = vel_ptrs(g, u, v, w);
  const double alpha = 0.5 * nu * dt;
  RhsParams P;
  P.g = g;
  P.scheme = scheme;
  P.has_prev = pu != nullptr;
  P.ix = 1.0 / (g.dx * g.dx);
  P.iy = 1.0 / (g.dy * g.dy);
  P.iz = 1.0 / (g.dz * g.dz);
  P.idx = 1.0 / g.dx;
  P.idy = 1.0 / g.dy;
  P.idz = 1.0 / g.dz;
  P.nu = nu;
  P.s_grad = -dt;
  P.s_dif = 0.5 * dt;
  P.s_adv1 = -1.5 * dt;
  P.s_prev = 0.5 * dt;
  P.s_adv0 = -dt;
  P.s_fc = alpha * alpha;
  P.fc = fc_params(g, variant, alpha);
  const double* prev[3] = {pu, pv, pw};
  double* adv[3] = {au, av, aw};
  double* rhs[3] = {ru, rv, rw};
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
    P.c = comp;
    ProfScope prof_(c, "k_predictor_rhs", true);
    if(variant == 0) {
      const unsigned gx = div_up(L.ax, kRX), gy = div_up(L.ay, kRY), gz = div_up(L.az, kRKZ);
      OFH_LAUNCH(k_predictor_rhs_tiled, dim3(gx * gy * gz), dim3(kRX, kRY), 0, c.stream, V, p,
                         make_lay(g, OF_CELL), prev[comp], adv[comp], rhs[comp], P);
    } else {
      OFH_LAUNCH(k_predictor_rhs, cl.grid, cl.block, 0, c.stream, V, p, make_lay(g, OF_CELL), prev[comp],
                         adv[comp], rhs[comp], P, cl.kz);
    }
  }
  OFH_CHECK(hipGetLastError());
}

void launch_correct(Context& c, const of_grid& g, const double* u, const double* v, const double* w,
                    const double* phi, double scale, double* ou, double* ov, double* ow) {
  const Lay Lp = make_lay(g, OF_CELL);
  const double* ins[3] = {u, v, w};
  double* outs[3] = {ou, ov, ow};
  for(int a = 0; a < 3; ++a) {
    const Lay L = make_lay(g, OF_FACE_X + a);
    const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
    {
      ProfScope prof_(c, "k_correct", true);
      OFH_LAUNCH(k_correct, cl.grid, cl.block, 0, c.stream, ins[a], L, phi, Lp, a, spacing(g, a), scale,
                         outs[a], cl.kz);
    }
  }
  OFH_CHECK(hipGetLastError());
}

}  // namespace ofh
