// Helmholtz predictor: (1 - a dxx)(1 - a dyy)(1 - a dzz) u* = rhs per velocity component,
// solved exactly as three batched sweeps of tridiagonal line solves
// (solver/projection.cpp:464-888,1028-1047).
//
// The line matrices have constant coefficients (-b, 1+2b, -b), b = a/h^2, with at most the
// first and last diagonal entries modified by the boundary relation, so the Thomas pivots,
// multipliers and the Sherman-Morrison vector of the periodic (cyclic) case depend only on
// (component, axis). They are computed once per sweep on the host with the reference's own
// recurrences (solve_tridiagonal :522-559, solve_cyclic_tridiagonal :561-617) and shared by
// every line; each device thread then runs the forward/backward substitution of one line.

#include "ofh.hpp"

#include <cmath>
#include <memory>
#include <mutex>
#include <vector>

namespace ofh {

void launch_component_bc(Context& c, const of_grid& g, const of_bc_set& b, int comp, double* f);
void launch_component_bc_adjoint(Context& c, const of_grid& g, int comp, double* f);
bool axis_periodic(const of_bc_set& b, int axis);

namespace {

constexpr double kSmallNumber = 1.0e-30;

struct LineSys {
  int b, e, n;           // unknowns are storage indices [b, e) along the axis
  int cyclic;            // Sherman-Morrison correction
  int copy_dup;          // copy index b to index e (periodic normal component)
  double lo;             // sub-diagonal (-beta)
  double add_lo, add_hi; // constants added to rhs[0], rhs[n-1]
  double uc, gamma;      // cyclic: upper corner and gamma
  double den;            // cyclic: Sherman-Morrison denominator
  double single_pivot;   // cyclic n == 1
};

// Thomas factors of one line matrix (solver/projection.cpp:522-559).
void thomas_factors(const std::vector<double>& lo, const std::vector<double>& di, const std::vector<double>& up,
                    std::vector<double>& piv, std::vector<double>& mu) {
  const size_t n = di.size();
  piv.assign(n, 0.0);
  mu.assign(n, 0.0);
  double p = di[0];
  if(std::abs(p) < kSmallNumber) throw RuntimeError("tridiagonal solve encountered a zero pivot");
  piv[0] = p;
  mu[0] = n > 1 ? up[0] / p : 0.0;
  for(size_t i = 1; i < n; ++i) {
    p = di[i] - lo[i] * mu[i - 1];
    if(std::abs(p) < kSmallNumber) throw RuntimeError("tridiagonal solve encountered a zero pivot");
    piv[i] = p;
    mu[i] = i + 1 < n ? up[i] / p : 0.0;
  }
}

std::vector<double> thomas_apply(const std::vector<double>& lo, const std::vector<double>& piv,
                                 const std::vector<double>& mu, const std::vector<double>& rhs) {
  const size_t n = piv.size();
  std::vector<double> mr(n), x(n);
  mr[0] = rhs[0] / piv[0];
  for(size_t i = 1; i < n; ++i) mr[i] = (rhs[i] - lo[i] * mr[i - 1]) / piv[i];
  x[n - 1] = mr[n - 1];
  for(size_t i = n - 1; i-- > 0;) x[i] = mr[i] - mu[i] * x[i + 1];
  return x;
}

// One thread per line. Forward elimination stores the modified right-hand side in place,
// back substitution overwrites it with the solution; the cyclic case then applies the
// rank-one correction sol = x - fac*z with fac from x[0], x[n-1].
__global__ void k_line_solve(double* f, Lay L, int axis, LineSys S, const double* __restrict__ piv,
                             const double* __restrict__ mu, const double* __restrict__ z) {
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  if(t0 >= L.a(o0) || t1 >= L.a(o1)) return;
  int c[3];
  c[axis] = 0;
  c[o0] = 1 + t0;
  c[o1] = 1 + t1;
  double* base = f + L.at(c[0], c[1], c[2]);
  const ptrdiff_t st = axis == 0 ? 1 : axis == 1 ? (ptrdiff_t)L.sx : (ptrdiff_t)L.sx * L.sy;
  const int n = S.n;
  if(S.cyclic && n == 1) {
    base[S.b * st] = base[S.b * st] / S.single_pivot;
  } else {
    // forward elimination
    double r0 = base[S.b * st];
    if(n == 1) {
      r0 = r0 + S.add_lo;
      r0 = r0 + S.add_hi;
    } else {
      r0 = r0 + S.add_lo;
    }
    double mr = r0 / piv[0];
    base[S.b * st] = mr;
    for(int t = 1; t < n; ++t) {
      double r = base[(S.b + t) * st];
      if(t == n - 1) r = r + S.add_hi;
      mr = (r - S.lo * mr) / piv[t];
      base[(S.b + t) * st] = mr;
    }
    // back substitution
    double x = mr;
    const double x_last = x;
    for(int t = n - 2; t >= 0; --t) {
      x = base[(S.b + t) * st] - mu[t] * x;
      base[(S.b + t) * st] = x;
    }
    if(S.cyclic) {
      const double fac = (x + S.uc * x_last / S.gamma) / S.den;
      for(int t = 0; t < n; ++t) base[(S.b + t) * st] = base[(S.b + t) * st] - fac * z[t];
    }
  }
  if(S.copy_dup) base[S.e * st] = base[S.b * st];
}

// x-lines (unit stride along the line): one wave owns 64 lines; the line data moves through
// LDS in chunks of 64 lines x kLX elements so that every global access is a coalesced row
// segment (loads issued in batches of 16), and each lane then runs the recurrence of its own
// line out of LDS. Same arithmetic as k_line_solve.
constexpr int kLT = 64;
constexpr int kLX = 32;
__global__ void __launch_bounds__(kLT) k_line_solve_x(double* f, Lay L, LineSys S, const double* __restrict__ piv,
                                                      const double* __restrict__ mu, const double* __restrict__ z) {
  __shared__ double T[kLT][kLX + 1];
  const int lane = threadIdx.x;
  const int nl1 = L.ay, nlines = L.ay * L.az;
  const int line0 = blockIdx.x * kLT;
  const int nr = min(kLT, nlines - line0);
  const int n = S.n;
  const bool mine = lane < nr;
  // storage offset of this lane's line (32-bit: storage sizes stay below 2^31 elements)
  const int my_off = mine ? (int)L.at(0, 1 + (line0 + lane) % nl1, 1 + (line0 + lane) / nl1) : 0;
  if(S.cyclic && n == 1) {
    if(mine) {
      f[my_off + S.b] = f[my_off + S.b] / S.single_pivot;
      if(S.copy_dup) f[my_off + S.e] = f[my_off + S.b];
    }
    return;
  }
  const int col = lane % kLX, half = lane / kLX;  // two lanes per row segment of kLX elements
  auto load = [&](int c0, int cn) {
    for(int r0 = 0; r0 < kLT; r0 += 32) {
      double v[16];
#pragma unroll
      for(int e = 0; e < 16; ++e) {
        const int r = r0 + 2 * e + half;
        const int off = __shfl(my_off, r, 64);
        v[e] = (r < nr && col < cn) ? f[off + S.b + c0 + col] : 0.0;
      }
#pragma unroll
      for(int e = 0; e < 16; ++e) T[r0 + 2 * e + half][col] = v[e];
    }
    __syncthreads();
  };
  auto store = [&](int c0, int cn) {
    __syncthreads();
    for(int r0 = 0; r0 < kLT; r0 += 32) {
#pragma unroll
      for(int e = 0; e < 16; ++e) {
        const int r = r0 + 2 * e + half;
        const int off = __shfl(my_off, r, 64);
        if(r < nr && col < cn) f[off + S.b + c0 + col] = T[r][col];
      }
    }
    __syncthreads();
  };
  double mr = 0.0;
  for(int c0 = 0; c0 < n; c0 += kLX) {  // forward elimination
    const int cn = min(kLX, n - c0);
    load(c0, cn);
    if(mine) {
      for(int t = 0; t < cn; ++t) {
        const int tg = c0 + t;
        double r = T[lane][t];
        if(tg == 0) {
          if(n == 1) {
            r = r + S.add_lo;
            r = r + S.add_hi;
          } else {
            r = r + S.add_lo;
          }
          mr = r / piv[0];
        } else {
          if(tg == n - 1) r = r + S.add_hi;
          mr = (r - S.lo * mr) / piv[tg];
        }
        T[lane][t] = mr;
      }
    }
    store(c0, cn);
  }
  double x = mr;  // back substitution; x[n-1] = mr[n-1] is stored already
  const double x_last = x;
  for(int c0 = ((n - 1) / kLX) * kLX; c0 >= 0; c0 -= kLX) {
    const int cn = min(kLX, n - c0);
    load(c0, cn);
    if(mine) {
      for(int t = cn - 1; t >= 0; --t) {
        const int tg = c0 + t;
        if(tg == n - 1) continue;
        x = T[lane][t] - mu[tg] * x;
        T[lane][t] = x;
      }
    }
    store(c0, cn);
  }
  if(S.cyclic) {  // Sherman-Morrison correction
    const double fac = (x + S.uc * x_last / S.gamma) / S.den;
    double first = 0.0;
    for(int c0 = 0; c0 < n; c0 += kLX) {
      const int cn = min(kLX, n - c0);
      load(c0, cn);
      if(mine) {
        for(int t = 0; t < cn; ++t) T[lane][t] = T[lane][t] - fac * z[c0 + t];
        if(c0 == 0) first = T[lane][0];
      }
      store(c0, cn);
    }
    if(S.copy_dup && mine) f[my_off + S.e] = first;
  }
}

// y- and z-lines (lanes along i: coalesced): one thread per line as in k_line_solve, with the
// line streamed through double-buffered register chunks so that the loads of the next chunk
// overlap the recurrence of the current one.
constexpr int kLC = 16;
__global__ void __launch_bounds__(256) k_line_solve_s(double* f, Lay L, int axis, LineSys S,
                                                      const double* __restrict__ piv, const double* __restrict__ mu,
                                                      const double* __restrict__ z) {
  const int o1 = axis == 2 ? 1 : 2;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  if(t0 >= L.ax || t1 >= L.a(o1)) return;
  int cc[3];
  cc[axis] = 0;
  cc[0] = 1 + t0;
  cc[o1] = 1 + t1;
  double* base = f + L.at(cc[0], cc[1], cc[2]);
  const ptrdiff_t st = axis == 1 ? (ptrdiff_t)L.sx : (ptrdiff_t)L.sx * L.sy;
  const int n = S.n;
  if(S.cyclic && n == 1) {
    base[S.b * st] = base[S.b * st] / S.single_pivot;
    if(S.copy_dup) base[S.e * st] = base[S.b * st];
    return;
  }
  auto ld = [&](int c0, double(&v)[kLC]) {
#pragma unroll
    for(int e = 0; e < kLC; ++e)
      if(c0 + e < n) v[e] = base[(ptrdiff_t)(S.b + c0 + e) * st];
  };
  auto stv = [&](int c0, const double(&v)[kLC]) {
#pragma unroll
    for(int e = 0; e < kLC; ++e)
      if(c0 + e < n) base[(ptrdiff_t)(S.b + c0 + e) * st] = v[e];
  };
  double cur[kLC], nxt[kLC];
  double mr = 0.0;
  ld(0, cur);
  for(int c0 = 0; c0 < n; c0 += kLC) {  // forward elimination
    if(c0 + kLC < n) ld(c0 + kLC, nxt);
#pragma unroll
    for(int e = 0; e < kLC; ++e) {
      const int tg = c0 + e;
      if(tg < n) {
        double r = cur[e];
        if(tg == 0) {
          if(n == 1) {
            r = r + S.add_lo;
            r = r + S.add_hi;
          } else {
            r = r + S.add_lo;
          }
          mr = r / piv[0];
        } else {
          if(tg == n - 1) r = r + S.add_hi;
          mr = (r - S.lo * mr) / piv[tg];
        }
        cur[e] = mr;
      }
    }
    stv(c0, cur);
#pragma unroll
    for(int e = 0; e < kLC; ++e) cur[e] = nxt[e];
  }
  double x = mr;  // back substitution
  const double x_last = x;
  const int clast = ((n - 1) / kLC) * kLC;
  ld(clast, cur);
  for(int c0 = clast; c0 >= 0; c0 -= kLC) {
    if(c0 > 0) ld(c0 - kLC, nxt);
#pragma unroll
    for(int e = kLC - 1; e >= 0; --e) {
      const int tg = c0 + e;
      if(tg < n - 1) {
        x = cur[e] - mu[tg] * x;
        cur[e] = x;
      }
    }
    stv(c0, cur);
#pragma unroll
    for(int e = 0; e < kLC; ++e) cur[e] = nxt[e];
  }
  if(S.cyclic) {
    const double fac = (x + S.uc * x_last / S.gamma) / S.den;
    double first = 0.0;
    ld(0, cur);
    for(int c0 = 0; c0 < n; c0 += kLC) {
      if(c0 + kLC < n) ld(c0 + kLC, nxt);
#pragma unroll
      for(int e = 0; e < kLC; ++e)
        if(c0 + e < n) cur[e] = cur[e] - fac * z[c0 + e];
      if(c0 == 0) first = cur[0];
      stv(c0, cur);
#pragma unroll
      for(int e = 0; e < kLC; ++e) cur[e] = nxt[e];
    }
    if(S.copy_dup) base[S.e * st] = first;
  }
}

inline bool dirichlet_velocity(const of_bc& c) {  // solver/projection.cpp:45-49
  return c.type == OF_NO_SLIP_WALL || c.type == OF_PRESCRIBED_VELOCITY || c.type == OF_SYMMETRY;
}

// solver/projection.cpp:60-80
void tangential_relation(const of_bc& c, int comp, double& coef, double& cst) {
  switch(c.type) {
    case OF_NO_SLIP_WALL:
    case OF_PRESCRIBED_VELOCITY:
      coef = -1.0;
      cst = 2.0 * c.velocity[comp];
      return;
    case OF_SYMMETRY:
    case OF_FIXED_PRESSURE:
      coef = 1.0;
      cst = 0.0;
      return;
    default:
      coef = 0.0;
      cst = 0.0;
  }
}

}  // namespace

namespace {

struct Pending {
  int comp, axis;
  LineSys S;
  size_t off;
};

// Line systems of every (component, axis) sweep with their Thomas / Sherman-Morrison factors,
// computed on the host with the reference's recurrences and staged in a device buffer.
struct AdiPlan {
  std::vector<Pending> sweeps;
  double* dev = nullptr;
  int nmax = 0;
  int line_solves = 0;
};

AdiPlan build_adi_plan(const of_grid& g, const of_bc_set& b, double alpha) {
  AdiPlan P;
  P.nmax = std::max(g.nx, std::max(g.ny, g.nz)) + 2;
  const int nmax = P.nmax;
  std::vector<double> host(3 * (size_t)nmax * 9, 0.0);
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    for(int axis = 0; axis < 3; ++axis) {
      const double h = spacing(g, axis);
      const double beta = alpha / (h * h);
      const int A = L.a(axis);
      LineSys S{};
      S.lo = -beta;
      std::vector<double> lo, di, up, piv, mu, zv;
      if(axis_periodic(b, axis)) {  // solve_component_sweep_periodic :619-720
        S.b = 1;
        S.e = comp == axis ? A : 1 + A;
        S.n = S.e - S.b;
        S.copy_dup = comp == axis;
        S.cyclic = 1;
        if(S.n <= 0) continue;
        const int n = S.n;
        lo.assign(n, -beta);
        di.assign(n, 1.0 + 2.0 * beta);
        up.assign(n, -beta);
        lo[0] = 0.0;
        up[n - 1] = 0.0;
        const double lc = -beta, uc = -beta;
        if(n == 1) {
          S.single_pivot = di[0] + lc + uc;
          if(std::abs(S.single_pivot) < kSmallNumber)
            throw RuntimeError("cyclic tridiagonal solve encountered a zero pivot");
          piv.assign(1, 1.0);
          mu.assign(1, 0.0);
          zv.assign(1, 0.0);
        } else {
          std::vector<double> dm = di;
          const double gamma = -di[0];
          if(std::abs(gamma) < kSmallNumber) throw RuntimeError("cyclic tridiagonal solve encountered a zero gamma");
          dm[0] = di[0] - gamma;
          dm[n - 1] = di[n - 1] - lc * uc / gamma;
          std::vector<double> lm = lo, um = up;
          lm[0] = 0.0;
          um[n - 1] = 0.0;
          thomas_factors(lm, dm, um, piv, mu);
          std::vector<double> uvec(n, 0.0);
          uvec[0] = gamma;
          uvec[n - 1] = lc;
          zv = thomas_apply(lm, piv, mu, uvec);
          S.den = 1.0 + zv[0] + uc * zv[n - 1] / gamma;
          if(std::abs(S.den) < kSmallNumber)
            throw RuntimeError("cyclic tridiagonal solve encountered a singular Sherman-Morrison update");
          S.uc = uc;
          S.gamma = gamma;
        }
      } else {  // solve_component_sweep :722-866
        int bb = 1, ee = 1 + A;
        if(comp == axis) {
          if(dirichlet_velocity(b.faces[2 * axis])) bb += 1;
          if(dirichlet_velocity(b.faces[2 * axis + 1])) ee -= 1;
        }
        S.b = bb;
        S.e = ee;
        S.n = ee - bb;
        S.cyclic = 0;
        S.copy_dup = 0;
        if(S.n <= 0) continue;
        const int n = S.n;
        lo.assign(n, -beta);
        di.assign(n, 1.0 + 2.0 * beta);
        up.assign(n, -beta);
        lo[0] = 0.0;
        up[n - 1] = 0.0;
        const of_bc& cl = b.faces[2 * axis];
        const of_bc& cu = b.faces[2 * axis + 1];
        if(comp == axis && dirichlet_velocity(cl)) {
          S.add_lo = beta * cl.velocity[axis];
        } else {
          double coef, cst;
          tangential_relation(cl, comp, coef, cst);
          di[0] -= beta * coef;
          S.add_lo = beta * cst;
        }
        if(comp == axis && dirichlet_velocity(cu)) {
          S.add_hi = beta * cu.velocity[axis];
        } else {
          double coef, cst;
          tangential_relation(cu, comp, coef, cst);
          di[n - 1] -= beta * coef;
          S.add_hi = beta * cst;
        }
        thomas_factors(lo, di, up, piv, mu);
        zv.assign(n, 0.0);
      }
      const size_t off = (size_t)(comp * 3 + axis) * 3 * nmax;
      std::copy(piv.begin(), piv.end(), host.begin() + off);
      std::copy(mu.begin(), mu.end(), host.begin() + off + nmax);
      std::copy(zv.begin(), zv.end(), host.begin() + off + 2 * nmax);
      const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
      P.line_solves += L.a(o0) * L.a(o1);
      P.sweeps.push_back(Pending{comp, axis, S, off});
    }
  }
  // a fresh buffer, filled by a blocking copy before any launch can read it
  OFH_CHECK(hipMalloc(&P.dev, host.size() * sizeof(double)));
  OFH_CHECK(hipMemcpy(P.dev, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
  return P;
}

// Plans depend only on (device, grid, boundary conditions, alpha); a stepper reuses one plan
// for every step, so the factors are built and uploaded once instead of once per solve.
struct AdiCacheEntry {
  int device;
  of_grid g;
  of_bc_set b;
  double alpha;
  AdiPlan plan;
};
constexpr size_t kAdiCacheCap = 16;
std::mutex g_adi_mutex;
std::vector<std::unique_ptr<AdiCacheEntry>> g_adi_cache;  // most recently used last

bool same_plan_key(const AdiCacheEntry& e, int device, const of_grid& g, const of_bc_set& b, double alpha) {
  if(e.device != device || e.alpha != alpha) return false;
  if(e.g.nx != g.nx || e.g.ny != g.ny || e.g.nz != g.nz || e.g.dx != g.dx || e.g.dy != g.dy || e.g.dz != g.dz)
    return false;
  for(int f = 0; f < 6; ++f) {
    const of_bc &x = e.b.faces[f], &y = b.faces[f];
    if(x.type != y.type) return false;
    for(int k = 0; k < 3; ++k)
      if(x.velocity[k] != y.velocity[k]) return false;
  }
  return true;
}

// Returns a copy of the cached plan (the sweep list and the device address of its factors).
AdiPlan plan_adi(Context& c, const of_grid& g, const of_bc_set& b, double alpha) {
  std::lock_guard<std::mutex> lock(g_adi_mutex);
  for(size_t i = 0; i < g_adi_cache.size(); ++i) {
    if(!same_plan_key(*g_adi_cache[i], c.device, g, b, alpha)) continue;
    std::unique_ptr<AdiCacheEntry> e = std::move(g_adi_cache[i]);
    g_adi_cache.erase(g_adi_cache.begin() + (ptrdiff_t)i);
    g_adi_cache.push_back(std::move(e));
    return g_adi_cache.back()->plan;
  }
  if(g_adi_cache.size() >= kAdiCacheCap) {
    // the least recently used plan may still be read by queued line solves on any stream
    OFH_CHECK(hipDeviceSynchronize());
    OFH_CHECK(hipFree(g_adi_cache.front()->plan.dev));
    g_adi_cache.erase(g_adi_cache.begin());
  }
  auto e = std::make_unique<AdiCacheEntry>();
  e->device = c.device;
  e->g = g;
  e->b = b;
  e->alpha = alpha;
  e->plan = build_adi_plan(g, b, alpha);
  g_adi_cache.push_back(std::move(e));
  return g_adi_cache.back()->plan;
}

void launch_sweep(Context& c, const Lay& L, const Pending& p, const AdiPlan& P, double* f, const LineSys& S) {
  const int o1 = p.axis == 2 ? 1 : 2;
  const double* piv = P.dev + p.off;
  const double* mu = P.dev + p.off + P.nmax;
  const double* zv = P.dev + p.off + 2 * P.nmax;
  if(p.axis == 0) {
    ProfScope prof_(c, "k_line_solve_x", true);
    hipLaunchKernelGGL(k_line_solve_x, dim3(div_up((size_t)L.ay * L.az, kLT)), dim3(kLT), 0, c.stream, f, L, S, piv,
                       mu, zv);
  } else {
    ProfScope prof_(c, "k_line_solve_s", true);
    const dim3 grid((unsigned)((L.ax + 63) / 64), (unsigned)((L.a(o1) + 3) / 4), 1);
    hipLaunchKernelGGL(k_line_solve_s, grid, dim3(64, 4, 1), 0, c.stream, f, L, p.axis, S, piv, mu, zv);
  }
}

// Transpose of the duplicated-face copy of a periodic normal sweep (x[e] := x[b]):
// x_bar[b] += x_bar[e], x_bar[e] = 0, for every line of the component along `axis`.
__global__ void k_dup_fold(double* f, Lay L, int axis, int b, int e) {
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  if(t0 >= L.a(o0) || t1 >= L.a(o1)) return;
  int cb[3], ce[3];
  cb[axis] = b;
  ce[axis] = e;
  cb[o0] = ce[o0] = 1 + t0;
  cb[o1] = ce[o1] = 1 + t1;
  const size_t ib = L.at(cb[0], cb[1], cb[2]), ie = L.at(ce[0], ce[1], ce[2]);
  f[ib] = f[ib] + f[ie];
  f[ie] = 0.0;
}

}  // namespace

// Returns the reference's line_solves count.
int predictor_adi(Context& c, const of_grid& g, const of_bc_set& b, double alpha, const double* ru, const double* rv,
                  const double* rw, double* ou, double* ov, double* ow) {
  if(alpha < 0.0) throw InvalidArgument("solve_predictor_adi expects a non-negative alpha");
  const double* rhs[3] = {ru, rv, rw};
  double* out[3] = {ou, ov, ow};
  const AdiPlan P = plan_adi(c, g, b, alpha);
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    OFH_CHECK(hipMemcpyAsync(out[comp], rhs[comp], L.size() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  }
  for(int comp = 0; comp < 3; ++comp) {
    launch_component_bc(c, g, b, comp, out[comp]);
    if(alpha == 0.0) continue;
    const Lay L = make_lay(g, OF_FACE_X + comp);
    for(const Pending& p : P.sweeps) {
      if(p.comp != comp) continue;
      launch_sweep(c, L, p, P, out[comp], p.S);
      launch_component_bc(c, g, b, comp, out[comp]);
    }
  }
  OFH_CHECK(hipGetLastError());
  return alpha == 0.0 ? 0 : P.line_solves;
}

// Transpose of predictor_adi as a linear map rhs -> out (periodic boundaries): the sweeps in
// reverse order with the transposed ghost fill between them; each cyclic line matrix is
// symmetric, so its transposed solve is the same solve once the duplicated face is folded back.
void predictor_adi_adjoint(Context& c, const of_grid& g, const of_bc_set& b, double alpha, const double* const* ybar,
                           double* const* rbar) {
  for(int a = 0; a < 3; ++a)
    if(!axis_periodic(b, a)) throw InvalidArgument("the ADI adjoint supports periodic boundaries");
  const AdiPlan P = plan_adi(c, g, b, alpha);
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    if(rbar[comp] != ybar[comp])
      OFH_CHECK(hipMemcpyAsync(rbar[comp], ybar[comp], L.size() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    launch_component_bc_adjoint(c, g, comp, rbar[comp]);
    if(alpha == 0.0) continue;
    for(auto it = P.sweeps.rbegin(); it != P.sweeps.rend(); ++it) {
      const Pending& p = *it;
      if(p.comp != comp) continue;
      LineSys S = p.S;
      if(S.copy_dup) {
        const int o0 = p.axis == 0 ? 1 : 0, o1 = p.axis == 2 ? 1 : 2;
        const dim3 block(64, 4, 1);
        const dim3 grid((unsigned)((L.a(o0) + 63) / 64), (unsigned)((L.a(o1) + 3) / 4), 1);
        hipLaunchKernelGGL(k_dup_fold, grid, block, 0, c.stream, rbar[comp], L, p.axis, S.b, S.e);
        S.copy_dup = 0;
      }
      launch_sweep(c, L, p, P, rbar[comp], S);
      launch_component_bc_adjoint(c, g, comp, rbar[comp]);
    }
  }
  OFH_CHECK(hipGetLastError());
}

}  // namespace ofh
