// ofref: FP64 CPU restatement of zfifteen/solver's incompressible projection step.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product in solver_amd/.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product never links, imports or calls it.
//
// Every function restates the behaviour of the reference function cited next to it
// (paths are relative to the reference repository root). It follows the reference's
// order of floating-point operations wherever that order is visible in the source, so
// that results agree with the reference to rounding level (the golden vectors under
// tests/golden/ are produced by the reference itself and pin this file).
//
// Storage contract (core/field_layout.hpp:69-290, core/grid.hpp:57-136): one ghost layer,
// i fastest, index = i + sx*(j + sy*k); cell-centred fields have active extent (nx,ny,nz),
// face-centred fields have one extra active entry along their normal axis (the periodic
// face is stored twice). Buffers are zero-initialised (core/aligned_buffer.hpp:60-69).
//
// Single-threaded, built with -O3 -ffp-contract=off and no fast-math (the reference's
// "deterministic" profile, CMakeLists.txt:40-50).

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace ofref {

struct Grid {
  int nx, ny, nz;
  double dx, dy, dz;
};

enum Loc : int { kCell = 0, kFaceX = 1, kFaceY = 2, kFaceZ = 3 };

// bc/boundary_conditions.hpp:10-22
enum PhysType : int { kNoSlip = 0, kPrescribed = 1, kSymmetry = 2, kFixedPressure = 3, kPeriodic = 4 };
enum PresType : int { kNeumann = 0, kDirichlet = 1, kPresPeriodic = 2 };

struct BC {
  int type;
  double velocity[3];
  double pressure;
  double pressure_gradient;
};
struct BCSet {
  BC f[6];
};
struct PBC {
  int type;
  double value;
  double gradient;
};
struct PBCSet {
  PBC f[6];
};

struct Range {
  int ib, ie, jb, je, kb, ke;
};

static inline double spacing(const Grid& g, int axis) { return axis == 0 ? g.dx : axis == 1 ? g.dy : g.dz; }
static inline int cells(const Grid& g, int axis) { return axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz; }

struct Layout {
  Grid g;
  Loc loc;
  int a[3];  // active extents
  int s[3];  // storage extents
  size_t size() const { return (size_t)s[0] * s[1] * s[2]; }
  size_t idx(int i, int j, int k) const { return (size_t)i + (size_t)s[0] * ((size_t)j + (size_t)s[1] * (size_t)k); }
  Range active() const { return Range{1, 1 + a[0], 1, 1 + a[1], 1, 1 + a[2]}; }
  // core/field_layout.hpp:185-194
  Range boundary_active(int face) const {
    Range r = active();
    const int axis = face / 2;
    const int first = (face % 2 == 0) ? 1 : a[axis];
    set_axis(r, axis, first, first + 1);
    return r;
  }
  // core/field_layout.hpp:196-213 (layer 0)
  Range ghost(int face) const {
    Range r = active();
    const int axis = face / 2;
    const int b = (face % 2 == 0) ? 0 : 1 + a[axis];
    set_axis(r, axis, b, b + 1);
    return r;
  }
  static void set_axis(Range& r, int axis, int b, int e) {
    if(axis == 0) { r.ib = b; r.ie = e; }
    else if(axis == 1) { r.jb = b; r.je = e; }
    else { r.kb = b; r.ke = e; }
  }
  int normal() const { return loc == kCell ? -1 : (int)loc - 1; }
};

// core/field_layout.hpp:71-86
static Layout make_layout(const Grid& g, Loc loc) {
  Layout L;
  L.g = g;
  L.loc = loc;
  L.a[0] = g.nx; L.a[1] = g.ny; L.a[2] = g.nz;
  if(loc != kCell) L.a[(int)loc - 1] += 1;
  for(int d = 0; d < 3; ++d) L.s[d] = L.a[d] + 2;
  return L;
}

struct Field {
  Layout L;
  double* p;
  std::vector<double> own;
  Field() : p(nullptr) {}
  Field(const Layout& l) : L(l), own(l.size(), 0.0) { p = own.data(); }
  Field(const Layout& l, double* ext) : L(l), p(ext) {}
  Field(const Field& o) : L(o.L), own(o.p, o.p + o.L.size()) { p = own.data(); }
  Field& operator=(const Field& o) {  // value copy of the whole storage
    std::memcpy(p, o.p, sizeof(double) * L.size());
    return *this;
  }
  double& operator()(int i, int j, int k) { return p[L.idx(i, j, k)]; }
  double operator()(int i, int j, int k) const { return p[L.idx(i, j, k)]; }
  void fill(double v) { std::fill(p, p + L.size(), v); }
};

struct Vel {
  Field x, y, z;
  Vel() {}
  explicit Vel(const Grid& g) : x(make_layout(g, kFaceX)), y(make_layout(g, kFaceY)), z(make_layout(g, kFaceZ)) {}
  Vel(const Grid& g, double* u, double* v, double* w)
      : x(make_layout(g, kFaceX), u), y(make_layout(g, kFaceY), v), z(make_layout(g, kFaceZ), w) {}
  Field& c(int a) { return a == 0 ? x : a == 1 ? y : z; }
  const Field& c(int a) const { return a == 0 ? x : a == 1 ? y : z; }
  Vel& operator=(const Vel& o) { x = o.x; y = o.y; z = o.z; return *this; }
  void fill(double v) { x.fill(v); y.fill(v); z.fill(v); }
};

template <typename F>
static inline void for_range(const Range& r, F&& f) {
  for(int k = r.kb; k < r.ke; ++k)
    for(int j = r.jb; j < r.je; ++j)
      for(int i = r.ib; i < r.ie; ++i) f(i, j, k);
}

static void copy_range(Field& dst, const Range& dr, const Field& src, const Range& sr) {
  const int ex = dr.ie - dr.ib, ey = dr.je - dr.jb, ez = dr.ke - dr.kb;
  for(int k = 0; k < ez; ++k)
    for(int j = 0; j < ey; ++j)
      for(int i = 0; i < ex; ++i) dst(dr.ib + i, dr.jb + j, dr.kb + k) = src(sr.ib + i, sr.jb + j, sr.kb + k);
}

// ---------------------------------------------------------------- reductions
// solver/projection.cpp:117-185 and linsolve/poisson_solver.cpp:51-162: sequential k,j,i sums.
static double active_mean(const Field& f) {
  double s = 0.0;
  for_range(f.L.active(), [&](int i, int j, int k) { s += f(i, j, k); });
  return s / (double)((size_t)f.L.a[0] * f.L.a[1] * f.L.a[2]);
}
static double subtract_active_mean(Field& f) {
  const double m = active_mean(f);
  for_range(f.L.active(), [&](int i, int j, int k) { f(i, j, k) -= m; });
  return m;
}
static double active_l2(const Field& f) {
  double s = 0.0;
  for_range(f.L.active(), [&](int i, int j, int k) { s += f(i, j, k) * f(i, j, k); });
  return std::sqrt(s / (double)((size_t)f.L.a[0] * f.L.a[1] * f.L.a[2]));
}
static double dot_active(const Field& a, const Field& b) {
  double s = 0.0;
  for_range(a.L.active(), [&](int i, int j, int k) { s += a(i, j, k) * b(i, j, k); });
  return s;
}
static void axpy_active(Field& d, const Field& s, double scale) {
  for_range(d.L.active(), [&](int i, int j, int k) { d(i, j, k) += scale * s(i, j, k); });
}
static void scale_active(Field& f, double scale) {
  for_range(f.L.active(), [&](int i, int j, int k) { f(i, j, k) *= scale; });
}
static void copy_active(const Field& s, Field& d) {
  for_range(s.L.active(), [&](int i, int j, int k) { d(i, j, k) = s(i, j, k); });
}

// ---------------------------------------------------------------- velocity BCs
static bool velocity_axis_periodic(const BCSet& b, int axis) {  // solver/projection.cpp:197-205
  const bool lo = b.f[2 * axis].type == kPeriodic, hi = b.f[2 * axis + 1].type == kPeriodic;
  if(lo != hi) throw std::invalid_argument("periodic velocity boundaries must be paired on both sides");
  return lo;
}
static bool pressure_axis_periodic(const PBCSet& b, int axis) {  // solver/projection.cpp:187-195
  const bool lo = b.f[2 * axis].type == kPresPeriodic, hi = b.f[2 * axis + 1].type == kPresPeriodic;
  if(lo != hi) throw std::invalid_argument("periodic pressure boundaries must be paired on both sides");
  return lo;
}

// solver/projection.cpp:207-231 (both branches select the same source slabs)
static void periodic_pair_structured(Field& f, int axis) {
  const Range act = f.L.active();
  Range last = act, first = act;
  Layout::set_axis(last, axis, (axis == 0 ? act.ie : axis == 1 ? act.je : act.ke) - 1,
                   axis == 0 ? act.ie : axis == 1 ? act.je : act.ke);
  const int b = axis == 0 ? act.ib : axis == 1 ? act.jb : act.kb;
  Layout::set_axis(first, axis, b, b + 1);
  copy_range(f, f.L.ghost(2 * axis), f, last);
  copy_range(f, f.L.ghost(2 * axis + 1), f, first);
}

// solver/projection.cpp:233-300
static void periodic_pair_face(Field& f, int axis) {
  if(f.L.normal() != axis) {
    periodic_pair_structured(f, axis);
    return;
  }
  const Range lo = f.L.boundary_active(2 * axis), hi = f.L.boundary_active(2 * axis + 1);
  const int ex = lo.ie - lo.ib, ey = lo.je - lo.jb, ez = lo.ke - lo.kb;
  for(int k = 0; k < ez; ++k)
    for(int j = 0; j < ey; ++j)
      for(int i = 0; i < ex; ++i) {
        const double shared = 0.5 * (f(lo.ib + i, lo.jb + j, lo.kb + k) + f(hi.ib + i, hi.jb + j, hi.kb + k));
        f(lo.ib + i, lo.jb + j, lo.kb + k) = shared;
        f(hi.ib + i, hi.jb + j, hi.kb + k) = shared;
      }
  const Range act = f.L.active();
  Range src_lo = act, src_hi = act;
  const int e = axis == 0 ? act.ie : axis == 1 ? act.je : act.ke;
  const int b = axis == 0 ? act.ib : axis == 1 ? act.jb : act.kb;
  Layout::set_axis(src_lo, axis, e - 2, e - 1);
  Layout::set_axis(src_hi, axis, b + 1, b + 2);
  copy_range(f, f.L.ghost(2 * axis), f, src_lo);
  copy_range(f, f.L.ghost(2 * axis + 1), f, src_hi);
}

static inline bool dirichlet_velocity(const BC& c) {  // solver/projection.cpp:45-49
  return c.type == kNoSlip || c.type == kPrescribed || c.type == kSymmetry;
}

// solver/projection.cpp:384-449
static void face_field_boundary(Field& f, int face, const BC& c) {
  if(c.type == kPeriodic) return;
  const int baxis = face / 2;
  const bool lower = face % 2 == 0;
  const int comp = f.L.normal();
  const Range ba = f.L.boundary_active(face), gr = f.L.ghost(face);
  const int ex = ba.ie - ba.ib, ey = ba.je - ba.jb, ez = ba.ke - ba.kb;
  const int di = baxis == 0, dj = baxis == 1, dk = baxis == 2;
  for(int k = 0; k < ez; ++k)
    for(int j = 0; j < ey; ++j)
      for(int i = 0; i < ex; ++i) {
        const int bi = ba.ib + i, bj = ba.jb + j, bk = ba.kb + k;
        const int gi = gr.ib + i, gj = gr.jb + j, gk = gr.kb + k;
        if(comp == baxis) {
          const int ii = lower ? bi + di : bi - di, jj = lower ? bj + dj : bj - dj, kk = lower ? bk + dk : bk - dk;
          if(c.type == kFixedPressure) {
            f(bi, bj, bk) = f(ii, jj, kk);
            f(gi, gj, gk) = f(bi, bj, bk);
            continue;
          }
          const double bv = c.type == kSymmetry ? 0.0 : c.velocity[comp];
          f(bi, bj, bk) = bv;
          f(gi, gj, gk) = 2.0 * bv - f(ii, jj, kk);
          continue;
        }
        if(c.type == kNoSlip || c.type == kPrescribed) {
          f(gi, gj, gk) = 2.0 * c.velocity[comp] - f(bi, bj, bk);
        } else if(c.type == kSymmetry || c.type == kFixedPressure) {
          f(gi, gj, gk) = f(bi, bj, bk);
        }
      }
}

// solver/projection.cpp:451-462
static void velocity_component_boundaries(const BCSet& b, Field& f) {
  for(int axis = 0; axis < 3; ++axis) {
    if(velocity_axis_periodic(b, axis)) {
      periodic_pair_face(f, axis);
      continue;
    }
    face_field_boundary(f, 2 * axis, b.f[2 * axis]);
    face_field_boundary(f, 2 * axis + 1, b.f[2 * axis + 1]);
  }
}

// solver/projection.cpp:966-971
static void apply_velocity_bc(const BCSet& b, Vel& v) {
  velocity_component_boundaries(b, v.x);
  velocity_component_boundaries(b, v.y);
  velocity_component_boundaries(b, v.z);
}

// solver/projection.cpp:302-337
static void pressure_face_boundary(Field& p, int face, const PBC& c) {
  const int axis = face / 2;
  const bool lower = face % 2 == 0;
  const Range ba = p.L.boundary_active(face), gr = p.L.ghost(face);
  const double h = spacing(p.L.g, axis);
  const int ex = ba.ie - ba.ib, ey = ba.je - ba.jb, ez = ba.ke - ba.kb;
  for(int k = 0; k < ez; ++k)
    for(int j = 0; j < ey; ++j)
      for(int i = 0; i < ex; ++i) {
        const double av = p(ba.ib + i, ba.jb + j, ba.kb + k);
        double& g = p(gr.ib + i, gr.jb + j, gr.kb + k);
        if(c.type == kNeumann) g = lower ? av - c.gradient * h : av + c.gradient * h;
        else if(c.type == kDirichlet) g = 2.0 * c.value - av;
      }
}

// solver/projection.cpp:973-984
static void apply_pressure_bc(const PBCSet& b, Field& p) {
  for(int axis = 0; axis < 3; ++axis) {
    if(pressure_axis_periodic(b, axis)) {
      periodic_pair_structured(p, axis);
      continue;
    }
    pressure_face_boundary(p, 2 * axis, b.f[2 * axis]);
    pressure_face_boundary(p, 2 * axis + 1, b.f[2 * axis + 1]);
  }
}

// solver/projection.cpp:339-382
static void total_pressure_gradient_face(Field& p, int face, const Field& src) {
  const int axis = face / 2;
  const bool lower = face % 2 == 0;
  const Range ba = p.L.boundary_active(face), gr = p.L.ghost(face), wa = src.L.active();
  const double h = spacing(p.L.g, axis);
  const int ex = ba.ie - ba.ib, ey = ba.je - ba.jb, ez = ba.ke - ba.kb;
  for(int k = 0; k < ez; ++k)
    for(int j = 0; j < ey; ++j)
      for(int i = 0; i < ex; ++i) {
        const int bi = ba.ib + i, bj = ba.jb + j, bk = ba.kb + k;
        int wi = bi, wj = bj, wk = bk;
        if(axis == 0) wi = lower ? wa.ib : wa.ie - 1;
        else if(axis == 1) wj = lower ? wa.jb : wa.je - 1;
        else wk = lower ? wa.kb : wa.ke - 1;
        const double grad = src(wi, wj, wk);
        const double av = p(bi, bj, bk);
        p(gr.ib + i, gr.jb + j, gr.kb + k) = lower ? av - grad * h : av + grad * h;
      }
}

// solver/projection.cpp:986-1026
static void apply_total_pressure_bc(const BCSet& b, const Vel& src, Field& p) {
  for(int axis = 0; axis < 3; ++axis) {
    if(velocity_axis_periodic(b, axis)) {
      periodic_pair_structured(p, axis);
      continue;
    }
    for(int face = 2 * axis; face <= 2 * axis + 1; ++face) {
      const BC& c = b.f[face];
      if(c.type == kNoSlip || c.type == kPrescribed || c.type == kSymmetry) {
        total_pressure_gradient_face(p, face, src.c(axis));
      } else if(c.type == kFixedPressure) {
        PBC pb{kDirichlet, c.pressure, 0.0};
        pressure_face_boundary(p, face, pb);
      }
    }
  }
}

// solver/projection.cpp:934-964
static PBCSet derive_pressure_correction_bc(const BCSet& b) {
  PBCSet m{};
  for(int f = 0; f < 6; ++f) {
    m.f[f] = PBC{kNeumann, 0.0, 0.0};
    switch(b.f[f].type) {
      case kNoSlip:
      case kPrescribed:
      case kSymmetry: m.f[f].type = kNeumann; break;
      case kFixedPressure: m.f[f].type = kDirichlet; break;
      case kPeriodic: m.f[f].type = kPresPeriodic; break;
    }
  }
  for(int a = 0; a < 3; ++a) pressure_axis_periodic(m, a);
  return m;
}

// ---------------------------------------------------------------- operators
// operators/discrete_operators.cpp:37-73
static void laplacian(const Field& in, Field& out) {
  const Grid& g = in.L.g;
  const double ix = 1.0 / (g.dx * g.dx), iy = 1.0 / (g.dy * g.dy), iz = 1.0 / (g.dz * g.dz);
  for_range(in.L.active(), [&](int i, int j, int k) {
    const double c = in(i, j, k);
    out(i, j, k) = (in(i + 1, j, k) - 2.0 * c + in(i - 1, j, k)) * ix +
                   (in(i, j + 1, k) - 2.0 * c + in(i, j - 1, k)) * iy +
                   (in(i, j, k + 1) - 2.0 * c + in(i, j, k - 1)) * iz;
  });
}

// operators/discrete_operators.cpp:77-109
static void gradient(const Field& p, Vel& g) {
  const Grid& gr = p.L.g;
  for_range(g.x.L.active(), [&](int i, int j, int k) { g.x(i, j, k) = (p(i, j, k) - p(i - 1, j, k)) / gr.dx; });
  for_range(g.y.L.active(), [&](int i, int j, int k) { g.y(i, j, k) = (p(i, j, k) - p(i, j - 1, k)) / gr.dy; });
  for_range(g.z.L.active(), [&](int i, int j, int k) { g.z(i, j, k) = (p(i, j, k) - p(i, j, k - 1)) / gr.dz; });
}

// operators/discrete_operators.cpp:111-126
static void divergence(const Vel& v, Field& d) {
  const Grid& g = d.L.g;
  for_range(d.L.active(), [&](int i, int j, int k) {
    const double a = (v.x(i + 1, j, k) - v.x(i, j, k)) / g.dx;
    const double b = (v.y(i, j + 1, k) - v.y(i, j, k)) / g.dy;
    const double c = (v.z(i, j, k + 1) - v.z(i, j, k)) / g.dz;
    d(i, j, k) = a + b + c;
  });
}

// ---------------------------------------------------------------- advection
constexpr double kSmallDenominator = 1.0e-14;  // solver/momentum_terms.cpp:62

// Forward-mode dual number (value, derivative) for the tangent-linear restatement. Branches
// test the value only, so every branch is frozen at the base state. |x| takes the derivative
// sign of -x at x = 0, which makes the van Leer limiter's derivative 0 at its kink r = 0.
struct Dual {
  double v, d;
};
static inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
static inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
static inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
static inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
static inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
static inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
static inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
static inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
static inline double value_of(double x) { return x; }
static inline double value_of(Dual x) { return x.v; }
static inline double abs_of(double x) { return std::abs(x); }
static inline Dual abs_of(Dual x) { return x.v > 0.0 ? x : -x; }
template <typename T>
static inline T zero_of() { return T{}; }

template <typename T>
static inline T van_leer(T r) { return (r + abs_of(r)) / (1.0 + abs_of(r)); }  // :72-79
template <typename T>
static inline T safe_ratio(T n, T d) { return std::abs(value_of(d)) < kSmallDenominator ? zero_of<T>() : n / d; }  // :81-86

// solver/momentum_terms.cpp:88-197: face value at the interface between storage index
// `right`-1 and `right` along `axis`; `at(o)` reads the field at offset o along the axis
// relative to `right`. scheme: 0 tvd, 1 upwind, 2 central. T is double, or Dual for the
// tangent (the same operations on (value, derivative) pairs).
template <typename T, typename At>
static inline T reconstruct(At&& at, int right, int max_index, T vel, int scheme) {
  const T left = at(-1), rgt = at(0);
  if(scheme == 2) return 0.5 * (left + rgt);
  if(scheme == 1 || std::abs(value_of(vel)) < kSmallDenominator) return value_of(vel) >= 0.0 ? left : rgt;
  if(value_of(vel) >= 0.0) {
    if(right - 2 < 0) return left;
    const T dd = rgt - left;
    const T r = safe_ratio(left - at(-2), dd);
    return left + 0.5 * van_leer(r) * dd;
  }
  if(right + 1 > max_index) return rgt;
  const T d = rgt - left;
  const T r = safe_ratio(at(1) - rgt, d);
  return rgt - 0.5 * van_leer(r) * d;
}

// Reads component c of the velocity at (i, j, k) as T: the value, or (value, perturbation).
template <typename T>
struct VelReader;
template <>
struct VelReader<double> {
  const Vel& v;
  double operator()(int c, int i, int j, int k) const { return v.c(c)(i, j, k); }
};
template <>
struct VelReader<Dual> {
  const Vel& v;
  const Vel& dv;
  Dual operator()(int c, int i, int j, int k) const { return Dual{v.c(c)(i, j, k), dv.c(c)(i, j, k)}; }
};

// solver/momentum_terms.cpp:199-312 (component c advected by the MAC velocity); `put` stores
// the T-valued result.
template <typename T, typename Put>
static void advection_component_t(const VelReader<T>& V, const Layout& Lq, int c, int scheme, Put&& put) {
  const Grid& g = Lq.g;
  auto rec = [&](int axis, int i, int j, int k, T vel) {
    const int right = axis == 0 ? i : axis == 1 ? j : k;
    return reconstruct<T>(
        [&](int o) { return V(c, i + (axis == 0 ? o : 0), j + (axis == 1 ? o : 0), k + (axis == 2 ? o : 0)); }, right,
        Lq.s[axis] - 1, vel, scheme);
  };
  for_range(Lq.active(), [&](int i, int j, int k) {
    T wv, ev, sv, nv, bv, fv;
    if(c == 0) {
      wv = 0.5 * (V(0, i - 1, j, k) + V(0, i, j, k));
      ev = 0.5 * (V(0, i, j, k) + V(0, i + 1, j, k));
      sv = 0.5 * (V(1, i - 1, j, k) + V(1, i, j, k));
      nv = 0.5 * (V(1, i - 1, j + 1, k) + V(1, i, j + 1, k));
      bv = 0.5 * (V(2, i - 1, j, k) + V(2, i, j, k));
      fv = 0.5 * (V(2, i - 1, j, k + 1) + V(2, i, j, k + 1));
    } else if(c == 1) {
      wv = 0.5 * (V(0, i, j - 1, k) + V(0, i, j, k));
      ev = 0.5 * (V(0, i + 1, j - 1, k) + V(0, i + 1, j, k));
      sv = 0.5 * (V(1, i, j - 1, k) + V(1, i, j, k));
      nv = 0.5 * (V(1, i, j, k) + V(1, i, j + 1, k));
      bv = 0.5 * (V(2, i, j - 1, k) + V(2, i, j, k));
      fv = 0.5 * (V(2, i, j - 1, k + 1) + V(2, i, j, k + 1));
    } else {
      wv = 0.5 * (V(0, i, j, k - 1) + V(0, i, j, k));
      ev = 0.5 * (V(0, i + 1, j, k - 1) + V(0, i + 1, j, k));
      sv = 0.5 * (V(1, i, j, k - 1) + V(1, i, j, k));
      nv = 0.5 * (V(1, i, j + 1, k - 1) + V(1, i, j + 1, k));
      bv = 0.5 * (V(2, i, j, k - 1) + V(2, i, j, k));
      fv = 0.5 * (V(2, i, j, k) + V(2, i, j, k + 1));
    }
    const T wf = wv * rec(0, i, j, k, wv);
    const T ef = ev * rec(0, i + 1, j, k, ev);
    const T sf = sv * rec(1, i, j, k, sv);
    const T nf = nv * rec(1, i, j + 1, k, nv);
    const T bf = bv * rec(2, i, j, k, bv);
    const T ff = fv * rec(2, i, j, k + 1, fv);
    put(i, j, k, (ef - wf) / g.dx + (nf - sf) / g.dy + (ff - bf) / g.dz);
  });
}

static void advection_component(const Vel& v, int c, int scheme, Field& out) {
  advection_component_t<double>(VelReader<double>{v}, v.c(c).L, c, scheme,
                                [&](int i, int j, int k, double a) { out(i, j, k) = a; });
}

// Tangent of the advection at v along dv with the branches frozen at v (dual numbers).
static void advection_tangent(const Vel& v, const Vel& dv, int scheme, Vel& out) {
  for(int c = 0; c < 3; ++c) {
    Field& o = out.c(c);
    advection_component_t<Dual>(VelReader<Dual>{v, dv}, v.c(c).L, c, scheme,
                                [&](int i, int j, int k, Dual a) { o(i, j, k) = a.d; });
  }
}

static void advection(const Vel& v, int scheme, Vel& out) {  // solver/momentum_terms.cpp:353-360
  for(int c = 0; c < 3; ++c) advection_component(v, c, scheme, out.c(c));
}

static void diffusion(const Vel& v, double nu, Vel& out) {  // solver/momentum_terms.cpp:362-371
  for(int c = 0; c < 3; ++c) {
    laplacian(v.c(c), out.c(c));
    scale_active(out.c(c), nu);
  }
}

struct Cfl {
  double dt, max_cfl, max_u, max_v, max_w;
};
static Cfl advective_cfl(const Vel& v, double dt) {  // solver/momentum_terms.cpp:373-401
  if(dt < 0.0) throw std::invalid_argument("compute_advective_cfl expects a non-negative dt");
  const Grid& g = v.x.L.g;
  Cfl d{dt, 0.0, 0.0, 0.0, 0.0};
  for_range(make_layout(g, kCell).active(), [&](int i, int j, int k) {
    const double uc = 0.5 * (v.x(i, j, k) + v.x(i + 1, j, k));
    const double vc = 0.5 * (v.y(i, j, k) + v.y(i, j + 1, k));
    const double wc = 0.5 * (v.z(i, j, k) + v.z(i, j, k + 1));
    d.max_u = std::max(d.max_u, std::abs(uc));
    d.max_v = std::max(d.max_v, std::abs(vc));
    d.max_w = std::max(d.max_w, std::abs(wc));
    const double local = dt * (std::abs(uc) / g.dx + std::abs(vc) / g.dy + std::abs(wc) / g.dz);
    d.max_cfl = std::max(d.max_cfl, local);
  });
  return d;
}

// ---------------------------------------------------------------- factorisation correction
// solver/taylor_green.cpp:132-194
static inline double sd_axis(const Field& f, int i, int j, int k, int axis) {
  const double h = spacing(f.L.g, axis);
  const double inv = 1.0 / (h * h);
  const int di = axis == 0, dj = axis == 1, dk = axis == 2;
  return (f(i + di, j + dj, k + dk) - 2.0 * f(i, j, k) + f(i - di, j - dj, k - dk)) * inv;
}
static inline double mixed_sd(const Field& f, int i, int j, int k, int outer, int inner) {
  const double h = spacing(f.L.g, outer);
  const double inv = 1.0 / (h * h);
  const int di = outer == 0, dj = outer == 1, dk = outer == 2;
  return (sd_axis(f, i + di, j + dj, k + dk, inner) - 2.0 * sd_axis(f, i, j, k, inner) +
          sd_axis(f, i - di, j - dj, k - dk, inner)) *
         inv;
}
static inline double triple_sd(const Field& f, int i, int j, int k) {
  const double inv = 1.0 / (f.L.g.dx * f.L.g.dx);
  return (mixed_sd(f, i + 1, j, k, 1, 2) - 2.0 * mixed_sd(f, i, j, k, 1, 2) + mixed_sd(f, i - 1, j, k, 1, 2)) * inv;
}

// solver/taylor_green.cpp:196-225
static void factorized_correction_tg(const Vel& v, double alpha, Vel& out) {
  out.fill(0.0);
  const bool three_d = v.x.L.g.nz != 1;
  for(int c = 0; c < 3; ++c) {
    const Field& f = v.c(c);
    Field& o = out.c(c);
    for_range(f.L.active(), [&](int i, int j, int k) {
      double corr = mixed_sd(f, i, j, k, 0, 1);
      if(three_d) {
        corr += mixed_sd(f, i, j, k, 0, 2);
        corr += mixed_sd(f, i, j, k, 1, 2);
        corr -= alpha * triple_sd(f, i, j, k);
      }
      o(i, j, k) = corr;
    });
  }
}

// solver/lid_driven_cavity.cpp:219-252
static void factorized_correction_cavity(const Vel& v, Vel& out) {
  out.fill(0.0);
  for(int c = 0; c < 3; ++c) {
    const Field& f = v.c(c);
    Field& o = out.c(c);
    const double dy2 = f.L.g.dy * f.L.g.dy;
    const double idx2 = 1.0 / (f.L.g.dx * f.L.g.dx);
    for_range(f.L.active(), [&](int i, int j, int k) {
      const double a = (f(i - 1, j + 1, k) - 2.0 * f(i - 1, j, k) + f(i - 1, j - 1, k)) / dy2;
      const double b = (f(i, j + 1, k) - 2.0 * f(i, j, k) + f(i, j - 1, k)) / dy2;
      const double c2 = (f(i + 1, j + 1, k) - 2.0 * f(i + 1, j, k) + f(i + 1, j - 1, k)) / dy2;
      o(i, j, k) = (c2 - 2.0 * b + a) * idx2;
    });
  }
}

static void axpy_vel(Vel& d, const Vel& s, double scale) {
  for(int c = 0; c < 3; ++c) axpy_active(d.c(c), s.c(c), scale);
}

// solver/taylor_green.cpp:227-256 (cavity variant solver/lid_driven_cavity.cpp:489-518)
// The terms after the advection, with the advection `adv` given. Linear in (cur, p_total,
// prev, adv), so it also assembles the tangent from the perturbations.
static void assemble_predictor_rhs(const Vel& cur, const Field& p_total, const Vel* prev, const Vel& adv, double nu,
                                   double dt, bool cavity, Vel& rhs);

static void predictor_rhs(const Vel& cur, const Field& p_total, const Vel* prev, int scheme, double nu, double dt,
                          bool cavity, Vel& adv, Vel& rhs) {
  advection(cur, scheme, adv);
  assemble_predictor_rhs(cur, p_total, prev, adv, nu, dt, cavity, rhs);
}

static void assemble_predictor_rhs(const Vel& cur, const Field& p_total, const Vel* prev, const Vel& adv, double nu,
                                   double dt, bool cavity, Vel& rhs) {
  const Grid& g = cur.x.L.g;
  Vel dif(g), grad(g), fc(g);
  diffusion(cur, nu, dif);
  gradient(p_total, grad);
  const double alpha = 0.5 * nu * dt;
  if(cavity) factorized_correction_cavity(cur, fc);
  else factorized_correction_tg(cur, alpha, fc);
  rhs = cur;
  axpy_vel(rhs, grad, -dt);
  axpy_vel(rhs, dif, 0.5 * dt);
  if(prev != nullptr) {
    axpy_vel(rhs, adv, -1.5 * dt);
    axpy_vel(rhs, *prev, 0.5 * dt);
  } else {
    axpy_vel(rhs, adv, -dt);
  }
  axpy_vel(rhs, fc, alpha * alpha);
}

// ---------------------------------------------------------------- ADI Helmholtz predictor
constexpr double kSmallNumber = 1.0e-30;

// solver/projection.cpp:522-559 (Thomas algorithm)
static void solve_tridiagonal(const std::vector<double>& lo, const std::vector<double>& di,
                              const std::vector<double>& up, const std::vector<double>& rhs, std::vector<double>& x) {
  const size_t n = di.size();
  if(n == 0) { x.clear(); return; }
  std::vector<double> mu(n, 0.0), mr(n, 0.0);
  x.assign(n, 0.0);
  double piv = di[0];
  if(std::abs(piv) < kSmallNumber) throw std::runtime_error("tridiagonal solve encountered a zero pivot");
  mu[0] = n > 1 ? up[0] / piv : 0.0;
  mr[0] = rhs[0] / piv;
  for(size_t i = 1; i < n; ++i) {
    piv = di[i] - lo[i] * mu[i - 1];
    if(std::abs(piv) < kSmallNumber) throw std::runtime_error("tridiagonal solve encountered a zero pivot");
    mu[i] = i + 1 < n ? up[i] / piv : 0.0;
    mr[i] = (rhs[i] - lo[i] * mr[i - 1]) / piv;
  }
  x[n - 1] = mr[n - 1];
  for(size_t i = n - 1; i-- > 0;) x[i] = mr[i] - mu[i] * x[i + 1];
}

// solver/projection.cpp:561-617 (Sherman-Morrison cyclic variant)
static void solve_cyclic_tridiagonal(const std::vector<double>& lo, const std::vector<double>& di,
                                     const std::vector<double>& up, double lc, double uc,
                                     const std::vector<double>& rhs, std::vector<double>& sol) {
  const size_t n = di.size();
  if(n == 0) { sol.clear(); return; }
  if(n == 1) {
    const double piv = di[0] + lc + uc;
    if(std::abs(piv) < kSmallNumber) throw std::runtime_error("cyclic tridiagonal solve encountered a zero pivot");
    sol.assign(1, rhs[0] / piv);
    return;
  }
  std::vector<double> dm = di;
  const double gamma = -di[0];
  if(std::abs(gamma) < kSmallNumber) throw std::runtime_error("cyclic tridiagonal solve encountered a zero gamma");
  dm[0] = di[0] - gamma;
  dm[n - 1] = di[n - 1] - lc * uc / gamma;
  std::vector<double> lm = lo, um = up;
  lm[0] = 0.0;
  um[n - 1] = 0.0;
  std::vector<double> x, z, u(n, 0.0);
  solve_tridiagonal(lm, dm, um, rhs, x);
  u[0] = gamma;
  u[n - 1] = lc;
  solve_tridiagonal(lm, dm, um, u, z);
  const double den = 1.0 + z[0] + uc * z[n - 1] / gamma;
  if(std::abs(den) < kSmallNumber)
    throw std::runtime_error("cyclic tridiagonal solve encountered a singular Sherman-Morrison update");
  const double fac = (x[0] + uc * x[n - 1] / gamma) / den;
  sol.resize(n);
  for(size_t i = 0; i < n; ++i) sol[i] = x[i] - fac * z[i];
}

static inline int axis_b(const Range& r, int a) { return a == 0 ? r.ib : a == 1 ? r.jb : r.kb; }
static inline int axis_e(const Range& r, int a) { return a == 0 ? r.ie : a == 1 ? r.je : r.ke; }

// Visit every line along `axis` of the active range; `fn(i,j,k)` gets the line origin.
// Loop order follows solver/projection.cpp:667-718 / 753-865.
template <typename F>
static void for_lines(const Range& act, int axis, F&& fn) {
  if(axis == 0) {
    for(int k = act.kb; k < act.ke; ++k)
      for(int j = act.jb; j < act.je; ++j) fn(0, j, k);
  } else if(axis == 1) {
    for(int k = act.kb; k < act.ke; ++k)
      for(int i = act.ib; i < act.ie; ++i) fn(i, 0, k);
  } else {
    for(int j = act.jb; j < act.je; ++j)
      for(int i = act.ib; i < act.ie; ++i) fn(i, j, 0);
  }
}

static inline double& line_at(Field& f, int i, int j, int k, int axis, int t) {
  return axis == 0 ? f(t, j, k) : axis == 1 ? f(i, t, k) : f(i, j, t);
}

// solver/projection.cpp:619-720
static void sweep_periodic(const Field& in, int axis, double alpha, Field& out, int& line_solves) {
  const double h = spacing(in.L.g, axis);
  const double beta = alpha / (h * h);
  const Range act = in.L.active();
  const int b = axis_b(act, axis);
  const int ae = axis_e(act, axis);
  const int e = in.L.normal() == axis ? ae - 1 : ae;
  const int n = e - b;
  if(n <= 0) return;
  std::vector<double> lo(n, -beta), di(n, 1.0 + 2.0 * beta), up(n, -beta), rhs(n, 0.0), sol;
  lo[0] = 0.0;
  up[n - 1] = 0.0;
  Field& inm = const_cast<Field&>(in);
  for_lines(act, axis, [&](int i, int j, int k) {
    for(int t = b; t < e; ++t) rhs[t - b] = line_at(inm, i, j, k, axis, t);
    solve_cyclic_tridiagonal(lo, di, up, -beta, -beta, rhs, sol);
    for(int t = b; t < e; ++t) line_at(out, i, j, k, axis, t) = sol[t - b];
    if(in.L.normal() == axis) line_at(out, i, j, k, axis, ae - 1) = line_at(out, i, j, k, axis, b);
    ++line_solves;
  });
}

struct GhostRelation {
  double coefficient, constant;
};
// solver/projection.cpp:60-80
static GhostRelation tangential_relation(const BC& c, int comp) {
  switch(c.type) {
    case kNoSlip:
    case kPrescribed: return GhostRelation{-1.0, 2.0 * c.velocity[comp]};
    case kSymmetry:
    case kFixedPressure: return GhostRelation{1.0, 0.0};
    default: return GhostRelation{0.0, 0.0};
  }
}

// solver/projection.cpp:464-520 + 722-866
static void sweep(const Field& in, int axis, const BCSet& bcs, double alpha, Field& out, int& line_solves) {
  if(velocity_axis_periodic(bcs, axis)) {
    sweep_periodic(in, axis, alpha, out, line_solves);
    return;
  }
  const double h = spacing(in.L.g, axis);
  const double beta = alpha / (h * h);
  const Range act = in.L.active();
  const int comp = in.L.normal();
  int b = axis_b(act, axis), e = axis_e(act, axis);
  if(comp == axis) {
    if(dirichlet_velocity(bcs.f[2 * axis])) b += 1;
    if(dirichlet_velocity(bcs.f[2 * axis + 1])) e -= 1;
  }
  const int n = e - b;
  if(n <= 0) return;
  std::vector<double> lo(n, -beta), di(n, 1.0 + 2.0 * beta), up(n, -beta), rhs(n, 0.0), sol;
  lo[0] = 0.0;
  up[n - 1] = 0.0;
  Field& inm = const_cast<Field&>(in);
  for_lines(act, axis, [&](int i, int j, int k) {
    std::fill(di.begin(), di.end(), 1.0 + 2.0 * beta);
    for(int t = b; t < e; ++t) rhs[t - b] = line_at(inm, i, j, k, axis, t);
    const BC& cl = bcs.f[2 * axis];
    const BC& cu = bcs.f[2 * axis + 1];
    if(comp == axis && dirichlet_velocity(cl)) {
      rhs[0] += beta * cl.velocity[axis];
    } else {
      const GhostRelation r = tangential_relation(cl, comp);
      di[0] -= beta * r.coefficient;
      rhs[0] += beta * r.constant;
    }
    if(comp == axis && dirichlet_velocity(cu)) {
      rhs[n - 1] += beta * cu.velocity[axis];
    } else {
      const GhostRelation r = tangential_relation(cu, comp);
      di[n - 1] -= beta * r.coefficient;
      rhs[n - 1] += beta * r.constant;
    }
    solve_tridiagonal(lo, di, up, rhs, sol);
    for(int t = b; t < e; ++t) line_at(out, i, j, k, axis, t) = sol[t - b];
    ++line_solves;
  });
}

// solver/projection.cpp:868-888
static void component_adi(Field& f, const BCSet& bcs, double alpha, int& line_solves) {
  velocity_component_boundaries(bcs, f);
  if(alpha == 0.0) return;
  Field working(f);
  Field next(f.L);
  for(int axis = 0; axis < 3; ++axis) {
    next = working;
    sweep(working, axis, bcs, alpha, next, line_solves);
    velocity_component_boundaries(bcs, next);
    working = next;
  }
  f = working;
}

// solver/projection.cpp:1028-1047
static int predictor_adi(const Vel& rhs, double alpha, const BCSet& bcs, Vel& pred) {
  if(alpha < 0.0) throw std::invalid_argument("solve_predictor_adi expects a non-negative alpha");
  pred = rhs;
  int line_solves = 0;
  for(int c = 0; c < 3; ++c) component_adi(pred.c(c), bcs, alpha, line_solves);
  return line_solves;
}

// ---------------------------------------------------------------- MGPCG pressure Poisson
struct PoissonDiag {
  int iterations;
  double initial_residual_l2, final_residual_l2, relative_residual, removed_rhs_mean;
  int multigrid_levels, coarse_unknowns, pre_smoothing_steps, post_smoothing_steps;
  int converged, zero_mean_enforced;
};

struct ProjOpts {
  double dt, density;
  int max_iterations;
  double tolerance;
};

static bool has_null_space(const PBCSet& b) {  // linsolve/poisson_solver.cpp:185-193
  for(int f = 0; f < 6; ++f)
    if(b.f[f].type == kDirichlet) return false;
  return true;
}

// linsolve/poisson_solver.cpp:195-269
static void poisson_apply(const Field& in, const PBCSet& b, Field& out) {
  const Grid& g = in.L.g;
  const Range act = in.L.active();
  const double h2[3] = {g.dx * g.dx, g.dy * g.dy, g.dz * g.dz};
  const double hs[3] = {g.dx, g.dy, g.dz};
  for_range(act, [&](int i, int j, int k) {
    const double c = in(i, j, k);
    double nb[3][2];
    const int co[3] = {i, j, k};
    for(int a = 0; a < 3; ++a) {
      const int lo_b = axis_b(act, a), hi_e = axis_e(act, a);
      const int di = a == 0, dj = a == 1, dk = a == 2;
      for(int s = 0; s < 2; ++s) {
        const bool upper = s == 1;
        if((!upper && co[a] > lo_b) || (upper && co[a] < hi_e - 1)) {
          nb[a][s] = upper ? in(i + di, j + dj, k + dk) : in(i - di, j - dj, k - dk);
          continue;
        }
        const PBC& bc = b.f[2 * a + (upper ? 1 : 0)];
        if(bc.type == kPresPeriodic) {
          const int t = upper ? lo_b : hi_e - 1;
          nb[a][s] = a == 0 ? in(t, j, k) : a == 1 ? in(i, t, k) : in(i, j, t);
        } else if(bc.type == kNeumann) {
          nb[a][s] = upper ? c + bc.gradient * hs[a] : c - bc.gradient * hs[a];
        } else {
          nb[a][s] = 2.0 * bc.value - c;
        }
      }
    }
    out(i, j, k) = -(nb[0][1] - 2.0 * c + nb[0][0]) / h2[0] - (nb[1][1] - 2.0 * c + nb[1][0]) / h2[1] -
                   (nb[2][1] - 2.0 * c + nb[2][0]) / h2[2];
  });
}

// linsolve/poisson_solver.cpp:271-314
static void operator_diagonal(const Layout& L, const PBCSet& b, Field& diag) {
  const Range act = L.active();
  for_range(act, [&](int i, int j, int k) {
    double v = 0.0;
    const int co[3] = {i, j, k};
    for(int a = 0; a < 3; ++a) {
      const double inv = 1.0 / std::pow(spacing(L.g, a), 2.0);
      double coef = 2.0 * inv;
      if(pressure_axis_periodic(b, a)) {
        v += coef;
        continue;
      }
      if(co[a] == axis_b(act, a)) coef += b.f[2 * a].type == kDirichlet ? inv : -inv;
      if(co[a] == axis_e(act, a) - 1) coef += b.f[2 * a + 1].type == kDirichlet ? inv : -inv;
      v += coef;
    }
    diag(i, j, k) = v;
  });
}

static inline bool axis_coarsens(int n) { return n > 1 && n % 2 == 0 && n > 2; }  // :316-345

struct Level {
  Grid g;
  Field x, b, r, s, d;  // correction, rhs, residual, scratch, diagonal
  explicit Level(const Grid& gg)
      : g(gg), x(make_layout(gg, kCell)), b(make_layout(gg, kCell)), r(make_layout(gg, kCell)),
        s(make_layout(gg, kCell)), d(make_layout(gg, kCell)) {}
};

// linsolve/poisson_solver.cpp:364-384
static std::vector<Level*> build_hierarchy(const Grid& fine, const PBCSet& hom) {
  std::vector<Level*> h;
  h.push_back(new Level(fine));
  operator_diagonal(h.back()->x.L, hom, h.back()->d);
  for(;;) {
    const Grid& g = h.back()->g;
    if(!(axis_coarsens(g.nx) || axis_coarsens(g.ny) || axis_coarsens(g.nz))) break;
    Grid c{axis_coarsens(g.nx) ? g.nx / 2 : g.nx, axis_coarsens(g.ny) ? g.ny / 2 : g.ny,
           axis_coarsens(g.nz) ? g.nz / 2 : g.nz, axis_coarsens(g.nx) ? 2.0 * g.dx : g.dx,
           axis_coarsens(g.ny) ? 2.0 * g.dy : g.dy, axis_coarsens(g.nz) ? 2.0 * g.dz : g.dz};
    if(c.nx == g.nx && c.ny == g.ny && c.nz == g.nz) break;
    h.push_back(new Level(c));
    operator_diagonal(h.back()->x.L, hom, h.back()->d);
  }
  return h;
}

// linsolve/poisson_solver.cpp:397-424 (average of the children)
static void restrict_avg(const Field& fine, Field& coarse) {
  const int rx = fine.L.a[0] == coarse.L.a[0] ? 1 : 2, ry = fine.L.a[1] == coarse.L.a[1] ? 1 : 2,
            rz = fine.L.a[2] == coarse.L.a[2] ? 1 : 2;
  const double inv = 1.0 / (double)(rx * ry * rz);
  for(int kc = 0; kc < coarse.L.a[2]; ++kc)
    for(int jc = 0; jc < coarse.L.a[1]; ++jc)
      for(int ic = 0; ic < coarse.L.a[0]; ++ic) {
        double s = 0.0;
        for(int dz = 0; dz < rz; ++dz)
          for(int dy = 0; dy < ry; ++dy)
            for(int dx = 0; dx < rx; ++dx) s += fine(1 + ic * rx + dx, 1 + jc * ry + dy, 1 + kc * rz + dz);
        coarse(1 + ic, 1 + jc, 1 + kc) = inv * s;
      }
}

// linsolve/poisson_solver.cpp:426-451 (piecewise-constant injection-add)
static void prolong_add(const Field& coarse, Field& fine) {
  const int rx = fine.L.a[0] == coarse.L.a[0] ? 1 : 2, ry = fine.L.a[1] == coarse.L.a[1] ? 1 : 2,
            rz = fine.L.a[2] == coarse.L.a[2] ? 1 : 2;
  for(int kc = 0; kc < coarse.L.a[2]; ++kc)
    for(int jc = 0; jc < coarse.L.a[1]; ++jc)
      for(int ic = 0; ic < coarse.L.a[0]; ++ic) {
        const double cv = coarse(1 + ic, 1 + jc, 1 + kc);
        for(int dz = 0; dz < rz; ++dz)
          for(int dy = 0; dy < ry; ++dy)
            for(int dx = 0; dx < rx; ++dx) fine(1 + ic * rx + dx, 1 + jc * ry + dy, 1 + kc * rz + dz) += cv;
      }
}

// linsolve/poisson_solver.cpp:482-528
static void gaussian_elimination(std::vector<double>& m, std::vector<double>& rhs) {
  const size_t n = rhs.size();
  for(size_t piv = 0; piv < n; ++piv) {
    size_t best = piv;
    double best_abs = std::abs(m[piv * n + piv]);
    for(size_t row = piv + 1; row < n; ++row) {
      const double cand = std::abs(m[row * n + piv]);
      if(cand > best_abs) { best = row; best_abs = cand; }
    }
    if(best_abs < kSmallNumber) throw std::runtime_error("direct coarse solve encountered a singular matrix");
    if(best != piv) {
      for(size_t c = 0; c < n; ++c) std::swap(m[piv * n + c], m[best * n + c]);
      std::swap(rhs[piv], rhs[best]);
    }
    const double ip = 1.0 / m[piv * n + piv];
    for(size_t row = piv + 1; row < n; ++row) {
      const double f = m[row * n + piv] * ip;
      if(std::abs(f) < kSmallNumber) continue;
      for(size_t c = piv; c < n; ++c) m[row * n + c] -= f * m[piv * n + c];
      rhs[row] -= f * rhs[piv];
    }
  }
  for(size_t row = n; row-- > 0;) {
    double s = rhs[row];
    for(size_t c = row + 1; c < n; ++c) s -= m[row * n + c] * rhs[c];
    rhs[row] = s / m[row * n + row];
  }
}

static std::vector<double> flatten_active(const Field& f) {
  std::vector<double> v;
  v.reserve((size_t)f.L.a[0] * f.L.a[1] * f.L.a[2]);
  for_range(f.L.active(), [&](int i, int j, int k) { v.push_back(f(i, j, k)); });
  return v;
}
static void unflatten_active(const std::vector<double>& v, Field& f) {
  size_t c = 0;
  for_range(f.L.active(), [&](int i, int j, int k) { f(i, j, k) = v[c++]; });
}

// linsolve/poisson_solver.cpp:530-567
static void direct_coarse_solve(Level& lv, const PBCSet& b, bool zero_mean) {
  const std::vector<double> rv = flatten_active(lv.b);
  const size_t n = rv.size();
  std::vector<double> m(n * n, 0.0), sr = rv;
  Field basis(lv.x.L), image(lv.x.L);
  for(size_t col = 0; col < n; ++col) {
    basis.fill(0.0);
    std::vector<double> bv(n, 0.0);
    bv[col] = 1.0;
    unflatten_active(bv, basis);
    poisson_apply(basis, b, image);
    const std::vector<double> iv = flatten_active(image);
    for(size_t row = 0; row < n; ++row) m[row * n + col] = iv[row];
  }
  if(zero_mean) {
    for(size_t col = 0; col < n; ++col) m[(n - 1) * n + col] = 1.0;
    sr[n - 1] = 0.0;
  }
  gaussian_elimination(m, sr);
  lv.x.fill(0.0);
  unflatten_active(sr, lv.x);
  if(zero_mean) subtract_active_mean(lv.x);
}

constexpr int kPre = 2, kPost = 2;  // linsolve/poisson_solver.hpp:17-23
constexpr double kOmega = 2.0 / 3.0;

// linsolve/poisson_solver.cpp:569-595
static void smooth_jacobi(Level& lv, const PBCSet& b, int iters, bool zero_mean) {
  for(int it = 0; it < iters; ++it) {
    poisson_apply(lv.x, b, lv.s);
    lv.r = lv.b;
    axpy_active(lv.r, lv.s, -1.0);
    for_range(lv.x.L.active(), [&](int i, int j, int k) {
      lv.s(i, j, k) = lv.x(i, j, k) + kOmega * lv.r(i, j, k) / lv.d(i, j, k);
    });
    copy_active(lv.s, lv.x);
    if(zero_mean) subtract_active_mean(lv.x);
  }
}

// linsolve/poisson_solver.cpp:597-633
static void v_cycle(std::vector<Level*>& h, size_t l, const PBCSet& b, bool zero_mean) {
  Level& lv = *h[l];
  lv.x.fill(0.0);
  if(l + 1 == h.size()) {
    direct_coarse_solve(lv, b, zero_mean);
    return;
  }
  smooth_jacobi(lv, b, kPre, zero_mean);
  poisson_apply(lv.x, b, lv.s);
  lv.r = lv.b;
  axpy_active(lv.r, lv.s, -1.0);
  if(zero_mean) subtract_active_mean(lv.r);
  Level& cl = *h[l + 1];
  cl.b.fill(0.0);
  cl.x.fill(0.0);
  restrict_avg(lv.r, cl.b);
  if(zero_mean) subtract_active_mean(cl.b);
  v_cycle(h, l + 1, b, zero_mean);
  prolong_add(cl.x, lv.x);
  if(zero_mean) subtract_active_mean(lv.x);
  smooth_jacobi(lv, b, kPost, zero_mean);
}

// linsolve/poisson_solver.cpp:678-809
static PoissonDiag solve_poisson(const Field& rhs, const PBCSet& bcs, const ProjOpts& o, Field& p,
                                 std::vector<double>* history) {
  if(o.max_iterations <= 0) throw std::invalid_argument("solve_pressure_poisson expects iterations > 0");
  if(o.tolerance <= 0.0) throw std::invalid_argument("solve_pressure_poisson expects tolerance > 0");
  PBCSet hom = bcs;
  for(int f = 0; f < 6; ++f) { hom.f[f].value = 0.0; hom.f[f].gradient = 0.0; }
  const bool zm = has_null_space(bcs);
  const Layout L = p.L;
  Field sys(L), r(L), d(L), q(L), z(L), src(L), zero(L);
  poisson_apply(zero, bcs, src);  // :635-642
  copy_active(rhs, sys);
  scale_active(sys, -1.0);
  axpy_active(sys, src, -1.0);
  PoissonDiag dg{};
  dg.zero_mean_enforced = zm;
  dg.pre_smoothing_steps = kPre;
  dg.post_smoothing_steps = kPost;
  std::vector<Level*> h = build_hierarchy(L.g, hom);
  dg.multigrid_levels = (int)h.size();
  dg.coarse_unknowns = h.back()->g.nx * h.back()->g.ny * h.back()->g.nz;
  struct Cleanup {
    std::vector<Level*>& h;
    ~Cleanup() { for(Level* l : h) delete l; }
  } cleanup{h};
  if(zm) {
    dg.removed_rhs_mean = subtract_active_mean(sys);
    subtract_active_mean(p);
  }
  poisson_apply(p, hom, q);
  r = sys;
  axpy_active(r, q, -1.0);
  if(zm) subtract_active_mean(r);
  dg.initial_residual_l2 = active_l2(r);
  dg.final_residual_l2 = dg.initial_residual_l2;
  dg.relative_residual = dg.initial_residual_l2 > 0.0 ? 1.0 : 0.0;
  if(history) history->push_back(dg.initial_residual_l2);
  if(dg.initial_residual_l2 <= o.tolerance) {
    dg.converged = 1;
    dg.relative_residual = 0.0;
    return dg;
  }
  h.front()->b = r;
  v_cycle(h, 0, hom, zm);
  z = h.front()->x;
  if(zm) subtract_active_mean(z);
  d = z;
  double rz = dot_active(r, z);
  for(int it = 0; it < o.max_iterations; ++it) {
    poisson_apply(d, hom, q);
    if(zm) subtract_active_mean(q);
    const double dq = dot_active(d, q);
    if(std::abs(dq) < kSmallNumber) break;
    const double alpha = rz / dq;
    axpy_active(p, d, alpha);
    if(zm) subtract_active_mean(p);
    axpy_active(r, q, -alpha);
    if(zm) subtract_active_mean(r);
    dg.iterations = it + 1;
    dg.final_residual_l2 = active_l2(r);
    dg.relative_residual = dg.initial_residual_l2 > 0.0 ? dg.final_residual_l2 / dg.initial_residual_l2 : 0.0;
    if(history) history->push_back(dg.final_residual_l2);
    if(dg.relative_residual <= o.tolerance) {
      dg.converged = 1;
      break;
    }
    h.front()->b = r;
    v_cycle(h, 0, hom, zm);
    z = h.front()->x;
    if(zm) subtract_active_mean(z);
    const double nd = dot_active(r, z);
    const double beta = nd / rz;
    for_range(d.L.active(), [&](int i, int j, int k) { d(i, j, k) = z(i, j, k) + beta * d(i, j, k); });
    if(zm) subtract_active_mean(d);
    rz = nd;
  }
  return dg;
}

// linsolve/poisson_solver.cpp:668-676
static void poisson_rhs_from_pressure(const Field& p, const PBCSet& b, Field& rhs) {
  Field sr(p.L);
  poisson_apply(p, b, sr);
  copy_active(sr, rhs);
  scale_active(rhs, -1.0);
}

// ---------------------------------------------------------------- projection
// solver/projection.cpp:1049-1064
static void build_pressure_rhs(const Vel& pred, const BCSet& bcs, const ProjOpts& o, Field& rhs) {
  if(o.dt <= 0.0) throw std::invalid_argument("build_pressure_rhs expects dt > 0");
  if(o.density <= 0.0) throw std::invalid_argument("build_pressure_rhs expects density > 0");
  Vel w(pred.x.L.g);
  w = pred;
  apply_velocity_bc(bcs, w);
  divergence(w, rhs);
  scale_active(rhs, o.density / o.dt);
}

// solver/projection.cpp:1066-1095
static void correct_velocity(const Vel& pred, const Field& phi, const BCSet& bcs, const ProjOpts& o, Vel& corr) {
  if(o.dt <= 0.0) throw std::invalid_argument("correct_velocity expects dt > 0");
  if(o.density <= 0.0) throw std::invalid_argument("correct_velocity expects density > 0");
  Field pg(phi);
  apply_pressure_bc(derive_pressure_correction_bc(bcs), pg);
  Vel grad(phi.L.g);
  gradient(pg, grad);
  corr = pred;
  for(int c = 0; c < 3; ++c) axpy_active(corr.c(c), grad.c(c), -o.dt / o.density);
  apply_velocity_bc(bcs, corr);
}

struct ProjDiag {
  double rhs_l2, divergence_l2_before, divergence_l2_after, pressure_mean;
  PoissonDiag poisson;
};

// solver/projection.cpp:1097-1131
static ProjDiag project_velocity(const Vel& pred, const BCSet& bcs, const ProjOpts& o, Field& phi, Vel& corr,
                                 Field* prhs, std::vector<double>* history) {
  ProjDiag dg{};
  const Grid& g = phi.L.g;
  const PBCSet pb = derive_pressure_correction_bc(bcs);
  Field rhs(make_layout(g, kCell)), div(make_layout(g, kCell));
  Vel w(g);
  w = pred;
  apply_velocity_bc(bcs, w);
  divergence(w, div);
  dg.divergence_l2_before = active_l2(div);
  build_pressure_rhs(w, bcs, o, rhs);
  dg.rhs_l2 = active_l2(rhs);
  dg.poisson = solve_poisson(rhs, pb, o, phi, history);
  correct_velocity(w, phi, bcs, o, corr);
  divergence(corr, div);
  dg.divergence_l2_after = active_l2(div);
  dg.pressure_mean = active_mean(phi);
  if(prhs) copy_active(rhs, *prhs);
  return dg;
}

// ---------------------------------------------------------------- drivers
struct StepMetrics {
  int step;
  double time, dt, max_cfl, max_velocity_change, divergence_l2, max_divergence_l2;
  int pressure_iterations;
  double pressure_relative_residual;
  double divergence_linf;  // not in the reference metrics; reported for the GPU comparison
};

static double max_velocity_change(const Vel& a, const Vel& b) {  // solver/taylor_green.cpp:100-118
  double ch = 0.0;
  for(int c = 0; c < 3; ++c)
    for_range(a.c(c).L.active(), [&](int i, int j, int k) { ch = std::max(ch, std::abs(a.c(c)(i, j, k) - b.c(c)(i, j, k))); });
  return ch;
}

static double divergence_linf(const Vel& v) {
  Field d(make_layout(v.x.L.g, kCell));
  divergence(v, d);
  double m = 0.0;
  for_range(d.L.active(), [&](int i, int j, int k) { m = std::max(m, std::abs(d(i, j, k))); });
  return m;
}

// One fractional step shared by solver/taylor_green.cpp:577-633 and
// solver/lid_driven_cavity.cpp:562-623.
static void fractional_step(const BCSet& bcs, double nu, double dt, int scheme, bool cavity, int max_it, double tol,
                            Vel& velocity, Vel& adv_prev, Field& p_total, int& has_prev, StepMetrics& m,
                            const char* name, bool want_linf, const Vel* force = nullptr) {
  const Grid& g = velocity.x.L.g;
  const double alpha = 0.5 * nu * dt;
  const ProjOpts o{dt, 1.0, max_it, tol};
  Vel cur(g), adv(g), rhs(g), pred(g), corr(g), dif(g);
  Field phi(make_layout(g, kCell)), prhs(make_layout(g, kCell));
  cur = velocity;
  apply_velocity_bc(bcs, cur);
  predictor_rhs(cur, p_total, has_prev ? &adv_prev : nullptr, scheme, nu, dt, cavity, adv, rhs);
  // body force of this project (the reference has none): rhs += dt f after the reference's terms
  if(force) axpy_vel(rhs, *force, dt);
  pred = rhs;
  predictor_adi(rhs, alpha, bcs, pred);
  phi.fill(0.0);
  const ProjDiag pd = project_velocity(pred, bcs, o, phi, corr, &prhs, nullptr);
  if(!pd.poisson.converged) {  // solver/projection.cpp:1133-1147
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "%s pressure solve did not converge at step %d (iterations=%d, relative_residual=%g, tolerance=%g)",
                  name, m.step + 1, pd.poisson.iterations, pd.poisson.relative_residual, tol);
    throw std::runtime_error(buf);
  }
  axpy_active(p_total, phi, 1.0);
  axpy_active(p_total, prhs, -0.5 * nu * dt);
  diffusion(corr, nu, dif);
  apply_total_pressure_bc(bcs, dif, p_total);
  const Cfl cfl = advective_cfl(corr, dt);
  const double delta = max_velocity_change(velocity, corr);
  if(!std::isfinite(delta) || !std::isfinite(pd.divergence_l2_after))
    throw std::runtime_error(std::string(name) + " solve produced a non-finite state");
  velocity = corr;
  adv_prev = adv;
  has_prev = 1;
  StepMetrics n{};
  n.step = m.step + 1;
  n.time = m.time + dt;
  n.dt = dt;
  n.max_cfl = cfl.max_cfl;
  n.max_velocity_change = delta;
  n.divergence_l2 = pd.divergence_l2_after;
  n.max_divergence_l2 = std::max(m.max_divergence_l2, pd.divergence_l2_after);
  n.pressure_iterations = pd.poisson.iterations;
  n.pressure_relative_residual = pd.poisson.relative_residual;
  n.divergence_linf = want_linf ? std::max(m.divergence_linf, divergence_linf(corr)) : 0.0;
  m = n;
}

// Tangent-linear model of fractional_step for periodic boundaries (every ghost relation is then
// homogeneous and linear): advances the perturbation (dvel, dadv_prev, dp) of the state along
// with a force perturbation dforce, with the advection linearised at the base state; the base
// state is then advanced by fractional_step. Every other stage is linear and is applied to the
// perturbation as it is to the state.
static void fractional_step_tl(const BCSet& bcs, double nu, double dt, int scheme, int max_it, double tol,
                               Vel& velocity, Vel& adv_prev, Field& p_total, int& has_prev, StepMetrics& m,
                               const Vel* force, Vel& dvel, Vel& dadv_prev, Field& dp, const Vel* dforce) {
  for(int f = 0; f < 6; ++f)
    if(bcs.f[f].type != kPeriodic) throw std::invalid_argument("the tangent-linear step needs periodic boundaries");
  const Grid& g = velocity.x.L.g;
  const double alpha = 0.5 * nu * dt;
  const ProjOpts o{dt, 1.0, max_it, tol};
  Vel cur(g), dcur(g), dadv(g), drhs(g), dpred(g), dcorr(g), dif(g);
  Field dphi(make_layout(g, kCell)), dprhs(make_layout(g, kCell));
  cur = velocity;
  apply_velocity_bc(bcs, cur);
  dcur = dvel;
  apply_velocity_bc(bcs, dcur);
  advection_tangent(cur, dcur, scheme, dadv);
  assemble_predictor_rhs(dcur, dp, has_prev ? &dadv_prev : nullptr, dadv, nu, dt, false, drhs);
  if(dforce) axpy_vel(drhs, *dforce, dt);
  dpred = drhs;
  predictor_adi(drhs, alpha, bcs, dpred);
  dphi.fill(0.0);
  const ProjDiag pd = project_velocity(dpred, bcs, o, dphi, dcorr, &dprhs, nullptr);
  if(!pd.poisson.converged) throw std::runtime_error("tangent-linear pressure solve did not converge");
  axpy_active(dp, dphi, 1.0);
  axpy_active(dp, dprhs, -0.5 * nu * dt);
  diffusion(dcorr, nu, dif);
  apply_total_pressure_bc(bcs, dif, dp);
  dvel = dcorr;
  dadv_prev = dadv;
  fractional_step(bcs, nu, dt, scheme, false, max_it, tol, velocity, adv_prev, p_total, has_prev, m, "Taylor-Green",
                  false, force);
}

static double pi() { return std::acos(-1.0); }

struct TGConfig {
  int nx, ny, nz;
  double viscosity, cfl_limit, final_time;
  int max_iterations;
  double tolerance;
};

static Grid tg_grid(const TGConfig& c) {  // solver/taylor_green.cpp:497-508
  const double L = 2.0 * pi();
  return Grid{c.nx, c.ny, c.nz, L / (double)c.nx, L / (double)c.ny, c.nz > 1 ? L / (double)c.nz : 1.0};
}
static BCSet tg_bcs(const TGConfig& c) {  // solver/taylor_green.cpp:510-517
  BCSet b{};
  for(int f = 0; f < 6; ++f) b.f[f] = BC{kPeriodic, {0, 0, 0}, 0, 0};
  if(c.nz <= 1) { b.f[4].type = kSymmetry; b.f[5].type = kSymmetry; }
  return b;
}
static double tg_dt(const TGConfig& c) {  // solver/taylor_green.cpp:397-401,519-525
  const Grid g = tg_grid(c);
  const double inv = 1.0 / g.dx + 1.0 / g.dy + (g.nz == 1 ? 0.0 : 1.0 / g.dz);
  const double upper = c.cfl_limit / inv;
  const int steps = std::max(1, (int)std::ceil(c.final_time / upper));
  return c.final_time / (double)steps;
}

static inline double coord(const Layout& L, int axis, int s) {  // core/field_layout.hpp:366-377
  const int sh = s - 1;
  if(L.loc != kCell && axis == L.normal()) return (double)sh * spacing(L.g, axis);
  return ((double)sh + 0.5) * spacing(L.g, axis);
}

// solver/taylor_green.cpp:335-395. convention 0: the reference's fields; convention 1: the
// same fields shifted by (pi/2, pi/2): u = sin x cos y, v = -cos x sin y, p = +1/4(...).
static void tg_init(const TGConfig& c, int convention, Vel& v, Field& p) {
  const bool d3 = c.nz > 1;
  const double decay = std::exp(-(d3 ? 3.0 : 2.0) * c.viscosity * 0.0);
  const double sgn = convention == 0 ? 1.0 : -1.0;
  for_range(v.x.L.active(), [&](int i, int j, int k) {
    const double x = coord(v.x.L, 0, i), y = coord(v.x.L, 1, j), z = coord(v.x.L, 2, k);
    const double zf = d3 ? std::cos(z) : 1.0;
    v.x(i, j, k) = convention == 0 ? -std::cos(x) * std::sin(y) * zf * decay : std::sin(x) * std::cos(y) * zf * decay;
  });
  for_range(v.y.L.active(), [&](int i, int j, int k) {
    const double x = coord(v.y.L, 0, i), y = coord(v.y.L, 1, j), z = coord(v.y.L, 2, k);
    const double zf = d3 ? std::cos(z) : 1.0;
    v.y(i, j, k) = convention == 0 ? std::sin(x) * std::cos(y) * zf * decay : -std::cos(x) * std::sin(y) * zf * decay;
  });
  v.z.fill(0.0);
  const double pdecay = std::exp((d3 ? -6.0 : -4.0) * c.viscosity * 0.0);
  for_range(p.L.active(), [&](int i, int j, int k) {
    const double x = coord(p.L, 0, i), y = coord(p.L, 1, j), z = coord(p.L, 2, k);
    if(d3) p(i, j, k) = -sgn * 0.0625 * (std::cos(2.0 * x) + std::cos(2.0 * y)) * (std::cos(2.0 * z) + 2.0) * pdecay;
    else p(i, j, k) = -sgn * 0.25 * (std::cos(2.0 * x) + std::cos(2.0 * y)) * pdecay;
  });
  const BCSet b = tg_bcs(c);
  apply_velocity_bc(b, v);
  Vel dif(v.x.L.g);
  diffusion(v, c.viscosity, dif);
  apply_total_pressure_bc(b, dif, p);
}

static double kinetic_energy(const Vel& v) {  // solver/taylor_green.cpp:269-288
  double s = 0.0;
  size_t n = 0;
  for_range(make_layout(v.x.L.g, kCell).active(), [&](int i, int j, int k) {
    const double uc = 0.5 * (v.x(i, j, k) + v.x(i + 1, j, k));
    const double vc = 0.5 * (v.y(i, j, k) + v.y(i, j + 1, k));
    const double wc = 0.5 * (v.z(i, j, k) + v.z(i, j, k + 1));
    s += 0.5 * (uc * uc + vc * vc + wc * wc);
    ++n;
  });
  return s / (double)n;
}

struct CavityConfig {
  int nx, ny;
  double reynolds, lid_velocity, cfl_limit;
  int max_steps, min_steps;
  double steady_tolerance;
  int max_iterations;
  double tolerance;
};
static Grid cavity_grid(const CavityConfig& c) {  // solver/lid_driven_cavity.cpp:372-380
  return Grid{c.nx, c.ny, 1, 1.0 / (double)c.nx, 1.0 / (double)c.ny, 1.0};
}
static BCSet cavity_bcs(const CavityConfig& c) {  // solver/lid_driven_cavity.cpp:362-370
  BCSet b{};
  for(int f = 0; f < 6; ++f) b.f[f] = BC{kNoSlip, {0, 0, 0}, 0, 0};
  b.f[3].type = kPrescribed;
  b.f[3].velocity[0] = c.lid_velocity;
  b.f[4].type = kSymmetry;
  b.f[5].type = kSymmetry;
  return b;
}
static double cavity_dt(const CavityConfig& c) {  // solver/lid_driven_cavity.cpp:95-98,386-388
  const Grid g = cavity_grid(c);
  return c.cfl_limit / (c.lid_velocity / g.dx + c.lid_velocity / g.dy);
}

// solver/channel_flow.hpp:10-28 ChannelFlowConfig (advection: TVD, the default options)
struct ChannelConfig {
  int poiseuille;  // ChannelFlowCase: 0 couette, 1 poiseuille
  int nx, ny;
  double viscosity, top_velocity, pressure_drop, cfl_limit;
  int steps, max_iterations;
  double tolerance;
};
static Grid channel_grid(const ChannelConfig& c) {  // solver/channel_flow.cpp:435-442
  return Grid{c.nx, c.ny, 1, 1.0 / (double)c.nx, 1.0 / (double)c.ny, 1.0};
}
static BCSet channel_bcs(const ChannelConfig& c) {  // solver/channel_flow.cpp:408-430
  BCSet b{};
  for(int f = 0; f < 6; ++f) b.f[f] = BC{kNoSlip, {0, 0, 0}, 0, 0};
  b.f[4].type = kSymmetry;
  b.f[5].type = kSymmetry;
  if(!c.poiseuille) {
    b.f[0].type = kPeriodic;
    b.f[1].type = kPeriodic;
    b.f[3].type = kPrescribed;
    b.f[3].velocity[0] = c.top_velocity;
  } else {
    b.f[0].type = kFixedPressure;
    b.f[0].pressure = c.pressure_drop;
    b.f[1].type = kFixedPressure;
    b.f[1].pressure = 0.0;
  }
  return b;
}
static double channel_dt(const ChannelConfig& c) {  // solver/channel_flow.cpp:142-158
  const double max_u = c.poiseuille ? c.pressure_drop / (8.0 * c.viscosity) : c.top_velocity;
  return c.cfl_limit / (max_u / channel_grid(c).dx);
}
static double channel_u(const ChannelConfig& c, double y) {  // solver/channel_flow.cpp:160-169
  return c.poiseuille ? 0.5 * c.pressure_drop * y * (1.0 - y) / c.viscosity : c.top_velocity * y;
}
static double channel_p(const ChannelConfig& c, double x) {  // solver/channel_flow.cpp:171-180
  return c.poiseuille ? c.pressure_drop * (1.0 - x) : 0.0;
}

}  // namespace ofref

// =====================================================================================
// C ABI for the Python test harness (ctypes). All field pointers address full ghosted
// storage in the layout above. Functions return 0 on success, 1 on an exception whose
// message is available from ofref_last_error().
// =====================================================================================
using namespace ofref;

static thread_local std::string g_error;

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch(const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

extern "C" {

const char* ofref_last_error() { return g_error.c_str(); }

int ofref_velocity_bc(const Grid* g, const BCSet* b, double* u, double* v, double* w) {
  return guarded([&] {
    Vel vel(*g, u, v, w);
    apply_velocity_bc(*b, vel);
  });
}

int ofref_pressure_bc(const Grid* g, const PBCSet* b, double* p) {
  return guarded([&] {
    Field f(make_layout(*g, kCell), p);
    apply_pressure_bc(*b, f);
  });
}

int ofref_total_pressure_bc(const Grid* g, const BCSet* b, double* su, double* sv, double* sw, double* p) {
  return guarded([&] {
    Vel s(*g, su, sv, sw);
    Field f(make_layout(*g, kCell), p);
    apply_total_pressure_bc(*b, s, f);
  });
}

int ofref_derive_pressure_bc(const BCSet* b, PBCSet* out) {
  return guarded([&] { *out = derive_pressure_correction_bc(*b); });
}

int ofref_gradient(const Grid* g, double* p, double* gx, double* gy, double* gz) {
  return guarded([&] {
    Field pf(make_layout(*g, kCell), p);
    Vel gr(*g, gx, gy, gz);
    gradient(pf, gr);
  });
}

int ofref_divergence(const Grid* g, double* u, double* v, double* w, double* d) {
  return guarded([&] {
    Vel vel(*g, u, v, w);
    Field df(make_layout(*g, kCell), d);
    divergence(vel, df);
  });
}

int ofref_laplacian(const Grid* g, int loc, double* in, double* out) {
  return guarded([&] {
    Field a(make_layout(*g, (Loc)loc), in), b(make_layout(*g, (Loc)loc), out);
    laplacian(a, b);
  });
}

int ofref_advection(const Grid* g, int scheme, double* u, double* v, double* w, double* au, double* av, double* aw) {
  return guarded([&] {
    Vel vel(*g, u, v, w), a(*g, au, av, aw);
    advection(vel, scheme, a);
  });
}

int ofref_diffusion(const Grid* g, double nu, double* u, double* v, double* w, double* du, double* dv, double* dw) {
  return guarded([&] {
    Vel vel(*g, u, v, w), d(*g, du, dv, dw);
    diffusion(vel, nu, d);
  });
}

int ofref_cfl(const Grid* g, double* u, double* v, double* w, double dt, double* out5) {
  return guarded([&] {
    Vel vel(*g, u, v, w);
    const Cfl c = advective_cfl(vel, dt);
    out5[0] = c.dt; out5[1] = c.max_cfl; out5[2] = c.max_u; out5[3] = c.max_v; out5[4] = c.max_w;
  });
}

int ofref_factorized_correction(const Grid* g, int cavity, double alpha, double* u, double* v, double* w, double* fu,
                                double* fv, double* fw) {
  return guarded([&] {
    Vel vel(*g, u, v, w), f(*g, fu, fv, fw);
    if(cavity) factorized_correction_cavity(vel, f);
    else factorized_correction_tg(vel, alpha, f);
  });
}

// cur_* must already carry BC ghosts (as after apply_velocity_boundary_conditions).
int ofref_predictor_rhs(const Grid* g, int cavity, int scheme, double nu, double dt, int has_prev, double* u,
                        double* v, double* w, double* p, double* pu, double* pv, double* pw, double* au, double* av,
                        double* aw, double* ru, double* rv, double* rw) {
  return guarded([&] {
    Vel cur(*g, u, v, w), prev(*g, pu, pv, pw), adv(*g, au, av, aw), rhs(*g, ru, rv, rw);
    Field pf(make_layout(*g, kCell), p);
    predictor_rhs(cur, pf, has_prev ? &prev : nullptr, scheme, nu, dt, cavity != 0, adv, rhs);
  });
}

int ofref_predictor_adi(const Grid* g, const BCSet* b, double alpha, double* ru, double* rv, double* rw, double* ou,
                        double* ov, double* ow, int* line_solves) {
  return guarded([&] {
    Vel rhs(*g, ru, rv, rw), out(*g, ou, ov, ow);
    *line_solves = predictor_adi(rhs, alpha, *b, out);
  });
}

int ofref_build_pressure_rhs(const Grid* g, const BCSet* b, const ProjOpts* o, double* u, double* v, double* w,
                             double* rhs) {
  return guarded([&] {
    Vel vel(*g, u, v, w);
    Field r(make_layout(*g, kCell), rhs);
    build_pressure_rhs(vel, *b, *o, r);
  });
}

int ofref_correct_velocity(const Grid* g, const BCSet* b, const ProjOpts* o, double* u, double* v, double* w,
                           double* phi, double* cu, double* cv, double* cw) {
  return guarded([&] {
    Vel pred(*g, u, v, w), corr(*g, cu, cv, cw);
    Field pf(make_layout(*g, kCell), phi);
    correct_velocity(pred, pf, *b, *o, corr);
  });
}

int ofref_poisson(const Grid* g, const PBCSet* b, const ProjOpts* o, double* rhs, double* p, PoissonDiag* diag,
                  double* history, int history_capacity, int* history_len) {
  return guarded([&] {
    Field r(make_layout(*g, kCell), rhs), pf(make_layout(*g, kCell), p);
    std::vector<double> h;
    *diag = solve_poisson(r, *b, *o, pf, &h);
    const int n = std::min<int>((int)h.size(), history_capacity);
    for(int i = 0; i < n; ++i) history[i] = h[i];
    *history_len = (int)h.size();
  });
}

int ofref_poisson_rhs_from_pressure(const Grid* g, const PBCSet* b, double* p, double* rhs) {
  return guarded([&] {
    Field pf(make_layout(*g, kCell), p), r(make_layout(*g, kCell), rhs);
    poisson_rhs_from_pressure(pf, *b, r);
  });
}

int ofref_poisson_apply(const Grid* g, const PBCSet* b, double* in, double* out) {
  return guarded([&] {
    Field a(make_layout(*g, kCell), in), c(make_layout(*g, kCell), out);
    poisson_apply(a, *b, c);
  });
}

int ofref_project_velocity(const Grid* g, const BCSet* b, const ProjOpts* o, double* u, double* v, double* w,
                           double* phi, double* cu, double* cv, double* cw, double* prhs, ProjDiag* diag) {
  return guarded([&] {
    Vel pred(*g, u, v, w), corr(*g, cu, cv, cw);
    Field pf(make_layout(*g, kCell), phi);
    Field rf(make_layout(*g, kCell), prhs);
    *diag = project_velocity(pred, *b, *o, pf, corr, prhs ? &rf : nullptr, nullptr);
  });
}

int ofref_tg_dt(const TGConfig* c, double* dt) {
  return guarded([&] { *dt = tg_dt(*c); });
}

int ofref_tg_init(const TGConfig* c, int convention, double* u, double* v, double* w, double* p) {
  return guarded([&] {
    const Grid g = tg_grid(*c);
    Vel vel(g, u, v, w);
    Field pf(make_layout(g, kCell), p);
    tg_init(*c, convention, vel, pf);
  });
}

// Advances a Taylor-Green state in place (solver/taylor_green.cpp:540-634). The state is
// (u, v, w, p_total, A_prev x3, has_prev, metrics); metrics are read and updated.
int ofref_tg_steps(const TGConfig* c, int steps, int want_linf, double* u, double* v, double* w, double* p,
                   double* au, double* av, double* aw, int* has_prev, StepMetrics* m) {
  return guarded([&] {
    if(steps < 0) throw std::invalid_argument("run_taylor_green_steps expects a non-negative step count");
    const Grid g = tg_grid(*c);
    const BCSet b = tg_bcs(*c);
    const double dt = tg_dt(*c);
    Vel vel(g, u, v, w), prev(g, au, av, aw);
    Field pf(make_layout(g, kCell), p);
    for(int s = 0; s < steps; ++s)
      fractional_step(b, c->viscosity, dt, 0, false, c->max_iterations, c->tolerance, vel, prev, pf, *has_prev, *m,
                      "Taylor-Green", want_linf != 0);
  });
}

// Forward run with an optional body force, and optionally the tangent-linear model alongside
// it. The initial state is prepared as the device adjoint driver prepares it: only active
// entries of (u, v, w, p) and of the perturbation (du, dv, dw, dp) are read, ghosts are filled,
// A_prev = 0. On return (u, v, w, p) and (du, dv, dw, dp) hold the final state and
// perturbation; ke_out / dke_out the final kinetic energy and its directional derivative.
int ofref_tg_run_tl(const TGConfig* c, int scheme, int steps, double* fu, double* fv, double* fw, double* u,
                    double* v, double* w, double* p, double* dfu, double* dfv, double* dfw, double* du, double* dv,
                    double* dw, double* dp, double* ke_out, double* dke_out) {
  return guarded([&] {
    if(steps < 0) throw std::invalid_argument("run_taylor_green_steps expects a non-negative step count");
    const Grid g = tg_grid(*c);
    const BCSet b = tg_bcs(*c);
    const PBCSet pb = derive_pressure_correction_bc(b);
    const double dt = tg_dt(*c);
    Vel vel(g, u, v, w), prev(g);
    Field pf(make_layout(g, kCell), p);
    auto keep_active = [](Field& f) {
      Field t(f.L);
      copy_active(f, t);
      f = t;
    };
    for(int a = 0; a < 3; ++a) keep_active(vel.c(a));
    keep_active(pf);
    apply_velocity_bc(b, vel);
    apply_pressure_bc(pb, pf);
    const bool forced = fu != nullptr;
    Vel force(g);
    if(forced) force = Vel(g, fu, fv, fw);
    int has_prev = 0;
    StepMetrics m{};
    m.dt = dt;
    const bool tl = du != nullptr;
    Vel dvel(g), dprev(g), dforce(g);
    Field dpf(make_layout(g, kCell));
    if(tl) {
      dvel = Vel(g, du, dv, dw);
      for(int a = 0; a < 3; ++a) keep_active(dvel.c(a));
      if(dp) {
        dpf = Field(make_layout(g, kCell), dp);
        keep_active(dpf);
      }
      apply_velocity_bc(b, dvel);
      apply_pressure_bc(pb, dpf);
      if(dfu) dforce = Vel(g, dfu, dfv, dfw);
    }
    for(int s = 0; s < steps; ++s) {
      if(tl)
        fractional_step_tl(b, c->viscosity, dt, scheme, c->max_iterations, c->tolerance, vel, prev, pf, has_prev, m,
                           forced ? &force : nullptr, dvel, dprev, dpf, dfu ? &dforce : nullptr);
      else
        fractional_step(b, c->viscosity, dt, scheme, false, c->max_iterations, c->tolerance, vel, prev, pf, has_prev,
                        m, "Taylor-Green", false, forced ? &force : nullptr);
    }
    if(ke_out) *ke_out = kinetic_energy(vel);
    if(tl && dke_out) {
      double s = 0.0;
      size_t n = 0;
      for_range(make_layout(g, kCell).active(), [&](int i, int j, int k) {
        const double uc = 0.5 * (vel.x(i, j, k) + vel.x(i + 1, j, k));
        const double vc = 0.5 * (vel.y(i, j, k) + vel.y(i, j + 1, k));
        const double wc = 0.5 * (vel.z(i, j, k) + vel.z(i, j, k + 1));
        const double duc = 0.5 * (dvel.x(i, j, k) + dvel.x(i + 1, j, k));
        const double dvc = 0.5 * (dvel.y(i, j, k) + dvel.y(i, j + 1, k));
        const double dwc = 0.5 * (dvel.z(i, j, k) + dvel.z(i, j, k + 1));
        s += uc * duc + vc * dvc + wc * dwc;
        ++n;
      });
      *dke_out = s / (double)n;
    }
    // (u, v, w, p) are updated in place through the views; the perturbation is copied back
    if(tl) {
      Vel dout(g, du, dv, dw);
      dout = dvel;
      if(dp) {
        Field dpo(make_layout(g, kCell), dp);
        dpo = dpf;
      }
    }
  });
}

// Linearised advection at (u, v, w) along (du, dv, dw) (dual numbers; branches frozen).
int ofref_advection_tangent(const Grid* g, int scheme, double* u, double* v, double* w, double* du, double* dv,
                            double* dw, double* au, double* av, double* aw) {
  return guarded([&] {
    Vel vel(*g, u, v, w), dvel(*g, du, dv, dw), out(*g, au, av, aw);
    advection_tangent(vel, dvel, scheme, out);
  });
}

int ofref_kinetic_energy(const Grid* g, double* u, double* v, double* w, double* out) {
  return guarded([&] {
    Vel vel(*g, u, v, w);
    *out = kinetic_energy(vel);
  });
}

int ofref_cavity_init(const CavityConfig* c, double* u, double* v, double* w, double* p, double* dt) {
  return guarded([&] {  // solver/lid_driven_cavity.cpp:390-409
    const Grid g = cavity_grid(*c);
    const BCSet b = cavity_bcs(*c);
    Vel vel(g, u, v, w);
    Field pf(make_layout(g, kCell), p);
    vel.fill(0.0);
    pf.fill(0.0);
    apply_velocity_bc(b, vel);
    Vel dif(g);
    diffusion(vel, c->lid_velocity / c->reynolds, dif);
    apply_total_pressure_bc(b, dif, pf);
    *dt = cavity_dt(*c);
  });
}

// solver/lid_driven_cavity.cpp:524-624; stop_on_steady mirrors run_lid_driven_cavity.
int ofref_cavity_steps(const CavityConfig* c, int steps, int stop_on_steady, double* u, double* v, double* w,
                       double* p, double* au, double* av, double* aw, int* has_prev, StepMetrics* m) {
  return guarded([&] {
    const Grid g = cavity_grid(*c);
    const BCSet b = cavity_bcs(*c);
    const double nu = c->lid_velocity / c->reynolds;
    const double dt = cavity_dt(*c);
    Vel vel(g, u, v, w), prev(g, au, av, aw);
    Field pf(make_layout(g, kCell), p);
    for(int s = 0; s < steps; ++s) {
      fractional_step(b, nu, dt, 0, true, c->max_iterations, c->tolerance, vel, prev, pf, *has_prev, *m,
                      "lid-driven cavity", false);
      if(stop_on_steady && m->step >= c->min_steps && m->max_velocity_change <= c->steady_tolerance) break;
    }
  });
}

// solver/channel_flow.cpp:182-224 initialize_steady_channel_state and the set-up of
// run_channel_flow (:465-466): analytic profile, velocity BCs, total-pressure BCs from nu lap u.
int ofref_channel_init(const ChannelConfig* c, double* u, double* v, double* w, double* p, double* dt) {
  return guarded([&] {
    const Grid g = channel_grid(*c);
    const BCSet b = channel_bcs(*c);
    Vel vel(g, u, v, w);
    Field pf(make_layout(g, kCell), p);
    vel.fill(0.0);
    pf.fill(0.0);
    for_range(vel.x.L.active(), [&](int i, int j, int k) { vel.x(i, j, k) = channel_u(*c, coord(vel.x.L, 1, j)); });
    for_range(pf.L.active(), [&](int i, int j, int k) { pf(i, j, k) = channel_p(*c, coord(pf.L, 0, i)); });
    apply_velocity_bc(b, vel);
    Vel dif(g);
    diffusion(vel, c->viscosity, dif);
    apply_total_pressure_bc(b, dif, pf);
    *dt = channel_dt(*c);
  });
}

// The step loop of run_channel_flow (solver/channel_flow.cpp:468-523); its predictor
// right-hand side is the cavity's (two-dimensional factorisation correction).
int ofref_channel_steps(const ChannelConfig* c, int steps, double* u, double* v, double* w, double* p, double* au,
                        double* av, double* aw, int* has_prev, StepMetrics* m) {
  return guarded([&] {
    const Grid g = channel_grid(*c);
    const BCSet b = channel_bcs(*c);
    const double dt = channel_dt(*c);
    Vel vel(g, u, v, w), prev(g, au, av, aw);
    Field pf(make_layout(g, kCell), p);
    for(int s = 0; s < steps; ++s)
      fractional_step(b, c->viscosity, dt, 0, true, c->max_iterations, c->tolerance, vel, prev, pf, *has_prev, *m,
                      "channel flow", false);
  });
}

// solver/channel_flow.cpp:258-301: x-averaged u per active row (first k plane) and its relative
// L2 distance to the analytic profile; profile receives ny values.
int ofref_channel_profile(const ChannelConfig* c, const double* u, double* profile, double* relative_l2) {
  return guarded([&] {
    const Grid g = channel_grid(*c);
    Field fu(make_layout(g, kFaceX), const_cast<double*>(u));
    const auto r = fu.L.active();
    double num = 0.0, den = 0.0;
    int n = 0;
    for(int j = r.jb; j < r.je; ++j, ++n) {
      double sum = 0.0;
      for(int i = r.ib; i < r.ie; ++i) sum += fu(i, j, r.kb);
      const double val = sum / (double)(r.ie - r.ib);
      profile[n] = val;
      const double ref = channel_u(*c, coord(fu.L, 1, j));
      const double d = val - ref;
      num += d * d;
      den += ref * ref;
    }
    *relative_l2 = den == 0.0 ? std::sqrt(num) : std::sqrt(num / den);
  });
}

// solver/operator_verification.cpp:60-111
int ofref_mms(int n, double* out3) {
  return guarded([&] {
    if(n <= 0) throw std::invalid_argument("run_operator_manufactured_solution_case expects resolution > 0");
    const double L = 2.0 * pi();
    const double h = L / (double)n;
    const Grid g{n, n, 1, h, h, 1.0};
    Field p(make_layout(g, kCell)), lap(make_layout(g, kCell)), div(make_layout(g, kCell));
    Vel ex(g), gr(g);
    auto fill = [](Field& f, auto fn) {
      for(int k = 0; k < f.L.s[2]; ++k)
        for(int j = 0; j < f.L.s[1]; ++j)
          for(int i = 0; i < f.L.s[0]; ++i) f(i, j, k) = fn(coord(f.L, 0, i), coord(f.L, 1, j), coord(f.L, 2, k));
    };
    auto err = [](const Field& f, auto fn) {
      double s = 0.0;
      size_t c = 0;
      for_range(f.L.active(), [&](int i, int j, int k) {
        const double d = f(i, j, k) - fn(coord(f.L, 0, i), coord(f.L, 1, j), coord(f.L, 2, k));
        s += d * d;
        ++c;
      });
      return std::sqrt(s / (double)c);
    };
    fill(p, [](double x, double y, double) { return std::sin(x) * std::cos(y); });
    fill(ex.x, [](double x, double y, double) { return std::sin(x) * std::cos(y); });
    fill(ex.y, [](double x, double y, double) { return std::cos(x) * std::sin(y); });
    fill(ex.z, [](double, double, double) { return 0.0; });
    gradient(p, gr);
    divergence(ex, div);
    laplacian(p, lap);
    const double gx = err(gr.x, [](double x, double y, double) { return std::cos(x) * std::cos(y); });
    const double gy = err(gr.y, [](double x, double y, double) { return -std::sin(x) * std::sin(y); });
    out3[0] = std::sqrt(0.5 * (gx * gx + gy * gy));
    out3[1] = err(div, [](double x, double y, double) { return 2.0 * std::cos(x) * std::cos(y); });
    out3[2] = err(lap, [](double x, double y, double) { return -2.0 * std::sin(x) * std::cos(y); });
  });
}

}  // extern "C"
