// Discrete adjoint and tangent-linear model of the periodic Taylor-Green projection step, with
// velocity checkpointing. The reference has no derivatives; the forward step is the device
// stepper's (solver/taylor_green.cpp:540-634 restated in stepper.hip). Every linear stage has a
// hand-written transpose here; the TVD advection (solver/momentum_terms.cpp:88-312) is
// linearised with its branches (velocity sign, upwind fallback, small-denominator guard,
// limiter side r > 0) frozen at the forward state; the adjoint pressure solve reuses the
// symmetric MGPCG.
//
// Forward step n (all boundaries periodic):
//   c    = BCv(u)                                 velocity ghosts (duplicate faces averaged)
//   A    = Adv(c);  rhs = c + s_g Grad(p) + s_d nu Lap(c) + s_a A + s_p A_prev + s_f FC(c) + dt f
//   u*   = ADI(rhs)
//   prhs = (1/dt) Div(u*);  phi = Solve(prhs) (zero mean);  phi ghosts = BCp(phi)
//   u'   = BCv(u* - dt Grad(phi));  p' = BCp(p + phi - (nu dt / 2) prhs);  A_prev' = A
// The adjoint step maps (u'_bar, p'_bar, A_prev'_bar) back to (u_bar, p_bar, A_prev_bar) and
// accumulates f_bar += dt rhs_bar. Edge and corner ghosts are constant (zero) in the forward
// state, so their adjoints are dropped. Solve is the map -P0 A^+ P0 (mean removed from the rhs
// and from the solution), which is symmetric, so its transpose is one more solve.

#include "adjoint.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <thread>
#include <vector>

namespace ofh {

namespace {

constexpr double kSmallAdj = 1.0e-14;  // solver/momentum_terms.cpp:62 (safe_ratio guard)

// ------------------------------------------------------------------ transposed ghost fills
// Periodic velocity ghosts of one component along one axis (forward: component_axis_bc in
// fields.hip): normal component: F1 = FA = (F1 + FA)/2, F0 = F(A-1), F(A+1) = F2;
// tangential: F0 = FA, F(A+1) = F1. Transposed per line, last assignment first.
__global__ void k_vbc_periodic_adj(double* f, Lay L, int comp, int axis) {
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  if(t0 >= L.a(o0) || t1 >= L.a(o1)) return;
  const int A = L.a(axis);
  auto at = [&](int along) {
    int cc[3];
    cc[axis] = along;
    cc[o0] = 1 + t0;
    cc[o1] = 1 + t1;
    return L.at(cc[0], cc[1], cc[2]);
  };
  const double g0 = f[at(0)], gA1 = f[at(A + 1)];
  f[at(0)] = 0.0;
  f[at(A + 1)] = 0.0;
  if(comp == axis) {
    f[at(A - 1)] += g0;
    f[at(2)] += gA1;
    const double shared = 0.5 * (f[at(1)] + f[at(A)]);
    f[at(1)] = shared;
    f[at(A)] = shared;
  } else {
    f[at(A)] += g0;
    f[at(1)] += gA1;
  }
}

// Periodic cell ghosts (pressure and total pressure): F0 = FA, F(A+1) = F1. Transposed.
__global__ void k_pbc_periodic_adj(double* p, Lay L, int axis) {
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
  if(t0 >= L.a(o0) || t1 >= L.a(o1)) return;
  const int A = L.a(axis);
  auto at = [&](int along) {
    int cc[3];
    cc[axis] = along;
    cc[o0] = 1 + t0;
    cc[o1] = 1 + t1;
    return L.at(cc[0], cc[1], cc[2]);
  };
  const double g0 = p[at(0)], gA1 = p[at(A + 1)];
  p[at(0)] = 0.0;
  p[at(A + 1)] = 0.0;
  p[at(A)] += g0;
  p[at(1)] += gA1;
}

dim3 slab_grid2(int e0, int e1) { return dim3((unsigned)((e0 + 63) / 64), (unsigned)((e1 + 3) / 4), 1); }

// ------------------------------------------------------------------ storage-wide helpers
__device__ inline bool active(const Lay& L, int i, int j, int k) {
  return i >= 1 && i <= L.ax && j >= 1 && j <= L.ay && k >= 1 && k <= L.az;
}

struct Idx3 {
  int i, j, k;
};
__device__ inline Idx3 decompose(const Lay& L, size_t id) {
  Idx3 r;
  r.i = (int)(id % (size_t)L.sx);
  const size_t t = id / (size_t)L.sx;
  r.j = (int)(t % (size_t)L.sy);
  r.k = (int)(t / (size_t)L.sy);
  return r;
}

unsigned stride_blocks(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 8192); }

// y = 0 outside the active range (edges, corners and ghost slabs)
__global__ void k_zero_inactive(double* y, Lay L) {
  for(size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < L.size(); id += (size_t)gridDim.x * blockDim.x) {
    const Idx3 q = decompose(L, id);
    if(!active(L, q.i, q.j, q.k)) y[id] = 0.0;
  }
}

// ------------------------------------------------------------------ transposed operators
// Grad^T: p_bar(cell) += s * sum_a (g_bar_a(face at the cell) - g_bar_a(face above)) / h_a over
// active faces (forward: g_a(f) = (p(f) - p(f - e_a)) / h_a on the active faces of component a).
__global__ void k_grad_adj(const double* gx, const double* gy, const double* gz, Lay Lu, Lay Lv, Lay Lw, Lay Lc,
                           of_grid g, double s, double* pbar) {
  for(size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < Lc.size();
      id += (size_t)gridDim.x * blockDim.x) {
    const Idx3 q = decompose(Lc, id);
    double acc = 0.0;
    if(active(Lu, q.i, q.j, q.k)) acc += gx[Lu.at(q.i, q.j, q.k)] / g.dx;
    if(active(Lu, q.i + 1, q.j, q.k)) acc -= gx[Lu.at(q.i + 1, q.j, q.k)] / g.dx;
    if(active(Lv, q.i, q.j, q.k)) acc += gy[Lv.at(q.i, q.j, q.k)] / g.dy;
    if(active(Lv, q.i, q.j + 1, q.k)) acc -= gy[Lv.at(q.i, q.j + 1, q.k)] / g.dy;
    if(active(Lw, q.i, q.j, q.k)) acc += gz[Lw.at(q.i, q.j, q.k)] / g.dz;
    if(active(Lw, q.i, q.j, q.k + 1)) acc -= gz[Lw.at(q.i, q.j, q.k + 1)] / g.dz;
    pbar[id] += s * acc;
  }
}

// Div^T: u_bar_a(face) += s * (d_bar(cell below) - d_bar(cell above)) / h_a over active cells
// (forward: d(c) = sum_a (u_a(c + e_a) - u_a(c)) / h_a on active cells).
__global__ void k_div_adj(const double* dbar, Lay Lc, Lay L, int axis, double h, double s, double* ubar) {
  const int di = axis == 0, dj = axis == 1, dk = axis == 2;
  for(size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < L.size(); id += (size_t)gridDim.x * blockDim.x) {
    const Idx3 q = decompose(L, id);
    double acc = 0.0;
    if(active(Lc, q.i - di, q.j - dj, q.k - dk)) acc += dbar[Lc.at(q.i - di, q.j - dj, q.k - dk)];
    if(active(Lc, q.i, q.j, q.k)) acc -= dbar[Lc.at(q.i, q.j, q.k)];
    ubar[id] += s * acc / h;
  }
}

// The Laplacian (7 points) and the factorisation correction (27 points: dx dy + dx dz + dy dz
// - alpha dx dy dz, each d a 1D second difference) are symmetric stencils, so the transpose of
// applying them on the active range is the same stencil applied to the output adjoint extended
// by zero outside the active range.
__global__ void k_lap_fc_adj(const double* obar, Lay L, of_grid g, double s_lap, double s_fc, double alpha,
                             double* fbar) {
  const double inv[3] = {1.0 / (g.dx * g.dx), 1.0 / (g.dy * g.dy), 1.0 / (g.dz * g.dz)};
  const double cw[3] = {1.0, -2.0, 1.0};
  for(size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < L.size(); id += (size_t)gridDim.x * blockDim.x) {
    const Idx3 q = decompose(L, id);
    auto O = [&](int a, int b, int c) {
      const int i = q.i + a, j = q.j + b, k = q.k + c;
      return active(L, i, j, k) ? obar[L.at(i, j, k)] : 0.0;
    };
    const double o0 = O(0, 0, 0);
    double lap = (O(1, 0, 0) - 2.0 * o0 + O(-1, 0, 0)) * inv[0];
    lap += (O(0, 1, 0) - 2.0 * o0 + O(0, -1, 0)) * inv[1];
    lap += (O(0, 0, 1) - 2.0 * o0 + O(0, 0, -1)) * inv[2];
    double fc = 0.0;
    for(int a = -1; a <= 1; ++a)
      for(int b = -1; b <= 1; ++b)
        for(int c = -1; c <= 1; ++c) {
          const double o = O(a, b, c);
          const double wx = cw[a + 1] * inv[0], wy = cw[b + 1] * inv[1], wz = cw[c + 1] * inv[2];
          double w = 0.0;
          if(c == 0) w += wx * wy;
          if(b == 0) w += wx * wz;
          if(a == 0) w += wy * wz;
          w -= alpha * wx * wy * wz;
          fc += w * o;
        }
    fbar[id] += s_lap * lap + s_fc * fc;
  }
}

// ------------------------------------------------------------------ linearised TVD advection
struct VP {
  const double* q[3];
  Lay L[3];
};

// Face value R(ll, l, r, rr) of solver/momentum_terms.cpp:88-197 and its partial derivatives
// dR[0..3] w.r.t. ll, l, r, rr with every branch frozen at the forward values. psi is the van
// Leer limiter (r + |r|)/(1 + |r|); its derivative is taken as 2/(1+r)^2 for r > 0 and 0
// otherwise (the one-sided value at the kink r = 0).
struct FaceLin {
  double R, vel;
  double dR[4];
};
__device__ inline FaceLin face_lin(double ll, double l, double r, double rr, int r_index, int max_index, double vel,
                                   int scheme) {
  FaceLin f;
  f.vel = vel;
  f.dR[0] = f.dR[1] = f.dR[2] = f.dR[3] = 0.0;
  if(scheme == 2) {
    f.R = 0.5 * (l + r);
    f.dR[1] = f.dR[2] = 0.5;
    return f;
  }
  if(scheme == 1 || fabs(vel) < kSmallAdj) {
    if(vel >= 0.0) {
      f.R = l;
      f.dR[1] = 1.0;
    } else {
      f.R = r;
      f.dR[2] = 1.0;
    }
    return f;
  }
  const double d = r - l;
  const bool guard = fabs(d) < kSmallAdj;
  auto psi = [](double x) { return (x + fabs(x)) / (1.0 + fabs(x)); };
  auto dpsi = [](double x) { return x > 0.0 ? 2.0 / ((1.0 + x) * (1.0 + x)) : 0.0; };
  if(vel >= 0.0) {
    if(r_index - 2 < 0) {
      f.R = l;
      f.dR[1] = 1.0;
      return f;
    }
    // R = l + psi(t) d / 2, t = (l - ll) / d
    const double num = l - ll;
    const double t = guard ? 0.0 : num / d;
    const double ps = psi(t), dp = guard ? 0.0 : dpsi(t);
    f.R = l + 0.5 * ps * d;
    const double t_ll = guard ? 0.0 : -1.0 / d, t_l = guard ? 0.0 : 1.0 / d + num / (d * d),
                 t_r = guard ? 0.0 : -num / (d * d);
    f.dR[0] = 0.5 * d * dp * t_ll;
    f.dR[1] = 1.0 - 0.5 * ps + 0.5 * d * dp * t_l;
    f.dR[2] = 0.5 * ps + 0.5 * d * dp * t_r;
    return f;
  }
  if(r_index + 1 > max_index) {
    f.R = r;
    f.dR[2] = 1.0;
    return f;
  }
  // R = r - psi(t) d / 2, t = (rr - r) / d
  const double num = rr - r;
  const double t = guard ? 0.0 : num / d;
  const double ps = psi(t), dp = guard ? 0.0 : dpsi(t);
  f.R = r - 0.5 * ps * d;
  const double t_l = guard ? 0.0 : num / (d * d), t_r = guard ? 0.0 : -1.0 / d - num / (d * d),
               t_rr = guard ? 0.0 : 1.0 / d;
  f.dR[1] = 0.5 * ps - 0.5 * d * dp * t_l;
  f.dR[2] = 1.0 - 0.5 * ps - 0.5 * d * dp * t_r;
  f.dR[3] = -0.5 * d * dp * t_rr;
  return f;
}

// Inputs of one face flux of component c: the two velocity samples averaged into the face
// velocity (component cv) and the four samples of q along the face's axis.
struct FaceIn {
  int cv;
  size_t va, vb;
  size_t q[4];
  bool qv[4];  // the sample lies in storage
  int r_index, max_index;
  double sign_h;  // +1/h for the upper face, -1/h for the lower face
};

// The six faces (w e s n b f) of the flux divergence of component c at storage point (i, j, k)
// (solver/momentum_terms.cpp:199-312).
__device__ inline void faces_of(const VP& V, int c, int i, int j, int k, const of_grid& g, FaceIn (&F)[6]) {
  const Lay& L = V.L[c];
  const Lay &Lu = V.L[0], &Lv = V.L[1], &Lw = V.L[2];
  int cv[6];
  size_t va[6], vb[6];
  if(c == 0) {
    cv[0] = 0; va[0] = Lu.at(i - 1, j, k);     vb[0] = Lu.at(i, j, k);
    cv[1] = 0; va[1] = Lu.at(i, j, k);         vb[1] = Lu.at(i + 1, j, k);
    cv[2] = 1; va[2] = Lv.at(i - 1, j, k);     vb[2] = Lv.at(i, j, k);
    cv[3] = 1; va[3] = Lv.at(i - 1, j + 1, k); vb[3] = Lv.at(i, j + 1, k);
    cv[4] = 2; va[4] = Lw.at(i - 1, j, k);     vb[4] = Lw.at(i, j, k);
    cv[5] = 2; va[5] = Lw.at(i - 1, j, k + 1); vb[5] = Lw.at(i, j, k + 1);
  } else if(c == 1) {
    cv[0] = 0; va[0] = Lu.at(i, j - 1, k);     vb[0] = Lu.at(i, j, k);
    cv[1] = 0; va[1] = Lu.at(i + 1, j - 1, k); vb[1] = Lu.at(i + 1, j, k);
    cv[2] = 1; va[2] = Lv.at(i, j - 1, k);     vb[2] = Lv.at(i, j, k);
    cv[3] = 1; va[3] = Lv.at(i, j, k);         vb[3] = Lv.at(i, j + 1, k);
    cv[4] = 2; va[4] = Lw.at(i, j - 1, k);     vb[4] = Lw.at(i, j, k);
    cv[5] = 2; va[5] = Lw.at(i, j - 1, k + 1); vb[5] = Lw.at(i, j, k + 1);
  } else {
    cv[0] = 0; va[0] = Lu.at(i, j, k - 1);     vb[0] = Lu.at(i, j, k);
    cv[1] = 0; va[1] = Lu.at(i + 1, j, k - 1); vb[1] = Lu.at(i + 1, j, k);
    cv[2] = 1; va[2] = Lv.at(i, j, k - 1);     vb[2] = Lv.at(i, j, k);
    cv[3] = 1; va[3] = Lv.at(i, j + 1, k - 1); vb[3] = Lv.at(i, j + 1, k);
    cv[4] = 2; va[4] = Lw.at(i, j, k - 1);     vb[4] = Lw.at(i, j, k);
    cv[5] = 2; va[5] = Lw.at(i, j, k);         vb[5] = Lw.at(i, j, k + 1);
  }
  const int co[3] = {i, j, k};
  const int smax[3] = {L.sx - 1, L.sy - 1, L.sz - 1};
  const double h[3] = {g.dx, g.dy, g.dz};
  for(int f = 0; f < 6; ++f) {
    const int axis = f / 2;
    const bool upper = (f & 1) != 0;
    const int r_index = co[axis] + (upper ? 1 : 0);  // storage index of `right` along the axis
    FaceIn& o = F[f];
    o.cv = cv[f];
    o.va = va[f];
    o.vb = vb[f];
    o.r_index = r_index;
    o.max_index = smax[axis];
    o.sign_h = (upper ? 1.0 : -1.0) / h[axis];
    for(int t = 0; t < 4; ++t) {
      const int along = r_index - 2 + t;  // ll, l, r, rr
      int cc[3] = {i, j, k};
      cc[axis] = along;
      o.qv[t] = along >= 0 && along <= smax[axis];
      o.q[t] = o.qv[t] ? L.at(cc[0], cc[1], cc[2]) : 0;
    }
  }
}

__device__ inline FaceLin lin_at(const VP& V, int c, const FaceIn& o, int scheme) {
  const double* q = V.q[c];
  const double* vq = V.q[o.cv];
  const double vel = 0.5 * (vq[o.va] + vq[o.vb]);
  double s[4];
  for(int t = 0; t < 4; ++t) s[t] = o.qv[t] ? q[o.q[t]] : 0.0;
  return face_lin(s[0], s[1], s[2], s[3], o.r_index, o.max_index, vel, scheme);
}

// Tangent-linear advection: dA_c = sum_faces sign/h * (R (dv_a + dv_b)/2 + vel dR . dq).
__global__ void k_advection_tl(VP V, VP D, int c, of_grid g, int scheme, double* out) {
  const Lay& L = V.L[c];
  for(size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < L.size(); id += (size_t)gridDim.x * blockDim.x) {
    const Idx3 p = decompose(L, id);
    if(!active(L, p.i, p.j, p.k)) continue;
    FaceIn F[6];
    faces_of(V, c, p.i, p.j, p.k, g, F);
    double acc = 0.0;
    for(int f = 0; f < 6; ++f) {
      const FaceLin fl = lin_at(V, c, F[f], scheme);
      const double* dv = D.q[F[f].cv];
      double t = 0.5 * fl.R * (dv[F[f].va] + dv[F[f].vb]);
      for(int s = 0; s < 4; ++s)
        if(F[f].qv[s] && fl.dR[s] != 0.0) t += fl.vel * fl.dR[s] * D.q[c][F[f].q[s]];
      acc += F[f].sign_h * t;
    }
    out[id] = acc;
  }
}

// One face of component c's flux divergence, normal to axis a, whose right sample has index
// m = co[a] along a: the upper face of point co - e_a and the lower face of point co (the two
// points compute its flux from the same operands). `lin` is its linearisation at the forward
// values and fbar the adjoint of its flux, (abar(co - e_a) - abar(co)) / h over whichever of the
// two points is active; `exists` is false when neither is an active point of c.
struct FaceAdj {
  FaceLin lin;
  double fbar;
  bool exists;
};
__device__ inline FaceAdj face_adj(const VP& V, const double* abar, int c, int a, const int (&co)[3],
                                   const of_grid& g, int scheme) {
  FaceAdj o{};
  const Lay& L = V.L[c];
  const int ext[3] = {L.ax, L.ay, L.az};
  for(int b = 0; b < 3; ++b)
    if(co[b] < 1 || co[b] > ext[b] + (b == a ? 1 : 0)) return o;
  o.exists = true;
  const int m = co[a];
  const int da[3] = {a == 0, a == 1, a == 2};
  const double h = a == 0 ? g.dx : a == 1 ? g.dy : g.dz;
  const size_t at = L.at(co[0], co[1], co[2]);
  const size_t below = L.at(co[0] - da[0], co[1] - da[1], co[2] - da[2]);
  // the lower point's upper face (+1/h) and the upper point's lower face (-1/h)
  const double t_lo = m >= 2 ? (1.0 / h) * abar[below] : 0.0;
  const double t_up = m <= ext[a] ? (-1.0 / h) * abar[at] : 0.0;
  o.fbar = t_lo + t_up;
  // samples ll, l, r, rr of q_c along a, inside storage
  const double* q = V.q[c];
  const ptrdiff_t step = a == 0 ? 1 : a == 1 ? (ptrdiff_t)L.sx : (ptrdiff_t)L.sx * L.sy;
  const int smax = L.s(a) - 1;
  double s4[4];
  for(int t = 0; t < 4; ++t) {
    const int along = m - 2 + t;
    s4[t] = along >= 0 && along <= smax ? q[(ptrdiff_t)at + (ptrdiff_t)(t - 2) * step] : 0.0;
  }
  // face velocity: component a at co and at co - e_a (c = a) or co - e_c (c != a)
  const Lay& Lv = V.L[a];
  const int dc[3] = {c == 0, c == 1, c == 2};
  const int* dv = c == a ? da : dc;
  const double* vq = V.q[a];
  const double vel = 0.5 * (vq[Lv.at(co[0] - dv[0], co[1] - dv[1], co[2] - dv[2])] + vq[Lv.at(co[0], co[1], co[2])]);
  o.lin = face_lin(s4[0], s4[1], s4[2], s4[3], m, smax, vel, scheme);
  return o;
}

// Adjoint advection in gather form, deterministic: for every storage point of component a,
// the sum of what the tangent's faces take from it, each face's flux adjoint formed once from
// the output adjoints on its two sides. A sample of u_a is (1) sample t of the reconstruction
// of the faces of component a normal to each axis b whose right sample lies at its index + 2 - t
// (vel dR_t fbar), and (2) a velocity sample of the faces of every component c normal to a at
// its own position and at the position one step along c (R fbar / 2). The contributions are
// added in a fixed order and accumulated into out.
__global__ void k_advection_adj(VP V, VP A, int a, of_grid g, int scheme, double* out) {
  const Lay& La = V.L[a];
  for(size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < La.size(); id += (size_t)gridDim.x * blockDim.x) {
    const Idx3 p = decompose(La, id);
    const int q[3] = {p.i, p.j, p.k};
    double acc = 0.0;
    for(int b = 0; b < 3; ++b)
      for(int t = 0; t < 4; ++t) {
        int co[3] = {q[0], q[1], q[2]};
        co[b] = q[b] + 2 - t;
        const FaceAdj f = face_adj(V, A.q[a], a, b, co, g, scheme);
        if(f.exists && f.lin.dR[t] != 0.0) acc += f.lin.vel * f.lin.dR[t] * f.fbar;
      }
    for(int c = 0; c < 3; ++c)
      for(int s = 0; s < 2; ++s) {
        int co[3] = {q[0], q[1], q[2]};
        co[c] += s;
        const FaceAdj f = face_adj(V, A.q[c], c, a, co, g, scheme);
        if(f.exists) acc += 0.5 * f.lin.R * f.fbar;
      }
    out[id] += acc;
  }
}

// d/du of KE = (1/N) sum_cells (uc^2 + vc^2 + wc^2)/2 with uc = (u(f) + u(f + e)) / 2
// (solver/taylor_green.cpp:269-288): u_bar(f) = (uc(cell below) + uc(cell above)) / (2N).
__global__ void k_ke_grad(const double* q, Lay L, Lay Lc, int axis, double invn, double* qbar) {
  const int di = axis == 0, dj = axis == 1, dk = axis == 2;
  for(size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < L.size(); id += (size_t)gridDim.x * blockDim.x) {
    const Idx3 p = decompose(L, id);
    double acc = 0.0;
    if(active(L, p.i, p.j, p.k)) {
      if(active(Lc, p.i - di, p.j - dj, p.k - dk))
        acc += 0.5 * (q[L.at(p.i - di, p.j - dj, p.k - dk)] + q[id]);
      if(active(Lc, p.i, p.j, p.k)) acc += 0.5 * (q[id] + q[L.at(p.i + di, p.j + dj, p.k + dk)]);
    }
    qbar[id] = 0.5 * invn * acc;
  }
}

// sum over cells of uc duc + vc dvc + wc dwc (N times the kinetic-energy tangent)
__global__ void k_ke_tangent(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                             const double* __restrict__ du, const double* __restrict__ dv,
                             const double* __restrict__ dw, Lay Lu, Lay Lv, Lay Lw, of_grid g, double* partials,
                             int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  double vals[1] = {0.0};
  if(i <= g.nx && j <= g.ny) {
    const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + g.nz);
    for(int k = kb; k < ke; ++k) {
      const double uc = 0.5 * (u[Lu.at(i, j, k)] + u[Lu.at(i + 1, j, k)]);
      const double vc = 0.5 * (v[Lv.at(i, j, k)] + v[Lv.at(i, j + 1, k)]);
      const double wc = 0.5 * (w[Lw.at(i, j, k)] + w[Lw.at(i, j, k + 1)]);
      const double duc = 0.5 * (du[Lu.at(i, j, k)] + du[Lu.at(i + 1, j, k)]);
      const double dvc = 0.5 * (dv[Lv.at(i, j, k)] + dv[Lv.at(i, j + 1, k)]);
      const double dwc = 0.5 * (dw[Lw.at(i, j, k)] + dw[Lw.at(i, j, k + 1)]);
      vals[0] += uc * duc + vc * dvc + wc * dwc;
    }
  }
  const int ops[1] = {kSum};
  store_partials<1>(vals, ops, partials);
}

[[noreturn]] void solve_failed(const char* what, const PoissonResult& r) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "the %s pressure solve did not converge (iterations=%d, relative_residual=%g)", what,
                r.diag.iterations, r.diag.relative_residual);
  throw RuntimeError(buf);
}

VP vp_of(const of_grid& g, const double* u, const double* v, const double* w) {
  VP V;
  V.q[0] = u;
  V.q[1] = v;
  V.q[2] = w;
  for(int a = 0; a < 3; ++a) V.L[a] = make_lay(g, OF_FACE_X + a);
  return V;
}

}  // namespace

// ------------------------------------------------------------------ launch helpers
void launch_component_bc_adjoint(Context& c, const of_grid& g, int comp, double* f) {
  const Lay L = make_lay(g, OF_FACE_X + comp);
  for(int axis = 2; axis >= 0; --axis) {  // the forward fills axes 0, 1, 2 in turn
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    OFH_LAUNCH(k_vbc_periodic_adj, slab_grid2(L.a(o0), L.a(o1)), dim3(64, 4), 0, c.stream, f, L, comp, axis);
  }
}

void launch_velocity_bc_adjoint(Context& c, const of_grid& g, double* const* v) {
  for(int comp = 0; comp < 3; ++comp) launch_component_bc_adjoint(c, g, comp, v[comp]);
}

void launch_pressure_bc_adjoint(Context& c, const of_grid& g, double* p) {
  const Lay L = make_lay(g, OF_CELL);
  for(int axis = 2; axis >= 0; --axis) {
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    OFH_LAUNCH(k_pbc_periodic_adj, slab_grid2(L.a(o0), L.a(o1)), dim3(64, 4), 0, c.stream, p, L, axis);
  }
}

void launch_gradient_adjoint(Context& c, const of_grid& g, double scale, const double* const* gbar, double* pbar) {
  const Lay Lc = make_lay(g, OF_CELL);
  OFH_LAUNCH(k_grad_adj, stride_blocks(Lc.size()), 256, 0, c.stream, gbar[0], gbar[1], gbar[2],
                     make_lay(g, OF_FACE_X), make_lay(g, OF_FACE_Y), make_lay(g, OF_FACE_Z), Lc, g, scale, pbar);
}

void launch_divergence_adjoint(Context& c, const of_grid& g, double scale, const double* dbar, double* const* ubar) {
  const Lay Lc = make_lay(g, OF_CELL);
  for(int a = 0; a < 3; ++a) {
    const Lay L = make_lay(g, OF_FACE_X + a);
    OFH_LAUNCH(k_div_adj, stride_blocks(L.size()), 256, 0, c.stream, dbar, Lc, L, a, spacing(g, a), scale,
                       ubar[a]);
  }
}

void launch_advection_tangent(Context& c, const of_grid& g, int scheme, const double* const* vel,
                              const double* const* dvel, double* const* out) {
  const VP V = vp_of(g, vel[0], vel[1], vel[2]);
  const VP D = vp_of(g, dvel[0], dvel[1], dvel[2]);
  for(int comp = 0; comp < 3; ++comp)
    OFH_LAUNCH(k_advection_tl, stride_blocks(V.L[comp].size()), 256, 0, c.stream, V, D, comp, g, scheme,
                       out[comp]);
}

void launch_advection_adjoint(Context& c, const of_grid& g, int scheme, const double* const* vel,
                              const double* const* abar, double* const* bar) {
  const VP V = vp_of(g, vel[0], vel[1], vel[2]);
  const VP A = vp_of(g, abar[0], abar[1], abar[2]);
  for(int comp = 0; comp < 3; ++comp)
    OFH_LAUNCH(k_advection_adj, stride_blocks(V.L[comp].size()), 256, 0, c.stream, V, A, comp, g, scheme, bar[comp]);
}

// ------------------------------------------------------------------ checkpointed driver
struct AdjointRun {
  Context& c;
  Stepper fwd;
  int every;
  int steps = 0;
  size_t lay[4];  // element counts: cell, u, v, w
  // state at steps 0, every, 2 every, ...: u, v, w, p, A_prev x3 (full storage) and has_prev
  struct Ckpt {
    double* f[7];
    int has_prev;
    of_step_metrics m;
  };
  std::vector<Ckpt> ckpts;
  std::vector<double*> window;  // velocity with ghosts of every step of one window, 3 arrays per step
  double* ufinal[3]{};          // velocity after the last forward step
  of_step_metrics final_metrics{};
  // adjoint state and work arrays
  double *ub[3]{}, *ab[3]{}, *pb = nullptr;  // adjoints of u^n, A_prev^n, p^n
  double *sb[3]{}, *rb[3]{}, *tb[3]{};       // adjoints of u*, rhs and the advection output
  double *phib = nullptr, *prhsb = nullptr, *tmpc = nullptr;
  double* fb[3]{};  // forcing adjoint
  double* red = nullptr;
  // tangent-linear state (allocated on first use)
  double *du[3]{}, *da[3]{}, *dp = nullptr, *dA[3]{}, *dr[3]{}, *dlap[3]{}, *dpred[3]{}, *dfc[3]{};
  double *dphi = nullptr, *dprhs = nullptr, *dg[3]{};
  // Overlapped reverse sweep (allocated on the first backward that has room for it): a second
  // stepper on a context of its own (stream, reduction buffers) recomputes window w - 1 on a
  // worker thread while this thread runs the adjoint steps of window w, the two windows in
  // alternating buffers. Both halves are latency-bound on a few compute units at the sizes
  // where the copy fits, so they overlap almost entirely.
  std::unique_ptr<Context> rctx;
  std::unique_ptr<Stepper> rec;
  std::vector<double*> window2;
  std::vector<int> last_hp;  // has_prev of the last window's steps, kept by the forward pass
  bool last_kept = false;    // `window` holds the last window's c^n (until a recompute overwrites it)
  // Solver of the adjoint and tangent-linear steps. Those solves are linear solves of derivative
  // fields, not part of the forward trajectory, so they need the tolerance, not the forward's
  // arithmetic: behind an FP64 forward run they use the FP32 V-cycle preconditioner (the mixed
  // mode's: FP64 outer CG, same relative tolerance), which needs about half the iterations.
  // SOLVER_AMD_ADJOINT_SOLVE=fp64, read when the run is created, keeps the forward's solver.
  PoissonSolver* dsolver = nullptr;
  bool own_dsolver = false;
  std::vector<double*> owned;

  AdjointRun(Context& ctx, const of_stepper_config& cfg, int every_) : c(ctx), fwd(ctx, cfg), every(every_) {
    if(every < 1) throw InvalidArgument("the checkpoint interval must be at least 1");
    if(cfg.flow_case != OF_CASE_TAYLOR_GREEN || cfg.nz < 2 || !fwd.all_periodic)
      throw InvalidArgument("the adjoint supports the three-dimensional periodic Taylor-Green case");
    const of_grid& g = fwd.g;
    lay[0] = make_lay(g, OF_CELL).size();
    for(int a = 0; a < 3; ++a) lay[1 + a] = make_lay(g, OF_FACE_X + a).size();
    for(int a = 0; a < 3; ++a) {
      ub[a] = alloc(lay[1 + a]);
      ab[a] = alloc(lay[1 + a]);
      sb[a] = alloc(lay[1 + a]);
      rb[a] = alloc(lay[1 + a]);
      tb[a] = alloc(lay[1 + a]);
      fb[a] = alloc(lay[1 + a]);
      ufinal[a] = alloc(lay[1 + a]);
    }
    pb = alloc(lay[0]);
    phib = alloc(lay[0]);
    prhsb = alloc(lay[0]);
    tmpc = alloc(lay[0]);
    const char* e = std::getenv("SOLVER_AMD_ADJOINT_SOLVE");
    const bool fp64_solves = e != nullptr && std::strcmp(e, "fp64") == 0;
    if(cfg.precision == OF_PRECISION_FP64 && !fp64_solves) {
      dsolver = poisson_create(c, g, fwd.pbcs, OF_PRECISION_MIXED);
      own_dsolver = true;
    } else {
      dsolver = fwd.poisson;
    }
    red = alloc(8);
    OFH_CHECK(hipStreamSynchronize(c.stream));
  }
  ~AdjointRun() {
    OFH_RELEASE(hipStreamSynchronize(c.stream));
    if(rctx) OFH_RELEASE(hipStreamSynchronize(rctx->stream));
    rec.reset();
    rctx.reset();
    if(own_dsolver) poisson_destroy(dsolver);
    for(double* p : owned) OFH_RELEASE(hipFree(p));
  }
  double* alloc(size_t n) {
    double* p = nullptr;
    OFH_CHECK(hipMalloc(&p, n * sizeof(double)));
    owned.push_back(p);
    OFH_CHECK(hipMemsetAsync(p, 0, n * sizeof(double), c.stream));
    return p;
  }
  Lay face(int a) const { return make_lay(fwd.g, OF_FACE_X + a); }
  Lay cell() const { return make_lay(fwd.g, OF_CELL); }
  void copy(double* dst, const double* src, size_t n, hipStream_t stream = nullptr) {
    OFH_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream ? stream : c.stream));
  }
  void zero(double* p, size_t n) { OFH_CHECK(hipMemsetAsync(p, 0, n * sizeof(double), c.stream)); }
  void zero_inactive(double* p, const Lay& L) {
    OFH_LAUNCH(k_zero_inactive, stride_blocks(L.size()), 256, 0, c.stream, p, L);
  }
  void axpy(const Lay& L, double a, const double* x, double* y) { launch_axpy_active(c, L, a, x, y); }
  static double* state(Stepper& s, int i) { return i < 3 ? s.vel[i] : i == 3 ? s.p_total : s.adv[i - 4]; }
  double* state(int i) { return state(fwd, i); }
  size_t state_size(int i) const { return i < 3 ? lay[1 + i] : i == 3 ? lay[0] : lay[1 + (i - 4)]; }

  // out = Solve(rhs), computed as 2^e Solve(2^-e rhs) with 2^e the power of two nearest above
  // max|rhs|. Solve is linear and its stopping test relative, but its breakdown guards are
  // absolute (1e-30 on r.z and d.q, linsolve/poisson_solver.cpp:752-762), which a derivative
  // field of small magnitude reaches before the relative tolerance; the exact rescaling keeps
  // the guards out of reach. rhs is scaled in place and restored.
  PoissonResult scaled_solve(double* rhs, double* out, const char* what) {
    const Lay Lc = cell();
    launch_norms(c, Lc, rhs, kScalRed0);
    double h[2] = {0.0, 0.0};
    OFH_CHECK(hipMemcpyAsync(h, c.scal + kScalRed0, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    OFH_CHECK(hipStreamSynchronize(c.stream));
    zero(out, lay[0]);
    PoissonResult pr{};
    if(!(h[1] > 0.0)) {  // zero right-hand side: zero solution
      pr.diag.converged = 1;
      return pr;
    }
    if(!std::isfinite(h[1])) throw RuntimeError(std::string("the ") + what + " pressure right-hand side is not finite");
    int e = 0;
    (void)std::frexp(h[1], &e);  // h[1] = m 2^e with m in [0.5, 1)
    const double down = std::ldexp(1.0, -e), up = std::ldexp(1.0, e);
    launch_scale_active(c, Lc, down, rhs, rhs);
    pr = poisson_solve(dsolver, rhs, out, fwd.cfg.poisson_max_iterations, fwd.cfg.poisson_tolerance, false,
                       true);  // out was zeroed above
    if(!pr.diag.converged) solve_failed(what, pr);
    launch_scale_active(c, Lc, up, rhs, rhs);
    launch_scale_active(c, Lc, up, out, out);
    return pr;
  }

  void save(Ckpt& k) {
    for(int i = 0; i < 7; ++i) copy(k.f[i], state(i), state_size(i));
    k.has_prev = fwd.has_prev;
    k.m = fwd.metrics;
  }
  void restore(const Ckpt& k, Stepper& s) {
    for(int i = 0; i < 7; ++i) copy(state(s, i), k.f[i], state_size(i), s.c.stream);
    s.has_prev = k.has_prev;
    s.metrics = k.m;
  }
  void restore(const Ckpt& k) { restore(k, fwd); }

  // Recomputes window w on stepper s (on its stream), keeping c^n = BCv(u^n) of every step in
  // buf and has_prev in hp; the window's copies may still be in flight on return.
  void recompute(Stepper& s, int w, std::vector<double*>& buf, std::vector<int>& hp) {
    const int n0 = w * every, n1 = std::min(steps, n0 + every);
    restore(ckpts[w], s);
    for(int n = n0; n < n1; ++n) {
      launch_velocity_bc(s.c, s.g, s.bcs, s.vel[0], s.vel[1], s.vel[2]);
      for(int a = 0; a < 3; ++a) copy(buf[3 * (n - n0) + a], s.vel[a], lay[1 + a], s.c.stream);
      hp[n - n0] = s.has_prev;
      if(n + 1 < n1) s.step();
    }
  }

  // Sets up the overlapped sweep when SOLVER_AMD_ADJOINT_OVERLAP is not "0" and the device has
  // room for the second stepper and window with a wide margin (at 512^3 it does not, and the
  // sweep stays serial). Returns whether the overlapped sweep is available.
  bool overlap_ready() {
    const char* e = std::getenv("SOLVER_AMD_ADJOINT_OVERLAP");
    if(e != nullptr && e[0] == '0') return false;
    if(rec) {
      while(window2.size() < window.size()) window2.push_back(alloc(lay[1 + window2.size() % 3]));
      return true;
    }
    size_t free_b = 0, total_b = 0;
    OFH_CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t face = lay[1] * sizeof(double), cell = lay[0] * sizeof(double);
    const size_t need = 18 * face + 8 * cell + ((size_t)4 << 30) + window.size() * face;
    if((double)free_b < 1.25 * (double)need + (double)((size_t)8 << 30)) return false;
    rctx = std::make_unique<Context>();
    context_init(*rctx, c.device, nullptr);
    rec = std::make_unique<Stepper>(*rctx, fwd.cfg);
    while(window2.size() < window.size()) window2.push_back(alloc(lay[1 + window2.size() % 3]));
    OFH_CHECK(hipStreamSynchronize(c.stream));
    return true;
  }

  double forward(const double* const* u0, const double* p0, int nsteps, of_step_metrics* m) {
    if(nsteps < 1) throw InvalidArgument("the adjoint run needs at least one step");
    for(int a = 0; a < 3; ++a)
      if(u0 == nullptr || u0[a] == nullptr) throw InvalidArgument("initial velocity must not be null");
    const int nck = (nsteps + every - 1) / every;
    while((int)ckpts.size() < nck) {
      Ckpt k{};
      for(int i = 0; i < 7; ++i) k.f[i] = alloc(state_size(i));
      ckpts.push_back(k);
    }
    const int wsteps = std::min(every, nsteps);
    while((int)window.size() < 3 * wsteps)
      for(int a = 0; a < 3; ++a) window.push_back(alloc(lay[1 + a]));
    last_hp.assign(every, 0);
    last_kept = false;
    steps = 0;
    // initial state: only active entries of the inputs are read
    for(int a = 0; a < 3; ++a) {
      copy(fwd.vel[a], u0[a], lay[1 + a]);
      zero_inactive(fwd.vel[a], face(a));
      zero(fwd.adv[a], lay[1 + a]);
    }
    if(p0) {
      copy(fwd.p_total, p0, lay[0]);
      zero_inactive(fwd.p_total, cell());
    } else {
      zero(fwd.p_total, lay[0]);
    }
    launch_velocity_bc(c, fwd.g, fwd.bcs, fwd.vel[0], fwd.vel[1], fwd.vel[2]);
    fwd.total_pressure_bc(fwd.vel);
    fwd.has_prev = 0;
    std::memset(&fwd.metrics, 0, sizeof fwd.metrics);
    fwd.metrics.dt = fwd.cfg.dt;
    // the last window's c^n = BCv(u^n) are kept as the forward pass makes them (the BC fill is
    // idempotent, and the step applies it again), so the reverse sweep starts without
    // recomputing that window
    const int last0 = (nck - 1) * every;
    for(int n = 0; n < nsteps; ++n) {
      if(n % every == 0) save(ckpts[n / every]);
      if(n >= last0) {
        launch_velocity_bc(c, fwd.g, fwd.bcs, fwd.vel[0], fwd.vel[1], fwd.vel[2]);
        for(int a = 0; a < 3; ++a) copy(window[3 * (n - last0) + a], fwd.vel[a], lay[1 + a]);
        last_hp[n - last0] = fwd.has_prev;
      }
      fwd.step();
    }
    for(int a = 0; a < 3; ++a) copy(ufinal[a], fwd.vel[a], lay[1 + a]);
    last_kept = true;
    final_metrics = fwd.metrics;
    if(m) *m = fwd.metrics;
    steps = nsteps;
    return fwd.kinetic_energy();
  }

  // One adjoint step of forward step n; cvel is c^n = BCv(u^n).
  void adjoint_step(const double* const* cvel, bool has_prev) {
    const of_grid& g = fwd.g;
    const double dt = fwd.cfg.dt, nu = fwd.cfg.viscosity;
    const Lay Lc = cell();
    // p' = BCp(p + phi - (nu dt/2) prhs)
    launch_pressure_bc_adjoint(c, g, pb);
    zero(phib, lay[0]);
    zero(prhsb, lay[0]);
    axpy(Lc, 1.0, pb, phib);
    axpy(Lc, -0.5 * nu * dt, pb, prhsb);
    // u' = BCv(u* - dt Grad(phi)): u*_bar = BCv^T(u'_bar) on active faces, phi_bar += -dt Grad^T
    launch_velocity_bc_adjoint(c, g, ub);
    for(int a = 0; a < 3; ++a) {
      zero(sb[a], lay[1 + a]);
      axpy(face(a), 1.0, ub[a], sb[a]);
    }
    const double* ubc[3] = {ub[0], ub[1], ub[2]};
    launch_gradient_adjoint(c, g, -dt, ubc, phib);
    // phi ghosts = BCp(phi); phi = Solve(prhs) with Solve symmetric
    launch_pressure_bc_adjoint(c, g, phib);
    scaled_solve(phib, tmpc, "adjoint");
    axpy(Lc, 1.0, tmpc, prhsb);
    // prhs = (1/dt) Div(u*)
    launch_divergence_adjoint(c, g, 1.0 / dt, prhsb, sb);
    // u* = ADI(rhs); only the active entries of rhs are read by the forward solve
    const double* sbc[3] = {sb[0], sb[1], sb[2]};
    predictor_adi_adjoint(c, g, fwd.bcs, 0.5 * nu * dt, sbc, rb);
    for(int a = 0; a < 3; ++a) zero_inactive(rb[a], face(a));
    // rhs = c + s_g Grad(p) + s_d nu Lap(c) + s_a A + s_p A_prev + s_f FC(c) + dt f, A = Adv(c)
    const double s_grad = -dt, s_dif = 0.5 * dt, s_adv = has_prev ? -1.5 * dt : -dt, s_prev = 0.5 * dt;
    const double alpha = 0.5 * nu * dt, s_fc = alpha * alpha;
    for(int a = 0; a < 3; ++a) {
      axpy(face(a), dt, rb[a], fb[a]);
      // advection adjoint input: s_a rhs_bar + A_bar (A becomes the next step's A_prev)
      zero(tb[a], lay[1 + a]);
      axpy(face(a), s_adv, rb[a], tb[a]);
      axpy(face(a), 1.0, ab[a], tb[a]);
      zero(ab[a], lay[1 + a]);
      if(has_prev) axpy(face(a), s_prev, rb[a], ab[a]);
    }
    // p_bar: the identity part stays; Grad^T reaches p's ghosts
    const double* rbc[3] = {rb[0], rb[1], rb[2]};
    launch_gradient_adjoint(c, g, s_grad, rbc, pb);
    // c_bar: identity + Lap^T + FC^T + Adv^T
    for(int a = 0; a < 3; ++a) {
      const Lay L = face(a);
      zero(ub[a], lay[1 + a]);
      axpy(L, 1.0, rb[a], ub[a]);
      OFH_LAUNCH(k_lap_fc_adj, stride_blocks(L.size()), 256, 0, c.stream, rb[a], L, g, s_dif * nu, s_fc, alpha,
                         ub[a]);
    }
    const double* tbc[3] = {tb[0], tb[1], tb[2]};
    launch_advection_adjoint(c, g, fwd.cfg.scheme, cvel, tbc, ub);
    // c = BCv(u)
    launch_velocity_bc_adjoint(c, g, ub);
    OFH_CHECK(hipGetLastError());
  }

  void backward(const double* const* seed, double* const* gu, double* gp, double* const* gf) {
    if(steps < 1) throw InvalidArgument("run the forward pass before the adjoint");
    const of_grid& g = fwd.g;
    for(int a = 0; a < 3; ++a) {
      if(seed && seed[a]) {
        copy(ub[a], seed[a], lay[1 + a]);
      } else {
        const double invn = 1.0 / ((double)g.nx * g.ny * g.nz);
        OFH_LAUNCH(k_ke_grad, stride_blocks(lay[1 + a]), 256, 0, c.stream, ufinal[a], face(a), cell(), a, invn,
                           ub[a]);
      }
      zero(ab[a], lay[1 + a]);
      zero(fb[a], lay[1 + a]);
    }
    zero(pb, lay[0]);
    const int nwin = (steps + every - 1) / every;
    auto adjoint_window = [&](int w, const std::vector<double*>& buf, const std::vector<int>& hp) {
      const int n0 = w * every, n1 = std::min(steps, n0 + every);
      for(int n = n1 - 1; n >= n0; --n) {
        const double* cv[3] = {buf[3 * (n - n0)], buf[3 * (n - n0) + 1], buf[3 * (n - n0) + 2]};
        adjoint_step(cv, hp[n - n0] != 0);
      }
    };
    // the last window is in `window` as the forward pass left it, unless an earlier sweep's
    // recomputes have written over it
    if(!last_kept) {
      recompute(fwd, nwin - 1, window, last_hp);
      last_kept = true;
    }
    if(!overlap_ready()) {
      std::vector<int> hp(every);
      for(int w = nwin - 1; w >= 0; --w) {
        if(w == nwin - 1) {
          adjoint_window(w, window, last_hp);
          continue;
        }
        last_kept = false;
        recompute(fwd, w, window, hp);
        adjoint_window(w, window, hp);
      }
    } else {
      for(int a = 0; a < 3; ++a) rec->force[a] = fwd.force[a];  // read-only, shared
      // window w in buffer (nwin - 1 - w) mod 2: the kept last window is in `window`
      std::vector<double*>* buf[2] = {&window, &window2};
      std::vector<int> hp[2] = {last_hp, std::vector<int>(every)};
      auto slot = [&](int w) { return (nwin - 1 - w) % 2; };
      for(int w = nwin - 1; w >= 0; --w) {
        // window w - 1 is recomputed on the worker thread while the adjoint steps of window w
        // run here; before the next round both streams are drained, so the buffer of window
        // w is free for window w - 2
        std::exception_ptr err;
        std::thread worker;
        if(w > 0 && slot(w - 1) == 0) last_kept = false;
        if(w > 0)
          worker = std::thread([&, w] {
            try {
              OFH_CHECK(hipSetDevice(c.device));
              recompute(*rec, w - 1, *buf[slot(w - 1)], hp[slot(w - 1)]);  // slot 0 is `window`
              OFH_CHECK(hipStreamSynchronize(rctx->stream));
            } catch(...) {
              err = std::current_exception();
            }
          });
        try {
          adjoint_window(w, *buf[slot(w)], hp[slot(w)]);
          OFH_CHECK(hipStreamSynchronize(c.stream));
        } catch(...) {
          if(worker.joinable()) worker.join();
          throw;
        }
        if(worker.joinable()) worker.join();
        if(err) std::rethrow_exception(err);
      }
    }
    // u^0 = BCv(active input), p^0 = BCp(active input)
    launch_velocity_bc_adjoint(c, g, ub);
    for(int a = 0; a < 3; ++a) {
      zero_inactive(ub[a], face(a));
      if(gu && gu[a]) copy(gu[a], ub[a], lay[1 + a]);
      if(gf && gf[a]) copy(gf[a], fb[a], lay[1 + a]);
    }
    if(gp) {
      launch_pressure_bc_adjoint(c, g, pb);
      zero_inactive(pb, cell());
      copy(gp, pb, lay[0]);
    }
    restore_final();
    OFH_CHECK(hipGetLastError());
    OFH_CHECK(hipStreamSynchronize(c.stream));
  }

  // the recompute leaves an intermediate state in the stepper; put the final velocity back
  void restore_final() {
    for(int a = 0; a < 3; ++a) copy(fwd.vel[a], ufinal[a], lay[1 + a]);
    fwd.metrics = final_metrics;
  }

  void tangent(const double* const* du0, const double* dp0, const double* const* df, double* const* du_out,
               double* dke) {
    if(steps < 1) throw InvalidArgument("run the forward pass before the tangent-linear model");
    const of_grid& g = fwd.g;
    const double dt = fwd.cfg.dt, nu = fwd.cfg.viscosity;
    const double alpha = 0.5 * nu * dt;
    if(dp == nullptr) {
      for(int a = 0; a < 3; ++a) {
        du[a] = alloc(lay[1 + a]);
        da[a] = alloc(lay[1 + a]);
        dA[a] = alloc(lay[1 + a]);
        dr[a] = alloc(lay[1 + a]);
        dlap[a] = alloc(lay[1 + a]);
        dpred[a] = alloc(lay[1 + a]);
        dfc[a] = alloc(lay[1 + a]);
        dg[a] = alloc(lay[1 + a]);
      }
      dp = alloc(lay[0]);
      dphi = alloc(lay[0]);
      dprhs = alloc(lay[0]);
    }
    for(int a = 0; a < 3; ++a) {
      if(du0 && du0[a]) {
        copy(du[a], du0[a], lay[1 + a]);
        zero_inactive(du[a], face(a));
      } else {
        zero(du[a], lay[1 + a]);
      }
      zero(da[a], lay[1 + a]);
    }
    if(dp0) {
      copy(dp, dp0, lay[0]);
      zero_inactive(dp, cell());
    } else {
      zero(dp, lay[0]);
    }
    const bool forced = df && df[0] && df[1] && df[2];
    launch_velocity_bc(c, g, fwd.bcs, du[0], du[1], du[2]);
    launch_pressure_bc(c, g, fwd.pbcs, dp);
    restore(ckpts[0]);
    for(int n = 0; n < steps; ++n) {
      // c^n and dc^n
      launch_velocity_bc(c, g, fwd.bcs, fwd.vel[0], fwd.vel[1], fwd.vel[2]);
      launch_velocity_bc(c, g, fwd.bcs, du[0], du[1], du[2]);
      const double* cv[3] = {fwd.vel[0], fwd.vel[1], fwd.vel[2]};
      const double* dc[3] = {du[0], du[1], du[2]};
      launch_advection_tangent(c, g, fwd.cfg.scheme, cv, dc, dA);
      launch_factorized_correction(c, g, 0, alpha, du[0], du[1], du[2], dfc[0], dfc[1], dfc[2]);
      launch_gradient(c, g, dp, dg[0], dg[1], dg[2]);
      const bool hp = fwd.has_prev != 0;
      for(int a = 0; a < 3; ++a) {
        const Lay L = face(a);
        launch_laplacian(c, g, OF_FACE_X + a, nu, du[a], dlap[a]);
        zero(dr[a], lay[1 + a]);
        axpy(L, 1.0, du[a], dr[a]);
        axpy(L, -dt, dg[a], dr[a]);
        axpy(L, 0.5 * dt, dlap[a], dr[a]);
        axpy(L, hp ? -1.5 * dt : -dt, dA[a], dr[a]);
        if(hp) axpy(L, 0.5 * dt, da[a], dr[a]);
        axpy(L, alpha * alpha, dfc[a], dr[a]);
        if(forced) axpy(L, dt, df[a], dr[a]);
      }
      predictor_adi(c, g, fwd.bcs, alpha, dr[0], dr[1], dr[2], dpred[0], dpred[1], dpred[2]);
      launch_divergence_scaled(c, g, dpred[0], dpred[1], dpred[2], 1.0 / dt, dprhs);
      scaled_solve(dprhs, dphi, "tangent-linear");
      launch_pressure_bc(c, g, fwd.pbcs, dphi);
      launch_correct(c, g, dpred[0], dpred[1], dpred[2], dphi, -dt, du[0], du[1], du[2]);
      launch_velocity_bc(c, g, fwd.bcs, du[0], du[1], du[2]);
      axpy(cell(), 1.0, dphi, dp);
      axpy(cell(), -0.5 * nu * dt, dprhs, dp);
      launch_pressure_bc(c, g, fwd.pbcs, dp);
      for(int a = 0; a < 3; ++a) copy(da[a], dA[a], lay[1 + a]);
      fwd.step();
    }
    for(int a = 0; a < 3; ++a)
      if(du_out && du_out[a]) copy(du_out[a], du[a], lay[1 + a]);
    if(dke) {
      const ColumnLaunch cl = column_launch(g.nx, g.ny, g.nz);
      OFH_LAUNCH(k_ke_tangent, cl.grid, cl.block, 0, c.stream, fwd.vel[0], fwd.vel[1], fwd.vel[2], du[0], du[1],
                         du[2], face(0), face(1), face(2), g, c.partials, cl.kz);
      launch_finalize(c.stream, c.partials, cl.grid, EpiStore<1>{red, 0u});
      double s = 0.0;
      OFH_CHECK(hipMemcpyAsync(&s, red, sizeof(double), hipMemcpyDeviceToHost, c.stream));
      OFH_CHECK(hipStreamSynchronize(c.stream));
      *dke = s / ((double)g.nx * (double)g.ny * (double)g.nz);
    }
    OFH_CHECK(hipGetLastError());
    OFH_CHECK(hipStreamSynchronize(c.stream));
  }
};

AdjointRun* adjoint_create(Context& c, const of_stepper_config& cfg, int every) {
  return new AdjointRun(c, cfg, every);
}
void adjoint_destroy(AdjointRun* a) { delete a; }
Stepper& adjoint_stepper(AdjointRun* a) { return a->fwd; }
void adjoint_set_forcing(AdjointRun* a, const double* const* f) { a->fwd.set_forcing(f); }
double adjoint_forward(AdjointRun* a, const double* const* u0, const double* p0, int steps, of_step_metrics* m) {
  return a->forward(u0, p0, steps, m);
}
void adjoint_final_velocity(AdjointRun* a, double* const* out) {
  if(a->steps < 1) throw InvalidArgument("run the forward pass first");
  for(int i = 0; i < 3; ++i)
    if(out[i]) a->copy(out[i], a->ufinal[i], a->lay[1 + i]);
  OFH_CHECK(hipStreamSynchronize(a->c.stream));
}
void adjoint_backward(AdjointRun* a, const double* const* seed, double* const* gu, double* gp, double* const* gf) {
  a->backward(seed, gu, gp, gf);
}
void adjoint_tangent(AdjointRun* a, const double* const* du0, const double* dp0, const double* const* df,
                     double* const* du_out, double* dke) {
  a->tangent(du0, dp0, df, du_out, dke);
}

}  // namespace ofh
