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

// ------------------------------------------------------------------ batched line solves
// One sweep direction of up to three velocity components in one launch. Each thread owns one
// line and solves it in two passes over HBM instead of three:
//  * pass 1 reads the right-hand side top-down and runs the forward elimination
//    mr_t = (r_t - lo*mr_{t-1}) / piv_t, keeping only mr at the end of every kSeg-element
//    segment (checkpoints, 1/kSeg of the line); for a cyclic line it also accumulates
//    x_0 = sum_t c_t mr_t with c_t = prod_{s<t} (-mu_s), the first element of the back
//    substitution, so the Sherman-Morrison factor fac = (x_0 + uc*x_{n-1}/gamma) / den is
//    known after this pass;
//  * pass 2 walks the segments bottom-up: it re-reads the segment's right-hand side,
//    recomputes its mr from the checkpoint (the same operations in the same order, so the
//    same values), runs the back substitution x_t = mr_t - mu_t*x_{t+1} and stores the
//    corrected solution x_t - fac*z_t once.
// HBM traffic is about 25 bytes per element (read, read, write, checkpoints) against 48 for
// separate forward, backward and correction passes. Only x_0 differs from the reference's
// recurrences (solver/projection.cpp:522-617): it is summed in the forward pass rather than
// produced by the back substitution, which changes its rounding, not its value.
constexpr int kSeg = 16;      // checkpoint spacing = register chunk of one pass
constexpr int kPlanVecs = 4;  // piv, mu, z, c per (component, axis) in the plan buffer
constexpr int kAdiCheckpointSlot = 20;  // Context scratch slot of the checkpoints

struct LineTask {
  double* f;
  double* ck;  // checkpoints [segment][line] of this task
  Lay L;
  LineSys S;
  const double* coef;  // piv | mu | z | c, nmax apart
  unsigned blocks;     // thread blocks of this task
  int nlines;
};
struct LineBatch {
  LineTask t[3];
  int ntasks;
  int nmax;
};
// Line coefficients are read-only during a sweep and indexed by the (uniform) element index:
// reading them through the constant address space lets the compiler use scalar loads.
typedef const double __attribute__((address_space(4)))* const_dptr;
__device__ inline const_dptr as_constant(const double* p) { return (const_dptr)p; }
struct LineCoef {
  const_dptr piv;
  const_dptr mu;
  const_dptr zv;
  const_dptr cx;
  __device__ LineCoef(const double* c, int nmax)
      : piv(as_constant(c)), mu(as_constant(c + nmax)), zv(as_constant(c + 2 * nmax)), cx(as_constant(c + 3 * nmax)) {}
};

// Element at a block-uniform base plus a per-lane byte offset below 2^32 (scalar base plus
// 32-bit vector offset addressing: no per-element vector address arithmetic).
__device__ inline double& at_bytes(double* base, unsigned boff) {
  return *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + boff);
}

// Task of a block (tasks occupy consecutive block ranges); b becomes the block within it.
// The task is assembled field by field from block-uniform selects, so it stays in scalar
// registers (a copy of a whole indexed element would be placed in per-thread memory).
__device__ inline LineTask line_task(const LineBatch& B, unsigned& b) {
  int ti = 0;
  if(B.ntasks > 1 && b >= B.t[0].blocks) {
    b -= B.t[0].blocks;
    ti = 1;
    if(B.ntasks > 2 && b >= B.t[1].blocks) {
      b -= B.t[1].blocks;
      ti = 2;
    }
  }
  LineTask T;
#define OFH_PICK(fld) T.fld = ti == 0 ? B.t[0].fld : ti == 1 ? B.t[1].fld : B.t[2].fld
  OFH_PICK(f);
  OFH_PICK(ck);
  OFH_PICK(L.ax);
  OFH_PICK(L.ay);
  OFH_PICK(L.az);
  OFH_PICK(L.sx);
  OFH_PICK(L.sy);
  OFH_PICK(L.sz);
  OFH_PICK(S.b);
  OFH_PICK(S.e);
  OFH_PICK(S.n);
  OFH_PICK(S.cyclic);
  OFH_PICK(S.copy_dup);
  OFH_PICK(S.lo);
  OFH_PICK(S.add_lo);
  OFH_PICK(S.add_hi);
  OFH_PICK(S.uc);
  OFH_PICK(S.gamma);
  OFH_PICK(S.den);
  OFH_PICK(S.single_pivot);
  OFH_PICK(coef);
  OFH_PICK(blocks);
  OFH_PICK(nlines);
#undef OFH_PICK
  return T;
}

// Forward elimination of elements [s0, s0 + kSeg) held in v (overwritten with mr); mr
// carries mr_{s0-1} in and mr of the segment's last element out. EDGE: the segment holds
// the first or the last element (boundary constants, end of line). ACC: accumulate x_0.
template <bool EDGE, bool ACC>
__device__ inline void line_forward(double (&v)[kSeg], int s0, int n, const LineSys& S, const LineCoef& C,
                                    double& mr, double& x0) {
#pragma unroll
  for(int e = 0; e < kSeg; ++e) {
    const int t = s0 + e;
    if(EDGE && t >= n) break;
    double r = v[e];
    if(EDGE && t == 0) {
      r = r + S.add_lo;
      if(n == 1) r = r + S.add_hi;
      mr = r / C.piv[0];
    } else {
      if(EDGE && t == n - 1) r = r + S.add_hi;
      mr = (r - S.lo * mr) / C.piv[t];
    }
    v[e] = mr;
    if(ACC) x0 = fma(C.cx[t], mr, x0);
  }
}

// Back substitution and correction of the segment: v holds mr in, the solution out; x
// carries x_{s0+kSeg} in and x_{s0} out. EDGE: the segment holds the last element.
template <bool EDGE>
__device__ inline void line_backward(double (&v)[kSeg], int s0, int n, const LineCoef& C, double fac, double& x) {
#pragma unroll
  for(int e = kSeg - 1; e >= 0; --e) {
    const int t = s0 + e;
    if(EDGE && t >= n) continue;
    if(EDGE && t == n - 1) x = v[e];
    else x = v[e] - C.mu[t] * x;
    v[e] = x - fac * C.zv[t];
  }
}

// Pass-1 step for one segment (forward elimination; mr leaves as the checkpoint).
__device__ inline void line_pass1(double (&v)[kSeg], int s0, int n, const LineSys& S, const LineCoef& C, double& mr,
                                  double& x0) {
  if(s0 > 0 && s0 + kSeg < n) {
    if(S.cyclic) line_forward<false, true>(v, s0, n, S, C, mr, x0);
    else line_forward<false, false>(v, s0, n, S, C, mr, x0);
  } else {
    if(S.cyclic) line_forward<true, true>(v, s0, n, S, C, mr, x0);
    else line_forward<true, false>(v, s0, n, S, C, mr, x0);
  }
}

// Pass-2 step for one segment: recompute mr from the checkpoint mr_{s0-1}, then back-substitute.
__device__ inline void line_pass2(double (&v)[kSeg], int s0, int n, const LineSys& S, const LineCoef& C, double mr_in,
                                  double fac, double& x) {
  double mr = mr_in, unused = 0.0;
  if(s0 > 0 && s0 + kSeg < n) {
    line_forward<false, false>(v, s0, n, S, C, mr, unused);
    line_backward<false>(v, s0, n, C, fac, x);
  } else {
    line_forward<true, false>(v, s0, n, S, C, mr, unused);
    if(s0 + kSeg >= n) line_backward<true>(v, s0, n, C, fac, x);
    else line_backward<false>(v, s0, n, C, fac, x);
  }
}

__device__ inline double sherman_morrison_factor(const LineSys& S, double x0, double x_last) {
  return S.cyclic ? (x0 + S.uc * x_last / S.gamma) / S.den : 0.0;
}

// y- and z-lines: lanes along i, so every load and store of a wave is a coalesced 512-byte
// row segment. The element address is a block-uniform pointer plus the lane's 32-bit offset,
// so the strided walk costs no vector registers. A pass double-buffers its segments in
// registers: a segment's loads overlap the recurrence of the previous one.
constexpr int kLSY = 4;
__global__ void __launch_bounds__(64 * kLSY) k_line_solve_s(LineBatch B, int axis) {
  unsigned b = blockIdx.x;
  const LineTask T = line_task(B, b);
  const Lay L = T.L;
  const LineSys S = T.S;
  const int o1 = axis == 2 ? 1 : 2;
  const unsigned gx = (unsigned)(L.ax + 63) / 64u;
  const int bx0 = (int)(b % gx) * 64, b10 = (int)(b / gx) * kLSY;
  const int t0 = bx0 + threadIdx.x, t1 = b10 + threadIdx.y;
  if(t0 >= L.ax || t1 >= L.a(o1)) return;  // no block-level synchronisation below
  const ptrdiff_t st = axis == 1 ? (ptrdiff_t)L.sx : (ptrdiff_t)L.sx * L.sy;
  const ptrdiff_t so1 = o1 == 1 ? (ptrdiff_t)L.sx : (ptrdiff_t)L.sx * L.sy;
  // block origin (uniform): i = 1 + bx0, the line axis at S.b, the other axis at 1 + b10
  double* const ub = T.f + L.at(1 + bx0, axis == 1 ? S.b : 1 + b10, axis == 1 ? 1 + b10 : S.b);
  const unsigned boff = 8u * (threadIdx.x + (unsigned)(threadIdx.y * so1));  // lane byte offset
  auto at = [&](int t) -> double& { return at_bytes(ub + (ptrdiff_t)t * st, boff); };
  const int n = S.n;
  if(S.cyclic && n == 1) {
    at(0) = at(0) / S.single_pivot;
    if(S.copy_dup) at(S.e - S.b) = at(0);
    return;
  }
  const LineCoef C(T.coef, B.nmax);
  double* const ckb = T.ck + (size_t)b10 * L.ax + bx0;  // checkpoints [segment][line]
  const unsigned cboff = 8u * (threadIdx.x + threadIdx.y * (unsigned)L.ax);
  auto ck = [&](int sg) -> double& { return at_bytes(ckb + (size_t)sg * T.nlines, cboff); };
  const int nseg = (n + kSeg - 1) / kSeg;
  auto ld = [&](double(&v)[kSeg], int sg) {
    const int s0 = sg * kSeg;
#pragma unroll
    for(int e = 0; e < kSeg; ++e) v[e] = at(min(s0 + e, n - 1));
  };
  auto sv = [&](const double(&v)[kSeg], int sg) {
    const int s0 = sg * kSeg;
    if(s0 + kSeg <= n) {
#pragma unroll
      for(int e = 0; e < kSeg; ++e) at(s0 + e) = v[e];
    } else {
#pragma unroll
      for(int e = 0; e < kSeg; ++e)
        if(s0 + e < n) at(s0 + e) = v[e];
    }
  };
  double va[kSeg], vb[kSeg];
  double mr = 0.0, x0 = 0.0;
  ld(va, 0);
  for(int sg = 0; sg < nseg; sg += 2) {  // pass 1
    if(sg + 1 < nseg) ld(vb, sg + 1);
    line_pass1(va, sg * kSeg, n, S, C, mr, x0);
    ck(sg) = mr;
    if(sg + 1 >= nseg) break;
    if(sg + 2 < nseg) ld(va, sg + 2);
    line_pass1(vb, (sg + 1) * kSeg, n, S, C, mr, x0);
    ck(sg + 1) = mr;
  }
  const double fac = sherman_morrison_factor(S, x0, mr);
  double x = 0.0, first = 0.0;
  ld(va, nseg - 1);
  for(int sg = nseg - 1; sg >= 0; sg -= 2) {  // pass 2
    if(sg >= 1) ld(vb, sg - 1);
    line_pass2(va, sg * kSeg, n, S, C, sg > 0 ? ck(sg - 1) : 0.0, fac, x);
    sv(va, sg);
    if(sg == 0) {
      first = va[0];
      break;
    }
    if(sg >= 2) ld(va, sg - 2);
    line_pass2(vb, (sg - 1) * kSeg, n, S, C, sg > 1 ? ck(sg - 2) : 0.0, fac, x);
    sv(vb, sg - 1);
    if(sg == 1) first = vb[0];
  }
  if(S.copy_dup) at(S.e - S.b) = first;
}

// x-lines (unit stride along the line): one wave owns 64 lines. Each segment moves between
// HBM and a 64 x kSeg LDS tile in row segments of kSeg elements (whole 128-byte lines), and
// every lane then runs its own line's recurrence on a register copy of its tile row. The
// next segment's loads are issued before the current segment is computed.
constexpr int kXR = 64 / kSeg;  // tile rows covered by one load instruction
__global__ void __launch_bounds__(64) k_line_solve_x(LineBatch B) {
  __shared__ double tile[64][kSeg + 1];
  unsigned b = blockIdx.x;
  const LineTask T = line_task(B, b);
  const Lay L = T.L;
  const LineSys S = T.S;
  const int nl1 = L.ay;
  const int line0 = (int)b * 64;
  const int nr = min(64, T.nlines - line0);
  const int lane = threadIdx.x;
  const int myl = line0 + min(lane, nr - 1);  // rows past the last line repeat it (loads only)
  // element offsets relative to the block's first line (the host bounds their span)
  const size_t base_off = L.at(S.b, 1 + line0 % nl1, 1 + line0 / nl1);
  const unsigned my_off = (unsigned)(L.at(S.b, 1 + myl % nl1, 1 + myl / nl1) - base_off);
  double* const f = T.f + base_off;
  const int n = S.n;
  if(S.cyclic && n == 1) {
    if(lane < nr) {
      f[my_off] = f[my_off] / S.single_pivot;
      if(S.copy_dup) f[my_off + (S.e - S.b)] = f[my_off];
    }
    return;
  }
  const LineCoef C(T.coef, B.nmax);
  double* const ckb = T.ck + line0;  // checkpoints [segment][line]
  auto ck = [&](int sg) -> double& { return at_bytes(ckb + (size_t)sg * T.nlines, 8u * lane); };
  const int nseg = (n + kSeg - 1) / kSeg;
  const int col = lane % kSeg, sub = lane / kSeg;
  unsigned roff[kSeg];  // byte offsets of the rows this lane moves
#pragma unroll
  for(int e = 0; e < kSeg; ++e) roff[e] = 8u * (unsigned)__shfl((int)my_off, kXR * e + sub, 64);
  double pre[kSeg];
  auto ld = [&](int sg) {
    const int s0 = sg * kSeg;
    double* const fs = f + s0 + min(col, n - 1 - s0);  // clamped: loads stay inside the line
#pragma unroll
    for(int e = 0; e < kSeg; ++e) pre[e] = at_bytes(fs, roff[e]);
  };
  auto to_tile = [&]() {
#pragma unroll
    for(int e = 0; e < kSeg; ++e) tile[kXR * e + sub][col] = pre[e];
  };
  auto store_tile = [&](int sg) {
    const int s0 = sg * kSeg;
    const bool cok = s0 + col < n;
    double* const fs = f + s0 + col;
#pragma unroll
    for(int e = 0; e < kSeg; ++e)
      if(kXR * e + sub < nr && cok) at_bytes(fs, roff[e]) = tile[kXR * e + sub][col];
  };
  double v[kSeg];
  double mr = 0.0, x0 = 0.0;
  ld(0);
  for(int sg = 0; sg < nseg; ++sg) {  // pass 1
    __syncthreads();
    to_tile();
    __syncthreads();
    if(sg + 1 < nseg) ld(sg + 1);
#pragma unroll
    for(int e = 0; e < kSeg; ++e) v[e] = tile[lane][e];
    line_pass1(v, sg * kSeg, n, S, C, mr, x0);
    if(lane < nr) ck(sg) = mr;
  }
  const double fac = sherman_morrison_factor(S, x0, mr);
  double x = 0.0, first = 0.0;
  ld(nseg - 1);
  for(int sg = nseg - 1; sg >= 0; --sg) {  // pass 2
    __syncthreads();
    to_tile();
    __syncthreads();
    if(sg >= 1) ld(sg - 1);
#pragma unroll
    for(int e = 0; e < kSeg; ++e) v[e] = tile[lane][e];
    line_pass2(v, sg * kSeg, n, S, C, sg > 0 && lane < nr ? ck(sg - 1) : 0.0, fac, x);
    if(sg == 0) first = v[0];
#pragma unroll
    for(int e = 0; e < kSeg; ++e) tile[lane][e] = v[e];
    __syncthreads();
    store_tile(sg);
  }
  if(S.copy_dup && lane < nr) f[my_off + (S.e - S.b)] = first;
}

// ------------------------------------------------------------------ periodic lines as convolutions
// On a periodic axis every line matrix of a sweep is the same circulant tridiag(-b, 1 + 2b, -b)
// of order n (the cells along the axis; the normal component's duplicated face is not an
// unknown). Its inverse is the circulant with first row
//   g_k = (rho^k + rho^(n-k)) / ((1 - rho^n) (b/rho - b rho)),  k = 0 .. n-1,
// where rho = 2b / ((1 + 2b) + sqrt((1 + 2b)^2 - 4 b^2)) is the smaller root of
// b r^2 - (1 + 2b) r + b = 0, so g decays like rho^|k| (rho ~ b for small b: 2e-3 on the 256^3
// Taylor-Green step). Where the taps beyond K weigh less than 1e-18 of g_0 the sweep is the
// (2K+1)-point convolution x_t = sum_{|m|<=K} g_|m| r_{t+m mod n}: the same solution as the
// cyclic Thomas / Sherman-Morrison recurrences of solver/projection.cpp:561-617 to rounding
// (about 1e-16 relative), computed with one read and one write per element and no serial
// recurrence. Sweeps with more than kConvMaxTaps taps, or lines shorter than 2K+1, keep the
// line solves.
constexpr int kConvMaxTaps = 16;
struct ConvTaps {
  double g[kConvMaxTaps + 1];  // g_0 .. g_K (zero beyond K)
  int k;
};

bool conv_taps(int n, double beta, ConvTaps& t) {
  if(!(beta > 0.0) || n < 3) return false;
  const long double b = beta;
  const long double d = 1.0L + 2.0L * b;
  const long double rho = 2.0L * b / (d + std::sqrt(d * d - 4.0L * b * b));
  const long double scale = 1.0L / ((1.0L - std::pow(rho, (long double)n)) * (b / rho - b * rho));
  std::vector<long double> g((size_t)n / 2 + 1);
  for(int k = 0; k <= n / 2; ++k) g[k] = scale * (std::pow(rho, (long double)k) + std::pow(rho, (long double)(n - k)));
  int K = -1;
  for(int k = 1; k <= std::min(kConvMaxTaps, n / 2); ++k) {
    long double tail = 0.0L;  // weight of the taps the convolution leaves out, both sides
    for(int m = k + 1; m <= n / 2; ++m) tail += (2 * m == n ? 1.0L : 2.0L) * g[m];
    if(tail <= 1e-18L * g[0]) {
      K = k;
      break;
    }
  }
  if(K < 0 || 2 * K + 1 > n) return false;
  for(int m = 0; m <= kConvMaxTaps; ++m) t.g[m] = m <= K ? (double)g[m] : 0.0;
  t.k = K;
  return true;
}

// x sweep of one component: one wave per grid row. The row's n unknowns (the first sweep's
// input, so the normal component's duplicated face pair is averaged as the periodic ghost fill
// does before the first sweep, solver/projection.cpp:233-253) go to an LDS buffer with K
// wrapped elements on each side; each lane then forms outputs i = lane, lane + 64, ... from
// the symmetric pairs, smallest taps first.
template <int K>
__global__ void __launch_bounds__(256) k_helm_conv_x(const double* __restrict__ in, double* __restrict__ out, Lay L,
                                                     int comp, int n, int nrows, ConvTaps taps) {
  extern __shared__ double conv_row[];
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int r = (int)blockIdx.x * 4 + wave;
  double* const buf = conv_row + (size_t)wave * (n + 2 * K);
  if(r < nrows) {
    const int j = 1 + r % L.ay, k = 1 + r / L.ay;
    const double* const row = in + L.at(0, j, k);
    // partner row of the averaged duplicate pair (components 1 and 2), or null
    const double* prow = nullptr;
    if(comp == 1 && (j == 1 || j == L.ay)) prow = in + L.at(0, j == 1 ? L.ay : 1, k);
    if(comp == 2 && (k == 1 || k == L.az)) prow = in + L.at(0, j, k == 1 ? L.az : 1);
    for(int m = lane; m < n + 2 * K; m += 64) {
      int t = m - K;
      t = t < 0 ? t + n : (t >= n ? t - n : t);
      double v = row[1 + t];
      if(comp == 0 && t == 0) v = 0.5 * (v + row[1 + n]);
      if(prow != nullptr) v = 0.5 * (v + prow[1 + t]);
      buf[m] = v;
    }
  }
  __syncthreads();
  if(r >= nrows) return;
  const int j = 1 + r % L.ay, k = 1 + r / L.ay;
  double* const orow = out + L.at(0, j, k);
  for(int i = lane; i < n; i += 64) {
    const double* c = buf + i + K;
    double acc = 0.0;
#pragma unroll
    for(int m = K; m >= 1; --m) acc = fma(taps.g[m], c[-m] + c[m], acc);
    acc = fma(taps.g[0], c[0], acc);
    orow[1 + i] = acc;
    if(comp == 0 && i == 0) orow[1 + n] = acc;  // the duplicated face
  }
}

// y or z sweep of one component: lanes along i (coalesced rows), each thread convolving a
// chunk of S consecutive elements of one line through a register window of 2K+1 inputs. Blocks
// are ordered chunk-fastest and grouped per XCD, so neighbouring chunks of a line, which share
// 2K halo elements, tend to run on one XCD and find those in its L2.
constexpr int kConvChunk = 32;
template <int K>
__global__ void __launch_bounds__(256) k_helm_conv_s(const double* __restrict__ in, double* __restrict__ out, Lay L,
                                                     int comp, int axis, int n, int nchunks, int ngx, ConvTaps taps) {
  constexpr int S = kConvChunk;
  const int o1 = axis == 2 ? 1 : 2;  // the other axis of the line's position
  unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int chunk = (int)(bid % (unsigned)nchunks);
  bid /= (unsigned)nchunks;
  const int bx = (int)(bid % (unsigned)ngx), by = (int)(bid / (unsigned)ngx);
  const int i = bx * 64 + (int)threadIdx.x;
  const int o = by * 4 + (int)threadIdx.y;
  if(i >= L.ax || o >= L.a(o1)) return;
  const size_t st = axis == 1 ? (size_t)L.sx : (size_t)L.sx * L.sy;
  const size_t base = axis == 1 ? L.at(1 + i, 1, 1 + o) : L.at(1 + i, 1 + o, 1);  // element t at base + t st
  const double* const src = in + base;
  double* const dst = out + base;
  const int t0 = chunk * S;
  auto wrap = [&](int t) { return t < 0 ? t + n : (t >= n ? t - n : t); };
  double v[S + 2 * K];
#pragma unroll
  for(int m = 0; m < S + 2 * K; ++m) v[m] = src[(size_t)wrap(min(t0 - K + m, n + K - 1)) * st];
#pragma unroll
  for(int s = 0; s < S; ++s) {
    const int t = t0 + s;
    if(t >= n) break;
    const double* c = v + s + K;
    double acc = 0.0;
#pragma unroll
    for(int m = K; m >= 1; --m) acc = fma(taps.g[m], c[-m] + c[m], acc);
    acc = fma(taps.g[0], c[0], acc);
    dst[(size_t)t * st] = acc;
    if(comp == axis && t == 0) dst[(size_t)n * st] = acc;  // the duplicated face
  }
}

template <int K>
void launch_conv_sweep(Context& c, const of_grid& g, int comp, int axis, const double* in, double* out,
                       const ConvTaps& taps) {
  const Lay L = make_lay(g, OF_FACE_X + comp);
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  if(axis == 0) {
    const int nrows = L.ay * L.az;
    const size_t lds = 4 * (size_t)(n + 2 * K) * sizeof(double);
    ProfScope prof_(c, "k_helm_conv_x", true);
    OFH_LAUNCH(k_helm_conv_x<K>, dim3((unsigned)div_up((size_t)nrows, 4)), dim3(64, 4), lds, c.stream, in, out, L,
               comp, n, nrows, taps);
  } else {
    const int o1 = axis == 2 ? 1 : 2;
    const int nchunks = (n + kConvChunk - 1) / kConvChunk;
    const int ngx = (L.ax + 63) / 64, ngy = (L.a(o1) + 3) / 4;
    ProfScope prof_(c, "k_helm_conv_s", true);
    OFH_LAUNCH(k_helm_conv_s<K>, dim3((unsigned)(nchunks * ngx * ngy)), dim3(64, 4), 0, c.stream, in, out, L, comp,
               axis, n, nchunks, ngx, taps);
  }
}

void conv_sweep(Context& c, const of_grid& g, int comp, int axis, const double* in, double* out, const ConvTaps& t) {
  if(t.k <= 8) launch_conv_sweep<8>(c, g, comp, axis, in, out, t);
  else launch_conv_sweep<16>(c, g, comp, axis, in, out, t);
}

// The convolution taps of the three axes when every axis is periodic and every sweep qualifies.
bool conv_plan(const of_grid& g, const of_bc_set& b, double alpha, ConvTaps (&t)[3]) {
  if(alpha <= 0.0) return false;
  const int n[3] = {g.nx, g.ny, g.nz};
  for(int a = 0; a < 3; ++a) {
    if(!axis_periodic(b, a)) return false;
    const double h = spacing(g, a);
    if(!conv_taps(n[a], alpha / (h * h), t[a])) return false;
    if(a == 0 && 4 * (size_t)(n[a] + 2 * kConvMaxTaps) * sizeof(double) > 64 * 1024) return false;
  }
  return true;
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
  std::vector<double> host((size_t)kPlanVecs * nmax * 9, 0.0);
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
      // per-lane byte offsets of the line kernels are 32-bit: a block's rows must span < 4 GiB
      const double span = 8.0 * L.sx * L.sy * (64.0 / L.ay + 2.0) + 8.0 * L.sx;
      if(span >= 4294967296.0) throw InvalidArgument("ADI grid planes are too large for the line kernels");
      // c_t = prod_{s<t} (-mu_s): x_0 of the back substitution is sum_t c_t mr_t
      std::vector<double> cx(S.n, 0.0);
      if(S.cyclic) {
        double cprod = 1.0;
        for(int t = 0; t < S.n; ++t) {
          cx[t] = cprod;
          cprod *= -mu[t];
        }
      }
      const size_t off = (size_t)(comp * 3 + axis) * kPlanVecs * nmax;
      std::copy(piv.begin(), piv.end(), host.begin() + off);
      std::copy(mu.begin(), mu.end(), host.begin() + off + nmax);
      std::copy(zv.begin(), zv.end(), host.begin() + off + 2 * nmax);
      std::copy(cx.begin(), cx.end(), host.begin() + off + 3 * nmax);
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

// Line solves of every component along `axis` (one launch); returns the mask of the
// components swept. no_dup drops the duplicated-face copy (the transposed solve folds it
// back beforehand).
int launch_axis_sweeps(Context& c, const of_grid& g, const AdiPlan& P, int axis, double* const out[3], bool no_dup) {
  LineBatch B{};
  B.nmax = P.nmax;
  unsigned blocks = 0;
  size_t ck_total = 0;
  size_t ck_off[3];
  int mask = 0;
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  for(const Pending& p : P.sweeps) {
    if(p.axis != axis) continue;
    LineTask& T = B.t[B.ntasks];
    T.f = out[p.comp];
    T.L = make_lay(g, OF_FACE_X + p.comp);
    T.S = p.S;
    if(no_dup) T.S.copy_dup = 0;
    T.coef = P.dev + p.off;
    T.nlines = T.L.a(o0) * T.L.a(o1);
    T.blocks = axis == 0 ? div_up((size_t)T.nlines, 64) : div_up((size_t)T.L.ax, 64) * div_up((size_t)T.L.a(o1), kLSY);
    ck_off[B.ntasks++] = ck_total;
    ck_total += (size_t)T.nlines * (size_t)((p.S.n + kSeg - 1) / kSeg);
    blocks += T.blocks;
    mask |= 1 << p.comp;
  }
  if(B.ntasks == 0) return 0;
  // checkpoint scratch of the context (grow-only; one launch at a time on its stream)
  double* ck = static_cast<double*>(c.scratch_get(kAdiCheckpointSlot, ck_total * sizeof(double)));
  for(int t = 0; t < B.ntasks; ++t) B.t[t].ck = ck + ck_off[t];
  if(axis == 0) {
    ProfScope prof_(c, "k_line_solve_x", true);
    OFH_LAUNCH(k_line_solve_x, dim3(blocks), dim3(64), 0, c.stream, B);
  } else {
    ProfScope prof_(c, "k_line_solve_s", true);
    OFH_LAUNCH(k_line_solve_s, dim3(blocks), dim3(64, kLSY), 0, c.stream, B, axis);
  }
  return mask;
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
                  const double* rw, double* ou, double* ov, double* ow, double* const* scratch, bool keep_storage) {
  if(alpha < 0.0) throw InvalidArgument("solve_predictor_adi expects a non-negative alpha");
  const double* rhs[3] = {ru, rv, rw};
  double* out[3] = {ou, ov, ow};
  const AdiPlan P = plan_adi(c, g, b, alpha);
  ConvTaps taps[3];
  if(conv_plan(g, b, alpha, taps)) {
    // every axis periodic: x sweep rhs -> out, y sweep out -> tmp, z sweep tmp -> out, then
    // the ghost fill; the sweeps read no ghosts, and the duplicated faces are equal after the
    // first sweep, so the reference's fills between sweeps change nothing they read
    double* tmp[3];
    for(int comp = 0; comp < 3; ++comp) {
      const Lay L = make_lay(g, OF_FACE_X + comp);
      if(keep_storage)  // the reference's result starts as a copy of the whole rhs storage
        OFH_CHECK(hipMemcpyAsync(out[comp], rhs[comp], L.size() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
      tmp[comp] = scratch != nullptr ? scratch[comp]
                                     : static_cast<double*>(c.scratch_get(kAdiCheckpointSlot + 1 + comp,
                                                                          L.size() * sizeof(double)));
    }
    for(int comp = 0; comp < 3; ++comp) conv_sweep(c, g, comp, 0, rhs[comp], out[comp], taps[0]);
    for(int comp = 0; comp < 3; ++comp) conv_sweep(c, g, comp, 1, out[comp], tmp[comp], taps[1]);
    for(int comp = 0; comp < 3; ++comp) conv_sweep(c, g, comp, 2, tmp[comp], out[comp], taps[2]);
    launch_velocity_bc(c, g, b, out[0], out[1], out[2]);
    OFH_CHECK(hipGetLastError());
    return P.line_solves;
  }
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    OFH_CHECK(hipMemcpyAsync(out[comp], rhs[comp], L.size() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  }
  // The components are independent, so the reference's per-component order (ghost fill, then
  // x, y, z sweeps each followed by a ghost fill) is run axis by axis for all of them.
  launch_velocity_bc(c, g, b, out[0], out[1], out[2]);
  if(alpha != 0.0) {
    for(int axis = 0; axis < 3; ++axis) {
      const int mask = launch_axis_sweeps(c, g, P, axis, out, false);
      if(mask) launch_component_bc_mask(c, g, b, out[0], out[1], out[2], mask);
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
  double* out[3];
  for(int comp = 0; comp < 3; ++comp) {
    const Lay L = make_lay(g, OF_FACE_X + comp);
    out[comp] = rbar[comp];
    if(rbar[comp] != ybar[comp])
      OFH_CHECK(hipMemcpyAsync(rbar[comp], ybar[comp], L.size() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    launch_component_bc_adjoint(c, g, comp, rbar[comp]);
  }
  if(alpha == 0.0) return;
  for(int axis = 2; axis >= 0; --axis) {
    for(const Pending& p : P.sweeps) {
      if(p.axis != axis || !p.S.copy_dup) continue;
      const Lay L = make_lay(g, OF_FACE_X + p.comp);
      const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
      const dim3 block(64, 4, 1);
      const dim3 grid((unsigned)((L.a(o0) + 63) / 64), (unsigned)((L.a(o1) + 3) / 4), 1);
      OFH_LAUNCH(k_dup_fold, grid, block, 0, c.stream, rbar[p.comp], L, axis, p.S.b, p.S.e);
    }
    const int mask = launch_axis_sweeps(c, g, P, axis, out, true);
    for(int comp = 0; comp < 3; ++comp)
      if(mask & (1 << comp)) launch_component_bc_adjoint(c, g, comp, rbar[comp]);
  }
  OFH_CHECK(hipGetLastError());
}

}  // namespace ofh
