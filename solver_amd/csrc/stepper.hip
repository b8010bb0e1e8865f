// Device-resident fractional-step driver: the step loop of run_taylor_green_steps
// (solver/taylor_green.cpp:540-634), advance_lid_driven_cavity_impl
// (solver/lid_driven_cavity.cpp:524-624) and run_channel_flow (solver/channel_flow.cpp:432-523)
// with every field kept in HBM. The cavity and the channels share the two-dimensional predictor
// right-hand side (solver/lid_driven_cavity.cpp:488-518, solver/channel_flow.cpp:226-256).

#include "stepper.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace ofh {

bool axis_periodic(const of_bc_set& b, int axis);

namespace {

// Step diagnostics over the new velocity (solver/taylor_green.cpp:607-616): divergence RMS and
// max (metric), CFL (solver/momentum_terms.cpp:373-401), max |u_new - u_old| over every active
// face (solver/taylor_green.cpp:100-118) and a non-finite count. Cell (i,j,k) owns the lower
// face of each component, and the last cell along an axis also owns the upper face.
__global__ void k_step_diag(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                            const double* __restrict__ uo, const double* __restrict__ vo,
                            const double* __restrict__ wo, Lay Lu, Lay Lv, Lay Lw, of_grid g, double dt,
                            double* partials, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  // 0 sum div^2, 1 max|div|, 2 max cfl, 3 max|uc|, 4 max|vc|, 5 max|wc|, 6 max|du|, 7 non-finite
  double vals[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if(i <= g.nx && j <= g.ny) {
    const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + g.nz);
    for(int k = kb; k < ke; ++k) {
      const double uw = u[Lu.at(i, j, k)], ue = u[Lu.at(i + 1, j, k)];
      const double vs = v[Lv.at(i, j, k)], vn = v[Lv.at(i, j + 1, k)];
      const double wb = w[Lw.at(i, j, k)], wf = w[Lw.at(i, j, k + 1)];
      const double a = (ue - uw) / g.dx;
      const double b = (vn - vs) / g.dy;
      const double c = (wf - wb) / g.dz;
      const double dv = a + b + c;
      vals[0] += dv * dv;
      vals[1] = fmax(vals[1], fabs(dv));
      const double uc = 0.5 * (uw + ue), vc = 0.5 * (vs + vn), wc = 0.5 * (wb + wf);
      vals[3] = fmax(vals[3], fabs(uc));
      vals[4] = fmax(vals[4], fabs(vc));
      vals[5] = fmax(vals[5], fabs(wc));
      vals[2] = fmax(vals[2], dt * (fabs(uc) / g.dx + fabs(vc) / g.dy + fabs(wc) / g.dz));
      double ch = fabs(uw - uo[Lu.at(i, j, k)]);
      ch = fmax(ch, fabs(vs - vo[Lv.at(i, j, k)]));
      ch = fmax(ch, fabs(wb - wo[Lw.at(i, j, k)]));
      if(i == g.nx) ch = fmax(ch, fabs(ue - uo[Lu.at(i + 1, j, k)]));
      if(j == g.ny) ch = fmax(ch, fabs(vn - vo[Lv.at(i, j + 1, k)]));
      if(k == g.nz) ch = fmax(ch, fabs(wf - wo[Lw.at(i, j, k + 1)]));
      vals[6] = fmax(vals[6], ch);
      if(!isfinite(dv) || !isfinite(ch) || !isfinite(uc) || !isfinite(vc) || !isfinite(wc)) vals[7] += 1.0;
    }
  }
  const int ops[8] = {kSum, kMax, kMax, kMax, kMax, kMax, kMax, kSum};
  store_partials<8>(vals, ops, partials);
}
constexpr unsigned kStepDiagMaxMask = 0x7eu;  // values 1..6 are maxima

// The end of a step on a grid periodic on every axis, in one pass over the cells: the corrected
// velocity on every active face (each cell's lower faces, and the upper face at the end of a
// line), the total-pressure update and the step diagnostics, each value by the expression of
// k_correct, k_pressure_update and k_step_diag (the maxima are those of k_step_diag; the sums
// differ by rounding, columns being longer here). On such a grid the two copies of a duplicated face
// are bitwise equal before the ghost fill (pred carries its filled ghosts, phi its periodic
// copies), so the fill that follows leaves every value read here unchanged. 13 array passes
// against 19 for the three kernels.
__global__ void k_finish_periodic(const double* __restrict__ pu, const double* __restrict__ pv,
                                  const double* __restrict__ pw, const double* __restrict__ phi,
                                  const double* __restrict__ phi_mean, const double* __restrict__ prhs, const double* __restrict__ uo,
                                  const double* __restrict__ vo, const double* __restrict__ wo,
                                  double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                                  double* __restrict__ pt, Lay Lu, Lay Lv, Lay Lw, Lay Lc, of_grid g, double scale,
                                  double sp, double dt, double* partials, int kz) {
  const uint3 bk = remapped_block3();
  const int i = 1 + (int)bk.x * blockDim.x + threadIdx.x;
  const int j = 1 + (int)bk.y * blockDim.y + threadIdx.y;
  double vals[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // the solve's final mean of phi, subtracted on every read of phi (its ghosts are copies of
  // active values taken before the subtraction): the values of the subtracted field, bitwise
  const double m = phi_mean != nullptr ? *phi_mean : 0.0;
  if(i <= g.nx && j <= g.ny) {
    const int kb = 1 + (int)bk.z * kz, ke = min(kb + kz, 1 + g.nz);
    // per-thread column offsets plus uniform plane strides (no 64-bit index products per point)
    const size_t sc = (size_t)Lc.sx * Lc.sy, su = (size_t)Lu.sx * Lu.sy, sv = (size_t)Lv.sx * Lv.sy,
                 sw = (size_t)Lw.sx * Lw.sy;
    const size_t oc = Lc.at(i, j, 0), ou = Lu.at(i, j, 0), ov = Lv.at(i, j, 0), ow = Lw.at(i, j, 0);
    const size_t cy = (size_t)Lc.sx, vy = (size_t)Lv.sx;
    const bool xe = i == g.nx, ye = j == g.ny;
    // phi of plane k and the lower w face of plane k carried along the column: wb of plane k is
    // wf of plane k - 1 (the same operands, so the same value)
    const double phm = phi[oc + sc * (size_t)(kb - 1)] - m;
    double ph0 = phi[oc + sc * (size_t)kb] - m;
    double wb = pw[ow + sw * (size_t)kb] + scale * ((ph0 - phm) / g.dz);
    double wob = wo[ow + sw * (size_t)kb];
    for(int k = kb; k < ke; ++k) {
      const size_t c0 = oc + sc * (size_t)k, u0 = ou + su * (size_t)k, v0 = ov + sv * (size_t)k;
      const size_t w1 = ow + sw * (size_t)(k + 1);
      // every load is unconditional (the upper-face values of the line ends are read by every
      // thread and used where they apply), so no branch join waits for a load
      const double php = phi[c0 + sc] - m, pxm = phi[c0 - 1] - m, pxp = phi[c0 + 1] - m,
                   pym = phi[c0 - cy] - m, pyp = phi[c0 + cy] - m;
      const double pu0 = pu[u0], pu1 = pu[u0 + 1], pv0 = pv[v0], pv1 = pv[v0 + vy], pw1 = pw[w1];
      const double uo0 = uo[u0], uo1 = uo[u0 + 1], vo0 = vo[v0], vo1 = vo[v0 + vy], wo1 = wo[w1];
      const double ptc = pt[c0], rc = prhs[c0];
      const double pc = ph0;
      // k_correct: out = pred + scale * (phi(i) - phi(i - 1)) / h
      const double uw = pu0 + scale * ((pc - pxm) / g.dx);
      const double ue = pu1 + scale * ((pxp - pc) / g.dx);
      const double vs = pv0 + scale * ((pc - pym) / g.dy);
      const double vn = pv1 + scale * ((pyp - pc) / g.dy);
      const double wf = pw1 + scale * ((php - pc) / g.dz);
      u[u0] = uw;
      v[v0] = vs;
      w[w1 - sw] = wb;
      if(xe) u[u0 + 1] = ue;
      if(ye) v[v0 + vy] = vn;
      if(k == g.nz) w[w1] = wf;
      // k_pressure_update
      double p = ptc + 1.0 * pc;
      p = p + sp * rc;
      pt[c0] = p;
      // k_step_diag
      const double a = (ue - uw) / g.dx;
      const double b = (vn - vs) / g.dy;
      const double c = (wf - wb) / g.dz;
      const double dv = a + b + c;
      vals[0] += dv * dv;
      vals[1] = fmax(vals[1], fabs(dv));
      const double uc = 0.5 * (uw + ue), vc = 0.5 * (vs + vn), wc = 0.5 * (wb + wf);
      vals[3] = fmax(vals[3], fabs(uc));
      vals[4] = fmax(vals[4], fabs(vc));
      vals[5] = fmax(vals[5], fabs(wc));
      vals[2] = fmax(vals[2], dt * (fabs(uc) / g.dx + fabs(vc) / g.dy + fabs(wc) / g.dz));
      double ch = fabs(uw - uo0);
      ch = fmax(ch, fabs(vs - vo0));
      ch = fmax(ch, fabs(wb - wob));
      if(xe) ch = fmax(ch, fabs(ue - uo1));
      if(ye) ch = fmax(ch, fabs(vn - vo1));
      if(k == g.nz) ch = fmax(ch, fabs(wf - wo1));
      vals[6] = fmax(vals[6], ch);
      if(!isfinite(dv) || !isfinite(ch) || !isfinite(uc) || !isfinite(vc) || !isfinite(wc)) vals[7] += 1.0;
      ph0 = php;
      wb = wf;
      wob = wo1;
    }
  }
  const int ops[8] = {kSum, kMax, kMax, kMax, kMax, kMax, kMax, kSum};
  store_partials<8>(vals, ops, partials);
}

// p_total += phi; p_total += (-nu dt / 2) rhs   (solver/taylor_green.cpp:607-608)
__global__ void k_pressure_update(double* __restrict__ pt, const double* __restrict__ phi,
                                  const double* __restrict__ rhs, Lay L, double s, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + L.az);
  for(int k = kb; k < ke; ++k) {
    const size_t at = L.at(i, j, k);
    double p = pt[at] + 1.0 * phi[at];
    p = p + s * rhs[at];
    pt[at] = p;
  }
}

// Taylor-Green initial field (solver/taylor_green.cpp:335-389) on the device.
__global__ void k_tg_init(double* f, Lay L, int loc, of_grid g, int three_d, int convention) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  for(int k = 1; k <= L.az; ++k) {
    // core/field_layout.hpp:366-377
    const double x = loc == OF_FACE_X ? (double)(i - 1) * g.dx : ((double)(i - 1) + 0.5) * g.dx;
    const double y = loc == OF_FACE_Y ? (double)(j - 1) * g.dy : ((double)(j - 1) + 0.5) * g.dy;
    const double z = loc == OF_FACE_Z ? (double)(k - 1) * g.dz : ((double)(k - 1) + 0.5) * g.dz;
    double val;
    if(loc == OF_FACE_X) {
      const double zf = three_d ? cos(z) : 1.0;
      val = convention == 0 ? -cos(x) * sin(y) * zf * 1.0 : sin(x) * cos(y) * zf * 1.0;
    } else if(loc == OF_FACE_Y) {
      const double zf = three_d ? cos(z) : 1.0;
      val = convention == 0 ? sin(x) * cos(y) * zf * 1.0 : -cos(x) * sin(y) * zf * 1.0;
    } else if(loc == OF_FACE_Z) {
      val = 0.0;
    } else {
      const double sg = convention == 0 ? 1.0 : -1.0;
      if(three_d)
        val = -sg * 0.0625 * (cos(2.0 * x) + cos(2.0 * y)) * (cos(2.0 * z) + 2.0) * 1.0;
      else
        val = -sg * 0.25 * (cos(2.0 * x) + cos(2.0 * y)) * 1.0;
    }
    f[L.at(i, j, k)] = val;
  }
}

// Steady channel profile (solver/channel_flow.cpp:160-224) on the active range: u(y) on the x
// faces (Couette U y, Poiseuille G y (1 - y) / (2 nu)), total pressure G (1 - x) for Poiseuille;
// v, w and the Couette pressure stay zero.
__global__ void k_channel_init(double* f, Lay L, int loc, of_grid g, int poiseuille, double top, double drop,
                               double nu) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if(i > L.ax || j > L.ay) return;
  for(int k = 1; k <= L.az; ++k) {
    double val;
    if(loc == OF_FACE_X) {
      const double y = ((double)(j - 1) + 0.5) * g.dy;  // core/field_layout.hpp:366-377
      val = poiseuille ? 0.5 * drop * y * (1.0 - y) / nu : top * y;
    } else {
      const double x = ((double)(i - 1) + 0.5) * g.dx;
      val = poiseuille ? drop * (1.0 - x) : 0.0;
    }
    f[L.at(i, j, k)] = val;
  }
}

// kinetic energy sum (solver/taylor_green.cpp:269-288)
__global__ void k_kinetic(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                          Lay Lu, Lay Lv, Lay Lw, of_grid g, double* partials, int kz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  double vals[1] = {0.0};
  if(i <= g.nx && j <= g.ny) {
    const int kb = 1 + blockIdx.z * kz, ke = min(kb + kz, 1 + g.nz);
    for(int k = kb; k < ke; ++k) {
      const double uc = 0.5 * (u[Lu.at(i, j, k)] + u[Lu.at(i + 1, j, k)]);
      const double vc = 0.5 * (v[Lv.at(i, j, k)] + v[Lv.at(i, j + 1, k)]);
      const double wc = 0.5 * (w[Lw.at(i, j, k)] + w[Lw.at(i, j, k + 1)]);
      vals[0] += 0.5 * (uc * uc + vc * vc + wc * wc);
    }
  }
  const int ops[1] = {kSum};
  store_partials<1>(vals, ops, partials);
}

}  // namespace

of_bc_set stepper_bcs(const of_stepper_config& c) {
  of_bc_set b{};
  if(c.flow_case == OF_CASE_TAYLOR_GREEN) {  // solver/taylor_green.cpp:510-517
    for(int f = 0; f < 6; ++f) b.faces[f].type = OF_PERIODIC;
    if(c.nz <= 1) b.faces[4].type = b.faces[5].type = OF_SYMMETRY;
  } else if(c.flow_case == OF_CASE_LID_DRIVEN_CAVITY) {  // solver/lid_driven_cavity.cpp:362-370
    for(int f = 0; f < 6; ++f) b.faces[f].type = OF_NO_SLIP_WALL;
    b.faces[3].type = OF_PRESCRIBED_VELOCITY;
    b.faces[3].velocity[0] = c.lid_velocity;
    b.faces[4].type = b.faces[5].type = OF_SYMMETRY;
  } else {  // solver/channel_flow.cpp:408-430
    for(int f = 0; f < 6; ++f) b.faces[f].type = OF_NO_SLIP_WALL;
    b.faces[4].type = b.faces[5].type = OF_SYMMETRY;
    if(c.flow_case == OF_CASE_CHANNEL_COUETTE) {
      b.faces[0].type = b.faces[1].type = OF_PERIODIC;
      b.faces[3].type = OF_PRESCRIBED_VELOCITY;
      b.faces[3].velocity[0] = c.lid_velocity;
    } else {
      b.faces[0].type = b.faces[1].type = OF_FIXED_PRESSURE;
      b.faces[0].pressure = c.pressure_drop;
      b.faces[1].pressure = 0.0;
    }
  }
  return b;
}

of_pressure_bc_set derive_pressure_bcs(const of_bc_set& b) {  // solver/projection.cpp:934-964
  of_pressure_bc_set m{};
  for(int f = 0; f < 6; ++f) {
    switch(b.faces[f].type) {
      case OF_FIXED_PRESSURE: m.faces[f].type = OF_PRESSURE_DIRICHLET; break;
      case OF_PERIODIC: m.faces[f].type = OF_PRESSURE_PERIODIC; break;
      default: m.faces[f].type = OF_PRESSURE_NEUMANN; break;
    }
    m.faces[f].value = 0.0;
    m.faces[f].gradient = 0.0;
  }
  for(int a = 0; a < 3; ++a) {
    const bool lo = m.faces[2 * a].type == OF_PRESSURE_PERIODIC, hi = m.faces[2 * a + 1].type == OF_PRESSURE_PERIODIC;
    if(lo != hi) throw InvalidArgument("periodic pressure boundaries must be paired on both sides");
  }
  return m;
}

Stepper::Stepper(Context& ctx, const of_stepper_config& cfg) : c(ctx), cfg(cfg) {
  if(cfg.nx <= 0 || cfg.ny <= 0 || cfg.nz <= 0) throw InvalidArgument("stepper expects positive nx, ny, nz");
  if(!(cfg.viscosity > 0.0)) throw InvalidArgument("stepper expects viscosity > 0");
  if(!(cfg.dt > 0.0)) throw InvalidArgument("stepper expects dt > 0");
  if(cfg.poisson_max_iterations <= 0 || !(cfg.poisson_tolerance > 0.0))
    throw InvalidArgument("stepper expects positive Poisson controls");
  const double L = 2.0 * std::acos(-1.0);
  if(cfg.flow_case == OF_CASE_TAYLOR_GREEN) {  // solver/taylor_green.cpp:497-508
    g = of_grid{cfg.nx, cfg.ny, cfg.nz, L / (double)cfg.nx, L / (double)cfg.ny, cfg.nz > 1 ? L / (double)cfg.nz : 1.0};
  } else if(cfg.flow_case == OF_CASE_LID_DRIVEN_CAVITY) {  // solver/lid_driven_cavity.cpp:372-380
    if(cfg.nz != 1) throw InvalidArgument("the lid-driven cavity is two-dimensional (nz = 1)");
    g = of_grid{cfg.nx, cfg.ny, 1, 1.0 / (double)cfg.nx, 1.0 / (double)cfg.ny, 1.0};
  } else if(cfg.flow_case == OF_CASE_CHANNEL_COUETTE ||
            cfg.flow_case == OF_CASE_CHANNEL_POISEUILLE) {  // solver/channel_flow.cpp:435-449
    if(cfg.nz != 1) throw InvalidArgument("channel flow is two-dimensional (nz = 1)");
    if(cfg.flow_case == OF_CASE_CHANNEL_COUETTE && !(cfg.lid_velocity > 0.0))
      throw InvalidArgument("Couette config expects top_velocity > 0");
    if(cfg.flow_case == OF_CASE_CHANNEL_POISEUILLE && !(cfg.pressure_drop > 0.0))
      throw InvalidArgument("Poiseuille config expects pressure_drop > 0");
    g = of_grid{cfg.nx, cfg.ny, 1, 1.0 / (double)cfg.nx, 1.0 / (double)cfg.ny, 1.0};
  } else {
    throw InvalidArgument("unknown flow case");
  }
  bcs = stepper_bcs(cfg);
  pbcs = derive_pressure_bcs(bcs);
  all_periodic = axis_periodic(bcs, 0) && axis_periodic(bcs, 1) && axis_periodic(bcs, 2);
  const size_t fu = make_lay(g, OF_FACE_X).size(), fv = make_lay(g, OF_FACE_Y).size(),
               fw = make_lay(g, OF_FACE_Z).size(), fc = make_lay(g, OF_CELL).size();
  const size_t sizes[3] = {fu, fv, fw};
  auto alloc = [&](double** p, size_t n) {
    OFH_CHECK(hipMalloc(p, n * sizeof(double)));
    OFH_CHECK(hipMemsetAsync(*p, 0, n * sizeof(double), c.stream));
    owned.push_back(*p);
  };
  for(int a = 0; a < 3; ++a) {
    alloc(&vel[a], sizes[a]);
    alloc(&adv[a], sizes[a]);
    alloc(&rhs[a], sizes[a]);
    alloc(&pred[a], sizes[a]);
    alloc(&corr[a], sizes[a]);
    if(!all_periodic) alloc(&dif[a], sizes[a]);
  }
  alloc(&p_total, fc);
  alloc(&phi, fc);
  alloc(&prhs, fc);
  OFH_CHECK(hipMalloc(&diag, 32 * sizeof(double)));
  owned.push_back(diag);
  poisson = poisson_create(c, g, pbcs, cfg.precision);
  std::memset(&metrics, 0, sizeof metrics);
  metrics.dt = cfg.dt;
  OFH_CHECK(hipStreamSynchronize(c.stream));
}

Stepper::~Stepper() {
  if(poisson) poisson_destroy(poisson);
  if(hcg) helmholtz_cg_destroy(hcg);
  for(double* p : owned) OFH_RELEASE(hipFree(p));
}

void Stepper::set_predictor(int kind, double tol, int maxit) {
  if(kind != OF_PREDICTOR_ADI && kind != OF_PREDICTOR_CG) throw InvalidArgument("unknown predictor kind");
  if(kind == OF_PREDICTOR_CG) {
    if(!all_periodic) throw InvalidArgument("the CG Helmholtz predictor supports periodic boundaries");
    if(!(tol > 0.0)) throw InvalidArgument("the CG Helmholtz predictor expects tolerance > 0");
    if(maxit <= 0) throw InvalidArgument("the CG Helmholtz predictor expects max_iterations > 0");
    if(!hcg) hcg = helmholtz_cg_create(c, g);
    predictor_tol = tol;
    predictor_maxit = maxit;
  }
  predictor_kind = kind;
}

void Stepper::total_pressure_bc(const double* const* v) {
  if(all_periodic) {
    launch_total_pressure_bc(c, g, bcs, v[0], v[1], v[2], p_total);  // periodic copies only
    return;
  }
  for(int a = 0; a < 3; ++a) launch_laplacian(c, g, OF_FACE_X + a, cfg.viscosity, v[a], dif[a]);
  launch_total_pressure_bc(c, g, bcs, dif[0], dif[1], dif[2], p_total);
}

void Stepper::set_forcing(const double* const* f) {
  if(f == nullptr || f[0] == nullptr) {
    for(int a = 0; a < 3; ++a) force[a] = nullptr;  // buffers stay owned until destruction
    return;
  }
  if(cfg.flow_case != OF_CASE_TAYLOR_GREEN) throw InvalidArgument("a body force is supported for Taylor-Green only");
  for(int a = 0; a < 3; ++a) {
    if(f[a] == nullptr) throw InvalidArgument("forcing needs all three components");
    const size_t n = make_lay(g, OF_FACE_X + a).size();
    if(force_buf[a] == nullptr) {
      OFH_CHECK(hipMalloc(&force_buf[a], n * sizeof(double)));
      owned.push_back(force_buf[a]);
    }
    OFH_CHECK(hipMemcpyAsync(force_buf[a], f[a], n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    force[a] = force_buf[a];
  }
}

void Stepper::initialize(int convention) {
  const int three_d = g.nz > 1;
  for(int a = 0; a < 3; ++a) OFH_CHECK(hipMemsetAsync(adv[a], 0, make_lay(g, OF_FACE_X + a).size() * 8, c.stream));
  for(int a = 0; a < 3; ++a) OFH_CHECK(hipMemsetAsync(vel[a], 0, make_lay(g, OF_FACE_X + a).size() * 8, c.stream));
  OFH_CHECK(hipMemsetAsync(p_total, 0, make_lay(g, OF_CELL).size() * 8, c.stream));
  if(cfg.flow_case == OF_CASE_TAYLOR_GREEN) {
    for(int loc = 0; loc < 4; ++loc) {
      const Lay L = make_lay(g, loc);
      double* f = loc == OF_CELL ? p_total : vel[loc - 1];
      const dim3 block(64, 4, 1);
      const dim3 grid((unsigned)((L.ax + 63) / 64), (unsigned)((L.ay + 3) / 4), 1);
      OFH_LAUNCH(k_tg_init, grid, block, 0, c.stream, f, L, loc, g, three_d, convention);
    }
  } else if(cfg.flow_case != OF_CASE_LID_DRIVEN_CAVITY) {
    const int pois = cfg.flow_case == OF_CASE_CHANNEL_POISEUILLE;
    for(int loc : {OF_FACE_X, OF_CELL}) {
      const Lay L = make_lay(g, loc);
      double* f = loc == OF_CELL ? p_total : vel[0];
      const dim3 block(64, 4, 1);
      const dim3 grid((unsigned)((L.ax + 63) / 64), (unsigned)((L.ay + 3) / 4), 1);
      OFH_LAUNCH(k_channel_init, grid, block, 0, c.stream, f, L, loc, g, pois, cfg.lid_velocity,
                         cfg.pressure_drop, cfg.viscosity);
    }
  }
  launch_velocity_bc(c, g, bcs, vel[0], vel[1], vel[2]);
  total_pressure_bc(vel);
  OFH_CHECK(hipGetLastError());
  has_prev = 0;
  std::memset(&metrics, 0, sizeof metrics);
  metrics.dt = cfg.dt;
  OFH_CHECK(hipStreamSynchronize(c.stream));
}

void Stepper::step() {
  const double dt = cfg.dt, nu = cfg.viscosity;
  const char* name = cfg.flow_case == OF_CASE_TAYLOR_GREEN        ? "Taylor-Green"
                     : cfg.flow_case == OF_CASE_LID_DRIVEN_CAVITY ? "lid-driven cavity"
                                                                  : "channel flow";
  // current velocity with BC ghosts (idempotent on a state that already carries them)
  launch_velocity_bc(c, g, bcs, vel[0], vel[1], vel[2]);
  // predictor RHS; A overwrites A_prev in place after it has been read
  launch_predictor_rhs(c, g, cfg.flow_case == OF_CASE_TAYLOR_GREEN ? 0 : 1, cfg.scheme, nu, dt, vel[0], vel[1], vel[2],
                       p_total, has_prev ? adv[0] : nullptr, has_prev ? adv[1] : nullptr,
                       has_prev ? adv[2] : nullptr, adv[0], adv[1], adv[2], rhs[0], rhs[1], rhs[2]);
  // body force, added after the reference's terms: rhs += dt f
  if(force[0])
    for(int a = 0; a < 3; ++a) launch_axpy_active(c, make_lay(g, OF_FACE_X + a), dt, force[a], rhs[a]);
  // Helmholtz predictor (exact line solves); output carries BC ghosts
  if(predictor_kind == OF_PREDICTOR_CG)
    predictor_iterations = predictor_cg(*hcg, g, bcs, 0.5 * nu * dt, rhs[0], rhs[1], rhs[2], pred[0], pred[1],
                                        pred[2], predictor_tol, predictor_maxit, nullptr);
  else
    predictor_adi(c, g, bcs, 0.5 * nu * dt, rhs[0], rhs[1], rhs[2], pred[0], pred[1], pred[2], rhs, false);
  // projection (solver/projection.cpp:1097-1131); pred already carries its BC ghosts
  launch_divergence_scaled(c, g, pred[0], pred[1], pred[2], 1.0 / dt, prhs);
  // phi starts from zero (solver/taylor_green.cpp:596): the zero-warm-start solve neither reads
  // it nor needs it cleared; it writes the active range and the BC fill the face ghosts (the
  // edge ghosts keep their zeros from allocation)
  // on a periodic grid the final mean of phi is subtracted where the step end reads phi
  const double* phi_mean = nullptr;
  const PoissonResult pr = poisson_solve(poisson, prhs, phi, cfg.poisson_max_iterations, cfg.poisson_tolerance, false,
                                         true, all_periodic ? &phi_mean : nullptr);
  if(!pr.diag.converged) {  // solver/projection.cpp:1133-1147
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "%s pressure solve did not converge at step %d (iterations=%d, relative_residual=%g, tolerance=%g)",
                  name, metrics.step + 1, pr.diag.iterations, pr.diag.relative_residual, cfg.poisson_tolerance);
    throw RuntimeError(buf);
  }
  launch_pressure_bc(c, g, pbcs, phi);
  if(all_periodic) {  // correction, pressure update and diagnostics in one pass, then the fills
    // columns of at least 16 planes: a quarter of the z re-reads of phi and w at block edges and
    // of the partials the finalize folds (the diagnostics' sums differ by rounding only)
    const ColumnLaunch cl = column_launch(g.nx, g.ny, g.nz, 16);
    {
      ProfScope prof_(c, "k_finish_periodic", true);
      OFH_LAUNCH(k_finish_periodic, cl.grid, cl.block, 0, c.stream, pred[0], pred[1], pred[2], phi, phi_mean, prhs,
                 vel[0],
                 vel[1], vel[2], corr[0], corr[1], corr[2], p_total, make_lay(g, OF_FACE_X), make_lay(g, OF_FACE_Y),
                 make_lay(g, OF_FACE_Z), make_lay(g, OF_CELL), g, -dt / 1.0, -0.5 * nu * dt, dt, c.partials, cl.kz);
    }
    launch_finalize(c.stream, c.partials, cl.grid, EpiStore<8>{diag, kStepDiagMaxMask});
    launch_velocity_bc(c, g, bcs, corr[0], corr[1], corr[2]);
    total_pressure_bc(corr);
  } else {
    launch_correct(c, g, pred[0], pred[1], pred[2], phi, -dt / 1.0, corr[0], corr[1], corr[2]);
    launch_velocity_bc(c, g, bcs, corr[0], corr[1], corr[2]);
    {
      const Lay L = make_lay(g, OF_CELL);
      const ColumnLaunch cl = column_launch(L.ax, L.ay, L.az);
      {
        ProfScope prof_(c, "k_pressure_update", true);
        OFH_LAUNCH(k_pressure_update, cl.grid, cl.block, 0, c.stream, p_total, phi, prhs, L, -0.5 * nu * dt, cl.kz);
      }
    }
    total_pressure_bc(corr);
    const ColumnLaunch cl = column_launch(g.nx, g.ny, g.nz);
    {
      ProfScope prof_(c, "k_step_diag", true);
      OFH_LAUNCH(k_step_diag, cl.grid, cl.block, 0, c.stream, corr[0], corr[1], corr[2], vel[0], vel[1], vel[2],
                 make_lay(g, OF_FACE_X), make_lay(g, OF_FACE_Y), make_lay(g, OF_FACE_Z), g, dt, c.partials, cl.kz);
    }
    launch_finalize(c.stream, c.partials, cl.grid, EpiStore<8>{diag, kStepDiagMaxMask});
  }
  OFH_CHECK(hipGetLastError());
  double h[8];
  OFH_CHECK(hipMemcpyAsync(h, diag, 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  OFH_CHECK(hipStreamSynchronize(c.stream));
  const double ncell = (double)g.nx * (double)g.ny * (double)g.nz;
  const double div_l2 = std::sqrt(h[0] / ncell);
  if(h[7] != 0.0 || !std::isfinite(h[6]) || !std::isfinite(div_l2))
    throw RuntimeError(std::string(name) + " solve produced a non-finite state");
  for(int a = 0; a < 3; ++a) std::swap(vel[a], corr[a]);
  has_prev = 1;
  of_step_metrics n{};
  n.step = metrics.step + 1;
  n.time = metrics.time + dt;
  n.dt = dt;
  n.max_cfl = h[2];
  n.max_velocity_change = h[6];
  n.divergence_l2 = div_l2;
  n.max_divergence_l2 = std::max(metrics.max_divergence_l2, div_l2);
  n.pressure_iterations = pr.diag.iterations;
  n.pressure_relative_residual = pr.diag.relative_residual;
  n.divergence_linf = h[1];
  n.max_divergence_linf = std::max(metrics.max_divergence_linf, h[1]);
  metrics = n;
}

double Stepper::kinetic_energy() {
  const ColumnLaunch cl = column_launch(g.nx, g.ny, g.nz);
  OFH_LAUNCH(k_kinetic, cl.grid, cl.block, 0, c.stream, vel[0], vel[1], vel[2], make_lay(g, OF_FACE_X),
                     make_lay(g, OF_FACE_Y), make_lay(g, OF_FACE_Z), g, c.partials, cl.kz);
  launch_finalize(c.stream, c.partials, cl.grid, EpiStore<1>{diag + 16, 0u});
  double e = 0.0;
  OFH_CHECK(hipMemcpyAsync(&e, diag + 16, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  OFH_CHECK(hipStreamSynchronize(c.stream));
  return e / ((double)g.nx * (double)g.ny * (double)g.nz);
}

}  // namespace ofh
