// Host-only checks of the C++ interface (no device work): layouts, descriptors, the .cfg
// surface and its error messages, the cavity envelope check and the channel profile check on
// synthetic profiles.
// Prints one line per check and exits non-zero on the first failure.
#include "core/fields.hpp"
#include "solver/channel_flow.hpp"
#include "solver/lid_driven_cavity.hpp"
#include "solver/taylor_green.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  std::cout << (ok ? "ok   " : "FAIL ") << what << '\n';
  if(!ok) ++failures;
}

template <typename E>
void expect_throw(const std::function<void()>& fn, const std::string& fragment, const std::string& what) {
  try {
    fn();
  } catch(const E& e) {
    expect(std::string(e.what()).find(fragment) != std::string::npos, what + " (" + e.what() + ")");
    return;
  } catch(const std::exception& e) {
    expect(false, what + " threw the wrong type: " + e.what());
    return;
  }
  expect(false, what + " did not throw");
}

std::string write_cfg(const std::string& name, const std::string& text) {
  const std::string path = std::string("/tmp/") + name;
  std::ofstream(path) << text;
  return path;
}

}  // namespace

int main() {
  using namespace solver;
  // layouts (core/field_layout.hpp:69-290)
  const Grid g{8, 6, 4, 0.5, 0.25, 0.125, 1};
  const FieldLayout c = FieldLayout::cell_centered(g);
  const FieldLayout u = FieldLayout::face_centered(Axis::x, g);
  const FieldLayout w = FieldLayout::face_centered(Axis::z, g);
  expect(c.storage_size() == 10u * 8u * 6u, "cell storage (nx+2)(ny+2)(nz+2)");
  expect(u.active_extent().nx == 9 && u.storage_extent().nx == 11, "x-face field has nx+1 active faces");
  expect(w.active_extent().nz == 5 && w.storage_extent().nz == 7, "z-face field has nz+1 active faces");
  expect(u.unchecked_index(2, 3, 4) == 2u + 11u * (3u + 8u * 4u), "i-fastest flat index");
  expect(u.coordinate_for_storage_index(Axis::x, 1) == 0.0, "face coordinate of the first active face");
  expect(c.coordinate_for_storage_index(Axis::x, 1) == 0.25, "cell-centre coordinate");
  const IndexRange3D gr = c.ghost_range(BoundaryFace::y_max);
  expect(gr.j_begin == 7 && gr.j_end == 8 && gr.i_begin == 1 && gr.i_end == 9, "ghost slab spans the active range");
  expect(!c.same_shape_as(u), "cell and face layouts differ");
  expect_throw<std::invalid_argument>([] { Grid bad{0, 1, 1, 1.0, 1.0, 1.0}; }, "grid dimensions must be positive",
                                      "grid validation");
  expect_throw<std::out_of_range>([&] { (void)c.index(-1, 0, 0); }, "storage index outside field extent",
                                  "checked index");
  StructuredField f{u};
  expect(f.is_aligned() && f.size() == u.storage_size(), "field storage is aligned and zero-sized to the layout");
  // descriptors
  const BoundaryConditionSet cav = BoundaryConditionSet::cavity();
  expect(cav[BoundaryFace::z_max].type == PhysicalBoundaryType::no_slip_wall, "cavity() is all no-slip");
  expect(to_string(PhysicalBoundaryType::fixed_pressure) == "fixed_pressure", "boundary type names");
  // Taylor-Green .cfg surface (solver/taylor_green.cpp:425-495)
  const std::string tg = write_cfg("tg_ok.cfg",
                                   "# comment\nnx = 16\nny = 12 # trailing\nnz = 8\nviscosity = 0.02\n"
                                   "final_time = 0.3\nbackend = hip\nprecision = mixed\nvalidate_energy = off\n");
  const TaylorGreenConfig tc = load_taylor_green_config(tg);
  expect(tc.nx == 16 && tc.ny == 12 && tc.nz == 8 && tc.viscosity == 0.02 && !tc.validate_energy &&
             tc.precision == SolvePrecision::mixed,
         "Taylor-Green cfg keys");
  expect(describe(tc).find("nx=16, ny=12, nz=8, viscosity=0.02") == 0, "describe() leads with the grid");
  expect(describe(tc).find("backend=hip, advection=scheme=tvd, limiter=van_leer, precision=mixed") != std::string::npos,
         "describe() names backend, advection and precision");
  const Grid tgg = make_taylor_green_grid(tc);
  expect(std::abs(tgg.dx - 2.0 * std::acos(-1.0) / 16.0) < 1e-15 && tgg.dz == 2.0 * std::acos(-1.0) / 8.0,
         "Taylor-Green grid spans [0, 2pi]^3");
  TaylorGreenConfig two_d = tc;
  two_d.nz = 1;
  expect(make_taylor_green_grid(two_d).dz == 1.0 &&
             make_taylor_green_boundary_conditions(two_d)[BoundaryFace::z_min].type == PhysicalBoundaryType::symmetry,
         "2D Taylor-Green: dz = 1, symmetry z faces");
  expect_throw<std::runtime_error>([] { (void)load_taylor_green_config(write_cfg("tg_bad.cfg", "mesh = 3\n")); },
                                   "unsupported Taylor-Green config key: mesh", "unknown key");
  expect_throw<std::runtime_error>([] { (void)load_taylor_green_config(write_cfg("tg_bad2.cfg", "nx 3\n")); },
                                   "invalid config line 1: nx 3", "line without '='");
  expect_throw<std::runtime_error>(
      [] { (void)load_taylor_green_config(write_cfg("tg_bad3.cfg", "validate_energy = maybe\n")); },
      "invalid boolean value: maybe", "boolean values");
  expect_throw<std::invalid_argument>([] { (void)load_taylor_green_config(write_cfg("tg_bad4.cfg", "viscosity = 0\n")); },
                                      "Taylor-Green config expects viscosity > 0", "config validation");
  expect_throw<std::runtime_error>([] { (void)load_taylor_green_config("/nonexistent/tg.cfg"); },
                                   "unable to open Taylor-Green config: /nonexistent/tg.cfg", "missing file");
  expect(parse_execution_backend("cpu") == ExecutionBackend::cpu, "cpu backend value still parses");
  TaylorGreenConfig cpu_run = tc;
  cpu_run.backend = ExecutionBackend::cpu;
  expect_throw<std::runtime_error>([&] { (void)run_taylor_green(cpu_run); }, "backend cpu is not available",
                                   "only the hip backend executes");
  // cavity .cfg surface and the Re = 100 envelope (solver/lid_driven_cavity.cpp:278-336,411-485)
  const LidDrivenCavityConfig cc = load_lid_driven_cavity_config(
      write_cfg("cav.cfg", "nx = 32\nny = 32\nreynolds = 100\nmin_steps = 4\nvalidate_reference = false\n"));
  expect(cc.nx == 32 && cc.min_steps == 4 && !cc.validate_reference, "cavity cfg keys");
  expect(describe(cc).find("nx=32, ny=32, Re=100, lid_velocity=1") == 0, "cavity describe()");
  const BoundaryConditionSet cb = make_lid_driven_cavity_boundary_conditions(cc);
  expect(cb[BoundaryFace::y_max].type == PhysicalBoundaryType::prescribed_velocity && cb[BoundaryFace::y_max].velocity[0] == 1.0,
         "lid on y_max");
  LidDrivenCavityResult res;
  for(int i = 0; i <= 64; ++i) {
    const double t = i / 64.0;
    res.u_vertical_centerline.coordinate.push_back(t);
    res.u_vertical_centerline.value.push_back(t * t * 0.84123);  // passes through the first point at t ~ 1
    res.v_horizontal_centerline.coordinate.push_back(t);
    res.v_horizontal_centerline.value.push_back(0.2);
  }
  const LidDrivenCavityValidation v = validate_lid_driven_cavity_re100(res);
  expect(v.reference_dataset == "Re100_centerline_literature_envelope", "envelope dataset name");
  const double s0 = v.points[0].sample_value, want = 0.9766 * 0.9766 * 0.84123;
  expect(std::abs(s0 - want) < 1e-12, "cubic interpolation is exact on a quadratic profile");
  expect(std::abs(v.points[2].relative_error - (0.2 - 0.17960) / 0.17960) < 1e-15, "miss against the envelope");
  expect(!v.pass, "a profile outside the envelope fails");
  // channel .cfg surface, boundary sets and profile check (solver/channel_flow.cpp:303-430)
  const ChannelFlowConfig ch = load_channel_flow_config(write_cfg(
      "chan.cfg", "case = couette\nnx = 32\nny = 32\nsteps = 4\nvalidate_profile = true\n"));
  expect(ch.case_kind == ChannelFlowCase::couette && ch.nx == 32 && ch.ny == 32 && ch.steps == 4 &&
             ch.validate_profile,
         "channel cfg keys");
  expect(describe(ch).find("case=couette, nx=32, ny=32, viscosity=0.1, top_velocity=1") == 0, "channel describe()");
  const BoundaryConditionSet cou = make_channel_flow_boundary_conditions(ch);
  expect(cou[BoundaryFace::x_min].type == PhysicalBoundaryType::periodic &&
             cou[BoundaryFace::x_max].type == PhysicalBoundaryType::periodic &&
             cou[BoundaryFace::y_min].type == PhysicalBoundaryType::no_slip_wall &&
             cou[BoundaryFace::y_max].type == PhysicalBoundaryType::prescribed_velocity &&
             cou[BoundaryFace::y_max].velocity[0] == ch.top_velocity &&
             cou[BoundaryFace::z_min].type == PhysicalBoundaryType::symmetry,
         "Couette boundaries: periodic x, moving upper wall, symmetry z");
  ChannelFlowConfig pz = ch;
  pz.case_kind = ChannelFlowCase::poiseuille;
  const BoundaryConditionSet poi = make_channel_flow_boundary_conditions(pz);
  expect(poi[BoundaryFace::x_min].type == PhysicalBoundaryType::fixed_pressure &&
             poi[BoundaryFace::x_min].pressure == pz.pressure_drop &&
             poi[BoundaryFace::x_max].type == PhysicalBoundaryType::fixed_pressure &&
             poi[BoundaryFace::x_max].pressure == 0.0 &&
             poi[BoundaryFace::y_max].type == PhysicalBoundaryType::no_slip_wall,
         "Poiseuille boundaries: fixed pressure drop along x, no-slip walls");
  expect(to_string(ChannelFlowCase::poiseuille) == "poiseuille", "channel case names");
  expect_throw<std::runtime_error>(
      [] { (void)load_channel_flow_config(write_cfg("chan_bad.cfg", "case = duct\n")); },
      "unsupported channel flow case: duct", "unknown channel case");
  expect_throw<std::runtime_error>(
      [] { (void)load_channel_flow_config(write_cfg("chan_bad2.cfg", "reynolds = 3\n")); },
      "unsupported channel flow config key: reynolds", "unknown channel key");
  expect_throw<std::invalid_argument>(
      [] { (void)load_channel_flow_config(write_cfg("chan_bad3.cfg", "case = poiseuille\npressure_drop = 0\n")); },
      "Poiseuille config expects pressure_drop > 0", "Poiseuille validation");
  expect_throw<std::invalid_argument>(
      [] { (void)load_channel_flow_config(write_cfg("chan_bad4.cfg", "steps = 0\n")); },
      "channel flow config expects steps > 0", "channel step validation");
  ChannelFlowResult cr;
  cr.config = pz;
  for(int j = 0; j < 32; ++j) {
    const double y = (j + 0.5) / 32.0;
    cr.streamwise_profile.coordinate.push_back(y);
    cr.streamwise_profile.value.push_back(0.5 * pz.pressure_drop * y * (1.0 - y) / pz.viscosity);
  }
  const ChannelFlowValidation cv = validate_channel_flow_profile(cr);
  expect(cv.reference_dataset == "analytic_poiseuille_profile" && cv.relative_l2_error == 0.0 && cv.pass,
         "the analytic Poiseuille profile passes with zero error");
  for(double& val : cr.streamwise_profile.value) val *= 1.01;
  const ChannelFlowValidation cv2 = validate_channel_flow_profile(cr);
  expect(std::abs(cv2.relative_l2_error - 0.01) < 1e-12 && !cv2.pass, "a 1% profile error fails the 0.5% check");
  std::cout << (failures == 0 ? "all host checks passed" : "host checks failed") << '\n';
  return failures == 0 ? 0 : 1;
}
