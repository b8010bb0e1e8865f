// io_check: host-only driver of the checkpoint and VTK writers for the tests (no device work).
//
//   io_check write-cavity <cfg> <fields.f64> <has_prev> <step> <time> <dt> <max_cfl>
//            <max_velocity_change> <divergence_l2> <pressure_iterations> <relative_residual> <out>
//   io_check load-cavity <cfg> <checkpoint> <fields_out.f64>
//   io_check write-tg <cfg> <fields.f64> <has_prev> <step> <time> <dt> <max_cfl>
//            <max_velocity_change> <divergence_l2> <pressure_iterations> <relative_residual>
//            <max_divergence_l2> <out>
//   io_check load-tg <cfg> <checkpoint> <fields_out.f64>
//   io_check write-vtk <nx> <ny> <nz> <dx> <dy> <dz> <fields.f64> <out.vtk>
//
// Field files hold whole ghosted storages, concatenated: u, v, w, p_total, then (state
// commands) adv_u, adv_v, adv_w. load-* print metadata and metrics as "key: value" lines
// (17 significant digits); errors go to stderr with exit status 1.
#include "io/checkpoint.hpp"
#include "io/vtk_export.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

std::vector<double> read_all(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if(!in) throw std::runtime_error("cannot open " + path);
  in.seekg(0, std::ios::end);
  const std::size_t n = (std::size_t)in.tellg() / sizeof(double);
  in.seekg(0);
  std::vector<double> v(n);
  in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
  return v;
}

struct Fields {
  std::vector<double> data;
  std::size_t at = 0;
  void into(solver::StructuredField& f) {
    if(at + f.size() > data.size()) throw std::runtime_error("field file too short");
    std::memcpy(f.data(), data.data() + at, f.size() * sizeof(double));
    at += f.size();
  }
};

void put(std::ofstream& out, const solver::StructuredField& f) {
  out.write(reinterpret_cast<const char*>(f.data()), (std::streamsize)(f.size() * sizeof(double)));
}

template <typename State>
void fill_state(State& s, const std::string& path) {
  Fields f{read_all(path)};
  f.into(s.velocity.x);
  f.into(s.velocity.y);
  f.into(s.velocity.z);
  f.into(s.pressure_total);
  f.into(s.advection_previous.x);
  f.into(s.advection_previous.y);
  f.into(s.advection_previous.z);
}

template <typename State>
void dump_state(const State& s, const std::string& path) {
  std::ofstream out(path, std::ios::binary);
  put(out, s.velocity.x);
  put(out, s.velocity.y);
  put(out, s.velocity.z);
  put(out, s.pressure_total);
  put(out, s.advection_previous.x);
  put(out, s.advection_previous.y);
  put(out, s.advection_previous.z);
}

template <typename M>
void set_metrics(M& m, char** a) {
  m.step = std::stoi(a[0]);
  m.time = std::stod(a[1]);
  m.dt = std::stod(a[2]);
  m.max_cfl = std::stod(a[3]);
  m.max_velocity_change = std::stod(a[4]);
  m.divergence_l2 = std::stod(a[5]);
  m.pressure_iterations = std::stoi(a[6]);
  m.pressure_relative_residual = std::stod(a[7]);
}

template <typename M>
void print_metrics(const M& m) {
  std::printf("step: %d\ntime: %.17g\ndt: %.17g\nmax_cfl: %.17g\nmax_velocity_change: %.17g\n"
              "divergence_l2: %.17g\npressure_iterations: %d\npressure_relative_residual: %.17g\n",
              m.step, m.time, m.dt, m.max_cfl, m.max_velocity_change, m.divergence_l2, m.pressure_iterations,
              m.pressure_relative_residual);
}

void print_metadata(const solver::io::CheckpointMetadata& md, bool has_prev) {
  std::printf("format_version: %u\nendianness: %s\nscalar_bytes: %u\nchecksum: %llu\nbuild_hash: %llu\n"
              "configuration_hash: %llu\nhas_previous_advection: %d\n",
              md.format_version, md.endianness.c_str(), md.scalar_bytes, (unsigned long long)md.checksum,
              (unsigned long long)md.build_hash, (unsigned long long)md.configuration_hash, has_prev ? 1 : 0);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if(argc < 2) throw std::runtime_error("usage: io_check <command> ...");
    const std::string cmd = argv[1];
    if(cmd == "write-cavity" && argc == 14) {
      const solver::LidDrivenCavityConfig c = solver::load_lid_driven_cavity_config(argv[2]);
      solver::LidDrivenCavityState s{solver::make_lid_driven_cavity_grid(c)};
      fill_state(s, argv[3]);
      s.has_previous_advection = std::stoi(argv[4]) != 0;
      set_metrics(s.metrics, argv + 5);
      solver::io::write_lid_driven_cavity_checkpoint(argv[13], c, s);
    } else if(cmd == "load-cavity" && argc == 5) {
      const solver::LidDrivenCavityConfig c = solver::load_lid_driven_cavity_config(argv[2]);
      const solver::io::LidDrivenCavityCheckpoint ck = solver::io::load_lid_driven_cavity_checkpoint(argv[3], c);
      print_metadata(ck.metadata, ck.state.has_previous_advection);
      print_metrics(ck.state.metrics);
      dump_state(ck.state, argv[4]);
    } else if(cmd == "write-tg" && argc == 15) {
      const solver::TaylorGreenConfig c = solver::load_taylor_green_config(argv[2]);
      solver::TaylorGreenState s{solver::make_taylor_green_grid(c)};
      fill_state(s, argv[3]);
      s.has_previous_advection = std::stoi(argv[4]) != 0;
      set_metrics(s.metrics, argv + 5);
      s.metrics.max_divergence_l2 = std::stod(argv[13]);
      solver::io::write_taylor_green_checkpoint(argv[14], c, s);
    } else if(cmd == "load-tg" && argc == 5) {
      const solver::TaylorGreenConfig c = solver::load_taylor_green_config(argv[2]);
      const solver::io::TaylorGreenCheckpoint ck = solver::io::load_taylor_green_checkpoint(argv[3], c);
      print_metadata(ck.metadata, ck.state.has_previous_advection);
      print_metrics(ck.state.metrics);
      std::printf("max_divergence_l2: %.17g\n", ck.state.metrics.max_divergence_l2);
      dump_state(ck.state, argv[4]);
    } else if(cmd == "write-vtk" && argc == 10) {
      const solver::Grid g{std::stoi(argv[2]), std::stoi(argv[3]), std::stoi(argv[4]), std::stod(argv[5]),
                           std::stod(argv[6]), std::stod(argv[7]), 1};
      solver::VelocityField v{g};
      solver::PressureField p{g};
      Fields f{read_all(argv[8])};
      f.into(v.x);
      f.into(v.y);
      f.into(v.z);
      f.into(p);
      solver::io::write_mac_fields_vtk(argv[9], v, p);
    } else {
      throw std::runtime_error("unknown command or wrong argument count: " + cmd);
    }
    return 0;
  } catch(const std::exception& e) {
    std::cerr << "io_check failed: " << e.what() << '\n';
    return 1;
  }
}
