// SLVCHKP1 checkpoint container (reference io/checkpoint.cpp:20-461), written and read in
// streaming passes.
//
// File: "SLVCHKP1", u32 format version (1), u64 payload bytes, u64 FNV-1a of the payload, then
// the payload: sections of {4-byte tag, u64 length, body}. All integers and doubles are
// little-endian. Bodies: META (scalar size, build hash, configuration hash, AB2 flag, step
// metrics), CONF (the run configuration), GRID, and one section per stored field
// (VELX/VELY/VELZ, ADVX/ADVY/ADVZ, PRST) holding the three storage extents and the whole
// ghosted storage in index order (i fastest).
//
// The writer streams sections to the file while hashing them and patches the length and
// checksum into the header at the end; the reader verifies the checksum in one streamed pass
// and decodes the sections in a second, reading field bodies straight into the state. Neither
// holds a serialised copy of the state, which keeps checkpoints of large device states (a 512^3
// Taylor-Green state is 7.5 GB) within host memory.
#include "io/checkpoint.hpp"

#include "of_hip.h"

#include <array>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "checkpoints store little-endian values in host order");

namespace solver::io {

std::uint64_t fnv1a_64(const void* data, const std::size_t size, std::uint64_t hash) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for(std::size_t n = 0; n < size; ++n) {
    hash ^= p[n];
    hash *= 1099511628211ull;
  }
  return hash;
}

namespace {

using Tag = std::array<char, 4>;
constexpr char kMagic[8] = {'S', 'L', 'V', 'C', 'H', 'K', 'P', '1'};
constexpr std::uint32_t kFormatVersion = 1;
constexpr std::size_t kHeaderBytes = sizeof kMagic + 4 + 8 + 8;
constexpr Tag kMeta{'M', 'E', 'T', 'A'}, kConf{'C', 'O', 'N', 'F'}, kGrid{'G', 'R', 'I', 'D'};
constexpr Tag kFieldTags[7] = {{'V', 'E', 'L', 'X'}, {'V', 'E', 'L', 'Y'}, {'V', 'E', 'L', 'Z'}, {'A', 'D', 'V', 'X'},
                               {'A', 'D', 'V', 'Y'}, {'A', 'D', 'V', 'Z'}, {'P', 'R', 'S', 'T'}};

std::uint64_t hash_text(const std::string& s) { return fnv1a_64(s.data(), s.size()); }

// Identity of the executing build: library build string, host compiler and numerics profile.
std::uint64_t build_hash() {
  const std::string d = std::string("solver_amd|") + of_build_info() + "|GNU " + __VERSION__ + "|deterministic|strict";
  return hash_text(d);
}

bool same_grid(const Grid& a, const Grid& b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx && a.dy == b.dy && a.dz == b.dz &&
         a.ghost_layers == b.ghost_layers;
}

// Little-endian section body under construction.
struct Body {
  std::vector<unsigned char> b;
  void raw(const void* p, std::size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    b.insert(b.end(), c, c + n);
  }
  void u32(std::uint32_t v) { raw(&v, 4); }
  void u64(std::uint64_t v) { raw(&v, 8); }
  void f64(double v) { raw(&v, 8); }
};

// Cursor over a decoded section body; running past the end raises the reference's messages.
struct Cursor {
  const std::vector<unsigned char>& b;
  std::size_t at = 0;
  void take(void* out, std::size_t n, const char* what) {
    if(at + n > b.size()) throw std::runtime_error(std::string("checkpoint truncated while reading ") + what);
    std::memcpy(out, b.data() + at, n);
    at += n;
  }
  std::uint32_t u32() {
    std::uint32_t v;
    take(&v, 4, "u32");
    return v;
  }
  std::uint64_t u64() {
    std::uint64_t v;
    take(&v, 8, "u64");
    return v;
  }
  double f64() {
    double v;
    take(&v, 8, "u64");
    return v;
  }
  void end(const char* section) const {
    if(at != b.size()) throw std::runtime_error(std::string("checkpoint ") + section + " section contains trailing bytes");
  }
};

// ------------------------------------------------------------------ per-case encodings
// Each case provides: the grid its config implies, the CONF body and its check, and the META
// body and its decoding (metrics differ between the cases).
struct CavityCase {
  using Config = LidDrivenCavityConfig;
  using State = LidDrivenCavityState;
  static Grid grid(const Config& c) { return make_lid_driven_cavity_grid(c); }
  static std::uint64_t config_hash(const Config& c) { return hash_text(describe(c)); }
  static Body conf(const Config& c) {  // reference io/checkpoint.cpp:122-138
    Body o;
    o.u32((std::uint32_t)c.nx);
    o.u32((std::uint32_t)c.ny);
    o.f64(c.reynolds);
    o.f64(c.lid_velocity);
    o.f64(c.cfl_limit);
    o.u32((std::uint32_t)c.max_steps);
    o.u32((std::uint32_t)c.min_steps);
    o.f64(c.steady_tolerance);
    o.u32((std::uint32_t)c.poisson_max_iterations);
    o.f64(c.poisson_tolerance);
    o.u32(c.validate_reference ? 1u : 0u);
    o.u32((std::uint32_t)c.advection.scheme);
    o.u32((std::uint32_t)c.advection.limiter);
    return o;
  }
  static bool conf_matches(Cursor& in, const Config& c) {
    const int nx = (int)in.u32(), ny = (int)in.u32();
    const double re = in.f64(), lid = in.f64(), cfl = in.f64();
    const int max_steps = (int)in.u32(), min_steps = (int)in.u32();
    const double steady = in.f64();
    const int maxit = (int)in.u32();
    const double tol = in.f64();
    const bool validate = in.u32() != 0u;
    const int scheme = (int)in.u32(), limiter = (int)in.u32();
    in.end("config");
    return nx == c.nx && ny == c.ny && re == c.reynolds && lid == c.lid_velocity && cfl == c.cfl_limit &&
           max_steps == c.max_steps && min_steps == c.min_steps && steady == c.steady_tolerance &&
           maxit == c.poisson_max_iterations && tol == c.poisson_tolerance && validate == c.validate_reference &&
           scheme == (int)c.advection.scheme && limiter == (int)c.advection.limiter;
  }
  static void metrics_out(Body& o, const State& s) {
    const SimulationStepMetrics& m = s.metrics;
    o.u32((std::uint32_t)m.step);
    o.f64(m.time);
    o.f64(m.dt);
    o.f64(m.max_cfl);
    o.f64(m.max_velocity_change);
    o.f64(m.divergence_l2);
    o.u32((std::uint32_t)m.pressure_iterations);
    o.f64(m.pressure_relative_residual);
  }
  static void metrics_in(Cursor& in, State& s) {
    SimulationStepMetrics& m = s.metrics;
    m.step = (int)in.u32();
    m.time = in.f64();
    m.dt = in.f64();
    m.max_cfl = in.f64();
    m.max_velocity_change = in.f64();
    m.divergence_l2 = in.f64();
    m.pressure_iterations = (int)in.u32();
    m.pressure_relative_residual = in.f64();
  }
  static const char* grid_error() { return "checkpoint state grid does not match the cavity config"; }
};

struct TaylorGreenCase {
  using Config = TaylorGreenConfig;
  using State = TaylorGreenState;
  static Grid grid(const Config& c) { return make_taylor_green_grid(c); }
  static std::uint64_t config_hash(const Config& c) { return hash_text(describe(c)); }
  static Body conf(const Config& c) {
    Body o;
    o.u32((std::uint32_t)c.nx);
    o.u32((std::uint32_t)c.ny);
    o.u32((std::uint32_t)c.nz);
    o.f64(c.viscosity);
    o.f64(c.cfl_limit);
    o.f64(c.final_time);
    o.u32((std::uint32_t)c.poisson_max_iterations);
    o.f64(c.poisson_tolerance);
    o.u32(c.validate_energy ? 1u : 0u);
    o.u32((std::uint32_t)c.backend);
    o.u32((std::uint32_t)c.advection.scheme);
    o.u32((std::uint32_t)c.advection.limiter);
    o.u32((std::uint32_t)c.precision);
    return o;
  }
  static bool conf_matches(Cursor& in, const Config& c) {
    const int nx = (int)in.u32(), ny = (int)in.u32(), nz = (int)in.u32();
    const double nu = in.f64(), cfl = in.f64(), final_time = in.f64();
    const int maxit = (int)in.u32();
    const double tol = in.f64();
    const bool validate = in.u32() != 0u;
    const int backend = (int)in.u32(), scheme = (int)in.u32(), limiter = (int)in.u32(), precision = (int)in.u32();
    in.end("config");
    return nx == c.nx && ny == c.ny && nz == c.nz && nu == c.viscosity && cfl == c.cfl_limit &&
           final_time == c.final_time && maxit == c.poisson_max_iterations && tol == c.poisson_tolerance &&
           validate == c.validate_energy && backend == (int)c.backend && scheme == (int)c.advection.scheme &&
           limiter == (int)c.advection.limiter && precision == (int)c.precision;
  }
  static void metrics_out(Body& o, const State& s) {
    const TaylorGreenStepMetrics& m = s.metrics;
    o.u32((std::uint32_t)m.step);
    o.f64(m.time);
    o.f64(m.dt);
    o.f64(m.max_cfl);
    o.f64(m.max_velocity_change);
    o.f64(m.divergence_l2);
    o.u32((std::uint32_t)m.pressure_iterations);
    o.f64(m.pressure_relative_residual);
    o.f64(m.max_divergence_l2);
  }
  static void metrics_in(Cursor& in, State& s) {
    TaylorGreenStepMetrics& m = s.metrics;
    m.step = (int)in.u32();
    m.time = in.f64();
    m.dt = in.f64();
    m.max_cfl = in.f64();
    m.max_velocity_change = in.f64();
    m.divergence_l2 = in.f64();
    m.pressure_iterations = (int)in.u32();
    m.pressure_relative_residual = in.f64();
    m.max_divergence_l2 = in.f64();
  }
  static const char* grid_error() { return "checkpoint state grid does not match the Taylor-Green config"; }
};

template <typename State>
const StructuredField& field_of(const State& s, int n) {
  switch(n) {
    case 0: return s.velocity.x;
    case 1: return s.velocity.y;
    case 2: return s.velocity.z;
    case 3: return s.advection_previous.x;
    case 4: return s.advection_previous.y;
    case 5: return s.advection_previous.z;
    default: return s.pressure_total;
  }
}

template <typename State>
StructuredField& field_of(State& s, int n) {
  return const_cast<StructuredField&>(field_of(static_cast<const State&>(s), n));
}

// ------------------------------------------------------------------ writing
class StreamWriter {
 public:
  explicit StreamWriter(const std::string& path) : path_(path), out_(path, std::ios::binary | std::ios::trunc) {
    if(!out_.is_open()) throw std::runtime_error("unable to open checkpoint for writing: " + path);
    const unsigned char zeros[kHeaderBytes] = {};
    out_.write(reinterpret_cast<const char*>(zeros), (std::streamsize)kHeaderBytes);  // patched by finish()
  }
  void section(const Tag& tag, const Body& body) {
    head(tag, body.b.size());
    payload(body.b.data(), body.b.size());
  }
  void field(const Tag& tag, const StructuredField& f) {
    const Extent3D e = f.layout().storage_extent();
    const std::uint32_t ext[3] = {(std::uint32_t)e.nx, (std::uint32_t)e.ny, (std::uint32_t)e.nz};
    head(tag, sizeof ext + f.size() * sizeof(double));
    payload(ext, sizeof ext);
    payload(f.data(), f.size() * sizeof(double));
  }
  void finish() {
    out_.seekp(0);
    out_.write(kMagic, sizeof kMagic);
    out_.write(reinterpret_cast<const char*>(&kFormatVersion), 4);
    out_.write(reinterpret_cast<const char*>(&bytes_), 8);
    out_.write(reinterpret_cast<const char*>(&hash_), 8);
    out_.flush();
    if(!out_.good()) throw std::runtime_error("failed while writing checkpoint: " + path_);
  }

 private:
  void head(const Tag& tag, std::uint64_t n) {
    payload(tag.data(), 4);
    payload(&n, 8);
  }
  void payload(const void* p, std::size_t n) {
    out_.write(static_cast<const char*>(p), (std::streamsize)n);
    hash_ = fnv1a_64(p, n, hash_);
    bytes_ += n;
  }
  std::string path_;
  std::ofstream out_;
  std::uint64_t hash_ = 14695981039346656037ull;
  std::uint64_t bytes_ = 0;
};

template <typename Case>
void write_checkpoint(const std::string& path, const typename Case::Config& c, const typename Case::State& s) {
  if(!same_grid(s.grid, Case::grid(c))) throw std::invalid_argument(Case::grid_error());
  Body meta;
  meta.u32((std::uint32_t)sizeof(double));
  meta.u64(build_hash());
  meta.u64(Case::config_hash(c));
  meta.u32(s.has_previous_advection ? 1u : 0u);
  Case::metrics_out(meta, s);
  Body grid;
  grid.u32((std::uint32_t)s.grid.nx);
  grid.u32((std::uint32_t)s.grid.ny);
  grid.u32((std::uint32_t)s.grid.nz);
  grid.f64(s.grid.dx);
  grid.f64(s.grid.dy);
  grid.f64(s.grid.dz);
  grid.u32((std::uint32_t)s.grid.ghost_layers);
  StreamWriter w(path);
  w.section(kMeta, meta);
  w.section(kConf, Case::conf(c));
  w.section(kGrid, grid);
  for(int n = 0; n < 7; ++n) w.field(kFieldTags[n], field_of(s, n));  // the reference's section order
  w.finish();
}

// ------------------------------------------------------------------ reading
class StreamReader {
 public:
  explicit StreamReader(const std::string& path) : path_(path), in_(path, std::ios::binary) {
    if(!in_.is_open()) throw std::runtime_error("unable to open checkpoint: " + path);
    in_.seekg(0, std::ios::end);
    const std::streamoff size = in_.tellg();
    if(size < 0) throw std::runtime_error("unable to read checkpoint size: " + path);
    size_ = (std::uint64_t)size;
    in_.seekg(0);
  }
  // Header checks in the reference's order; returns the format version and the checksum.
  void header(std::uint32_t& version, std::uint64_t& checksum) {
    if(size_ < kHeaderBytes) throw std::runtime_error("checkpoint file is too small: " + path_);
    char magic[sizeof kMagic];
    read(magic, sizeof magic);
    if(std::memcmp(magic, kMagic, sizeof kMagic) != 0) throw std::runtime_error("checkpoint magic mismatch: " + path_);
    read(&version, 4);
    if(version != kFormatVersion)
      throw std::runtime_error("unsupported checkpoint format version: " + std::to_string(version));
    read(&payload_, 8);
    read(&checksum, 8);
    if(kHeaderBytes + payload_ != size_)
      throw std::runtime_error("checkpoint payload length does not match the file size");
    std::uint64_t h = 14695981039346656037ull;
    std::vector<char> chunk(std::size_t(8) << 20);
    for(std::uint64_t left = payload_; left > 0;) {
      const std::size_t n = (std::size_t)std::min<std::uint64_t>(left, chunk.size());
      read(chunk.data(), n);
      h = fnv1a_64(chunk.data(), n, h);
      left -= n;
    }
    if(h != checksum) throw std::runtime_error("checkpoint checksum verification failed");
    in_.seekg((std::streamoff)kHeaderBytes);
    at_ = kHeaderBytes;
  }
  bool more() const { return at_ < size_; }
  // Next section header; the body is then read by body() or field().
  std::uint64_t next(Tag& tag) {
    if(at_ + 12 > size_) throw std::runtime_error("checkpoint section header is truncated");
    read(tag.data(), 4);
    std::uint64_t n;
    read(&n, 8);
    if(n > size_ - at_) throw std::runtime_error("checkpoint section payload is truncated");
    return n;
  }
  std::vector<unsigned char> body(std::uint64_t n) {
    std::vector<unsigned char> b((std::size_t)n);
    read(b.data(), b.size());
    return b;
  }
  void field(std::uint64_t n, StructuredField& f) {
    const Extent3D e = f.layout().storage_extent();
    if(n < 12) {
      skip(n);
      throw std::runtime_error("checkpoint truncated while reading u32");
    }
    std::uint32_t ext[3];
    read(ext, sizeof ext);
    if((int)ext[0] != e.nx || (int)ext[1] != e.ny || (int)ext[2] != e.nz)
      throw std::runtime_error("checkpoint field extent does not match the current layout");
    const std::uint64_t want = f.size() * sizeof(double);
    if(n - 12 < want) throw std::runtime_error("checkpoint truncated while reading u64");
    if(n - 12 > want) throw std::runtime_error("checkpoint field section contains trailing bytes");
    read(f.data(), (std::size_t)want);
  }
  void skip(std::uint64_t n) {
    in_.seekg((std::streamoff)n, std::ios::cur);
    at_ += n;
  }

 private:
  void read(void* p, std::size_t n) {
    if(!in_.read(static_cast<char*>(p), (std::streamsize)n))
      throw std::runtime_error("unable to read checkpoint payload: " + path_);
    at_ += n;
  }
  std::string path_;
  std::ifstream in_;
  std::uint64_t size_ = 0, payload_ = 0, at_ = 0;
};

template <typename Case, typename Checkpoint>
Checkpoint load_checkpoint(const std::string& path, const typename Case::Config& c) {
  const Grid grid = Case::grid(c);
  StreamReader r(path);
  Checkpoint ck{grid};
  r.header(ck.metadata.format_version, ck.metadata.checksum);
  bool seen_meta = false, seen_conf = false, seen_grid = false, seen_field[7] = {};
  while(r.more()) {
    Tag tag;
    const std::uint64_t n = r.next(tag);
    if(tag == kMeta) {
      const std::vector<unsigned char> b = r.body(n);
      Cursor in{b};
      ck.metadata.scalar_bytes = in.u32();
      ck.metadata.build_hash = in.u64();
      ck.metadata.configuration_hash = in.u64();
      ck.state.has_previous_advection = in.u32() != 0u;
      Case::metrics_in(in, ck.state);
      in.end("metadata");
      if(ck.metadata.scalar_bytes != sizeof(double))
        throw std::runtime_error("checkpoint precision does not match the current solver precision");
      if(ck.metadata.build_hash != build_hash())
        throw std::runtime_error("checkpoint build hash does not match the current executable");
      if(ck.metadata.configuration_hash != Case::config_hash(c))
        throw std::runtime_error("checkpoint configuration hash does not match the requested config");
      seen_meta = true;
    } else if(tag == kConf) {
      const std::vector<unsigned char> b = r.body(n);
      Cursor in{b};
      if(!Case::conf_matches(in, c))
        throw std::runtime_error("checkpoint config does not match the requested run configuration");
      seen_conf = true;
    } else if(tag == kGrid) {
      const std::vector<unsigned char> b = r.body(n);
      Cursor in{b};
      Grid g = grid;
      g.nx = (int)in.u32();
      g.ny = (int)in.u32();
      g.nz = (int)in.u32();
      g.dx = in.f64();
      g.dy = in.f64();
      g.dz = in.f64();
      g.ghost_layers = (int)in.u32();
      in.end("grid");
      if(!same_grid(g, grid)) throw std::runtime_error("checkpoint grid does not match the requested run configuration");
      seen_grid = true;
    } else {
      int which = -1;
      for(int f = 0; f < 7; ++f)
        if(tag == kFieldTags[f]) which = f;
      if(which < 0) {
        r.skip(n);  // unknown sections are skipped
        continue;
      }
      r.field(n, field_of(ck.state, which));
      seen_field[which] = true;
    }
  }
  bool all = seen_meta && seen_conf && seen_grid;
  for(const bool s : seen_field) all = all && s;
  if(!all) throw std::runtime_error("checkpoint is missing one or more required sections");
  return ck;
}

}  // namespace

void write_lid_driven_cavity_checkpoint(const std::string& path, const LidDrivenCavityConfig& config,
                                        const LidDrivenCavityState& state) {
  write_checkpoint<CavityCase>(path, config, state);
}

LidDrivenCavityCheckpoint load_lid_driven_cavity_checkpoint(const std::string& path,
                                                            const LidDrivenCavityConfig& expected_config) {
  return load_checkpoint<CavityCase, LidDrivenCavityCheckpoint>(path, expected_config);
}

void write_taylor_green_checkpoint(const std::string& path, const TaylorGreenConfig& config,
                                   const TaylorGreenState& state) {
  write_checkpoint<TaylorGreenCase>(path, config, state);
}

TaylorGreenCheckpoint load_taylor_green_checkpoint(const std::string& path, const TaylorGreenConfig& expected_config) {
  return load_checkpoint<TaylorGreenCase, TaylorGreenCheckpoint>(path, expected_config);
}

}  // namespace solver::io
