// SLVCHKP1 checkpoints (reference io/checkpoint.hpp:1-35): the lid-driven cavity state with the
// reference's byte layout, and the same container for the Taylor-Green state (this build's
// generalisation: its CONF section holds the Taylor-Green config, its META section appends
// max_divergence_l2). Files are written and verified in a streaming pass, so a state never needs
// a second in-memory copy of its serialised form.
#pragma once

#include "solver/lid_driven_cavity.hpp"
#include "solver/taylor_green.hpp"

#include <cstdint>
#include <string>

namespace solver::io {

struct CheckpointMetadata {
  std::uint32_t format_version = 0;
  std::string endianness = "little";
  std::uint32_t scalar_bytes = 0;
  std::uint64_t checksum = 0;
  std::uint64_t build_hash = 0;
  std::uint64_t configuration_hash = 0;
};

struct LidDrivenCavityCheckpoint {
  CheckpointMetadata metadata{};
  LidDrivenCavityState state;

  explicit LidDrivenCavityCheckpoint(const Grid& grid) : state(grid) {}
};

void write_lid_driven_cavity_checkpoint(const std::string& path, const LidDrivenCavityConfig& config,
                                        const LidDrivenCavityState& state);

[[nodiscard]] LidDrivenCavityCheckpoint load_lid_driven_cavity_checkpoint(const std::string& path,
                                                                          const LidDrivenCavityConfig& expected_config);

struct TaylorGreenCheckpoint {
  CheckpointMetadata metadata{};
  TaylorGreenState state;

  explicit TaylorGreenCheckpoint(const Grid& grid) : state(grid) {}
};

void write_taylor_green_checkpoint(const std::string& path, const TaylorGreenConfig& config,
                                   const TaylorGreenState& state);

[[nodiscard]] TaylorGreenCheckpoint load_taylor_green_checkpoint(const std::string& path,
                                                                 const TaylorGreenConfig& expected_config);

// FNV-1a (64-bit) of a byte range: the payload checksum and the build/config hashes.
[[nodiscard]] std::uint64_t fnv1a_64(const void* data, std::size_t size,
                                     std::uint64_t seed = 14695981039346656037ull);

}  // namespace solver::io
