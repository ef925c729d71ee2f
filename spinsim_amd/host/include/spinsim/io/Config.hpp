#pragma once
// A complete description of one experiment (include/spinsim/io/Config.hpp:32-58
// of the reference) and its TOML configuration files (loadConfig /
// parseConfig, :60-66; the reader is this build's own, src/config.cpp).

#include <cstdint>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "spinsim/ensemble/Ensemble.hpp"
#include "spinsim/sequence/Sequence.hpp"

namespace spinsim {

class ConfigError : public std::runtime_error {
 public:
  ConfigError(std::string_view file, std::string_view key, std::string_view problem);
};

struct RunConfig {
  std::size_t nspins = 0;
  std::vector<Field> baseFields;
  InitialStateKind initialState = InitialStateKind::Random;
  std::size_t basisIndex = 0;

  std::size_t realizations = 1;
  std::uint64_t seed = 0;
  bool generateGeometry = true;
  double dipolarScale = 1.0;
  double fieldDisorder = 0.0;
  /// Geometry options with no counterpart in the reference's RunConfig
  /// (runRealization always uses the generator's defaults there).
  double minSeparation = 5.0e-2;
  bool secularAngle = false;
  /// Fixed couplings applied to every realisation, with no counterpart in
  /// the reference's RunConfig (there, generateGeometry = false leaves the
  /// system uncoupled). Only allowed with generateGeometry = false; they use
  /// no random draws, so the realisation streams are unchanged.
  std::vector<Coupling> couplings;

  Axis referenceAxis = Axis::Z;
  std::vector<std::string> observables;

  double epsilon = 1.0e-7;

  Sequence sequence{{}};

  void validate(std::string_view origin = "<config>") const;
};

/// Reads a TOML configuration file.
[[nodiscard]] RunConfig loadConfig(const std::filesystem::path& path);

/// Parses TOML held in memory. `origin` only appears in error messages.
[[nodiscard]] RunConfig parseConfig(std::string_view text, std::string_view origin);

}  // namespace spinsim
