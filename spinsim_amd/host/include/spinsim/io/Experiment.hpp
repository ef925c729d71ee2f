#pragma once
// Per-realisation recipe and ensemble entry points
// (include/spinsim/io/Experiment.hpp:11-31 of the reference), plus the
// batched device runner that replaces the per-realisation job when every
// realisation fits in LDS.

#include <filesystem>
#include <iosfwd>
#include <vector>

#include "spinsim/ensemble/EnsembleRunner.hpp"
#include "spinsim/io/Config.hpp"

namespace spinsim {

void runRealization(const RunConfig& config, std::size_t index, VectorRecorder& out);

[[nodiscard]] EnsembleResult runExperiment(const RunConfig& config, EnsembleRunner& runner);

/// All realisations in one persistent device launch (libspx batch runner);
/// host threads draw the realisation inputs, the fold is in index order.
[[nodiscard]] EnsembleResult runExperimentBatched(const RunConfig& config,
                                                  std::size_t hostThreads = 0);

/// CSV: time, then mean/sd/stderr per column, %.10g / %.12g
/// (src/io/Experiment.cpp:74-98 of the reference).
void writeCsv(const EnsembleResult& result, std::ostream& out);
void writeCsv(const EnsembleResult& result, const std::filesystem::path& path);

}  // namespace spinsim
