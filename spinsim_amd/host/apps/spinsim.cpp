// The spinsim command line over the GPU drop-in (apps/spinsim/main.cpp:17-158
// of the reference): the same commands (run, validate, describe), options and
// outputs, so scripts written for the reference keep working. Two options are
// this build's own: --backend (only "hip" exists here) and --batched, which
// runs the ensemble through the device batch runner instead of one
// realisation per host thread.
#include <charconv>
#include <chrono>
#include <cstdio>
#include <exception>
#include <filesystem>
#include <iostream>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "spinsim/io/Config.hpp"
#include "spinsim/io/Experiment.hpp"

namespace {

using namespace spinsim;

constexpr const char* kUsage =
    "spinsim - Chebyshev spin dynamics on the GPU\n"
    "\n"
    "Usage:\n"
    "  spinsim run <config.toml> [--threads N] [--out FILE] [--realizations N]\n"
    "              [--backend hip] [--batched]\n"
    "  spinsim validate <config.toml>\n"
    "  spinsim describe <config.toml>\n"
    "\n"
    "Options:\n"
    "  --threads N       host threads, each with its own device stream;\n"
    "                    0 or omitted uses every core\n"
    "  --out FILE        write the CSV to FILE instead of standard output\n"
    "  --realizations N  replace ensemble.realizations from the config\n"
    "  --backend NAME    compute backend; this build has only \"hip\"\n"
    "  --batched         step realisations together on the device (random\n"
    "                    initial states only); agrees with the default\n"
    "                    runner to 1e-10\n"
    "\n"
    "Each value is averaged over the ensemble and reported with its standard\n"
    "deviation and the standard error of the mean.\n";

struct Options {
  std::string command;
  std::filesystem::path config;
  std::size_t threads = 0;
  std::optional<std::size_t> realizations;
  std::optional<std::filesystem::path> output;
  bool batched = false;
};

std::size_t count_argument(std::string_view text, std::string_view flag) {
  std::size_t value = 0;
  const char* end = text.data() + text.size();
  const auto [stop, ec] = std::from_chars(text.data(), end, value);
  if (text.empty() || ec != std::errc{} || stop != end)
    throw std::invalid_argument(std::string(flag) + " expects a non-negative integer, got '" +
                                std::string(text) + "'");
  return value;
}

Options parse_arguments(const std::vector<std::string_view>& args) {
  if (args.size() < 2) throw std::invalid_argument("no command given");
  Options o;
  o.command = args[0];
  o.config = args[1];
  for (std::size_t i = 2; i < args.size(); ++i) {
    const std::string_view flag = args[i];
    if (flag == "--batched") {
      o.batched = true;
      continue;
    }
    const bool takes_value =
        flag == "--threads" || flag == "--out" || flag == "--realizations" || flag == "--backend";
    if (!takes_value) throw std::invalid_argument("unknown option '" + std::string(flag) + "'");
    if (i + 1 >= args.size()) throw std::invalid_argument(std::string(flag) + " needs a value");
    const std::string_view value = args[++i];
    if (flag == "--threads")
      o.threads = count_argument(value, flag);
    else if (flag == "--realizations")
      o.realizations = count_argument(value, flag);
    else if (flag == "--out")
      o.output = std::filesystem::path(value);
    else if (value != "hip")
      throw std::invalid_argument("backend '" + std::string(value) + "' is not built here; the only backend is hip");
  }
  return o;
}

const char* initial_state_name(InitialStateKind k) {
  switch (k) {
    case InitialStateKind::Random: return "random";
    case InitialStateKind::AllUp: return "up";
    case InitialStateKind::AllDown: return "down";
    case InitialStateKind::BasisState: return "basis";
  }
  return "?";
}

// Same labels and layout as the reference's describe, one setting a line.
void describe(const RunConfig& c) {
  std::printf("spins           %zu\n", c.nspins);
  std::printf("state space     %zu amplitudes\n", stateCount(c.nspins));
  std::printf("initial state   %s\n", initial_state_name(c.initialState));
  std::printf("realizations    %zu\n", c.realizations);
  std::printf("seed            %llu\n", static_cast<unsigned long long>(c.seed));
  if (c.generateGeometry)
    std::printf("geometry        dipolar 2D, per realization\n");
  else if (c.couplings.empty())
    std::printf("geometry        from config\n");
  else
    std::printf("geometry        from config, %zu fixed couplings\n", c.couplings.size());
  std::printf("field disorder  %g\n", c.fieldDisorder);
  std::printf("reference axis  %s\n", std::string(axisName(c.referenceAxis)).c_str());
  std::printf("epsilon         %g\n", c.epsilon);
  std::printf("measurements    %zu per realization\n", c.sequence.measurementCount());
  std::printf("observables    ");
  for (const std::string& spec : c.observables) std::printf(" %s", spec.c_str());
  std::printf("\n");
}

int run(const Options& o) {
  RunConfig config = loadConfig(o.config);
  if (o.realizations) {
    config.realizations = *o.realizations;
    config.validate(o.config.string());
  }
  if (o.command == "validate") {
    std::printf("%s: ok\n", o.config.string().c_str());
    return 0;
  }
  if (o.command == "describe") {
    describe(config);
    return 0;
  }
  if (o.command != "run") throw std::invalid_argument("unknown command '" + o.command + "'");

  LocalRunner runner(o.threads);
  const auto started = std::chrono::steady_clock::now();
  const EnsembleResult result =
      o.batched ? runExperimentBatched(config, o.threads) : runExperiment(config, runner);
  const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
  if (o.output) {
    writeCsv(result, *o.output);
    std::fprintf(stderr, "%zu realizations on %zu threads%s in %.2fs -> %s\n", config.realizations,
                 runner.threads(), o.batched ? " (batched)" : "", seconds, o.output->string().c_str());
  } else {
    writeCsv(result, std::cout);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::vector<std::string_view> args(argv + 1, argv + argc);
  if (args.empty() || args[0] == "-h" || args[0] == "--help") {
    std::fputs(kUsage, stdout);
    return args.empty() ? 2 : 0;
  }
  try {
    return run(parse_arguments(args));
  } catch (const std::exception& e) {
    std::fflush(stdout);
    std::fprintf(stderr, "spinsim: %s\n", e.what());
    return 1;
  }
}
