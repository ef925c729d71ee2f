// Sequence interpreter, RNG, realisation recipe, statistics and ensemble
// runners of the spinsim drop-in.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <exception>
#include <fstream>
#include <mutex>
#include <numbers>
#include <ostream>
#include <stdexcept>
#include <thread>

#include "spinsim/core/SpinOperators.hpp"
#include "spinsim/io/Experiment.hpp"

namespace spinsim {

// ------------------------------------------------------------------ Sequence

namespace {

void requireContext(const RunContext& ctx) {
  if (ctx.hamiltonian == nullptr || ctx.state == nullptr || ctx.reference == nullptr ||
      ctx.propagator == nullptr)
    throw std::invalid_argument("run context is missing a required component");
}

std::vector<Field> swapFields(Hamiltonian& h, const std::vector<Field>& fields) {
  if (fields.empty()) return {};
  if (fields.size() != h.nspins())
    throw std::invalid_argument("step field table does not match the number of spins");
  std::vector<Field> previous = h.fields();
  for (SpinIndex s = 0; s < h.nspins(); ++s) h.setField(s, fields[s]);
  return previous;
}

void restoreFields(Hamiltonian& h, const std::vector<Field>& previous) {
  if (previous.empty()) return;
  for (SpinIndex s = 0; s < h.nspins(); ++s) h.setField(s, previous[s]);
}

void runStep(const Step& step, RunContext& ctx);

// Sequence::runEvolve (src/sequence/Sequence.cpp:35-48): state and reference
// see the same evolution; here both go through one launch per order.
void runEvolve(const Evolve& step, RunContext& ctx) {
  const std::vector<Field> previous = swapFields(*ctx.hamiltonian, step.fields);
  try {
    ctx.propagator->evolvePair(*ctx.hamiltonian, step.mode, *ctx.state, *ctx.reference, step.tau,
                               step.nsteps, step.autonormalize);
  } catch (...) {
    restoreFields(*ctx.hamiltonian, previous);
    throw;
  }
  restoreFields(*ctx.hamiltonian, previous);
  ctx.time += step.tau * static_cast<double>(step.nsteps);
}

void runMeasure(RunContext& ctx) {
  if (ctx.recorder == nullptr) return;
  MeasureContext mc;
  mc.time = ctx.time;
  mc.state = ctx.state;
  mc.reference = ctx.reference;
  mc.hamiltonian = ctx.hamiltonian;
  mc.subsystemSize = ctx.subsystemSize;
  std::vector<double> values;
  for (const auto& observable : ctx.observables) {
    const std::size_t at = values.size();
    values.resize(at + observable->width());
    observable->measure(mc, std::span<double>(values).subspan(at));
  }
  ctx.recorder->record(ctx.time, values);
}

void runStep(const Step& step, RunContext& ctx) {
  std::visit(
      [&ctx](const auto& s) {
        using T = std::decay_t<decltype(s)>;
        if constexpr (std::is_same_v<T, Evolve>) {
          runEvolve(s, ctx);
        } else if constexpr (std::is_same_v<T, Pulse>) {
          s.rotation.apply(*ctx.state);
          s.rotation.apply(*ctx.reference);
        } else if constexpr (std::is_same_v<T, Measure>) {
          runMeasure(ctx);
        } else {
          for (std::size_t i = 0; i < s.count; ++i)
            for (const Step& inner : s.body) runStep(inner, ctx);
        }
      },
      step);
}

std::size_t countMeasurements(const std::vector<Step>& steps) {
  std::size_t total = 0;
  for (const Step& step : steps) {
    if (std::holds_alternative<Measure>(step)) ++total;
    else if (const auto* r = std::get_if<Repeat>(&step)) total += r->count * countMeasurements(r->body);
  }
  return total;
}

}  // namespace

void Sequence::run(RunContext& ctx) const {
  requireContext(ctx);
  for (const Step& step : steps_) runStep(step, ctx);
}

std::size_t Sequence::measurementCount() const { return countMeasurements(steps_); }

std::vector<std::string> resultColumns(std::span<const std::unique_ptr<Observable>> observables) {
  std::vector<std::string> columns{"time"};
  for (const auto& o : observables)
    for (std::string& name : o->columns()) columns.push_back(std::move(name));
  return columns;
}

// ----------------------------------------------------------------------- Rng

double Rng::gaussian() {
  if (hasSpare_) {
    hasSpare_ = false;
    return spare_;
  }
  constexpr double kFloor = 1.0e-15;
  const double u = kFloor + (1.0 - 2.0 * kFloor) * uniform();
  const double v = uniform();
  const double radius = std::sqrt(-2.0 * std::log(u));
  const double phase = 2.0 * std::numbers::pi * v;
  spare_ = radius * std::sin(phase);
  hasSpare_ = true;
  return radius * std::cos(phase);
}

// ------------------------------------------------------------------ Ensemble

InitialStateKind parseInitialStateKind(std::string_view name) {
  if (name == "random") return InitialStateKind::Random;
  if (name == "up") return InitialStateKind::AllUp;
  if (name == "down") return InitialStateKind::AllDown;
  if (name == "basis") return InitialStateKind::BasisState;
  throw std::invalid_argument("unknown initial state '" + std::string(name) +
                              "'; expected random, up, down or basis");
}

namespace {

// The random typical state of prepareInitialState (src/ensemble/Ensemble.cpp:20-45)
// on the host: (re, im) Gaussians per amplitude, serial norm, 1/norm scale.
void randomState(std::vector<double>& v, std::size_t nspins, Rng& rng) {
  const std::size_t n = stateCount(nspins);
  v.resize(2 * n);
  for (std::size_t k = 0; k < n; ++k) {
    v[2 * k] = rng.gaussian();
    v[2 * k + 1] = rng.gaussian();
  }
  double sum = 0.0;
  for (const double x : v) sum += x * x;
  const double nrm = std::sqrt(sum);
  if (nrm == 0.0) throw std::domain_error("cannot normalize a zero state");
  const double inv = 1.0 / nrm;
  for (double& x : v) x *= inv;
}

}  // namespace

void prepareInitialState(StateVector& psi, InitialStateKind kind, Rng& rng, std::size_t basisIndex) {
  switch (kind) {
    case InitialStateKind::AllUp: psi.setBasisState(psi.size() - 1); return;
    case InitialStateKind::AllDown: psi.setBasisState(0); return;
    case InitialStateKind::BasisState: psi.setBasisState(basisIndex); return;
    case InitialStateKind::Random: break;
  }
  std::vector<double> v;
  randomState(v, psi.nspins(), rng);
  check(spx_state_upload(psi.device(true), v.data()));
}

void prepareReferenceState(const StateVector& psi, Axis axis, StateVector& reference) {
  if (reference.size() != psi.size() || reference.nspins() != psi.nspins())
    reference = StateVector(psi.nspins());
  check(spx_apply_total_spin(static_cast<int>(axis), psi.device(false), reference.device(true)));
}

std::vector<Coupling> generateDipolarLattice2D(std::size_t nspins, Rng& rng, double scale,
                                               double minSeparation, bool secularAngle) {
  if (nspins == 0) throw std::invalid_argument("need at least one spin");
  const double halfWidth = 0.5 * std::sqrt(static_cast<double>(nspins));
  std::vector<double> x(nspins, 0.0), y(nspins, 0.0);
  std::vector<Coupling> couplings;
  couplings.reserve(nspins * (nspins - 1) / 2);
  for (std::size_t i = 1; i < nspins; ++i) {
    double px = 0.0, py = 0.0;
    bool placed = false;
    while (!placed) {
      px = halfWidth * (2.0 * rng.uniform() - 1.0);
      py = halfWidth * (2.0 * rng.uniform() - 1.0);
      placed = true;
      for (std::size_t j = 0; j < i; ++j)
        if (std::abs(px - x[j]) < minSeparation && std::abs(py - y[j]) < minSeparation) {
          placed = false;
          break;
        }
    }
    x[i] = px;
    y[i] = py;
    for (std::size_t j = 0; j < i; ++j) {
      const double dx = px - x[j], dy = py - y[j];
      const double r2 = dx * dx + dy * dy;
      double jz = scale / (r2 * std::sqrt(r2));
      if (secularAngle) jz = jz * (1.0 - 3.0 * ((dx * dx) / r2));
      couplings.push_back(Coupling{SpinPair{j, i}, -0.5 * jz, -0.5 * jz, jz});
    }
  }
  return couplings;
}

void applyFieldDisorder(Hamiltonian& hamiltonian, Rng& rng, double sigma) {
  for (SpinIndex s = 0; s < hamiltonian.nspins(); ++s) {
    Field f = hamiltonian.field(s);
    f.z += sigma * rng.gaussian();
    hamiltonian.setField(s, f);
  }
}

Accumulator::Accumulator(std::size_t width) : mean_(width, 0.0), m2_(width, 0.0) {}

void Accumulator::add(std::span<const double> sample) {
  if (sample.size() != mean_.size())
    throw std::invalid_argument("sample width does not match the accumulator");
  ++count_;
  const double n = static_cast<double>(count_);
  for (std::size_t k = 0; k < sample.size(); ++k) {
    const double delta = sample[k] - mean_[k];
    mean_[k] += delta / n;
    m2_[k] += delta * (sample[k] - mean_[k]);
  }
}

void Accumulator::merge(const Accumulator& other) {
  if (other.mean_.size() != mean_.size())
    throw std::invalid_argument("cannot merge accumulators of different width");
  if (other.count_ == 0) return;
  if (count_ == 0) {
    *this = other;
    return;
  }
  const double na = static_cast<double>(count_), nb = static_cast<double>(other.count_);
  const double total = na + nb;
  for (std::size_t k = 0; k < mean_.size(); ++k) {
    const double delta = other.mean_[k] - mean_[k];
    mean_[k] += delta * nb / total;
    m2_[k] += other.m2_[k] + delta * delta * na * nb / total;
  }
  count_ += other.count_;
}

std::vector<double> Accumulator::standardDeviation() const {
  std::vector<double> out(mean_.size(), 0.0);
  if (count_ < 2) return out;
  const double denom = static_cast<double>(count_ - 1);
  for (std::size_t k = 0; k < out.size(); ++k) out[k] = std::sqrt(m2_[k] / denom);
  return out;
}

std::vector<double> Accumulator::standardError() const {
  std::vector<double> out = standardDeviation();
  if (count_ < 2) return out;
  const double scale = 1.0 / std::sqrt(static_cast<double>(count_));
  for (double& v : out) v *= scale;
  return out;
}

// -------------------------------------------------------------------- Runner

void VectorRecorder::record(double time, std::span<const double> values) {
  times_.push_back(time);
  values_.insert(values_.end(), values.begin(), values.end());
}

void VectorRecorder::clear() {
  times_.clear();
  values_.clear();
}

LocalRunner::LocalRunner(std::size_t threads)
    : threads_(threads > 0 ? threads : std::thread::hardware_concurrency()) {
  if (threads_ == 0) threads_ = 1;
}

EnsembleResult foldInOrder(std::vector<VectorRecorder>& per, std::span<const std::string> columns) {
  EnsembleResult result;
  result.times = per.front().times();
  result.columns.assign(columns.begin(), columns.end());
  const std::size_t points = result.times.size();
  const std::size_t width = points * result.columns.size();
  result.statistics = Accumulator(width);
  for (std::size_t i = 0; i < per.size(); ++i) {
    const VectorRecorder& r = per[i];
    if (r.rows() != points || r.values().size() != width) {
      char msg[160];
      std::snprintf(msg, sizeof msg, "realisation %zu recorded %zu points and %zu values; expected %zu and %zu",
                    i, r.rows(), r.values().size(), points, width);
      throw std::runtime_error(msg);
    }
    result.statistics.add(r.values());
  }
  return result;
}

EnsembleResult LocalRunner::run(std::size_t realizations, std::span<const std::string> columns,
                                const RealizationJob& job) {
  if (realizations == 0) throw std::invalid_argument("an ensemble needs at least one realisation");
  std::vector<VectorRecorder> per(realizations);
  std::atomic<std::size_t> next{0};
  std::mutex failureMutex;
  std::exception_ptr failure;
  const std::size_t workers = std::min(threads_, realizations);
  {
    std::vector<std::jthread> pool;
    pool.reserve(workers);
    for (std::size_t w = 0; w < workers; ++w) {
      pool.emplace_back([&] {
        for (;;) {
          const std::size_t index = next.fetch_add(1);
          if (index >= realizations) return;
          try {
            job(index, per[index]);
          } catch (...) {
            const std::lock_guard lock(failureMutex);
            if (!failure) failure = std::current_exception();
            next.store(realizations);
            return;
          }
        }
      });
    }
  }
  if (failure) std::rethrow_exception(failure);
  return foldInOrder(per, columns);
}

// -------------------------------------------------------------------- Config

ConfigError::ConfigError(std::string_view file, std::string_view key, std::string_view problem)
    : std::runtime_error(std::string(file) + ": '" + std::string(key) + "' " + std::string(problem)) {}

void RunConfig::validate(std::string_view origin) const {
  if (nspins == 0 || nspins > kMaxSpins)
    throw ConfigError(origin, "system.spins", "must be between 1 and " + std::to_string(kMaxSpins));
  if (baseFields.size() != nspins) throw ConfigError(origin, "field", "does not cover every spin");
  if (realizations == 0) throw ConfigError(origin, "ensemble.realizations", "must be at least 1");
  if (!(epsilon > 0.0)) throw ConfigError(origin, "numerics.epsilon", "must be positive");
  if (observables.empty()) throw ConfigError(origin, "measurement.observables", "must not be empty");
  if (sequence.measurementCount() == 0)
    throw ConfigError(origin, "step", "the sequence records nothing; add a measure step");
  if (initialState == InitialStateKind::BasisState && basisIndex >= stateCount(nspins))
    throw ConfigError(origin, "system.basis_index", "outside the 2^L state space");
  if (!couplings.empty() && generateGeometry)
    throw ConfigError(origin, "coupling", "fixed couplings need ensemble.dipolar_geometry = false");
  for (std::size_t p = 0; p < couplings.size(); ++p) {
    const SpinPair pair = couplings[p].pair;
    if (pair.high() >= nspins || pair.low() == pair.high())
      throw ConfigError(origin, "coupling", "pair " + std::to_string(p + 1) + " names a spin outside the system");
    for (std::size_t q = 0; q < p; ++q)
      if (couplings[q].pair == pair)
        throw ConfigError(origin, "coupling", "pair " + std::to_string(p + 1) + " repeats pair " +
                                                  std::to_string(q + 1));
  }
  for (const std::string& spec : observables) {
    try {
      (void)makeObservable(spec);
    } catch (const std::invalid_argument& e) {
      throw ConfigError(origin, "measurement.observables", e.what());
    }
  }
}

// ---------------------------------------------------------------- Experiment

namespace {

/// Fields, couplings and initial state of realisation `index`, drawn in
/// runRealization's order (src/io/Experiment.cpp:10-55).
struct Inputs {
  std::vector<Field> fields;
  std::vector<Coupling> couplings;
  std::vector<double> psi0;   // only for a random initial state
};

Inputs drawInputs(const RunConfig& cfg, std::size_t index) {
  Rng rng = Rng::forRealization(cfg.seed, index);
  Inputs in;
  in.fields = cfg.baseFields;
  if (cfg.generateGeometry)
    in.couplings = generateDipolarLattice2D(cfg.nspins, rng, cfg.dipolarScale, cfg.minSeparation,
                                            cfg.secularAngle);
  else
    in.couplings = cfg.couplings;
  if (cfg.fieldDisorder != 0.0)
    for (Field& f : in.fields) f.z += cfg.fieldDisorder * rng.gaussian();
  if (cfg.initialState == InitialStateKind::Random) randomState(in.psi0, cfg.nspins, rng);
  return in;
}

}  // namespace

void runRealization(const RunConfig& config, std::size_t index, VectorRecorder& out) {
  Rng rng = Rng::forRealization(config.seed, index);
  Hamiltonian hamiltonian(config.nspins);
  for (SpinIndex s = 0; s < config.nspins; ++s) hamiltonian.setField(s, config.baseFields[s]);
  if (config.generateGeometry)
    for (const Coupling& c : generateDipolarLattice2D(config.nspins, rng, config.dipolarScale,
                                                      config.minSeparation, config.secularAngle))
      hamiltonian.setCoupling(c.pair, c.jx, c.jy, c.jz);
  else
    for (const Coupling& c : config.couplings) hamiltonian.setCoupling(c.pair, c.jx, c.jy, c.jz);
  if (config.fieldDisorder != 0.0) applyFieldDisorder(hamiltonian, rng, config.fieldDisorder);
  StateVector state(config.nspins);
  prepareInitialState(state, config.initialState, rng, config.basisIndex);
  StateVector reference(config.nspins);
  prepareReferenceState(state, config.referenceAxis, reference);
  std::vector<std::unique_ptr<Observable>> observables;
  for (const std::string& spec : config.observables) observables.push_back(makeObservable(spec));
  Propagator propagator(config.nspins, config.epsilon);
  RunContext ctx;
  ctx.hamiltonian = &hamiltonian;
  ctx.state = &state;
  ctx.reference = &reference;
  ctx.propagator = &propagator;
  ctx.observables = observables;
  ctx.recorder = &out;
  ctx.time = 0.0;
  config.sequence.run(ctx);
}

EnsembleResult runExperiment(const RunConfig& config, EnsembleRunner& runner) {
  std::vector<std::unique_ptr<Observable>> probes;
  for (const std::string& spec : config.observables) probes.push_back(makeObservable(spec));
  std::vector<std::string> columns = resultColumns(probes);
  columns.erase(columns.begin());
  return runner.run(config.realizations, columns, [&config](std::size_t index, VectorRecorder& out) {
    runRealization(config, index, out);
  });
}

namespace {

void flatten(spx_program* p, const std::vector<Step>& steps) {
  for (const Step& step : steps) {
    if (const auto* e = std::get_if<Evolve>(&step)) {
      std::vector<double> f;
      for (const Field& fl : e->fields) {
        f.push_back(fl.x);
        f.push_back(fl.y);
        f.push_back(fl.z);
      }
      check(spx_program_evolve(p, e->tau, e->nsteps, e->mode == TimeMode::Real ? 0 : 1,
                               e->autonormalize ? 1 : 0, f.empty() ? nullptr : f.data(),
                               e->fields.size()));
    } else if (const auto* pu = std::get_if<Pulse>(&step)) {
      const SpinRotation& r = pu->rotation;
      std::vector<uint32_t> t(r.targets().begin(), r.targets().end());
      check(spx_program_pulse(p, r.axis().x, r.axis().y, r.axis().z, r.angle(), t.data(), t.size()));
    } else if (std::holds_alternative<Measure>(step)) {
      check(spx_program_measure(p));
    } else {
      const Repeat& rp = std::get<Repeat>(step);
      check(spx_program_repeat(p, rp.count));
      flatten(p, rp.body);
      check(spx_program_end(p));
    }
  }
}

}  // namespace

EnsembleResult runExperimentBatched(const RunConfig& config, std::size_t hostThreads) {
  config.validate();
  if (config.initialState != InitialStateKind::Random)
    throw std::invalid_argument("the batched runner draws random initial states only");
  std::vector<std::unique_ptr<Observable>> probes;
  std::vector<int> codes, spins;
  for (const std::string& spec : config.observables) {
    probes.push_back(makeObservable(spec));
    codes.push_back(probes.back()->code());
    spins.push_back(probes.back()->codeSpin());
  }
  std::vector<std::string> columns = resultColumns(probes);
  columns.erase(columns.begin());

  spx_program* prog = nullptr;
  check(spx_program_create(&prog));
  std::unique_ptr<spx_program, int (*)(spx_program*)> progGuard(prog, spx_program_destroy);
  flatten(prog, config.sequence.steps());
  auto ctx = DeviceContext::current();
  spx_batch* batch = nullptr;
  check(spx_batch_create(ctx->get(), prog, config.epsilon, static_cast<int>(config.referenceAxis),
                         codes.data(), spins.data(), codes.size(), &batch));
  std::unique_ptr<spx_batch, int (*)(spx_batch*)> batchGuard(batch, spx_batch_destroy);

  // realisation inputs drawn on host threads, added in index order
  const std::size_t n = config.realizations;
  std::vector<Inputs> inputs(n);
  std::size_t threads = hostThreads ? hostThreads : std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, n);
  {
    std::atomic<std::size_t> next{0};
    std::vector<std::jthread> pool;
    for (std::size_t w = 0; w < threads; ++w)
      pool.emplace_back([&] {
        for (std::size_t i; (i = next.fetch_add(1)) < n;) inputs[i] = drawInputs(config, i);
      });
  }
  for (std::size_t i = 0; i < n; ++i) {
    const Inputs& in = inputs[i];
    std::vector<double> f;
    for (const Field& fl : in.fields) {
      f.push_back(fl.x);
      f.push_back(fl.y);
      f.push_back(fl.z);
    }
    std::vector<uint32_t> lo, hi;
    std::vector<double> jx, jy, jz;
    for (const Coupling& c : in.couplings) {
      lo.push_back(static_cast<uint32_t>(c.pair.low()));
      hi.push_back(static_cast<uint32_t>(c.pair.high()));
      jx.push_back(c.jx);
      jy.push_back(c.jy);
      jz.push_back(c.jz);
    }
    check(spx_batch_add(batch, config.nspins, f.data(), lo.size(), lo.data(), hi.data(), jx.data(),
                        jy.data(), jz.data(), in.psi0.data()));
  }
  const std::size_t points = config.sequence.measurementCount();
  std::size_t width = 0;
  check(spx_batch_width(batch, &width));
  std::vector<double> times(points), values(n * points * width);
  check(spx_batch_run(batch, times.data(), values.data()));
  std::vector<VectorRecorder> per(n);
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t p = 0; p < points; ++p)
      per[i].record(times[p], std::span<const double>(values.data() + (i * points + p) * width, width));
  return foldInOrder(per, columns);
}

void writeCsv(const EnsembleResult& result, std::ostream& out) {
  const std::size_t ncols = result.columns.size();
  if (result.statistics.width() != result.times.size() * ncols)
    throw std::runtime_error("result shape does not match its column names");
  out << "time";
  for (const std::string& name : result.columns)
    out << ',' << name << "_mean," << name << "_sd," << name << "_stderr";
  out << '\n';
  const std::vector<double>& mean = result.statistics.mean();
  const std::vector<double> sd = result.statistics.standardDeviation();
  const std::vector<double> se = result.statistics.standardError();
  char buf[64];
  for (std::size_t point = 0; point < result.times.size(); ++point) {
    std::snprintf(buf, sizeof buf, "%.10g", result.times[point]);
    out << buf;
    for (std::size_t c = 0; c < ncols; ++c) {
      const std::size_t k = point * ncols + c;
      std::snprintf(buf, sizeof buf, ",%.12g", mean[k]);
      out << buf;
      std::snprintf(buf, sizeof buf, ",%.12g", sd[k]);
      out << buf;
      std::snprintf(buf, sizeof buf, ",%.12g", se[k]);
      out << buf;
    }
    out << '\n';
  }
}

void writeCsv(const EnsembleResult& result, const std::filesystem::path& path) {
  if (path.has_parent_path()) std::filesystem::create_directories(path.parent_path());
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot open '" + path.string() + "' for writing");
  writeCsv(result, out);
}

}  // namespace spinsim
