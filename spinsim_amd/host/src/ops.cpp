// Propagator, Pauli kernels, gates and observables of the spinsim drop-in.
#include <charconv>
#include <cmath>
#include <numbers>
#include <stdexcept>
#include <string>

#include "spinsim/core/Gate.hpp"
#include "spinsim/core/Observable.hpp"
#include "spinsim/core/Propagator.hpp"
#include "spinsim/core/SpinOperators.hpp"

namespace spinsim {

// ---------------------------------------------------------------- Propagator

Propagator::Propagator(std::size_t nspins, double epsilon)
    : nspins_(nspins), epsilon_(epsilon), ctx_(DeviceContext::current()) {
  if (!(epsilon > 0.0)) throw std::invalid_argument("epsilon must be positive");
  check(spx_prop_create(ctx_->get(), nspins, epsilon, &dev_));
}

Propagator::~Propagator() {
  if (dev_) spx_prop_destroy(dev_);
}

void Propagator::evolve(const Hamiltonian& hamiltonian, TimeMode mode, StateVector& psi, double tau,
                        std::size_t nsteps, bool autonormalize) {
  if (psi.nspins() != hamiltonian.nspins())
    throw std::invalid_argument("state and Hamiltonian disagree on spin count");
  if (psi.nspins() != nspins_) throw std::invalid_argument("state does not match the propagator's size");
  if (nsteps == 0) return;
  check(spx_prop_evolve(dev_, hamiltonian.device(), mode == TimeMode::Real ? 0 : 1, psi.device(true),
                        tau, nsteps, autonormalize ? 1 : 0));
  check(spx_prop_last_order(dev_, &lastOrder_));
}

void Propagator::evolvePair(const Hamiltonian& hamiltonian, TimeMode mode, StateVector& a,
                            StateVector& b, double tau, std::size_t nsteps, bool autonormalize) {
  for (const StateVector* s : {&a, &b}) {
    if (s->nspins() != hamiltonian.nspins())
      throw std::invalid_argument("state and Hamiltonian disagree on spin count");
    if (s->nspins() != nspins_) throw std::invalid_argument("state does not match the propagator's size");
  }
  if (nsteps == 0) return;
  spx_state* both[2] = {a.device(true), b.device(true)};
  check(spx_prop_evolve_many(dev_, hamiltonian.device(), mode == TimeMode::Real ? 0 : 1, both, 2,
                             tau, nsteps, autonormalize ? 1 : 0));
  check(spx_prop_last_order(dev_, &lastOrder_));
}

// ------------------------------------------------------------ Pauli kernels

namespace {

void sigma(int axis, double c, const StateVector& in, SpinIndex spin, StateVector& out, bool acc) {
  if (&in == &out) throw std::invalid_argument("kernels require distinct input and output");
  check(spx_apply_sigma(axis, c, in.device(false), spin, out.device(true), acc ? 1 : 0));
}

}  // namespace

void axpySigmaX(double c, const StateVector& in, SpinIndex s, StateVector& out) { sigma(0, c, in, s, out, true); }
void axpySigmaY(double c, const StateVector& in, SpinIndex s, StateVector& out) { sigma(1, c, in, s, out, true); }
void axpySigmaZ(double c, const StateVector& in, SpinIndex s, StateVector& out) { sigma(2, c, in, s, out, true); }
void applySigmaX(const StateVector& in, SpinIndex s, StateVector& out) { sigma(0, 1.0, in, s, out, false); }
void applySigmaY(const StateVector& in, SpinIndex s, StateVector& out) { sigma(1, 1.0, in, s, out, false); }
void applySigmaZ(const StateVector& in, SpinIndex s, StateVector& out) { sigma(2, 1.0, in, s, out, false); }

// --------------------------------------------------------------------- Gates

Axis3 Axis3::normalized() const {
  const double length = std::sqrt(x * x + y * y + z * z);
  if (!(length > 0.0)) throw std::invalid_argument("rotation axis must have non-zero length");
  return Axis3{x / length, y / length, z / length};
}

SpinRotation::SpinRotation(Axis3 axis, double angle, std::vector<SpinIndex> targets)
    : raw_(axis), axis_(axis.normalized()), angle_(angle), targets_(std::move(targets)) {
  if (targets_.empty()) throw std::invalid_argument("a rotation must name at least one spin");
}

SpinRotation SpinRotation::onAllSpins(Axis3 axis, double angle, std::size_t nspins) {
  std::vector<SpinIndex> all(nspins);
  for (std::size_t s = 0; s < nspins; ++s) all[s] = s;
  return SpinRotation(axis, angle, std::move(all));
}

void SpinRotation::apply(StateVector& psi) const {
  std::vector<uint32_t> t(targets_.begin(), targets_.end());
  for (SpinIndex s : targets_)
    if (s >= psi.nspins()) throw std::out_of_range("rotation names a spin outside the system");
  check(spx_rotate(psi.device(true), raw_.x, raw_.y, raw_.z, angle_, t.data(), t.size()));
}

TwoSpinGate::TwoSpinGate(Matrix matrix, SpinPair pair) : matrix_(matrix), pair_(pair) {}

void TwoSpinGate::apply(StateVector& psi) const {
  if (pair_.high() >= psi.nspins()) throw std::out_of_range("gate names a spin outside the system");
  double u[32];
  for (int i = 0; i < 16; ++i) {
    u[2 * i] = matrix_[i].real();
    u[2 * i + 1] = matrix_[i].imag();
  }
  check(spx_two_spin_gate(psi.device(true), u, pair_.low(), pair_.high()));
}

double TwoSpinGate::unitarityDefect() const {
  double worst = 0.0;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      std::complex<double> acc{0.0, 0.0};
      for (int k = 0; k < 4; ++k) acc += std::conj(matrix_[k * 4 + r]) * matrix_[k * 4 + c];
      const std::complex<double> want = r == c ? 1.0 : 0.0;
      worst = std::max(worst, std::abs(acc - want));
    }
  return worst;
}

namespace {

std::vector<std::string_view> split(std::string_view spec) {
  std::vector<std::string_view> parts;
  std::size_t start = 0;
  for (;;) {
    const std::size_t sep = spec.find(':', start);
    if (sep == std::string_view::npos) {
      parts.push_back(spec.substr(start));
      return parts;
    }
    parts.push_back(spec.substr(start, sep - start));
    start = sep + 1;
  }
}

std::string quoted(std::string_view s) { return "'" + std::string(s) + "'"; }

}  // namespace

SpinRotation makeRotation(std::string_view spec, std::size_t nspins) {
  const auto parts = split(spec);
  if (parts.size() != 2 && parts.size() != 3)
    throw std::invalid_argument("malformed pulse " + quoted(spec) +
                                "; expected <angle>:<axis> or <angle>:<axis>:<spin>");
  double angle = 0.0;
  if (parts[0] == "pi") angle = std::numbers::pi;
  else if (parts[0] == "pi/2") angle = 0.5 * std::numbers::pi;
  else if (parts[0] == "pi/4") angle = 0.25 * std::numbers::pi;
  else throw std::invalid_argument("unknown pulse angle " + quoted(parts[0]) + "; expected pi, pi/2 or pi/4");
  Axis3 axis;
  if (parts[1] == "x") axis = Axis3{1.0, 0.0, 0.0};
  else if (parts[1] == "y") axis = Axis3{0.0, 1.0, 0.0};
  else if (parts[1] == "z") axis = Axis3{0.0, 0.0, 1.0};
  else throw std::invalid_argument("unknown rotation axis " + quoted(parts[1]) + "; expected x, y or z");
  if (parts.size() == 2) return SpinRotation::onAllSpins(axis, angle, nspins);
  unsigned long index = 0;
  const char* first = parts[2].data();
  const char* last = first + parts[2].size();
  const auto [ptr, ec] = std::from_chars(first, last, index);
  if (ec != std::errc{} || ptr != last || index == 0 || index > nspins)
    throw std::invalid_argument("pulse " + quoted(spec) + ": spin must be between 1 and " +
                                std::to_string(nspins));
  return SpinRotation(axis, angle, {static_cast<SpinIndex>(index - 1)});
}

// --------------------------------------------------------------- Observables

Axis parseAxis(std::string_view name) {
  if (name == "x" || name == "X") return Axis::X;
  if (name == "y" || name == "Y") return Axis::Y;
  if (name == "z" || name == "Z") return Axis::Z;
  throw std::invalid_argument("unknown axis " + quoted(name) + "; expected x, y or z");
}

std::string_view axisName(Axis axis) {
  switch (axis) {
    case Axis::X: return "x";
    case Axis::Y: return "y";
    case Axis::Z: return "z";
  }
  return "?";
}

namespace {

void requireStates(const MeasureContext& ctx, bool needsReference) {
  if (ctx.state == nullptr) throw std::invalid_argument("measurement context has no state");
  if (needsReference && ctx.reference == nullptr)
    throw std::invalid_argument("this observable needs a reference state");
}

}  // namespace

std::string TotalSpinCorrelator::name() const { return "corr:" + std::string(axisName(axis_)); }

std::vector<std::string> TotalSpinCorrelator::columns() const {
  const std::string a(axisName(axis_));
  return {"corr_" + a + "_re", "corr_" + a + "_im"};
}

void TotalSpinCorrelator::measure(const MeasureContext& ctx, std::span<double> out) const {
  requireStates(ctx, true);
  check(spx_measure_corr(static_cast<int>(axis_), ctx.state->device(false),
                         ctx.reference->device(false), &out[0], &out[1]));
}

std::string SingleSpinMagnetization::name() const {
  return "mag:" + std::to_string(spin_ + 1) + ":" + std::string(axisName(axis_));
}

std::vector<std::string> SingleSpinMagnetization::columns() const {
  return {"mag_" + std::to_string(spin_ + 1) + "_" + std::string(axisName(axis_))};
}

void SingleSpinMagnetization::measure(const MeasureContext& ctx, std::span<double> out) const {
  requireStates(ctx, false);
  if (spin_ >= ctx.state->nspins()) throw std::out_of_range("observable names a spin outside the system");
  check(spx_measure_mag(static_cast<int>(axis_), spin_, ctx.state->device(false), &out[0]));
}

std::vector<std::string> StateNorms::columns() const { return {"norm_state", "norm_reference"}; }

void StateNorms::measure(const MeasureContext& ctx, std::span<double> out) const {
  requireStates(ctx, true);
  out[0] = ctx.state->norm();
  out[1] = ctx.reference->norm();
}

std::vector<std::string> Energy::columns() const { return {"energy"}; }

void Energy::measure(const MeasureContext& ctx, std::span<double> out) const {
  requireStates(ctx, false);
  if (ctx.hamiltonian == nullptr) throw std::invalid_argument("energy needs a Hamiltonian in the context");
  check(spx_measure_energy(ctx.hamiltonian->device(), ctx.state->device(false), &out[0]));
}

std::unique_ptr<Observable> makeObservable(std::string_view spec) {
  const auto parts = split(spec);
  const auto reject = [&]() -> std::unique_ptr<Observable> {
    throw std::invalid_argument("unknown observable " + quoted(spec) +
                                "; expected one of corr:<axis>, mag:<spin>:<axis>, norm, energy");
  };
  if (parts.empty()) return reject();
  if (parts[0] == "norm" && parts.size() == 1) return std::make_unique<StateNorms>();
  if (parts[0] == "energy" && parts.size() == 1) return std::make_unique<Energy>();
  if (parts[0] == "corr" && parts.size() == 2) return std::make_unique<TotalSpinCorrelator>(parseAxis(parts[1]));
  if (parts[0] == "mag" && parts.size() == 3) {
    unsigned long index = 0;
    const char* first = parts[1].data();
    const char* last = first + parts[1].size();
    const auto [ptr, ec] = std::from_chars(first, last, index);
    if (ec != std::errc{} || ptr != last || index == 0)
      throw std::invalid_argument("observable " + quoted(spec) + ": spin must be a positive integer");
    return std::make_unique<SingleSpinMagnetization>(static_cast<SpinIndex>(index - 1), parseAxis(parts[2]));
  }
  return reject();
}

}  // namespace spinsim
