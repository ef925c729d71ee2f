#pragma once
// Ideal pulses (include/spinsim/core/Gate.hpp:14-93 of the reference).

#include <array>
#include <complex>
#include <string_view>
#include <vector>

#include "spinsim/core/Bits.hpp"
#include "spinsim/core/StateVector.hpp"

namespace spinsim {

struct Axis3 {
  double x = 0.0;
  double y = 0.0;
  double z = 1.0;
  [[nodiscard]] Axis3 normalized() const;
};

/// U = exp(-i (theta/2) n.sigma) on each target spin.
class SpinRotation {
 public:
  SpinRotation(Axis3 axis, double angle, std::vector<SpinIndex> targets);
  static SpinRotation onAllSpins(Axis3 axis, double angle, std::size_t nspins);
  void apply(StateVector& psi) const;
  [[nodiscard]] const std::vector<SpinIndex>& targets() const noexcept { return targets_; }
  [[nodiscard]] double angle() const noexcept { return angle_; }
  [[nodiscard]] Axis3 axis() const noexcept { return axis_; }

 private:
  Axis3 raw_;   // as given; the device normalises it as Axis3::normalized does
  Axis3 axis_;
  double angle_;
  std::vector<SpinIndex> targets_;
};

/// A 4x4 unitary on a pair, basis (down-down, low-up, high-up, up-up).
class TwoSpinGate {
 public:
  using Matrix = std::array<std::complex<double>, 16>;
  TwoSpinGate(Matrix matrix, SpinPair pair);
  void apply(StateVector& psi) const;
  [[nodiscard]] double unitarityDefect() const;

 private:
  Matrix matrix_;
  SpinPair pair_;
};

/// "pi:x", "pi/2:y", "pi:x:3" (1-based spin); std::invalid_argument otherwise.
[[nodiscard]] SpinRotation makeRotation(std::string_view spec, std::size_t nspins);

}  // namespace spinsim
