#pragma once
// Realisation recipe pieces and the running statistics
// (include/spinsim/ensemble/Ensemble.hpp:21-96 of the reference).

#include <cstddef>
#include <span>
#include <string_view>
#include <vector>

#include "spinsim/core/Hamiltonian.hpp"
#include "spinsim/core/Observable.hpp"
#include "spinsim/core/StateVector.hpp"
#include "spinsim/ensemble/Rng.hpp"

namespace spinsim {

enum class InitialStateKind { Random, AllUp, AllDown, BasisState };

[[nodiscard]] InitialStateKind parseInitialStateKind(std::string_view name);

/// Random: 2^L complex Gaussians drawn on the host in (re, im) order from
/// `rng`, normalised on the device.
void prepareInitialState(StateVector& psi, InitialStateKind kind, Rng& rng,
                         std::size_t basisIndex = 0);

/// reference = sum_i sigma_axis(i) psi, on the device.
void prepareReferenceState(const StateVector& psi, Axis axis, StateVector& reference);

/// Random 2D placement with the both-coordinates rejection and Jz = scale/r^3,
/// Jx = Jy = -Jz/2, couplings in (j < i) order. `secularAngle` multiplies Jz
/// by (1 - 3 cos^2 theta), theta the angle to an in-plane x field (no
/// counterpart in the reference generator; the draws are unchanged).
[[nodiscard]] std::vector<Coupling> generateDipolarLattice2D(std::size_t nspins, Rng& rng,
                                                             double scale = 1.0,
                                                             double minSeparation = 5.0e-2,
                                                             bool secularAngle = false);

void applyFieldDisorder(Hamiltonian& hamiltonian, Rng& rng, double sigma);

class Accumulator {
 public:
  explicit Accumulator(std::size_t width);
  void add(std::span<const double> sample);
  void merge(const Accumulator& other);
  [[nodiscard]] std::size_t count() const noexcept { return count_; }
  [[nodiscard]] std::size_t width() const noexcept { return mean_.size(); }
  [[nodiscard]] const std::vector<double>& mean() const noexcept { return mean_; }
  [[nodiscard]] std::vector<double> standardDeviation() const;
  [[nodiscard]] std::vector<double> standardError() const;

 private:
  std::size_t count_ = 0;
  std::vector<double> mean_;
  std::vector<double> m2_;
};

}  // namespace spinsim
