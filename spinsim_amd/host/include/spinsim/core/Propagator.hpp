#pragma once
// Chebyshev propagator (include/spinsim/core/Propagator.hpp:21-55 of the
// reference): each call builds the expansion on the host and runs the fused
// device orders (T_n recurrence + Bessel-weighted accumulation in one pass).

#include <cstddef>
#include <memory>

#include "spinsim/core/Chebyshev.hpp"
#include "spinsim/core/Hamiltonian.hpp"
#include "spinsim/core/StateVector.hpp"

namespace spinsim {

class Propagator {
 public:
  explicit Propagator(std::size_t nspins, double epsilon = 1.0e-7);
  ~Propagator();
  Propagator(const Propagator&) = delete;
  Propagator& operator=(const Propagator&) = delete;

  void evolve(const Hamiltonian& hamiltonian, TimeMode mode, StateVector& psi, double tau,
              std::size_t nsteps, bool autonormalize = false);

  /// The same step on two states under one Hamiltonian (a state and its
  /// reference, Sequence::runEvolve), one launch per order for both.
  void evolvePair(const Hamiltonian& hamiltonian, TimeMode mode, StateVector& a, StateVector& b,
                  double tau, std::size_t nsteps, bool autonormalize = false);

  [[nodiscard]] std::size_t lastOrder() const noexcept { return lastOrder_; }
  [[nodiscard]] double epsilon() const noexcept { return epsilon_; }

 private:
  std::size_t nspins_;
  double epsilon_;
  std::size_t lastOrder_ = 0;
  std::shared_ptr<DeviceContext> ctx_;
  spx_prop* dev_ = nullptr;
};

}  // namespace spinsim
