#pragma once
// H = sum_i (Hx Sx + Hy Sy + Hz Sz)_i + sum_<ij> (Jx SxSx + Jy SySy + Jz SzSz),
// S = sigma/2 (include/spinsim/core/Hamiltonian.hpp:11-110 of the reference).
// Fields and couplings are kept on the host; libspx holds the device
// sigma_z diagonal and flip-term table and rebuilds them when they change.

#include <vector>

#include "spinsim/core/Bits.hpp"
#include "spinsim/core/Device.hpp"
#include "spinsim/core/StateVector.hpp"

namespace spinsim {

struct Field {
  double x = 0.0;
  double y = 0.0;
  double z = 0.0;
};

struct Coupling {
  SpinPair pair;
  double jx = 0.0;
  double jy = 0.0;
  double jz = 0.0;
};

class Hamiltonian {
 public:
  explicit Hamiltonian(std::size_t nspins);
  Hamiltonian(const Hamiltonian& other);
  Hamiltonian& operator=(const Hamiltonian& other);
  ~Hamiltonian();

  [[nodiscard]] std::size_t nspins() const noexcept { return nspins_; }
  void setField(SpinIndex spin, Field field);
  [[nodiscard]] const Field& field(SpinIndex spin) const;
  void setCoupling(SpinPair pair, double jx, double jy, double jz);
  [[nodiscard]] const std::vector<Coupling>& couplings() const noexcept { return couplings_; }
  [[nodiscard]] const std::vector<Field>& fields() const noexcept { return fields_; }
  [[nodiscard]] double spectralBound() const noexcept;

  void applyPhysical(const StateVector& in, StateVector& out) const;
  void applyTwiceNormalized(double bound, const StateVector& in, StateVector& out) const;
  void accumulateTwiceNormalized(double bound, const StateVector& in, StateVector& out) const;

  [[nodiscard]] spx_ham* device() const noexcept { return dev_; }

 private:
  void create();
  std::size_t nspins_;
  std::vector<Field> fields_;
  std::vector<Coupling> couplings_;
  std::shared_ptr<DeviceContext> ctx_;
  spx_ham* dev_ = nullptr;
};

}  // namespace spinsim
