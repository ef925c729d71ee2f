#pragma once
// StateVector (include/spinsim/core/StateVector.hpp:82-140 of the reference):
// 2^L interleaved complex FP64 amplitudes. Here the amplitudes live in HBM;
// data()/re()/im() expose a host mirror that is synchronised on demand, so
// code written against the reference's host accessors keeps working.

#include <cstddef>
#include <memory>
#include <span>
#include <utility>
#include <vector>

#include "spinsim/core/Bits.hpp"
#include "spinsim/core/Device.hpp"

namespace spinsim {

class StateVector {
 public:
  StateVector() = default;
  explicit StateVector(std::size_t nspins);
  StateVector(const StateVector& other);
  StateVector& operator=(const StateVector& other);
  StateVector(StateVector&& other) noexcept;
  StateVector& operator=(StateVector&& other) noexcept;
  ~StateVector();

  [[nodiscard]] std::size_t nspins() const noexcept { return nspins_; }
  [[nodiscard]] std::size_t size() const noexcept { return stateCount(nspins_) * (dev_ ? 1 : 0); }

  /// Host view of the interleaved store (2 * size() doubles). The mutable
  /// overloads mark the host copy as the newest; the next device operation
  /// uploads it.
  [[nodiscard]] double* data();
  [[nodiscard]] const double* data() const;
  [[nodiscard]] std::span<double> amplitudes();
  [[nodiscard]] std::span<const double> amplitudes() const;
  [[nodiscard]] double& re(std::size_t k) { return data()[2 * k]; }
  [[nodiscard]] double& im(std::size_t k) { return data()[2 * k + 1]; }
  [[nodiscard]] double re(std::size_t k) const { return data()[2 * k]; }
  [[nodiscard]] double im(std::size_t k) const { return data()[2 * k + 1]; }

  void setZero();
  void scale(double factor);
  void negate();
  void addScaled(double cr, double ci, const StateVector& other);
  void setBasisState(std::size_t index);
  [[nodiscard]] double norm() const;
  double normalize();
  [[nodiscard]] std::pair<double, double> innerProduct(const StateVector& other) const;
  void swap(StateVector& other) noexcept;

  /// The device vector, current (host edits uploaded). `willWrite` marks the
  /// device copy as the newest afterwards.
  [[nodiscard]] spx_state* device(bool willWrite) const;
  [[nodiscard]] const std::shared_ptr<DeviceContext>& context() const noexcept { return ctx_; }

 private:
  enum class Fresh { Both, Host, Device };
  void pull() const;

  std::size_t nspins_ = 0;
  std::shared_ptr<DeviceContext> ctx_;
  spx_state* dev_ = nullptr;
  mutable std::vector<double> host_;
  mutable Fresh fresh_ = Fresh::Both;
};

}  // namespace spinsim
