#pragma once
// Measurement observables (include/spinsim/core/Observable.hpp:16-134 of the
// reference), evaluated by fixed-order device reductions.

#include <memory>
#include <span>
#include <string>
#include <string_view>
#include <vector>

#include "spinsim/core/Bits.hpp"
#include "spinsim/core/Hamiltonian.hpp"
#include "spinsim/core/StateVector.hpp"

namespace spinsim {

enum class Axis { X, Y, Z };

[[nodiscard]] Axis parseAxis(std::string_view name);
[[nodiscard]] std::string_view axisName(Axis axis);

struct MeasureContext {
  double time = 0.0;
  const StateVector* state = nullptr;
  const StateVector* reference = nullptr;
  const Hamiltonian* hamiltonian = nullptr;
  std::size_t subsystemSize = 0;
};

class Observable {
 public:
  virtual ~Observable() = default;
  [[nodiscard]] virtual std::string name() const = 0;
  [[nodiscard]] virtual std::vector<std::string> columns() const = 0;
  virtual void measure(const MeasureContext& ctx, std::span<double> out) const = 0;
  [[nodiscard]] std::size_t width() const { return columns().size(); }
  /// The libspx observable code (corr 0-2, mag 3-5, norm 6, energy 7) and
  /// its spin, for the batched runner.
  [[nodiscard]] virtual int code() const = 0;
  [[nodiscard]] virtual int codeSpin() const { return 0; }
};

/// <reference| sum_i sigma_axis(i) |state> (re, im): the echo.
class TotalSpinCorrelator final : public Observable {
 public:
  explicit TotalSpinCorrelator(Axis axis) : axis_(axis) {}
  [[nodiscard]] std::string name() const override;
  [[nodiscard]] std::vector<std::string> columns() const override;
  void measure(const MeasureContext& ctx, std::span<double> out) const override;
  [[nodiscard]] int code() const override { return static_cast<int>(axis_); }

 private:
  Axis axis_;
};

/// Re <state| sigma_axis(spin) |state>.
class SingleSpinMagnetization final : public Observable {
 public:
  SingleSpinMagnetization(SpinIndex spin, Axis axis) : spin_(spin), axis_(axis) {}
  [[nodiscard]] std::string name() const override;
  [[nodiscard]] std::vector<std::string> columns() const override;
  void measure(const MeasureContext& ctx, std::span<double> out) const override;
  [[nodiscard]] int code() const override { return 3 + static_cast<int>(axis_); }
  [[nodiscard]] int codeSpin() const override { return static_cast<int>(spin_); }

 private:
  SpinIndex spin_;
  Axis axis_;
};

/// (||state||, ||reference||).
class StateNorms final : public Observable {
 public:
  [[nodiscard]] std::string name() const override { return "norm"; }
  [[nodiscard]] std::vector<std::string> columns() const override;
  void measure(const MeasureContext& ctx, std::span<double> out) const override;
  [[nodiscard]] int code() const override { return 6; }
};

/// Re <state|H|state>.
class Energy final : public Observable {
 public:
  [[nodiscard]] std::string name() const override { return "energy"; }
  [[nodiscard]] std::vector<std::string> columns() const override;
  void measure(const MeasureContext& ctx, std::span<double> out) const override;
  [[nodiscard]] int code() const override { return 7; }
};

[[nodiscard]] std::unique_ptr<Observable> makeObservable(std::string_view spec);

}  // namespace spinsim
