#pragma once
// Spin indexing (include/spinsim/core/Bits.hpp:13-55 of the reference):
// spin s owns bit s of a basis index, set bit = up.

#include <cassert>
#include <cstddef>
#include <cstdint>

namespace spinsim {

inline constexpr std::size_t kMaxSpins = 30;

using SpinIndex = std::size_t;

[[nodiscard]] constexpr std::uint64_t spinMask(SpinIndex spin) noexcept {
  return std::uint64_t{1} << spin;
}

[[nodiscard]] constexpr std::size_t stateCount(std::size_t nspins) noexcept {
  return std::size_t{1} << nspins;
}

/// Unordered pair of distinct spins, stored low < high.
class SpinPair {
 public:
  constexpr SpinPair(SpinIndex a, SpinIndex b) noexcept
      : low_(a < b ? a : b), high_(a < b ? b : a) {
    assert(a != b && "a spin cannot couple to itself");
  }
  [[nodiscard]] constexpr SpinIndex low() const noexcept { return low_; }
  [[nodiscard]] constexpr SpinIndex high() const noexcept { return high_; }
  friend constexpr bool operator==(SpinPair, SpinPair) noexcept = default;

 private:
  SpinIndex low_;
  SpinIndex high_;
};

}  // namespace spinsim
