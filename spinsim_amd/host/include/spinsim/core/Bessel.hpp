#pragma once
// Bessel sequences (include/spinsim/core/Bessel.hpp:9-45 of the reference),
// computed by libspx's host numerics.

#include <cstddef>
#include <vector>

namespace spinsim {

struct BesselSequence {
  std::vector<double> values;
  std::size_t nonzero = 0;
};

/// J_0(x)..J_{count-1}(x); std::invalid_argument for x < 0 or count == 0.
[[nodiscard]] BesselSequence besselJ(double x, std::size_t count);

/// I_0(x)..I_{count-1}(x), unscaled; std::overflow_error above x = 709.
[[nodiscard]] BesselSequence besselI(double x, std::size_t count);

}  // namespace spinsim
