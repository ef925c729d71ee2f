#pragma once
// Pauli kernels (include/spinsim/core/SpinOperators.hpp:22-45 of the
// reference), on the device. `in` and `out` must be distinct.

#include "spinsim/core/Bits.hpp"
#include "spinsim/core/StateVector.hpp"

namespace spinsim {

void axpySigmaX(double c, const StateVector& in, SpinIndex spin, StateVector& out);
void axpySigmaY(double c, const StateVector& in, SpinIndex spin, StateVector& out);
void axpySigmaZ(double c, const StateVector& in, SpinIndex spin, StateVector& out);
void applySigmaX(const StateVector& in, SpinIndex spin, StateVector& out);
void applySigmaY(const StateVector& in, SpinIndex spin, StateVector& out);
void applySigmaZ(const StateVector& in, SpinIndex spin, StateVector& out);

}  // namespace spinsim
