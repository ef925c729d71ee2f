// Test inputs of the reference's own test-suite, reproduced with the same
// standard-library engines: randomState (tests/DenseReference.cpp:204-214)
// fills re then im of each amplitude from std::normal_distribution over
// std::mt19937_64(seed) and normalises. TEST INFRASTRUCTURE ONLY; the
// values depend on libstdc++'s normal_distribution, which is what the
// reference's CI (g++ / libstdc++) uses too.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <random>

extern "C" void std_random_state(std::size_t nspins, std::uint64_t seed, double* out) {
  const std::size_t n = std::size_t{1} << nspins;
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> gauss(0.0, 1.0);
  for (std::size_t k = 0; k < n; ++k) {
    out[2 * k] = gauss(rng);
    out[2 * k + 1] = gauss(rng);
  }
  double sum = 0.0;
  for (std::size_t k = 0; k < 2 * n; ++k) sum += out[k] * out[k];
  const double inv = 1.0 / std::sqrt(sum);
  for (std::size_t k = 0; k < 2 * n; ++k) out[k] *= inv;
}
