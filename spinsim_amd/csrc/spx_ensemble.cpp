// Disorder draws of the ensemble loop, on the host: the per-realisation
// random stream (reference include/spinsim/ensemble/Rng.hpp:25-58 and
// src/ensemble/Rng.cpp:8-23), the 2D dipolar lattice
// (src/ensemble/Ensemble.cpp:60-104) and the Gaussian Sz disorder
// (src/ensemble/Ensemble.cpp:106-112), drawn in runRealization's order
// (src/io/Experiment.cpp:10-55) so that realisation (seed, index) is the
// reference's. Many realisations are drawn on a pool of host threads, each
// into its own slot, so the output does not depend on the thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <numbers>
#include <random>
#include <thread>
#include <vector>

#include "spx_common.hpp"
#include "spx_ensemble.hpp"

// The reference builds with -ffp-contract=off (CMakeLists.txt:16-22); the
// draws and the normalisation here must round as it does, so no multiply-add
// is fused in this file.
#pragma clang fp contract(off)

namespace spx {
namespace {

constexpr size_t kMaxPlacementAttempts = size_t(1) << 24;

uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Stream {
  std::mt19937_64 engine;
  double spare = 0.0;
  bool has_spare = false;

  Stream(uint64_t seed, uint64_t index)
      : engine(splitmix(splitmix(seed ^ (splitmix(index + 1) * 0x9E3779B97F4A7C15ULL)))) {}

  double uniform() { return static_cast<double>(engine() >> 11) * 0x1.0p-53; }

  double gaussian() {
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    constexpr double kFloor = 1.0e-15;
    const double u = kFloor + (1.0 - 2.0 * kFloor) * uniform();
    const double v = uniform();
    const double radius = std::sqrt(-2.0 * std::log(u));
    const double phase = 2.0 * std::numbers::pi * v;
    // one sincos call for both, as GCC compiles the reference's separate
    // std::sin and std::cos of one argument; glibc's sincos and its sin / cos
    // can differ in the last bit
    double sn = 0.0, cs = 0.0;
    ::sincos(phase, &sn, &cs);
    spare = radius * sn;
    has_spare = true;
    return radius * cs;
  }
};

}  // namespace

namespace {

// generateDipolarLattice2D then applyFieldDisorder on one stream
// (src/io/Experiment.cpp:17-27 draw order).
int draw_lattice(const DipolarSpec& s, Stream& r, uint32_t* lo, uint32_t* hi, double* jz, double* hz) {
  const size_t L = s.nspins;
  const double half = 0.5 * std::sqrt(static_cast<double>(L));
  double x[kMaxSpins] = {}, y[kMaxSpins] = {};
  size_t p = 0;
  for (size_t i = 1; i < L; ++i) {
    double px = 0.0, py = 0.0;
    size_t attempts = 0;
    for (bool placed = false; !placed;) {
      if (++attempts > kMaxPlacementAttempts) return SPX_ECONVERGE;
      px = half * (2.0 * r.uniform() - 1.0);
      py = half * (2.0 * r.uniform() - 1.0);
      placed = true;
      for (size_t j = 0; j < i && placed; ++j)
        placed = !(std::abs(px - x[j]) < s.min_separation && std::abs(py - y[j]) < s.min_separation);
    }
    x[i] = px;
    y[i] = py;
    for (size_t j = 0; j < i; ++j, ++p) {
      const double dx = px - x[j], dy = py - y[j];
      const double r2 = dx * dx + dy * dy;
      double c = s.scale / (r2 * std::sqrt(r2));
      if (s.angular) c = c * (1.0 - 3.0 * ((dx * dx) / r2));
      lo[p] = static_cast<uint32_t>(j);
      hi[p] = static_cast<uint32_t>(i);
      jz[p] = c;
    }
  }
  for (size_t i = 0; i < L; ++i) hz[i] = s.disorder != 0.0 ? s.disorder * r.gaussian() : 0.0;
  return SPX_OK;
}

}  // namespace

int draw_dipolar(const DipolarSpec& s, uint64_t index, uint32_t* lo, uint32_t* hi, double* jz, double* hz) {
  Stream r(s.seed, index);
  return draw_lattice(s, r, lo, hi, jz, hz);
}

// The same draws, then prepareInitialState's typical state on the same
// stream (src/ensemble/Ensemble.cpp:20-45): a Gaussian real then imaginary
// part per amplitude, normalised as StateVector::normalize does (one serial
// sum of squares over the interleaved values, then a scale by 1 / norm,
// src/core/StateVector.cpp:43-56).
int draw_realisation(const DipolarSpec& s, uint64_t index, uint32_t* lo, uint32_t* hi, double* jz, double* hz,
                     double* psi) {
  Stream r(s.seed, index);
  SPX_TRY(draw_lattice(s, r, lo, hi, jz, hz));
  const size_t n = static_cast<size_t>(2) << s.nspins;
  for (size_t k = 0; k < n; ++k) psi[k] = r.gaussian();
  double sum = 0.0;
  for (size_t k = 0; k < n; ++k) sum += psi[k] * psi[k];
  const double norm = std::sqrt(sum);
  if (norm == 0.0) return SPX_EDOMAIN;
  const double f = 1.0 / norm;
  for (size_t k = 0; k < n; ++k) psi[k] *= f;
  return SPX_OK;
}

int check_spec(const DipolarSpec& s) {
  if (s.nspins == 0 || s.nspins > kMaxSpins) return fail(SPX_EINVAL, "nspins must be in [1, 30]");
  if (!(s.min_separation >= 0.0)) return fail(SPX_EINVAL, "min_separation must be >= 0");
  return SPX_OK;
}

int draw_graphs(const DipolarSpec& s, uint64_t first, size_t count, double* J, double* h, float* Jf, float* hf,
                int threads) {
  const size_t L = s.nspins, np = L * (L - 1) / 2;
  if (threads <= 0) threads = static_cast<int>(std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency())));
  threads = static_cast<int>(std::min<size_t>(static_cast<size_t>(threads), std::max<size_t>(1, count / 64)));
  std::atomic<size_t> next{0};
  std::atomic<int> status{SPX_OK};
  std::atomic<uint64_t> bad{0};
  auto work = [&]() {
    std::vector<uint32_t> lo(np), hi(np);
    std::vector<double> jz(np), hz(L);
    for (;;) {
      const size_t g = next.fetch_add(1);
      if (g >= count || status.load() != SPX_OK) return;
      const int rc = draw_dipolar(s, first + g, lo.data(), hi.data(), jz.data(), hz.data());
      if (rc != SPX_OK) {
        status.store(rc);
        bad.store(first + g);
        return;
      }
      if (J) {
        double* Jg = J + g * L * L;
        std::fill(Jg, Jg + L * L, 0.0);
        for (size_t p = 0; p < np; ++p) Jg[lo[p] * L + hi[p]] = Jg[hi[p] * L + lo[p]] = jz[p];
      }
      if (Jf) {
        float* Jg = Jf + g * L * L;
        std::fill(Jg, Jg + L * L, 0.f);
        for (size_t p = 0; p < np; ++p) Jg[lo[p] * L + hi[p]] = Jg[hi[p] * L + lo[p]] = static_cast<float>(jz[p]);
      }
      for (size_t i = 0; i < L; ++i) {
        if (h) h[g * L + i] = hz[i];
        if (hf) hf[g * L + i] = static_cast<float>(hz[i]);
      }
    }
  };
  if (threads == 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(work);
    for (auto& t : pool) t.join();
  }
  if (status.load() != SPX_OK)
    return fail(status.load(), "could not place the spins of realisation " + std::to_string(bad.load()) +
                                   " at the minimum separation");
  return SPX_OK;
}

}  // namespace spx

using namespace spx;

extern "C" {

int spx_dipolar_realisation(uint64_t seed, size_t index, size_t nspins, double scale, double min_separation,
                            int angular, double disorder, uint32_t* lo, uint32_t* hi, double* jx, double* jy,
                            double* jz, double* hz) {
  const DipolarSpec s{seed, nspins, scale, min_separation, angular != 0, disorder};
  SPX_TRY(check_spec(s));
  if (!lo || !hi || !jx || !jy || !jz || !hz) return fail(SPX_EINVAL, "null output");
  const int rc = draw_dipolar(s, index, lo, hi, jz, hz);
  if (rc != SPX_OK) return fail(rc, "could not place the spins at the minimum separation");
  for (size_t p = 0; p < nspins * (nspins - 1) / 2; ++p) jx[p] = jy[p] = -0.5 * jz[p];
  return SPX_OK;
}

int spx_dipolar_state(uint64_t seed, size_t index, size_t nspins, double scale, double min_separation,
                      int angular, double disorder, double* psi) {
  const DipolarSpec s{seed, nspins, scale, min_separation, angular != 0, disorder};
  SPX_TRY(check_spec(s));
  if (!psi) return fail(SPX_EINVAL, "null output");
  const size_t np = nspins * (nspins - 1) / 2;
  std::vector<uint32_t> lo(np + 1), hi(np + 1);
  std::vector<double> jz(np + 1), hz(nspins);
  const int rc = draw_realisation(s, index, lo.data(), hi.data(), jz.data(), hz.data(), psi);
  if (rc == SPX_EDOMAIN) return fail(rc, "cannot normalize a zero state");
  if (rc != SPX_OK) return fail(rc, "could not place the spins at the minimum separation");
  return SPX_OK;
}

int spx_dipolar_graphs(uint64_t seed, size_t first, size_t count, size_t nspins, double scale,
                       double min_separation, int angular, double disorder, double* J, double* h) {
  const DipolarSpec s{seed, nspins, scale, min_separation, angular != 0, disorder};
  SPX_TRY(check_spec(s));
  if (!J || !h) return fail(SPX_EINVAL, "null output");
  return draw_graphs(s, first, count, J, h, nullptr, nullptr, 0);
}

}  // extern "C"
