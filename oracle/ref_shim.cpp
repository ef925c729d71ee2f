// C-ABI shim over the reference's own sources, for validating the oracle.
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) against
// the reference files that compile here unmodified:
//   src/ensemble/Rng.cpp, src/core/{Bessel,StateVector,SpinOperators,Hamiltonian}.cpp
// The rest of the reference path (Chebyshev.cpp, Gate.cpp, Observable.cpp,
// Ensemble.cpp, EnsembleRunner.cpp, Experiment.cpp) includes <format>, which
// libstdc++ 11 in this image lacks, so it is not built; see the oracle header
// for how those parts are pinned instead.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <thread>
#include <vector>

#include "spinsim/core/Bessel.hpp"
#include "spinsim/core/Hamiltonian.hpp"
#include "spinsim/core/SpinOperators.hpp"
#include "spinsim/core/StateVector.hpp"
#include "spinsim/ensemble/Rng.hpp"

using namespace spinsim;

namespace {

StateVector load(std::size_t L, const double* v) {
  StateVector s(L);
  std::memcpy(s.data(), v, 2 * s.size() * sizeof(double));
  return s;
}

void store(const StateVector& s, double* v) {
  std::memcpy(v, s.data(), 2 * s.size() * sizeof(double));
}

Hamiltonian build(std::size_t L, const double* fields, std::size_t np, const uint32_t* lo,
                  const uint32_t* hi, const double* jx, const double* jy, const double* jz) {
  Hamiltonian h(L);
  for (std::size_t s = 0; s < L; ++s)
    h.setField(s, Field{fields[3 * s], fields[3 * s + 1], fields[3 * s + 2]});
  for (std::size_t p = 0; p < np; ++p) h.setCoupling(SpinPair{lo[p], hi[p]}, jx[p], jy[p], jz[p]);
  return h;
}

}  // namespace

extern "C" {

void ref_rng_draws(uint64_t base, uint64_t index, std::size_t n, int gaussian, double* out) {
  Rng r = Rng::forRealization(base, index);
  for (std::size_t i = 0; i < n; ++i) out[i] = gaussian ? r.gaussian() : r.uniform();
}

void ref_rng_seeded(uint64_t seed, std::size_t n, int gaussian, double* out) {
  Rng r(seed);
  for (std::size_t i = 0; i < n; ++i) out[i] = gaussian ? r.gaussian() : r.uniform();
}

int ref_bessel(int kind, double x, std::size_t count, double* out, std::size_t* nonzero) {
  try {
    const BesselSequence s = kind == 0 ? besselJ(x, count) : besselI(x, count);
    std::memcpy(out, s.values.data(), count * sizeof(double));
    *nonzero = s.nonzero;
    return 0;
  } catch (const std::invalid_argument&) {
    return -1;
  } catch (const std::overflow_error&) {
    return -2;
  } catch (...) {
    return -3;
  }
}

// mode 0: applyPhysical, 1: applyTwiceNormalized, 2: accumulateTwiceNormalized
int ref_ham_apply(int mode, std::size_t L, const double* fields, std::size_t np,
                  const uint32_t* lo, const uint32_t* hi, const double* jx, const double* jy,
                  const double* jz, double bound, const double* in, double* out) {
  try {
    const Hamiltonian h = build(L, fields, np, lo, hi, jx, jy, jz);
    const StateVector vin = load(L, in);
    StateVector vout = load(L, out);
    if (mode == 0) h.applyPhysical(vin, vout);
    else if (mode == 1) h.applyTwiceNormalized(bound, vin, vout);
    else h.accumulateTwiceNormalized(bound, vin, vout);
    store(vout, out);
    return 0;
  } catch (...) {
    return -7;
  }
}

double ref_ham_bound(std::size_t L, const double* fields, std::size_t np, const uint32_t* lo,
                     const uint32_t* hi, const double* jx, const double* jy, const double* jz) {
  return build(L, fields, np, lo, hi, jx, jy, jz).spectralBound();
}

// kernel: 0..2 axpySigmaX/Y/Z(c0, spin), 3..5 applySigmaX/Y/Z(spin),
// 6 axpyCoupling(c0,c1,c2, lo=spin, hi=spin2), 7 axpyTransverse(c0,c1, spin),
// 8 axpyFlipFlop(c0, lo, hi), 9 axpyFlipBoth(c0, lo, hi), 10 axpyDiagonal(c0, diag=aux)
void ref_kernel(int kernel, std::size_t L, double c0, double c1, double c2, std::size_t spin,
                std::size_t spin2, const double* aux, const double* in, double* out) {
  const StateVector vin = load(L, in);
  StateVector vout = load(L, out);
  switch (kernel) {
    case 0: axpySigmaX(c0, vin, spin, vout); break;
    case 1: axpySigmaY(c0, vin, spin, vout); break;
    case 2: axpySigmaZ(c0, vin, spin, vout); break;
    case 3: applySigmaX(vin, spin, vout); break;
    case 4: applySigmaY(vin, spin, vout); break;
    case 5: applySigmaZ(vin, spin, vout); break;
    case 6: axpyCoupling(c0, c1, c2, vin, SpinPair{spin, spin2}, vout); break;
    case 7: axpyTransverse(c0, c1, vin, spin, vout); break;
    case 8: axpyFlipFlop(c0, vin, SpinPair{spin, spin2}, vout); break;
    case 9: axpyFlipBoth(c0, vin, SpinPair{spin, spin2}, vout); break;
    case 10: {
      std::vector<double> d(aux, aux + vin.size());
      axpyDiagonal(c0, d, vin, vout);
      break;
    }
    default: break;
  }
  store(vout, out);
}

void ref_state_ops(std::size_t L, const double* a, const double* b, double cr, double ci,
                   double* added, double* norm, double* ip_re, double* ip_im) {
  const StateVector va = load(L, a);
  const StateVector vb = load(L, b);
  StateVector sum = va;
  sum.addScaled(cr, ci, vb);
  store(sum, added);
  *norm = va.norm();
  const auto [re, im] = va.innerProduct(vb);
  *ip_re = re;
  *ip_im = im;
}

// One Chebyshev step whose every state operation is the reference's own
// StateVector / Hamiltonian code, driven with precomputed coefficients in the
// sequence of Propagator::applyOneStep (src/core/Propagator.cpp:46-73).
int ref_one_step(int mode, std::size_t L, const double* fields, std::size_t np,
                 const uint32_t* lo, const uint32_t* hi, const double* jx, const double* jy,
                 const double* jz, const double* a, std::size_t order, double bound,
                 std::size_t nsteps, double* psi_io) {
  try {
    const Hamiltonian h = build(L, fields, np, lo, hi, jx, jy, jz);
    StateVector psi = load(L, psi_io);
    StateVector current(L), previous(L);
    for (std::size_t step = 0; step < nsteps; ++step) {
      current = psi;
      psi.scale(a[0]);
      previous.setZero();
      double wre = mode == 0 ? 0.0 : -1.0, wim = mode == 0 ? -1.0 : 0.0;
      for (std::size_t n = 1; n <= order; ++n) {
        previous.negate();
        h.accumulateTwiceNormalized(bound, current, previous);
        if (n == 1) previous.scale(0.5);
        previous.swap(current);
        psi.addScaled(2.0 * a[n] * wre, 2.0 * a[n] * wim, current);
        if (mode == 0) {
          const double pr = wre;
          wre = wim;
          wim = -pr;
        } else {
          wre = -wre;
          wim = -wim;
        }
      }
    }
    store(psi, psi_io);
    return 0;
  } catch (...) {
    return -7;
  }
}

// CPU baseline: `orders` Chebyshev orders of the reference's own kernels
// (negate + accumulateTwiceNormalized + addScaled, Propagator.cpp:58-72) on
// `threads` independent states of one Hamiltonian, like LocalRunner's
// realisation-parallel pool. Returns wall seconds.
double ref_time_orders(std::size_t L, const double* fields, std::size_t np, const uint32_t* lo,
                       const uint32_t* hi, const double* jx, const double* jy, const double* jz,
                       std::size_t orders, std::size_t threads) {
  const Hamiltonian proto = build(L, fields, np, lo, hi, jx, jy, jz);
  const double bound = proto.spectralBound();
  std::vector<Hamiltonian> hs(threads == 0 ? 1 : threads, proto);
  // warm the lazily built diagonal outside the clock
  for (auto& h : hs) {
    StateVector a(L), b(L);
    h.accumulateTwiceNormalized(bound, a, b);
  }
  const auto t0 = std::chrono::steady_clock::now();
  {
    std::vector<std::jthread> pool;
    for (std::size_t t = 0; t < hs.size(); ++t) {
      pool.emplace_back([&, t] {
        StateVector psi(L), cur(L), prev(L);
        for (std::size_t k = 0; k < psi.size(); ++k) {
          psi.re(k) = 1e-3 * static_cast<double>(k % 7);
          cur.re(k) = psi.re(k);
        }
        for (std::size_t o = 0; o < orders; ++o) {
          prev.negate();
          hs[t].accumulateTwiceNormalized(bound, cur, prev);
          prev.swap(cur);
          psi.addScaled(0.1, -0.2, cur);
        }
      });
    }
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
