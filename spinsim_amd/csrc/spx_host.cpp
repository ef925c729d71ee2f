// Host-side numerics of libspx: error state, Bessel sequences and the
// Chebyshev expansion (O(order) scalar work per expansion, so it stays on the
// host in FP64, as in the reference).
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "spx_common.hpp"

namespace spx {

namespace {
thread_local std::string g_error;
thread_local double g_conv[4] = {0, 0, 0, 0};
}  // namespace

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

int ok() { return SPX_OK; }

void set_convergence(double bound, double tau, double eps, size_t order) {
  g_conv[0] = bound;
  g_conv[1] = tau;
  g_conv[2] = eps;
  g_conv[3] = static_cast<double>(order);
}

const char* last_error() { return g_error.c_str(); }

void last_convergence(double* b, double* t, double* e, size_t* o) {
  if (b) *b = g_conv[0];
  if (t) *t = g_conv[1];
  if (e) *e = g_conv[2];
  if (o) *o = static_cast<size_t>(g_conv[3]);
}

// Miller's backward recurrence, normalised by the sum rule
// J_0 + 2 sum J_2k = 1 or I_0 + 2 sum I_k = e^x (src/core/Bessel.cpp:20-149).
// The start index max(count, ceil x) + 40 + 6 sqrt(top), rounded up to even,
// and the 1e250 rescale are the reference's, so the sequences agree to the
// last bit with it.
int bessel(int kind, double x, size_t count, double* out, size_t* nonzero) {
  if (count == 0) return fail(SPX_EINVAL, "Bessel sequence length must be positive");
  if (!(x >= 0.0)) return fail(SPX_EINVAL, "Bessel argument must be non-negative");
  for (size_t i = 0; i < count; ++i) out[i] = 0.0;
  if (x == 0.0) {
    out[0] = 1.0;
    if (nonzero) *nonzero = 1;
    return SPX_OK;
  }
  if (kind != 0 && x > 709.0) return fail(SPX_EOVERFLOW, "unscaled modified Bessel I overflows");

  const double top = std::max(static_cast<double>(count), std::ceil(x));
  const size_t start0 = static_cast<size_t>(top + (40.0 + 6.0 * std::sqrt(top)));
  const size_t start = start0 + (start0 % 2);

  const double sign = kind == 0 ? -1.0 : 1.0;
  double higher = 0.0, current = 1.0e-30, norm = 0.0;
  for (size_t k = start; k >= 1; --k) {
    const double lower = (2.0 * static_cast<double>(k) / x) * current + sign * higher;
    higher = current;
    current = lower;
    if (kind != 0 || k % 2 == 1) norm += 2.0 * current;
    if (std::abs(current) > 1.0e250) {
      current *= 1.0e-250;
      higher *= 1.0e-250;
      norm *= 1.0e-250;
      for (size_t i = 0; i < count; ++i) out[i] *= 1.0e-250;
    }
    if (k - 1 < count) out[k - 1] = current;
  }
  norm -= current;
  if (norm == 0.0 || !std::isfinite(norm))
    return fail(SPX_ERUNTIME, "Bessel recurrence failed to normalise");
  const double scale = (kind == 0 ? 1.0 : std::exp(x)) / norm;
  for (size_t i = 0; i < count; ++i) out[i] *= scale;
  if (nonzero) {
    size_t nz = count;
    while (nz > 0 && out[nz - 1] == 0.0) --nz;
    *nonzero = nz;
  }
  return SPX_OK;
}

// ChebyshevExpansion::build (src/core/Chebyshev.cpp:54-107): identity below
// alpha = 2 eps, one Bessel sequence of min(1.5 alpha + 64, maxOrder + 4)
// terms, truncation before the first run of three coefficients that are all
// <= eps and strictly decreasing, odd orders negated for tau < 0.
int cheb_build(int mode, double bound, double tau, double eps, size_t max_order,
               std::vector<double>& coeffs, double& eff_bound) {
  if (!(bound >= 0.0)) return fail(SPX_EINVAL, "spectral bound must be non-negative");
  if (!(eps > 0.0)) return fail(SPX_EINVAL, "epsilon must be positive");
  const double alpha = bound * std::abs(tau);
  if (alpha < 2.0 * eps) {
    coeffs.assign(1, 1.0);
    eff_bound = 1.0;
    return SPX_OK;
  }
  const size_t want = static_cast<size_t>(1.5 * alpha) + 64;
  const size_t count = std::min(want, max_order + 3 + 1);
  std::vector<double> seq(count);
  SPX_TRY(bessel(mode == 0 ? 0 : 1, alpha, count, seq.data(), nullptr));

  size_t tail = static_cast<size_t>(-1);
  if (count >= 4) {
    for (size_t i = 1; i + 2 < count; ++i) {
      const double a = std::abs(seq[i]), b = std::abs(seq[i + 1]), c = std::abs(seq[i + 2]);
      if (a <= eps && b <= eps && c <= eps && b < a && c < b) {
        tail = i;
        break;
      }
    }
  }
  if (tail == static_cast<size_t>(-1) || tail - 1 > max_order) {
    set_convergence(bound, tau, eps, count);
    char msg[320];
    std::snprintf(msg, sizeof msg,
                  "Chebyshev expansion did not converge: |H| <= %g, tau = %g, alpha = %g, "
                  "epsilon = %g, order tried = %zu. Shorten the time step or raise epsilon.",
                  bound, tau, alpha, eps, count);
    return fail(SPX_ECONVERGE, msg);
  }
  coeffs.assign(seq.begin(), seq.begin() + static_cast<std::ptrdiff_t>(tail));
  eff_bound = bound;
  if (tau < 0.0)
    for (size_t n = 1; n < coeffs.size(); n += 2) coeffs[n] = -coeffs[n];
  return SPX_OK;
}

}  // namespace spx
