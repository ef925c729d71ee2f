/*
 * CPU oracle: plain-C restatement of the reference spinsim hot path.
 * TEST INFRASTRUCTURE ONLY (see spinsim_oracle.h for what pins each part).
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#define _GNU_SOURCE
#include "spinsim_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SO_PI 3.141592653589793238462643383279502884
#define SO_NEGLIGIBLE 1.0e-14 /* src/core/Hamiltonian.cpp:12-16 */

/* ===================================================================== RNG */

/* SplitMix64 finaliser: include/spinsim/ensemble/Rng.hpp:48-53. */
uint64_t so_splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* The 64-bit Mersenne Twister as specified by the C++ standard
 * [rand.eng.mers] / [rand.predef] (mt19937_64): w=64 n=312 m=156 r=31. The
 * reference uses std::mt19937_64 (include/spinsim/ensemble/Rng.hpp:56). */
#define MT_N 312
#define MT_M 156
#define MT_A 0xB5026F5AA96619E9ULL
#define MT_UPPER 0xFFFFFFFF80000000ULL
#define MT_LOWER 0x000000007FFFFFFFULL

static void mt_seed(so_rng* r, uint64_t seed) {
  r->mt[0] = seed;
  for (size_t i = 1; i < MT_N; ++i) {
    const uint64_t prev = r->mt[i - 1];
    r->mt[i] = 6364136223846793005ULL * (prev ^ (prev >> 62)) + (uint64_t)i;
  }
  r->mti = MT_N;
}

static void mt_twist(so_rng* r) {
  for (size_t i = 0; i < MT_N; ++i) {
    const uint64_t y = (r->mt[i] & MT_UPPER) | (r->mt[(i + 1) % MT_N] & MT_LOWER);
    uint64_t v = r->mt[(i + MT_M) % MT_N] ^ (y >> 1);
    if (y & 1ULL) v ^= MT_A;
    r->mt[i] = v;
  }
  r->mti = 0;
}

uint64_t so_mt_next(so_rng* r) {
  if (r->mti >= MT_N) mt_twist(r);
  uint64_t y = r->mt[r->mti++];
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  y ^= (y >> 43);
  return y;
}

/* Rng(seed): engine_(mix(seed)) -- include/spinsim/ensemble/Rng.hpp:27. */
void so_rng_seed(so_rng* r, uint64_t seed) {
  mt_seed(r, so_splitmix(seed));
  r->spare = 0.0;
  r->has_spare = 0;
}

/* Rng::forRealization -- include/spinsim/ensemble/Rng.hpp:30-35. */
void so_rng_for_realization(so_rng* r, uint64_t base, uint64_t index) {
  so_rng_seed(r, so_splitmix(base ^ (so_splitmix(index + 1) * 0x9E3779B97F4A7C15ULL)));
}

/* Rng::uniform -- include/spinsim/ensemble/Rng.hpp:38-41. */
double so_rng_uniform(so_rng* r) { return (double)(so_mt_next(r) >> 11) * 0x1.0p-53; }

/* Rng::gaussian -- src/ensemble/Rng.cpp:8-23 (Box-Muller, spare cached). */
double so_rng_gaussian(so_rng* r) {
  if (r->has_spare) {
    r->has_spare = 0;
    return r->spare;
  }
  const double floor_ = 1.0e-15;
  const double u = floor_ + (1.0 - 2.0 * floor_) * so_rng_uniform(r);
  const double v = so_rng_uniform(r);
  const double radius = sqrt(-2.0 * log(u));
  const double phase = 2.0 * SO_PI * v;
  r->spare = radius * sin(phase);
  r->has_spare = 1;
  return radius * cos(phase);
}

void so_mt_raw(uint64_t seed, size_t n, uint64_t* out) {
  so_rng r;
  mt_seed(&r, seed);
  for (size_t i = 0; i < n; ++i) out[i] = so_mt_next(&r);
}

void so_rng_draws(uint64_t base, uint64_t index, size_t n, int gaussian, double* out) {
  so_rng r;
  so_rng_for_realization(&r, base, index);
  for (size_t i = 0; i < n; ++i) out[i] = gaussian ? so_rng_gaussian(&r) : so_rng_uniform(&r);
}

/* ================================================================== Bessel */

/* startIndex -- src/core/Bessel.cpp:20-26. */
static size_t bessel_start(double x, size_t count) {
  const double c = (double)count;
  const double cx = ceil(x);
  const double top = c > cx ? c : cx;
  const double margin = 40.0 + 6.0 * sqrt(top);
  const size_t start = (size_t)(top + margin);
  return start + (start % 2);
}

static size_t count_nonzero(const double* v, size_t n) {
  while (n > 0 && v[n - 1] == 0.0) --n;
  return n;
}

/* besselJ -- src/core/Bessel.cpp:45-96. */
int so_bessel_j(double x, size_t count, double* out, size_t* nonzero) {
  if (count == 0 || !(x >= 0.0)) return -1;
  for (size_t i = 0; i < count; ++i) out[i] = 0.0;
  if (x == 0.0) {
    out[0] = 1.0;
    if (nonzero) *nonzero = 1;
    return 0;
  }
  const size_t start = bessel_start(x, count);
  double higher = 0.0, current = 1.0e-30, normalization = 0.0;
  for (size_t k = start; k >= 1; --k) {
    const double lower = (2.0 * (double)k / x) * current - higher;
    higher = current;
    current = lower;
    if (k % 2 == 1) normalization += 2.0 * current;
    if (fabs(current) > 1.0e250) {
      current *= 1.0e-250;
      higher *= 1.0e-250;
      normalization *= 1.0e-250;
      for (size_t i = 0; i < count; ++i) out[i] *= 1.0e-250;
    }
    if (k - 1 < count) out[k - 1] = current;
  }
  normalization -= current;
  if (normalization == 0.0 || !isfinite(normalization)) return -3;
  const double scale = 1.0 / normalization;
  for (size_t i = 0; i < count; ++i) out[i] *= scale;
  if (nonzero) *nonzero = count_nonzero(out, count);
  return 0;
}

/* besselI -- src/core/Bessel.cpp:98-149. */
int so_bessel_i(double x, size_t count, double* out, size_t* nonzero) {
  if (count == 0 || !(x >= 0.0)) return -1;
  for (size_t i = 0; i < count; ++i) out[i] = 0.0;
  if (x == 0.0) {
    out[0] = 1.0;
    if (nonzero) *nonzero = 1;
    return 0;
  }
  if (x > 709.0) return -2;
  const size_t start = bessel_start(x, count);
  double higher = 0.0, current = 1.0e-30, normalization = 0.0;
  for (size_t k = start; k >= 1; --k) {
    const double lower = (2.0 * (double)k / x) * current + higher;
    higher = current;
    current = lower;
    normalization += 2.0 * current;
    if (fabs(current) > 1.0e250) {
      current *= 1.0e-250;
      higher *= 1.0e-250;
      normalization *= 1.0e-250;
      for (size_t i = 0; i < count; ++i) out[i] *= 1.0e-250;
    }
    if (k - 1 < count) out[k - 1] = current;
  }
  normalization -= current;
  if (normalization == 0.0 || !isfinite(normalization)) return -3;
  const double scale = exp(x) / normalization;
  for (size_t i = 0; i < count; ++i) out[i] *= scale;
  if (nonzero) *nonzero = count_nonzero(out, count);
  return 0;
}

/* =============================================================== Chebyshev */

/* findTailStart -- src/core/Chebyshev.cpp:24-35 (run of three). */
static size_t find_tail(const double* a, size_t n, double eps) {
  if (n < 4) return (size_t)-1;
  for (size_t i = 1; i + 2 < n; ++i) {
    const double x = fabs(a[i]), y = fabs(a[i + 1]), z = fabs(a[i + 2]);
    if (x <= eps && y <= eps && z <= eps && y < x && z < y) return i;
  }
  return (size_t)-1;
}

/* ChebyshevExpansion::build -- src/core/Chebyshev.cpp:54-107. */
int so_cheb_build(int mode, double bound, double tau, double eps, size_t max_order,
                  double* coeffs, size_t* order, double* eff_bound) {
  if (!(bound >= 0.0) || !(eps > 0.0)) return -1;
  const double alpha = bound * fabs(tau);
  if (alpha < 2.0 * eps) {
    coeffs[0] = 1.0;
    *order = 0;
    *eff_bound = 1.0;
    return 0;
  }
  const size_t want = (size_t)(1.5 * alpha) + 64;
  const size_t cap = max_order + 3 + 1;
  const size_t count = want < cap ? want : cap;
  double* seq = (double*)malloc(count * sizeof(double));
  if (!seq) return -1;
  const int rc = mode == 0 ? so_bessel_j(alpha, count, seq, NULL)
                           : so_bessel_i(alpha, count, seq, NULL);
  if (rc != 0) {
    free(seq);
    return rc;
  }
  const size_t tail = find_tail(seq, count, eps);
  if (tail == (size_t)-1 || tail - 1 > max_order) {
    free(seq);
    return -4;
  }
  for (size_t i = 0; i < tail; ++i) coeffs[i] = seq[i];
  free(seq);
  *order = tail - 1;
  *eff_bound = bound;
  if (tau < 0.0)
    for (size_t n = 1; n < tail; n += 2) coeffs[n] = -coeffs[n];
  return 0;
}

/* ============================================================ StateVector */

void so_state_zero(double* psi, size_t n) { memset(psi, 0, 2 * n * sizeof(double)); }

void so_state_scale(double* psi, size_t n, double f) {
  for (size_t k = 0; k < 2 * n; ++k) psi[k] *= f;
}

void so_state_negate(double* psi, size_t n) {
  for (size_t k = 0; k < 2 * n; ++k) psi[k] = -psi[k];
}

/* StateVector::addScaled -- src/core/StateVector.cpp:21-32. */
void so_state_add_scaled(double* dst, size_t n, double cr, double ci, const double* src) {
  for (size_t k = 0; k < 2 * n; k += 2) {
    const double orr = src[k], oi = src[k + 1];
    dst[k] += cr * orr - ci * oi;
    dst[k + 1] += cr * oi + ci * orr;
  }
}

/* StateVector::norm -- src/core/StateVector.cpp:43-49 (serial sum). */
double so_state_norm(const double* psi, size_t n) {
  double sum = 0.0;
  for (size_t k = 0; k < 2 * n; ++k) sum += psi[k] * psi[k];
  return sqrt(sum);
}

int so_state_normalize(double* psi, size_t n, double* norm_out) {
  const double nrm = so_state_norm(psi, n);
  if (norm_out) *norm_out = nrm;
  if (nrm == 0.0) return -5;
  so_state_scale(psi, n, 1.0 / nrm);
  return 0;
}

/* StateVector::innerProduct -- src/core/StateVector.cpp:58-70: <other|self>. */
void so_state_inner(const double* a, const double* b, size_t n, double* re, double* im) {
  double real = 0.0, imag = 0.0;
  for (size_t k = 0; k < 2 * n; k += 2) {
    real += b[k] * a[k] + b[k + 1] * a[k + 1];
    imag += b[k] * a[k + 1] - b[k + 1] * a[k];
  }
  *re = real;
  *im = imag;
}

/* ================================================================= kernels */

/* axpySigmaX/Y/Z -- src/core/SpinOperators.cpp:90-137; pair walk :30-39. */
void so_axpy_sigma(int axis, double c, const double* i, size_t L, size_t spin, double* o) {
  const size_t n = (size_t)1 << L, stride = (size_t)1 << spin, block = stride << 1;
  for (size_t base = 0; base < n; base += block) {
    for (size_t lo = base; lo < base + stride; ++lo) {
      const size_t l = 2 * lo, h = 2 * (lo + stride);
      if (axis == 0) {
        o[h] += c * i[l];
        o[h + 1] += c * i[l + 1];
        o[l] += c * i[h];
        o[l + 1] += c * i[h + 1];
      } else if (axis == 1) {
        o[h] += c * i[l + 1];
        o[h + 1] -= c * i[l];
        o[l] -= c * i[h + 1];
        o[l + 1] += c * i[h];
      } else {
        o[h] += c * i[h];
        o[h + 1] += c * i[h + 1];
        o[l] -= c * i[l];
        o[l + 1] -= c * i[l + 1];
      }
    }
  }
}

/* applySigmaX/Y/Z -- src/core/SpinOperators.cpp:139-177. */
void so_apply_sigma(int axis, const double* i, size_t L, size_t spin, double* o) {
  const size_t n = (size_t)1 << L, stride = (size_t)1 << spin, block = stride << 1;
  for (size_t base = 0; base < n; base += block) {
    for (size_t lo = base; lo < base + stride; ++lo) {
      const size_t l = 2 * lo, h = 2 * (lo + stride);
      if (axis == 0) {
        o[h] = i[l];
        o[h + 1] = i[l + 1];
        o[l] = i[h];
        o[l + 1] = i[h + 1];
      } else if (axis == 1) {
        o[h] = i[l + 1];
        o[h + 1] = -i[l];
        o[l] = -i[h + 1];
        o[l + 1] = i[h];
      } else {
        o[h] = i[h];
        o[h + 1] = i[h + 1];
        o[l] = -i[l];
        o[l + 1] = -i[l + 1];
      }
    }
  }
}

/* axpyCoupling -- src/core/SpinOperators.cpp:179-209 (four-corner walk). */
void so_axpy_coupling(double jx, double jy, double jz, const double* i, size_t L,
                      size_t lo_s, size_t hi_s, double* o) {
  const size_t n = (size_t)1 << L;
  const size_t hiMask = (size_t)1 << hi_s, loMask = (size_t)1 << lo_s;
  const size_t hiBlock = hiMask << 1, loBlock = loMask << 1;
  const double flipBoth = jx - jy, flipOne = jx + jy;
  for (size_t outer = 0; outer < n; outer += hiBlock)
    for (size_t mid = outer; mid < outer + hiMask; mid += loBlock)
      for (size_t dd = mid; dd < mid + loMask; ++dd) {
        const size_t d = 2 * dd, a = 2 * (dd + hiMask), b = 2 * (dd + loMask),
                     u = 2 * (dd + hiMask + loMask);
        o[u] += flipBoth * i[d] + jz * i[u];
        o[u + 1] += flipBoth * i[d + 1] + jz * i[u + 1];
        o[d] += flipBoth * i[u] + jz * i[d];
        o[d + 1] += flipBoth * i[u + 1] + jz * i[d + 1];
        o[a] += flipOne * i[b] - jz * i[a];
        o[a + 1] += flipOne * i[b + 1] - jz * i[a + 1];
        o[b] += flipOne * i[a] - jz * i[b];
        o[b + 1] += flipOne * i[a + 1] - jz * i[b + 1];
      }
}

/* axpyDiagonal -- src/core/SpinOperators.cpp:211-226. */
void so_axpy_diagonal(double c, const double* d, const double* i, size_t L, double* o) {
  const size_t n = (size_t)1 << L;
  for (size_t k = 0; k < n; ++k) {
    const double f = c * d[k];
    o[2 * k] += f * i[2 * k];
    o[2 * k + 1] += f * i[2 * k + 1];
  }
}

/* axpyTransverse -- src/core/SpinOperators.cpp:228-252. */
void so_axpy_transverse(double cx, double cy, const double* i, size_t L, size_t spin,
                        double* o) {
  const size_t n = (size_t)1 << L, stride = (size_t)1 << spin, block = stride << 1;
  for (size_t base = 0; base < n; base += block)
    for (size_t lo = base; lo < base + stride; ++lo) {
      const size_t l = 2 * lo, h = 2 * (lo + stride);
      const double ar = i[l], ai = i[l + 1], br = i[h], bi = i[h + 1];
      o[l] += cx * br - cy * bi;
      o[l + 1] += cx * bi + cy * br;
      o[h] += cx * ar + cy * ai;
      o[h + 1] += cx * ai - cy * ar;
    }
}

/* Two-corner channel walk -- src/core/SpinOperators.cpp:54-69. */
static void channel_walk(double c, const double* i, size_t L, size_t lo_s, size_t hi_s,
                         size_t off1, size_t off2, double* o) {
  const size_t n = (size_t)1 << L;
  const size_t hiMask = (size_t)1 << hi_s, loMask = (size_t)1 << lo_s;
  const size_t hiBlock = hiMask << 1, loBlock = loMask << 1;
  for (size_t outer = 0; outer < n; outer += hiBlock)
    for (size_t mid = outer; mid < outer + hiMask; mid += loBlock)
      for (size_t dd = mid; dd < mid + loMask; ++dd) {
        const size_t a = 2 * (dd + off1), b = 2 * (dd + off2);
        o[a] += c * i[b];
        o[a + 1] += c * i[b + 1];
        o[b] += c * i[a];
        o[b + 1] += c * i[a + 1];
      }
}

/* axpyFlipFlop -- src/core/SpinOperators.cpp:254-276. */
void so_axpy_flip_flop(double c, const double* i, size_t L, size_t lo_s, size_t hi_s,
                       double* o) {
  channel_walk(c, i, L, lo_s, hi_s, (size_t)1 << hi_s, (size_t)1 << lo_s, o);
}

/* axpyFlipBoth -- src/core/SpinOperators.cpp:278-291. */
void so_axpy_flip_both(double c, const double* i, size_t L, size_t lo_s, size_t hi_s,
                       double* o) {
  channel_walk(c, i, L, lo_s, hi_s, 0, ((size_t)1 << hi_s) + ((size_t)1 << lo_s), o);
}

/* ============================================================ Hamiltonian */

so_ham* so_ham_new(size_t L) {
  if (L == 0 || L > 30) return NULL;
  so_ham* h = (so_ham*)calloc(1, sizeof(so_ham));
  h->L = L;
  h->hx = (double*)calloc(L, sizeof(double));
  h->hy = (double*)calloc(L, sizeof(double));
  h->hz = (double*)calloc(L, sizeof(double));
  return h;
}

void so_ham_free(so_ham* h) {
  if (!h) return;
  free(h->hx); free(h->hy); free(h->hz);
  free(h->lo); free(h->hi); free(h->jx); free(h->jy); free(h->jz);
  free(h->diag);
  free(h);
}

int so_ham_set_field(so_ham* h, size_t s, double x, double y, double z) {
  if (s >= h->L) return -6;
  h->hx[s] = x; h->hy[s] = y; h->hz[s] = z;
  h->diag_valid = 0;
  return 0;
}

/* setCoupling -- src/core/Hamiltonian.cpp:38-52 (overwrites a repeated pair). */
int so_ham_set_coupling(so_ham* h, size_t a, size_t b, double jx, double jy, double jz) {
  if (a == b) return -1;
  const size_t lo = a < b ? a : b, hi = a < b ? b : a;
  if (hi >= h->L) return -6;
  for (size_t p = 0; p < h->npairs; ++p)
    if (h->lo[p] == lo && h->hi[p] == hi) {
      h->jx[p] = jx; h->jy[p] = jy; h->jz[p] = jz;
      h->diag_valid = 0;
      return 0;
    }
  if (h->npairs == h->cap) {
    const size_t cap = h->cap ? 2 * h->cap : 64;
    h->lo = (uint32_t*)realloc(h->lo, cap * sizeof(uint32_t));
    h->hi = (uint32_t*)realloc(h->hi, cap * sizeof(uint32_t));
    h->jx = (double*)realloc(h->jx, cap * sizeof(double));
    h->jy = (double*)realloc(h->jy, cap * sizeof(double));
    h->jz = (double*)realloc(h->jz, cap * sizeof(double));
    h->cap = cap;
  }
  const size_t p = h->npairs++;
  h->lo[p] = (uint32_t)lo; h->hi[p] = (uint32_t)hi;
  h->jx[p] = jx; h->jy[p] = jy; h->jz[p] = jz;
  h->diag_valid = 0;
  return 0;
}

/* spectralBound -- src/core/Hamiltonian.cpp:82-93. */
double so_ham_bound(const so_ham* h) {
  double bound = 0.0;
  for (size_t s = 0; s < h->L; ++s) bound += 0.5 * (fabs(h->hx[s]) + fabs(h->hy[s]) + fabs(h->hz[s]));
  for (size_t p = 0; p < h->npairs; ++p)
    bound += 0.25 * (fabs(h->jx[p]) + fabs(h->jy[p]) + fabs(h->jz[p]));
  return bound;
}

/* rebuildDiagonal -- src/core/Hamiltonian.cpp:54-80. */
void so_ham_rebuild_diag(so_ham* h) {
  const size_t n = (size_t)1 << h->L;
  if (!h->diag) h->diag = (double*)malloc(n * sizeof(double));
  for (size_t k = 0; k < n; ++k) h->diag[k] = 0.0;
  for (size_t s = 0; s < h->L; ++s) {
    const double c = 0.5 * h->hz[s];
    if (fabs(c) <= SO_NEGLIGIBLE) continue;
    const size_t mask = (size_t)1 << s;
    for (size_t k = 0; k < n; ++k) h->diag[k] += (k & mask) ? c : -c;
  }
  for (size_t p = 0; p < h->npairs; ++p) {
    const double c = 0.25 * h->jz[p];
    if (fabs(c) <= SO_NEGLIGIBLE) continue;
    const size_t hm = (size_t)1 << h->hi[p], lm = (size_t)1 << h->lo[p];
    for (size_t k = 0; k < n; ++k) {
      const int aligned = ((k & hm) != 0) == ((k & lm) != 0);
      h->diag[k] += aligned ? c : -c;
    }
  }
  h->diag_valid = 1;
}

const double* so_ham_diag(so_ham* h) {
  if (!h->diag_valid) so_ham_rebuild_diag(h);
  return h->diag;
}

/* accumulate -- src/core/Hamiltonian.cpp:95-127. */
void so_ham_accumulate(so_ham* h, double scale, const double* in, double* out) {
  if (!h->diag_valid) so_ham_rebuild_diag(h);
  so_axpy_diagonal(scale, h->diag, in, h->L, out);
  for (size_t s = 0; s < h->L; ++s) {
    const double cx = scale * 0.5 * h->hx[s];
    const double cy = scale * 0.5 * h->hy[s];
    if (fabs(cx) > SO_NEGLIGIBLE || fabs(cy) > SO_NEGLIGIBLE)
      so_axpy_transverse(cx, cy, in, h->L, s, out);
  }
  for (size_t p = 0; p < h->npairs; ++p) {
    const double quarter = scale * 0.25;
    const double ff = quarter * (h->jx[p] + h->jy[p]);
    const double fb = quarter * (h->jx[p] - h->jy[p]);
    if (fabs(ff) > SO_NEGLIGIBLE) so_axpy_flip_flop(ff, in, h->L, h->lo[p], h->hi[p], out);
    if (fabs(fb) > SO_NEGLIGIBLE) so_axpy_flip_both(fb, in, h->L, h->lo[p], h->hi[p], out);
  }
}

void so_ham_apply_physical(so_ham* h, const double* in, double* out) {
  so_state_zero(out, (size_t)1 << h->L);
  so_ham_accumulate(h, 1.0, in, out);
}

int so_ham_accumulate_twice_normalized(so_ham* h, double bound, const double* in, double* out) {
  if (!(bound > 0.0)) return -7;
  so_ham_accumulate(h, -2.0 / bound, in, out);
  return 0;
}

int so_ham_apply_twice_normalized(so_ham* h, double bound, const double* in, double* out) {
  so_state_zero(out, (size_t)1 << h->L);
  return so_ham_accumulate_twice_normalized(h, bound, in, out);
}

/* ============================================================= Propagator */

/* applyOneStep -- src/core/Propagator.cpp:46-73, with TermWeight :14-36. */
void so_apply_one_step(so_ham* h, const double* a, size_t order, double bound, int mode,
                       double* psi, double* cur, double* prev) {
  const size_t n = (size_t)1 << h->L;
  memcpy(cur, psi, 2 * n * sizeof(double));
  so_state_scale(psi, n, a[0]);
  so_state_zero(prev, n);
  double wre = mode == 0 ? 0.0 : -1.0, wim = mode == 0 ? -1.0 : 0.0;
  for (size_t k = 1; k <= order; ++k) {
    so_state_negate(prev, n);
    so_ham_accumulate(h, -2.0 / bound, cur, prev);
    if (k == 1) so_state_scale(prev, n, 0.5);
    double* t = prev; prev = cur; cur = t;   /* previous_.swap(current_) */
    so_state_add_scaled(psi, n, 2.0 * a[k] * wre, 2.0 * a[k] * wim, cur);
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

/* Propagator::evolve -- src/core/Propagator.cpp:75-100. */
int so_evolve(so_ham* h, int mode, double* psi, double tau, size_t nsteps, int autonorm,
              double eps, size_t* order_out) {
  if (!(eps > 0.0)) return -1;
  if (nsteps == 0) return 0;
  double* coeffs = (double*)malloc((6000 + 4) * sizeof(double));
  size_t order = 0;
  double bound = 0.0;
  int rc = so_cheb_build(mode, so_ham_bound(h), tau, eps, 6000, coeffs, &order, &bound);
  if (rc != 0) {
    free(coeffs);
    return rc;
  }
  if (order_out) *order_out = order;
  if (order == 0) {
    free(coeffs);
    return 0;
  }
  const size_t n = (size_t)1 << h->L;
  double* cur = (double*)malloc(2 * n * sizeof(double));
  double* prev = (double*)malloc(2 * n * sizeof(double));
  for (size_t s = 0; s < nsteps; ++s) {
    so_apply_one_step(h, coeffs, order, bound, mode, psi, cur, prev);
    if (autonorm) {
      rc = so_state_normalize(psi, n, NULL);
      if (rc) break;
    }
  }
  free(cur);
  free(prev);
  free(coeffs);
  return rc;
}

/* =================================================================== Gates */

/* Rotation2x2::make + SpinRotation::apply -- src/core/Gate.cpp:20-85. */
int so_rotate(double* d, size_t L, double ax, double ay, double az, double angle,
              const uint32_t* targets, size_t ntargets) {
  const double length = sqrt(ax * ax + ay * ay + az * az);
  if (!(length > 0.0)) return -1;
  if (ntargets == 0) return -1;
  const double nx = ax / length, ny = ay / length, nz = az / length;
  const double c = cos(0.5 * angle), s = sin(0.5 * angle);
  const double llr = c, lli = s * nz, lhr = s * ny, lhi = -s * nx;
  const double hlr = -s * ny, hli = -s * nx, hhr = c, hhi = -s * nz;
  const size_t n = (size_t)1 << L;
  for (size_t t = 0; t < ntargets; ++t) {
    const size_t spin = targets[t];
    if (spin >= L) return -6;
    const size_t stride = (size_t)1 << spin, block = stride << 1;
    for (size_t base = 0; base < n; base += block)
      for (size_t lo = base; lo < base + stride; ++lo) {
        const size_t hi = lo + stride;
        const double ar = d[2 * lo], ai = d[2 * lo + 1], br = d[2 * hi], bi = d[2 * hi + 1];
        /* std::complex products (a+bi)(c+di) = (ac - bd) + (ad + bc)i */
        const double p1r = llr * ar - lli * ai, p1i = llr * ai + lli * ar;
        const double p2r = lhr * br - lhi * bi, p2i = lhr * bi + lhi * br;
        const double q1r = hlr * ar - hli * ai, q1i = hlr * ai + hli * ar;
        const double q2r = hhr * br - hhi * bi, q2i = hhr * bi + hhi * br;
        d[2 * lo] = p1r + p2r;
        d[2 * lo + 1] = p1i + p2i;
        d[2 * hi] = q1r + q2r;
        d[2 * hi + 1] = q1i + q2i;
      }
  }
  return 0;
}

/* TwoSpinGate::apply -- src/core/Gate.cpp:90-121; u16 row-major complex,
 * local basis (dd, lo-up, hi-up, uu). */
int so_two_spin_gate(double* d, size_t L, const double* u, size_t lo_s, size_t hi_s) {
  if (lo_s == hi_s) return -1;
  if (lo_s > hi_s) { const size_t t = lo_s; lo_s = hi_s; hi_s = t; }
  if (hi_s >= L) return -6;
  const size_t n = (size_t)1 << L;
  const size_t hiMask = (size_t)1 << hi_s, loMask = (size_t)1 << lo_s;
  for (size_t outer = 0; outer < n; outer += hiMask << 1)
    for (size_t mid = outer; mid < outer + hiMask; mid += loMask << 1)
      for (size_t dd = mid; dd < mid + loMask; ++dd) {
        const size_t idx[4] = {dd, dd + loMask, dd + hiMask, dd + hiMask + loMask};
        double inr[4], ini[4];
        for (int k = 0; k < 4; ++k) { inr[k] = d[2 * idx[k]]; ini[k] = d[2 * idx[k] + 1]; }
        for (int r = 0; r < 4; ++r) {
          double accr = 0.0, acci = 0.0;
          for (int c = 0; c < 4; ++c) {
            const double mr = u[2 * (r * 4 + c)], mi = u[2 * (r * 4 + c) + 1];
            accr += mr * inr[c] - mi * ini[c];
            acci += mr * ini[c] + mi * inr[c];
          }
          d[2 * idx[r]] = accr;
          d[2 * idx[r] + 1] = acci;
        }
      }
  return 0;
}

/* ============================================================= Observables */

/* applyTotalSpin -- src/core/Observable.cpp:13-22. */
void so_apply_total_spin(int axis, const double* in, size_t L, double* out) {
  so_state_zero(out, (size_t)1 << L);
  for (size_t s = 0; s < L; ++s) so_axpy_sigma(axis, 1.0, in, L, s, out);
}

/* TotalSpinCorrelator::measure -- src/core/Observable.cpp:66-74. */
void so_corr(int axis, const double* state, const double* ref, size_t L, double* scratch,
             double* re, double* im) {
  so_apply_total_spin(axis, state, L, scratch);
  so_state_inner(scratch, ref, (size_t)1 << L, re, im);
}

/* SingleSpinMagnetization::measure -- src/core/Observable.cpp:84-97. */
double so_mag(int axis, size_t spin, const double* state, size_t L, double* scratch) {
  double re, im;
  so_apply_sigma(axis, state, L, spin, scratch);
  so_state_inner(scratch, state, (size_t)1 << L, &re, &im);
  return re;
}

/* Energy::measure -- src/core/Observable.cpp:112-120. */
double so_energy(so_ham* h, const double* state, double* scratch) {
  double re, im;
  so_ham_apply_physical(h, state, scratch);
  so_state_inner(state, scratch, (size_t)1 << h->L, &re, &im);
  return re;
}

/* ================================================================ Ensemble */

/* prepareInitialState -- src/ensemble/Ensemble.cpp:20-45. */
int so_prepare_initial_state(double* psi, size_t L, int kind, so_rng* r, size_t basis) {
  const size_t n = (size_t)1 << L;
  if (kind == 1 || kind == 2 || kind == 3) {
    const size_t idx = kind == 1 ? n - 1 : kind == 2 ? 0 : basis;
    if (idx >= n) return -6;
    so_state_zero(psi, n);
    psi[2 * idx] = 1.0;
    return 0;
  }
  for (size_t k = 0; k < n; ++k) {
    psi[2 * k] = so_rng_gaussian(r);
    psi[2 * k + 1] = so_rng_gaussian(r);
  }
  return so_state_normalize(psi, n, NULL);
}

/* prepareReferenceState -- src/ensemble/Ensemble.cpp:47-58. */
void so_prepare_reference(const double* psi, size_t L, int axis, double* ref) {
  so_apply_total_spin(axis, psi, L, ref);
}

/* generateDipolarLattice2D -- src/ensemble/Ensemble.cpp:60-104. */
int so_dipolar_lattice(size_t L, so_rng* r, double scale, double min_sep, int angular,
                       uint32_t* plo, uint32_t* phi, double* pjx, double* pjy, double* pjz) {
  if (L == 0) return -1;
  const double hw = 0.5 * sqrt((double)L);
  double* x = (double*)calloc(L, sizeof(double));
  double* y = (double*)calloc(L, sizeof(double));
  size_t p = 0;
  for (size_t i = 1; i < L; ++i) {
    double px = 0.0, py = 0.0;
    int placed = 0;
    while (!placed) {
      px = hw * (2.0 * so_rng_uniform(r) - 1.0);
      py = hw * (2.0 * so_rng_uniform(r) - 1.0);
      placed = 1;
      for (size_t j = 0; j < i; ++j)
        if (fabs(px - x[j]) < min_sep && fabs(py - y[j]) < min_sep) {
          placed = 0;
          break;
        }
    }
    x[i] = px;
    y[i] = py;
    for (size_t j = 0; j < i; ++j) {
      const double dx = px - x[j], dy = py - y[j];
      const double r2 = dx * dx + dy * dy;
      double jz = scale / (r2 * sqrt(r2));
      if (angular) {
        /* secular dipolar (1 - 3 cos^2 theta), theta from the in-plane x field */
        const double cos2 = (dx * dx) / r2;
        jz = jz * (1.0 - 3.0 * cos2);
      }
      plo[p] = (uint32_t)j; phi[p] = (uint32_t)i;
      pjx[p] = -0.5 * jz; pjy[p] = -0.5 * jz; pjz[p] = jz;
      ++p;
    }
  }
  free(x);
  free(y);
  return 0;
}

/* applyFieldDisorder -- src/ensemble/Ensemble.cpp:106-112. */
void so_field_disorder(so_ham* h, so_rng* r, double sigma) {
  for (size_t s = 0; s < h->L; ++s) {
    const double z = h->hz[s] + sigma * so_rng_gaussian(r);
    so_ham_set_field(h, s, h->hx[s], h->hy[s], z);
  }
}

/* ================================================================ Program */

typedef struct {
  int kind;
  double tau;
  size_t nsteps;
  int mode, autonorm;
  double* fields;   /* [3L] or NULL */
  size_t nf;
  double axis[3], angle;
  uint32_t* targets;
  size_t ntargets;
  size_t count;
  size_t match;     /* REPEAT <-> END partner */
} so_op;

struct so_program {
  so_op* ops;
  size_t n, cap;
  size_t stack[64];
  size_t depth;
  int error;
};

so_program* so_program_new(void) { return (so_program*)calloc(1, sizeof(so_program)); }

void so_program_free(so_program* p) {
  if (!p) return;
  for (size_t i = 0; i < p->n; ++i) { free(p->ops[i].fields); free(p->ops[i].targets); }
  free(p->ops);
  free(p);
}

static so_op* push_op(so_program* p, int kind) {
  if (p->n == p->cap) {
    p->cap = p->cap ? 2 * p->cap : 16;
    p->ops = (so_op*)realloc(p->ops, p->cap * sizeof(so_op));
  }
  so_op* op = &p->ops[p->n++];
  memset(op, 0, sizeof(*op));
  op->kind = kind;
  return op;
}

int so_program_evolve(so_program* p, double tau, size_t nsteps, int mode, int autonorm,
                      const double* fields, size_t L) {
  so_op* op = push_op(p, SO_OP_EVOLVE);
  op->tau = tau; op->nsteps = nsteps; op->mode = mode; op->autonorm = autonorm;
  if (fields) {
    op->fields = (double*)malloc(3 * L * sizeof(double));
    memcpy(op->fields, fields, 3 * L * sizeof(double));
    op->nf = L;
  }
  return 0;
}

int so_program_pulse(so_program* p, double ax, double ay, double az, double angle,
                     const uint32_t* targets, size_t ntargets) {
  so_op* op = push_op(p, SO_OP_PULSE);
  op->axis[0] = ax; op->axis[1] = ay; op->axis[2] = az; op->angle = angle;
  op->targets = (uint32_t*)malloc((ntargets ? ntargets : 1) * sizeof(uint32_t));
  if (ntargets) memcpy(op->targets, targets, ntargets * sizeof(uint32_t));
  op->ntargets = ntargets;
  return 0;
}

int so_program_measure(so_program* p) { push_op(p, SO_OP_MEASURE); return 0; }

int so_program_repeat(so_program* p, size_t count) {
  if (p->depth >= 64) return -1;
  so_op* op = push_op(p, SO_OP_REPEAT);
  op->count = count;
  p->stack[p->depth++] = p->n - 1;
  return 0;
}

int so_program_end(so_program* p) {
  if (p->depth == 0) return -1;
  const size_t open = p->stack[--p->depth];
  so_op* op = push_op(p, SO_OP_END);
  op->match = open;
  p->ops[open].match = p->n - 1;
  return 0;
}

/* Sequence::measurementCount -- src/sequence/Sequence.cpp:89-101. */
static size_t count_range(const so_program* p, size_t b, size_t e) {
  size_t total = 0;
  for (size_t i = b; i < e; ++i) {
    const so_op* op = &p->ops[i];
    if (op->kind == SO_OP_MEASURE) ++total;
    else if (op->kind == SO_OP_REPEAT) {
      total += op->count * count_range(p, i + 1, op->match);
      i = op->match;
    }
  }
  return total;
}

size_t so_program_measurements(const so_program* p) { return count_range(p, 0, p->n); }

/* ============================================================ Realisation */

size_t so_config_width(const so_config* c) {
  size_t w = 0;
  for (size_t o = 0; o < c->nobs; ++o) {
    const int code = c->obs_codes[o];
    w += (code <= 2 || code == SO_OBS_NORM) ? 2 : 1;
  }
  return w;
}

typedef struct {
  const so_config* c;
  so_ham* h;
  double *state, *ref, *scratch, *coeffs, *cur, *prev;
  double time;
  double* times;
  double* values;
  size_t row, width;
} run_ctx;

static int measure_row(run_ctx* r) {
  const so_config* c = r->c;
  const size_t L = c->L, n = (size_t)1 << L;
  double* out = r->values + r->row * r->width;
  size_t at = 0;
  for (size_t o = 0; o < c->nobs; ++o) {
    const int code = c->obs_codes[o];
    if (code <= 2) {
      so_corr(code, r->state, r->ref, L, r->scratch, &out[at], &out[at + 1]);
      at += 2;
    } else if (code <= 5) {
      const int spin = c->obs_spins[o];
      if (spin < 0 || (size_t)spin >= L) return -6;
      out[at++] = so_mag(code - 3, (size_t)spin, r->state, L, r->scratch);
    } else if (code == SO_OBS_NORM) {
      out[at++] = so_state_norm(r->state, n);
      out[at++] = so_state_norm(r->ref, n);
    } else {
      out[at++] = so_energy(r->h, r->state, r->scratch);
    }
  }
  r->times[r->row] = r->time;
  ++r->row;
  return 0;
}

/* One evolve of a single vector with a prebuilt expansion (evolve's loop). */
static int evolve_vec(run_ctx* r, double* v, const double* coeffs, size_t order,
                      double bound, int mode, size_t nsteps, int autonorm) {
  const size_t n = (size_t)1 << r->c->L;
  for (size_t s = 0; s < nsteps; ++s) {
    so_apply_one_step(r->h, coeffs, order, bound, mode, v, r->cur, r->prev);
    if (autonorm && so_state_normalize(v, n, NULL)) return -5;
  }
  return 0;
}

/* runStep / runEvolve -- src/sequence/Sequence.cpp:35-87. */
static int run_range(run_ctx* r, size_t b, size_t e) {
  const so_program* p = r->c->program;
  const size_t L = r->c->L;
  for (size_t i = b; i < e; ++i) {
    const so_op* op = &p->ops[i];
    int rc = 0;
    if (op->kind == SO_OP_EVOLVE) {
      double saved[3 * 30];
      if (op->fields) {
        if (op->nf != L) return -1;
        for (size_t s = 0; s < L; ++s) {
          saved[3 * s] = r->h->hx[s]; saved[3 * s + 1] = r->h->hy[s]; saved[3 * s + 2] = r->h->hz[s];
          so_ham_set_field(r->h, s, op->fields[3 * s], op->fields[3 * s + 1], op->fields[3 * s + 2]);
        }
      }
      if (op->nsteps > 0) {
        size_t order = 0;
        double bound = 0.0;
        rc = so_cheb_build(op->mode, so_ham_bound(r->h), op->tau, r->c->epsilon, 6000,
                           r->coeffs, &order, &bound);
        if (rc == 0 && order > 0) {
          rc = evolve_vec(r, r->state, r->coeffs, order, bound, op->mode, op->nsteps, op->autonorm);
          if (rc == 0)
            rc = evolve_vec(r, r->ref, r->coeffs, order, bound, op->mode, op->nsteps, op->autonorm);
        }
      }
      if (op->fields)
        for (size_t s = 0; s < L; ++s)
          so_ham_set_field(r->h, s, saved[3 * s], saved[3 * s + 1], saved[3 * s + 2]);
      if (rc) return rc;
      r->time += op->tau * (double)op->nsteps;
    } else if (op->kind == SO_OP_PULSE) {
      rc = so_rotate(r->state, L, op->axis[0], op->axis[1], op->axis[2], op->angle, op->targets, op->ntargets);
      if (rc == 0)
        rc = so_rotate(r->ref, L, op->axis[0], op->axis[1], op->axis[2], op->angle, op->targets, op->ntargets);
      if (rc) return rc;
    } else if (op->kind == SO_OP_MEASURE) {
      rc = measure_row(r);
      if (rc) return rc;
    } else if (op->kind == SO_OP_REPEAT) {
      for (size_t k = 0; k < op->count; ++k) {
        rc = run_range(r, i + 1, op->match);
        if (rc) return rc;
      }
      i = op->match;
    }
  }
  return 0;
}

/* runRealization -- src/io/Experiment.cpp:10-55 (draw order preserved). */
int so_run_realization(const so_config* c, uint64_t index, double* times, double* values) {
  const size_t L = c->L, n = (size_t)1 << L;
  so_rng rng;
  so_rng_for_realization(&rng, c->seed, index);
  so_ham* h = so_ham_new(L);
  if (!h) return -1;
  for (size_t s = 0; s < L; ++s)
    if (c->base_fields)
      so_ham_set_field(h, s, c->base_fields[3 * s], c->base_fields[3 * s + 1], c->base_fields[3 * s + 2]);
  int rc = 0;
  if (c->generate_geometry) {
    const size_t np = L * (L - 1) / 2;
    uint32_t* lo = (uint32_t*)malloc((np + 1) * sizeof(uint32_t));
    uint32_t* hi = (uint32_t*)malloc((np + 1) * sizeof(uint32_t));
    double* jx = (double*)malloc((np + 1) * sizeof(double));
    double* jy = (double*)malloc((np + 1) * sizeof(double));
    double* jz = (double*)malloc((np + 1) * sizeof(double));
    so_dipolar_lattice(L, &rng, c->dipolar_scale, c->min_separation, c->angular, lo, hi, jx, jy, jz);
    for (size_t p = 0; p < np; ++p) so_ham_set_coupling(h, lo[p], hi[p], jx[p], jy[p], jz[p]);
    free(lo); free(hi); free(jx); free(jy); free(jz);
  }
  if (c->field_disorder != 0.0) so_field_disorder(h, &rng, c->field_disorder);

  run_ctx r;
  memset(&r, 0, sizeof(r));
  r.c = c;
  r.h = h;
  r.state = (double*)calloc(2 * n, sizeof(double));
  r.ref = (double*)calloc(2 * n, sizeof(double));
  r.scratch = (double*)calloc(2 * n, sizeof(double));
  r.cur = (double*)calloc(2 * n, sizeof(double));
  r.prev = (double*)calloc(2 * n, sizeof(double));
  r.coeffs = (double*)malloc((6000 + 4) * sizeof(double));
  r.times = times;
  r.values = values;
  r.width = so_config_width(c);
  rc = so_prepare_initial_state(r.state, L, c->initial_state, &rng, c->basis_index);
  if (rc == 0) {
    so_prepare_reference(r.state, L, c->reference_axis, r.ref);
    rc = run_range(&r, 0, c->program->n);
  }
  free(r.state); free(r.ref); free(r.scratch); free(r.cur); free(r.prev); free(r.coeffs);
  so_ham_free(h);
  return rc;
}

/* ============================================================== Ensemble run */

typedef struct {
  const so_config* c;
  size_t n, points, width;
  double* times_all;
  double* per;
  size_t next;
  int failure;
  pthread_mutex_t mu;
} pool_t;

static void* pool_worker(void* arg) {
  pool_t* pl = (pool_t*)arg;
  for (;;) {
    pthread_mutex_lock(&pl->mu);
    const size_t idx = pl->next;
    if (idx >= pl->n) { pthread_mutex_unlock(&pl->mu); return NULL; }
    pl->next++;
    pthread_mutex_unlock(&pl->mu);
    const int rc = so_run_realization(pl->c, idx, pl->times_all + idx * pl->points,
                                      pl->per + idx * pl->points * pl->width);
    if (rc) {
      pthread_mutex_lock(&pl->mu);
      if (!pl->failure) pl->failure = rc;
      pl->next = pl->n;  /* stop handing out work (EnsembleRunner.cpp:61-64) */
      pthread_mutex_unlock(&pl->mu);
      return NULL;
    }
  }
}

/* LocalRunner::run -- src/ensemble/EnsembleRunner.cpp:27-88. */
int so_run_ensemble(const so_config* c, size_t n, size_t threads, double* times, double* per,
                    double* mean, double* sd, double* se) {
  if (n == 0) return -1;
  const size_t points = so_program_measurements(c->program), width = so_config_width(c);
  pool_t pl;
  memset(&pl, 0, sizeof(pl));
  pl.c = c; pl.n = n; pl.points = points; pl.width = width;
  pl.times_all = (double*)malloc((n * points + 1) * sizeof(double));
  pl.per = per ? per : (double*)malloc((n * points * width + 1) * sizeof(double));
  pthread_mutex_init(&pl.mu, NULL);
  if (threads == 0) threads = 1;
  if (threads > n) threads = n;
  pthread_t* th = (pthread_t*)malloc(threads * sizeof(pthread_t));
  for (size_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, pool_worker, &pl);
  for (size_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&pl.mu);
  int rc = pl.failure;
  if (rc == 0) {
    for (size_t p = 0; p < points; ++p) times[p] = pl.times_all[p];
    so_acc* acc = so_acc_new(points * width);
    for (size_t i = 0; i < n; ++i) so_acc_add(acc, pl.per + i * points * width);
    so_acc_stats(acc, mean, sd, se);
    so_acc_free(acc);
  }
  free(pl.times_all);
  if (!per) free(pl.per);
  return rc;
}

/* ============================================================ Accumulator */

so_acc* so_acc_new(size_t width) {
  so_acc* a = (so_acc*)calloc(1, sizeof(so_acc));
  a->width = width;
  a->mean = (double*)calloc(width + 1, sizeof(double));
  a->m2 = (double*)calloc(width + 1, sizeof(double));
  return a;
}

void so_acc_free(so_acc* a) {
  if (!a) return;
  free(a->mean);
  free(a->m2);
  free(a);
}

/* Accumulator::add -- src/ensemble/Ensemble.cpp:116-127 (Welford). */
int so_acc_add(so_acc* a, const double* x) {
  ++a->count;
  const double n = (double)a->count;
  for (size_t k = 0; k < a->width; ++k) {
    const double delta = x[k] - a->mean[k];
    a->mean[k] += delta / n;
    a->m2[k] += delta * (x[k] - a->mean[k]);
  }
  return 0;
}

/* Accumulator::merge -- src/ensemble/Ensemble.cpp:129-148 (Chan). */
int so_acc_merge(so_acc* a, const so_acc* b) {
  if (a->width != b->width) return -1;
  if (b->count == 0) return 0;
  if (a->count == 0) {
    a->count = b->count;
    memcpy(a->mean, b->mean, a->width * sizeof(double));
    memcpy(a->m2, b->m2, a->width * sizeof(double));
    return 0;
  }
  const double na = (double)a->count, nb = (double)b->count, total = na + nb;
  for (size_t k = 0; k < a->width; ++k) {
    const double delta = b->mean[k] - a->mean[k];
    a->mean[k] += delta * nb / total;
    a->m2[k] += b->m2[k] + delta * delta * na * nb / total;
  }
  a->count += b->count;
  return 0;
}

/* standardDeviation / standardError -- src/ensemble/Ensemble.cpp:150-166. */
void so_acc_stats(const so_acc* a, double* mean, double* sd, double* se) {
  for (size_t k = 0; k < a->width; ++k) {
    mean[k] = a->mean[k];
    sd[k] = 0.0;
    se[k] = 0.0;
  }
  if (a->count < 2) return;
  const double denom = (double)(a->count - 1);
  const double scale = 1.0 / sqrt((double)a->count);
  for (size_t k = 0; k < a->width; ++k) {
    sd[k] = sqrt(a->m2[k] / denom);
    se[k] = sd[k] * scale;
  }
}

/* ============================================================ CPU timing */

typedef struct {
  so_ham* h;
  size_t orders;
  double bound;
} time_arg;

static void* time_worker(void* arg) {
  time_arg* t = (time_arg*)arg;
  const size_t n = (size_t)1 << t->h->L;
  double* psi = (double*)malloc(2 * n * sizeof(double));
  double* cur = (double*)malloc(2 * n * sizeof(double));
  double* prev = (double*)malloc(2 * n * sizeof(double));
  for (size_t k = 0; k < 2 * n; ++k) { psi[k] = 1e-3 * (double)(k % 7); cur[k] = psi[k]; prev[k] = 0.0; }
  for (size_t o = 0; o < t->orders; ++o) {
    so_state_negate(prev, n);
    so_ham_accumulate(t->h, -2.0 / t->bound, cur, prev);
    double* sw = prev; prev = cur; cur = sw;
    so_state_add_scaled(psi, n, 0.1, -0.2, cur);
  }
  free(psi); free(cur); free(prev);
  return NULL;
}

double so_time_orders(so_ham* h, size_t orders, size_t threads) {
  so_ham_rebuild_diag(h);
  const double bound = so_ham_bound(h);
  if (threads == 0) threads = 1;
  pthread_t* th = (pthread_t*)malloc(threads * sizeof(pthread_t));
  time_arg arg = {h, orders, bound > 0 ? bound : 1.0};
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (size_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, time_worker, &arg);
  for (size_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
