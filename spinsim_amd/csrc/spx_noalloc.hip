// Propagator::evolve for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops): spx_prop_evolve_noalloc.
//
// spx_prop_evolve keeps its two Chebyshev scratch vectors in the Propagator
// (allocated when it is created) and reads the Hamiltonian's device tables
// (built, and allocated, on first use). This variant allocates nothing: the
// scratch vectors T_{n-1}, T_{n-2} and the sigma_z diagonal live in one
// file-scope device array, `scratch`, sized for kNoAllocMaxSpins spins, and
// the Hamiltonian's weights travel as kernel arguments, read from its host
// side (fields and couplings as set through spx_ham_set_*). Every call on a
// device shares that device's array: calls that overlap on one device race.
//
// The arithmetic is the reference's Propagator::applyOneStep
// (src/core/Propagator.cpp:58-72): T_1 = Hhat T_0, T_n = 2 Hhat T_{n-1} -
// T_{n-2}, psi = a_0 psi + sum_n 2 a_n w_n T_n with w_n = (-i)^n in real
// time and (-1)^n in imaginary time, Hhat = -H / E, the terms of
// Hamiltonian::accumulate (src/core/Hamiltonian.cpp:95-127) with the same
// 1e-14 skip rule; one fused launch per order, one amplitude per thread.
#include <mutex>
#include <vector>

#include "spx_common.hpp"

namespace spx {

constexpr int kNoAllocMaxSpins = 16;
constexpr size_t kNoAllocAmps = size_t(1) << kNoAllocMaxSpins;

// T_{n-1} / T_{n-2} (2^L each) and the diagonal (2^L doubles = 2^(L-1)
// double2): 2.5 MiB at 16 spins.
static __device__ __align__(256) double2 scratch[2 * kNoAllocAmps + kNoAllocAmps / 2];

namespace {

constexpr int kNoAllocMaxPairs = kNoAllocMaxSpins * (kNoAllocMaxSpins - 1) / 2;
constexpr int kNoAllocMaxTerms = kNoAllocMaxPairs + kNoAllocMaxSpins;
constexpr int kNoAllocBlock = 256;

// The scratch array's device address, looked up once per device with that
// device current (each device holds its own copy of the array).
std::mutex g_scratch_mutex;
double2* g_scratch_ptr[64] = {};

int scratch_on(int device, double2** out) {
  if (device < 0 || device >= 64) return fail(SPX_ERANGE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SPX_EDEVICE, std::string("hipGetSymbolAddress of the no-alloc scratch: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double2*>(p);
  }
  *out = g_scratch_ptr[device];
  return SPX_OK;
}

// Off-diagonal terms at the step's scale: pairs (mask of both spins; anti
// and aligned weights) then transverse fields (mask of the spin; cx, cy).
struct NaTerms {
  int nterms;
  int npairs;
  double sdiag;                       // the diagonal's scale
  uint32_t mask[kNoAllocMaxTerms];
  double w0[kNoAllocMaxTerms];        // pair: anti;    field: cx
  double w1[kNoAllocMaxTerms];        // pair: aligned; field: cy
};

// The sigma_z diagonal's terms (rebuildDiagonal, src/core/Hamiltonian.cpp:54-80):
// fields 0.5 hz on one spin, pairs 0.25 jz on two.
struct NaDiag {
  int nfield, npair;
  uint32_t spin[kNoAllocMaxSpins];
  double cf[kNoAllocMaxSpins];
  uint32_t mask[kNoAllocMaxPairs];
  double cp[kNoAllocMaxPairs];
};

struct NaStep {
  double a0, c1r, c1i, cr, ci;
};

enum { NA_ORDER1 = 0, NA_ORDER2 = 1, NA_ORDERN = 2 };

__global__ __launch_bounds__(kNoAllocBlock) void k_na_diag(NaDiag t, double* __restrict__ d, uint32_t n) {
  const uint32_t k = blockIdx.x * kNoAllocBlock + threadIdx.x;
  if (k >= n) return;
  double e = 0.0;
  for (int s = 0; s < t.nfield; ++s) e += ((k >> t.spin[s]) & 1u) ? t.cf[s] : -t.cf[s];
  for (int p = 0; p < t.npair; ++p) {
    const uint32_t b = k & t.mask[p];
    e += (b == 0u || b == t.mask[p]) ? t.cp[p] : -t.cp[p];
  }
  d[k] = e;
}

// One fused Chebyshev order (the epilogues of the streaming order kernels):
//   NA_ORDER1: out = Hhat in                    (T_1; psi holds T_0)
//   NA_ORDER2: out = 2 Hhat in - psi            (T_2), psi = a0 psi + c1 in + c2 T_2
//   NA_ORDERN: out = 2 Hhat in - out            (T_n over T_{n-2}), psi += c T_n
template <int KIND>
__global__ __launch_bounds__(kNoAllocBlock) void k_na_order(NaTerms t, NaStep c, const double2* __restrict__ in,
                                                             double2* __restrict__ out, double2* __restrict__ psi,
                                                             const double* __restrict__ diag, uint32_t n) {
  const uint32_t k = blockIdx.x * kNoAllocBlock + threadIdx.x;
  if (k >= n) return;
  const double2 x = in[k];
  const double f = t.sdiag * diag[k];
  double2 acc = make_double2(f * x.x, f * x.y);
  for (int i = 0; i < t.npairs; ++i) {
    const uint32_t m = t.mask[i];
    const double w = (__popc(k & m) == 1) ? t.w0[i] : t.w1[i];
    if (w != 0.0) {
      const double2 p = in[k ^ m];
      acc.x = fma(w, p.x, acc.x);
      acc.y = fma(w, p.y, acc.y);
    }
  }
  for (int i = t.npairs; i < t.nterms; ++i) {
    // (cx + i cy) p with the spin down at k, (cx - i cy) p with it up
    const uint32_t m = t.mask[i];
    const double2 p = in[k ^ m];
    const double sy = (k & m) ? -t.w1[i] : t.w1[i];
    acc.x = fma(t.w0[i], p.x, acc.x);
    acc.x = fma(-sy, p.y, acc.x);
    acc.y = fma(t.w0[i], p.y, acc.y);
    acc.y = fma(sy, p.x, acc.y);
  }
  if constexpr (KIND == NA_ORDER1) {
    out[k] = make_double2(0.5 * acc.x, 0.5 * acc.y);
  } else if constexpr (KIND == NA_ORDER2) {
    const double2 t0 = psi[k];
    const double2 t2 = make_double2(acc.x - t0.x, acc.y - t0.y);
    out[k] = t2;
    double2 q = make_double2(c.a0 * t0.x, c.a0 * t0.y);
    q.x += c.c1r * x.x - c.c1i * x.y;
    q.y += c.c1r * x.y + c.c1i * x.x;
    q.x += c.cr * t2.x - c.ci * t2.y;
    q.y += c.cr * t2.y + c.ci * t2.x;
    psi[k] = q;
  } else {
    const double2 tm2 = out[k];
    const double2 tn = make_double2(acc.x - tm2.x, acc.y - tm2.y);
    out[k] = tn;
    double2 q = psi[k];
    q.x += c.cr * tn.x - c.ci * tn.y;
    q.y += c.cr * tn.y + c.ci * tn.x;
    psi[k] = q;
  }
}

// psi = a0 psi + c1 T_1: a one-order expansion.
__global__ __launch_bounds__(kNoAllocBlock) void k_na_first(NaStep c, const double2* __restrict__ t1,
                                                             double2* __restrict__ psi, uint32_t n) {
  const uint32_t k = blockIdx.x * kNoAllocBlock + threadIdx.x;
  if (k >= n) return;
  const double2 t0 = psi[k], x = t1[k];
  double2 q = make_double2(c.a0 * t0.x, c.a0 * t0.y);
  q.x += c.c1r * x.x - c.c1i * x.y;
  q.y += c.c1r * x.y + c.c1i * x.x;
  psi[k] = q;
}

}  // namespace
}  // namespace spx

using namespace spx;

extern "C" {

int spx_prop_evolve_noalloc(spx_ham* h, int mode, spx_state* psi, double tau, size_t nsteps, double epsilon,
                            size_t* order) {
  if (!h || !psi || !psi->d) return fail(SPX_EINVAL, "null argument");
  if (psi->ctx != h->ctx) return fail(SPX_EINVAL, "state belongs to another context");
  if (psi->nspins != h->nspins) return fail(SPX_EINVAL, "state and Hamiltonian disagree on spin count");
  if (h->nspins > static_cast<size_t>(kNoAllocMaxSpins))
    return fail(SPX_ERANGE, "the allocation-free evolve holds at most 16 spins (2^16 amplitudes of static scratch)");
  if (mode != 0 && mode != 1) return fail(SPX_EINVAL, "unknown time mode");
  if (!(epsilon > 0.0)) return fail(SPX_EINVAL, "epsilon must be positive");
  if (h->lo.size() > static_cast<size_t>(kNoAllocMaxPairs)) return fail(SPX_EINVAL, "too many couplings");
  const size_t L = h->nspins;
  const uint32_t n = static_cast<uint32_t>(size_t(1) << L);

  double sb = 0.0;
  SPX_TRY(spx_ham_spectral_bound(h, &sb));
  std::vector<double> a;
  double bound = 1.0;
  SPX_TRY(cheb_build(mode, sb, tau, epsilon, 6000, a, bound));
  const size_t K = a.size() - 1;
  if (order) *order = K;
  if (nsteps == 0 || K == 0) return SPX_OK;   // identity: psi untouched
  if (!(bound > 0.0)) return fail(SPX_EDOMAIN, "spectral bound must be positive");

  SPX_HIP(hipSetDevice(h->ctx->device));
  double2* ptr = nullptr;
  SPX_TRY(scratch_on(h->ctx->device, &ptr));
  double2* A = ptr;
  double2* B = ptr + n;
  double* diag = reinterpret_cast<double*>(ptr + 2 * size_t(n));
  hipStream_t st = h->ctx->stream;

  // terms at the step's scale -2/E, after the 1e-14 skip rule
  const double scale = -2.0 / bound;
  NaTerms t{};
  NaDiag dg{};
  t.sdiag = scale;
  for (size_t p = 0; p < h->lo.size(); ++p) {
    double ff = scale * 0.25 * (h->jx[p] + h->jy[p]);
    double fb = scale * 0.25 * (h->jx[p] - h->jy[p]);
    if (!(std::abs(ff) > kNegligible)) ff = 0.0;
    if (!(std::abs(fb) > kNegligible)) fb = 0.0;
    const uint32_t m = (1u << h->lo[p]) | (1u << h->hi[p]);
    if (ff != 0.0 || fb != 0.0) {
      t.mask[t.nterms] = m;
      t.w0[t.nterms] = ff;
      t.w1[t.nterms] = fb;
      ++t.nterms;
    }
    const double cz = 0.25 * h->jz[p];
    if (std::abs(cz) > kNegligible) {
      dg.mask[dg.npair] = m;
      dg.cp[dg.npair] = cz;
      ++dg.npair;
    }
  }
  t.npairs = t.nterms;
  for (size_t s = 0; s < L; ++s) {
    const double cx = scale * 0.5 * h->hx[s], cy = scale * 0.5 * h->hy[s];
    if (std::abs(cx) > kNegligible || std::abs(cy) > kNegligible) {
      t.mask[t.nterms] = 1u << s;
      t.w0[t.nterms] = cx;
      t.w1[t.nterms] = cy;
      ++t.nterms;
    }
    const double cz = 0.5 * h->hz[s];
    if (std::abs(cz) > kNegligible) {
      dg.spin[dg.nfield] = static_cast<uint32_t>(s);
      dg.cf[dg.nfield] = cz;
      ++dg.nfield;
    }
  }
  const dim3 grid((n + kNoAllocBlock - 1) / kNoAllocBlock);
  k_na_diag<<<grid, kNoAllocBlock, 0, st>>>(dg, diag, n);
  SPX_HIP(hipGetLastError());

  // w_n: (-i)^n in real time, (-1)^n in imaginary time
  auto weight = [&](size_t k, double& wr, double& wi) {
    if (mode == 1) {
      wr = (k % 2) ? -1.0 : 1.0;
      wi = 0.0;
      return;
    }
    static const double re[4] = {1.0, 0.0, -1.0, 0.0}, im[4] = {0.0, -1.0, 0.0, 1.0};
    wr = re[k % 4];
    wi = im[k % 4];
  };
  for (size_t step = 0; step < nsteps; ++step) {
    NaStep c{};
    double wr = 0.0, wi = 0.0;
    c.a0 = a[0];
    weight(1, wr, wi);
    c.c1r = 2.0 * a[1] * wr;
    c.c1i = 2.0 * a[1] * wi;
    k_na_order<NA_ORDER1><<<grid, kNoAllocBlock, 0, st>>>(t, c, psi->d, B, nullptr, diag, n);
    SPX_HIP(hipGetLastError());
    if (K == 1) {
      k_na_first<<<grid, kNoAllocBlock, 0, st>>>(c, B, psi->d, n);
      SPX_HIP(hipGetLastError());
      continue;
    }
    weight(2, wr, wi);
    c.cr = 2.0 * a[2] * wr;
    c.ci = 2.0 * a[2] * wi;
    k_na_order<NA_ORDER2><<<grid, kNoAllocBlock, 0, st>>>(t, c, B, A, psi->d, diag, n);
    SPX_HIP(hipGetLastError());
    for (size_t k = 3; k <= K; ++k) {
      weight(k, wr, wi);
      c.cr = 2.0 * a[k] * wr;
      c.ci = 2.0 * a[k] * wi;
      const bool odd = (k % 2) == 1;
      k_na_order<NA_ORDERN><<<grid, kNoAllocBlock, 0, st>>>(t, c, odd ? A : B, odd ? B : A, psi->d, diag, n);
      SPX_HIP(hipGetLastError());
    }
  }
  return SPX_OK;
}

}  // extern "C"
