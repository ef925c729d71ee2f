// Device kernels of libspx for gfx950 (CDNA4, wave64), FP64 throughout.
//
// The Hamiltonian is applied matrix-free and gather-style: each output
// amplitude k collects its diagonal term and one partner amplitude per flip
// term (k ^ mask), so a whole H.psi is ONE pass over the state instead of the
// reference's one pass per coupled pair (src/core/Hamiltonian.cpp:95-127).
// A workgroup owns a tile of 2^tb consecutive amplitudes (the low tb bits of
// k) and stages its operand tile in LDS: partners that differ only in tile
// bits are read from LDS, the rest from global memory (L2 / Infinity Cache;
// the L=18 working set is far below the 256 MB MALL). The Chebyshev
// three-term recurrence and the Bessel-weighted accumulation into psi are
// fused into the same pass (src/core/Propagator.cpp:58-72), so an order moves
// T_{n-1}, T_{n-2}, T_n, the diagonal and psi exactly once through HBM.
#include <cstdlib>

#include "spx_common.hpp"

namespace spx {

namespace {

constexpr int kBlock = 256;

struct KArgs {
  const double2* in[kMaxMany];
  double2* out[kMaxMany];
  double2* psi[kMaxMany];
  const double* diag;
  const FlipTerm* terms;  // [nfield field terms][nlocal local pairs][nremote remote pairs]
  int nfield, nlocal, nremote;
  int tb;
  double sdiag;
  double a0, c1r, c1i, cr, ci;
};

enum { K_APPLY = 0, K_ACCUM = 1, K_ORDER1 = 2, K_ORDER2 = 3, K_ORDERN = 4 };

__device__ __forceinline__ void cmad(double2& acc, double c, double2 p) {
  acc.x = fma(c, p.x, acc.x);
  acc.y = fma(c, p.y, acc.y);
}

// (cx + i cy) p when the field spin is down at k, (cx - i cy) p when up
// (axpyTransverse, src/core/SpinOperators.cpp:228-252).
__device__ __forceinline__ void field_term(double2& acc, bool up, double cx, double cy, double2 p) {
  const double sy = up ? -cy : cy;
  acc.x = fma(cx, p.x, acc.x);
  acc.x = fma(-sy, p.y, acc.x);
  acc.y = fma(cx, p.y, acc.y);
  acc.y = fma(sy, p.x, acc.y);
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_horder(KArgs a) {
  extern __shared__ double2 tile[];
  const int v = blockIdx.y;
  const double2* __restrict__ in = a.in[v];
  const uint32_t tsize = 1u << a.tb;
  const uint32_t lmask = tsize - 1u;
  const size_t base = static_cast<size_t>(blockIdx.x) << a.tb;
  for (uint32_t j = threadIdx.x; j < tsize; j += kBlock) tile[j] = in[base + j];
  __syncthreads();

  const FlipTerm* __restrict__ fterm = a.terms;
  const FlipTerm* __restrict__ lterm = a.terms + a.nfield;
  const FlipTerm* __restrict__ rterm = a.terms + a.nfield + a.nlocal;

  for (uint32_t j = threadIdx.x; j < tsize; j += kBlock) {
    const size_t k = base + j;
    const double2 x = tile[j];
    const double f = a.sdiag * a.diag[k];
    double2 acc = make_double2(f * x.x, f * x.y);

    for (int t = 0; t < a.nfield; ++t) {
      const FlipTerm ft = fterm[t];
      const double2 p = (ft.mask & ~lmask) ? in[k ^ ft.mask] : tile[j ^ ft.mask];
      field_term(acc, (k & ft.mask) != 0, ft.anti, ft.aligned, p);
    }
#pragma unroll 4
    for (int t = 0; t < a.nlocal; ++t) {
      const FlipTerm ft = lterm[t];
      const bool anti = __popc(static_cast<uint32_t>(k) & ft.mask) == 1;
      const double c = anti ? ft.anti : ft.aligned;
      if (c != 0.0) cmad(acc, c, tile[j ^ ft.mask]);
    }
#pragma unroll 4
    for (int t = 0; t < a.nremote; ++t) {
      const FlipTerm ft = rterm[t];
      const bool anti = __popc(static_cast<uint32_t>(k) & ft.mask) == 1;
      const double c = anti ? ft.anti : ft.aligned;
      if (c != 0.0) cmad(acc, c, in[k ^ ft.mask]);
    }

    if constexpr (KIND == K_APPLY) {
      a.out[v][k] = acc;
    } else if constexpr (KIND == K_ACCUM) {
      double2 o = a.out[v][k];
      o.x += acc.x;
      o.y += acc.y;
      a.out[v][k] = o;
    } else if constexpr (KIND == K_ORDER1) {
      // T_1 = Hhat T_0: the general recurrence's 2 Hhat T_0 halved.
      a.out[v][k] = make_double2(0.5 * acc.x, 0.5 * acc.y);
    } else if constexpr (KIND == K_ORDER2) {
      // psi still holds T_0; psi = a0 T_0 + c1 T_1 + c2 T_2.
      const double2 t0 = a.psi[v][k];
      const double2 t2 = make_double2(acc.x - t0.x, acc.y - t0.y);
      a.out[v][k] = t2;
      double2 p = make_double2(a.a0 * t0.x, a.a0 * t0.y);
      p.x += a.c1r * x.x - a.c1i * x.y;
      p.y += a.c1r * x.y + a.c1i * x.x;
      p.x += a.cr * t2.x - a.ci * t2.y;
      p.y += a.cr * t2.y + a.ci * t2.x;
      a.psi[v][k] = p;
    } else {
      const double2 tm2 = a.out[v][k];
      const double2 tn = make_double2(acc.x - tm2.x, acc.y - tm2.y);
      a.out[v][k] = tn;
      double2 p = a.psi[v][k];
      p.x += a.cr * tn.x - a.ci * tn.y;
      p.y += a.cr * tn.y + a.ci * tn.x;
      a.psi[v][k] = p;
    }
  }
}

// rebuildDiagonal (src/core/Hamiltonian.cpp:54-80): per element, the same
// terms in the same order from 0.0, so the result is bitwise the reference's.
__global__ __launch_bounds__(kBlock) void k_build_diag(double* __restrict__ d, size_t n,
                                                       const double* __restrict__ tab, int nf,
                                                       int np) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= n) return;
  // tab: [nf] (bit, c) pairs then [np] (mask, c) pairs, bit/mask as double
  double s = 0.0;
  for (int i = 0; i < nf; ++i) {
    const uint32_t bit = static_cast<uint32_t>(tab[2 * i]);
    const double c = tab[2 * i + 1];
    s += ((k >> bit) & 1u) ? c : -c;
  }
  for (int i = 0; i < np; ++i) {
    const uint32_t m = static_cast<uint32_t>(tab[2 * (nf + i)]);
    const double c = tab[2 * (nf + i) + 1];
    const bool aligned = __popc(static_cast<uint32_t>(k) & m) != 1;
    s += aligned ? c : -c;
  }
  d[k] = s;
}

__global__ __launch_bounds__(kBlock) void k_fill(double2* y, size_t n) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) y[k] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(kBlock) void k_scale(double2* y, double f, size_t n) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) {
    double2 v = y[k];
    y[k] = make_double2(v.x * f, v.y * f);
  }
}

__global__ __launch_bounds__(kBlock) void k_negate(double2* y, size_t n) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) {
    double2 v = y[k];
    y[k] = make_double2(-v.x, -v.y);
  }
}

// addScaled (src/core/StateVector.cpp:21-32).
__global__ __launch_bounds__(kBlock) void k_add_scaled(double2* __restrict__ y, double cr, double ci,
                                                       const double2* __restrict__ x, size_t n) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) {
    const double2 o = x[k];
    double2 d = y[k];
    d.x += cr * o.x - ci * o.y;
    d.y += cr * o.y + ci * o.x;
    y[k] = d;
  }
}

// y = a y + (br + i bi) x  (closing a one-order expansion)
__global__ __launch_bounds__(kBlock) void k_axpby(double2* __restrict__ y, double a,
                                                  const double2* __restrict__ x, double br,
                                                  double bi, size_t n) {
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) {
    const double2 o = x[k];
    double2 d = y[k];
    d = make_double2(a * d.x, a * d.y);
    d.x += br * o.x - bi * o.y;
    d.y += br * o.y + bi * o.x;
    y[k] = d;
  }
}

__device__ __forceinline__ size_t insert_zero(size_t q, uint32_t bit) {
  const size_t low = q & ((static_cast<size_t>(1) << bit) - 1);
  return ((q >> bit) << (bit + 1)) | low;
}

// axpySigma* / applySigma* (src/core/SpinOperators.cpp:90-177), one thread
// per (down, up) index pair of `spin`.
__global__ __launch_bounds__(kBlock) void k_sigma(int axis, double c, const double2* __restrict__ in,
                                                  uint32_t spin, double2* __restrict__ out,
                                                  size_t half, int accumulate) {
  const size_t q = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= half) return;
  const size_t l = insert_zero(q, spin);
  const size_t h = l | (static_cast<size_t>(1) << spin);
  const double2 il = in[l], ih = in[h];
  double2 nl, nh;  // sigma applied, unscaled
  if (axis == 0) {
    nh = il;
    nl = ih;
  } else if (axis == 1) {
    nh = make_double2(il.y, -il.x);
    nl = make_double2(-ih.y, ih.x);
  } else {
    nh = ih;
    nl = make_double2(-il.x, -il.y);
  }
  if (accumulate) {
    double2 ol = out[l], oh = out[h];
    ol.x += c * nl.x;
    ol.y += c * nl.y;
    oh.x += c * nh.x;
    oh.y += c * nh.y;
    out[l] = ol;
    out[h] = oh;
  } else {
    out[l] = nl;
    out[h] = nh;
  }
}

// sum_s sigma_axis(s) applied at amplitude k, spins summed in index order as
// applyTotalSpin does (src/core/Observable.cpp:13-22).
__device__ __forceinline__ double2 total_spin_at(int axis, const double2* __restrict__ in, size_t k,
                                                 uint32_t L) {
  double2 w = make_double2(0.0, 0.0);
  if (axis == 2) {
    const double2 x = in[k];
    for (uint32_t s = 0; s < L; ++s) {
      if ((k >> s) & 1u) {
        w.x += x.x;
        w.y += x.y;
      } else {
        w.x -= x.x;
        w.y -= x.y;
      }
    }
    return w;
  }
  for (uint32_t s = 0; s < L; ++s) {
    const double2 p = in[k ^ (static_cast<size_t>(1) << s)];
    if (axis == 0) {
      w.x += p.x;
      w.y += p.y;
    } else if ((k >> s) & 1u) {  // up: -i * psi(down)
      w.x += p.y;
      w.y -= p.x;
    } else {  // down: +i * psi(up)
      w.x -= p.y;
      w.y += p.x;
    }
  }
  return w;
}

__global__ __launch_bounds__(kBlock) void k_total_spin(int axis, const double2* __restrict__ in,
                                                       double2* __restrict__ out, uint32_t L) {
  const size_t n = static_cast<size_t>(1) << L;
  const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k < n) out[k] = total_spin_at(axis, in, k, L);
}

struct Rot {
  double m[8];  // ll, lh, hl, hh as (re, im)
};

// SpinRotation::apply for one target (src/core/Gate.cpp:60-85).
__global__ __launch_bounds__(kBlock) void k_rotate(double2* __restrict__ s, size_t half, uint32_t spin,
                                                   Rot u) {
  const size_t q = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= half) return;
  const size_t l = insert_zero(q, spin);
  const size_t h = l | (static_cast<size_t>(1) << spin);
  const double2 a = s[l], b = s[h];
  const double llr = u.m[0], lli = u.m[1], lhr = u.m[2], lhi = u.m[3];
  const double hlr = u.m[4], hli = u.m[5], hhr = u.m[6], hhi = u.m[7];
  s[l] = make_double2((llr * a.x - lli * a.y) + (lhr * b.x - lhi * b.y),
                      (llr * a.y + lli * a.x) + (lhr * b.y + lhi * b.x));
  s[h] = make_double2((hlr * a.x - hli * a.y) + (hhr * b.x - hhi * b.y),
                      (hlr * a.y + hli * a.x) + (hhr * b.y + hhi * b.x));
}

struct Gate4 {
  double m[32];
};

// TwoSpinGate::apply (src/core/Gate.cpp:90-121).
__global__ __launch_bounds__(kBlock) void k_two_spin(double2* __restrict__ s, size_t quarter,
                                                     uint32_t lo, uint32_t hi, Gate4 u) {
  const size_t q = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= quarter) return;
  const size_t dd = insert_zero(insert_zero(q, lo), hi);
  const size_t idx[4] = {dd, dd | (static_cast<size_t>(1) << lo), dd | (static_cast<size_t>(1) << hi),
                         dd | (static_cast<size_t>(1) << lo) | (static_cast<size_t>(1) << hi)};
  double2 v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = s[idx[i]];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double ar = 0.0, ai = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double mr = u.m[2 * (4 * r + c)], mi = u.m[2 * (4 * r + c) + 1];
      ar += mr * v[c].x - mi * v[c].y;
      ai += mr * v[c].y + mi * v[c].x;
    }
    s[idx[r]] = make_double2(ar, ai);
  }
}

// ------------------------------------------------------------- reductions
// Fixed partition of [0, n) into gridDim.x chunks (gridDim.x a function of n
// only), fixed per-thread strides and a fixed LDS tree: the result is
// bitwise reproducible from run to run and across launch timing.

struct OpNorm {
  const double2* x;
  __device__ double2 operator()(size_t k) const {
    const double2 v = x[k];
    return make_double2(v.x * v.x + v.y * v.y, 0.0);
  }
};

struct OpInner {  // <other|self>
  const double2* a;
  const double2* b;
  __device__ double2 operator()(size_t k) const {
    const double2 u = a[k], w = b[k];
    return make_double2(w.x * u.x + w.y * u.y, w.x * u.y - w.y * u.x);
  }
};

struct OpCorr {  // <ref| sum sigma |state>
  int axis;
  uint32_t L;
  const double2* state;
  const double2* ref;
  __device__ double2 operator()(size_t k) const {
    const double2 w = total_spin_at(axis, state, k, L);
    const double2 r = ref[k];
    return make_double2(r.x * w.x + r.y * w.y, r.x * w.y - r.y * w.x);
  }
};

struct OpMag {  // Re <state|sigma(spin)|state>
  int axis;
  uint32_t spin;
  const double2* state;
  __device__ double2 operator()(size_t k) const {
    const double2 x = state[k];
    const bool up = (k >> spin) & 1u;
    double2 w;
    if (axis == 2) {
      w = up ? x : make_double2(-x.x, -x.y);
    } else {
      const double2 p = state[k ^ (static_cast<size_t>(1) << spin)];
      if (axis == 0) w = p;
      else w = up ? make_double2(p.y, -p.x) : make_double2(-p.y, p.x);
    }
    return make_double2(x.x * w.x + x.y * w.y, 0.0);
  }
};

struct OpEnergy {  // Re <H state|state>
  const double2* x;
  const double* diag;
  const FlipTerm* terms;
  int nfield, npair;
  __device__ double2 operator()(size_t k) const {
    const double2 v = x[k];
    const double f = diag[k];
    double2 acc = make_double2(f * v.x, f * v.y);
    for (int t = 0; t < nfield; ++t) {
      const FlipTerm ft = terms[t];
      field_term(acc, (k & ft.mask) != 0, ft.anti, ft.aligned, x[k ^ ft.mask]);
    }
    for (int t = 0; t < npair; ++t) {
      const FlipTerm ft = terms[nfield + t];
      const bool anti = __popc(static_cast<uint32_t>(k) & ft.mask) == 1;
      const double c = anti ? ft.anti : ft.aligned;
      if (c != 0.0) cmad(acc, c, x[k ^ ft.mask]);
    }
    return make_double2(acc.x * v.x + acc.y * v.y, 0.0);
  }
};

__device__ __forceinline__ void block_tree(double* lr, double* li, double sr, double si) {
  const int t = threadIdx.x;
  lr[t] = sr;
  li[t] = si;
  __syncthreads();
  for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      lr[t] += lr[t + s];
      li[t] += li[t + s];
    }
    __syncthreads();
  }
}

template <class Op>
__global__ __launch_bounds__(kReduceThreads) void k_reduce1(Op op, size_t n, double2* partials) {
  __shared__ double lr[kReduceThreads], li[kReduceThreads];
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const size_t beg = static_cast<size_t>(blockIdx.x) * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  double sr = 0.0, si = 0.0;
  for (size_t k = beg + threadIdx.x; k < end; k += kReduceThreads) {
    const double2 v = op(k);
    sr += v.x;
    si += v.y;
  }
  block_tree(lr, li, sr, si);
  if (threadIdx.x == 0) partials[blockIdx.x] = make_double2(lr[0], li[0]);
}

__global__ __launch_bounds__(kReduceThreads) void k_reduce2(const double2* partials, int nb,
                                                            double* result) {
  __shared__ double lr[kReduceThreads], li[kReduceThreads];
  double sr = 0.0, si = 0.0;
  for (int b = threadIdx.x; b < nb; b += kReduceThreads) {
    sr += partials[b].x;
    si += partials[b].y;
  }
  block_tree(lr, li, sr, si);
  if (threadIdx.x == 0) {
    result[0] = lr[0];
    result[1] = li[0];
  }
}

inline unsigned blocks_for(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

template <class Op>
int reduce(spx_ctx* c, Op op, size_t n, double* re, double* im) {
  size_t nb = (n + kReduceThreads - 1) / kReduceThreads;
  if (nb > static_cast<size_t>(kReduceBlocks)) nb = kReduceBlocks;
  if (nb == 0) nb = 1;
  double2* partials = reinterpret_cast<double2*>(c->d_partials);
  k_reduce1<Op><<<static_cast<unsigned>(nb), kReduceThreads, 0, c->stream>>>(op, n, partials);
  k_reduce2<<<1, kReduceThreads, 0, c->stream>>>(partials, static_cast<int>(nb), c->d_result);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipMemcpyAsync(c->h_result, c->d_result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  SPX_HIP(hipStreamSynchronize(c->stream));
  if (re) *re = c->h_result[0];
  if (im) *im = c->h_result[1];
  return SPX_OK;
}

}  // namespace

// --------------------------------------------------------------- launchers

int tile_bits_for(size_t L) {
  static int forced = [] {
    const char* e = std::getenv("SPX_TILE_BITS");
    return e ? std::atoi(e) : 0;
  }();
  int tb = forced > 0 ? forced : 10;
  if (tb > 12) tb = 12;
  if (tb < 6) tb = 6;
  if (static_cast<size_t>(tb) > L) tb = static_cast<int>(L);
  return tb;
}

// Diagonal (when hz or the couplings changed) and the flip table at `scale`,
// with the reference's skip rule applied to the scaled coefficients.
int ham_prepare(spx_ham* h, double scale) {
  spx_ctx* c = h->ctx;
  const size_t L = h->nspins, n = static_cast<size_t>(1) << L;
  if (!h->diag_valid || h->diag_hz != h->hz || h->diag_jz != h->jz || h->diag_lo != h->lo ||
      h->diag_hi != h->hi) {
    std::vector<double> tab;
    int nf = 0, np = 0;
    for (size_t s = 0; s < L; ++s) {
      const double cz = 0.5 * h->hz[s];
      if (std::abs(cz) <= kNegligible) continue;
      tab.push_back(static_cast<double>(s));
      tab.push_back(cz);
      ++nf;
    }
    for (size_t p = 0; p < h->lo.size(); ++p) {
      const double cz = 0.25 * h->jz[p];
      if (std::abs(cz) <= kNegligible) continue;
      tab.push_back(static_cast<double>((1u << h->lo[p]) | (1u << h->hi[p])));
      tab.push_back(cz);
      ++np;
    }
    if (!h->d_diag) SPX_HIP(hipMalloc(&h->d_diag, n * sizeof(double)));
    const size_t tbytes = (tab.size() + 2) * sizeof(double);
    if (tbytes > h->tab_cap) {
      if (h->d_tab) SPX_HIP(hipFree(h->d_tab));
      SPX_HIP(hipMalloc(&h->d_tab, tbytes));
      h->tab_cap = tbytes;
    }
    if (!tab.empty())
      SPX_HIP(hipMemcpyAsync(h->d_tab, tab.data(), tab.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    k_build_diag<<<blocks_for(n), kBlock, 0, c->stream>>>(h->d_diag, n, h->d_tab, nf, np);
    SPX_HIP(hipGetLastError());
    h->diag_valid = true;
    h->diag_hz = h->hz;
    h->diag_jz = h->jz;
    h->diag_lo = h->lo;
    h->diag_hi = h->hi;
    h->terms_valid = false;
  }
  if (h->terms_valid && h->terms_scale == scale) return SPX_OK;

  const int tb = tile_bits_for(L);
  const uint32_t lmask = (1u << tb) - 1u;
  std::vector<FlipTerm> fields, local, remote;
  for (size_t s = 0; s < L; ++s) {
    const double cx = scale * 0.5 * h->hx[s];
    const double cy = scale * 0.5 * h->hy[s];
    if (std::abs(cx) > kNegligible || std::abs(cy) > kNegligible)
      fields.push_back(FlipTerm{1u << s, 1u, cx, cy, 0.0});
  }
  for (size_t p = 0; p < h->lo.size(); ++p) {
    const double quarter = scale * 0.25;
    double ff = quarter * (h->jx[p] + h->jy[p]);
    double fb = quarter * (h->jx[p] - h->jy[p]);
    if (!(std::abs(ff) > kNegligible)) ff = 0.0;
    if (!(std::abs(fb) > kNegligible)) fb = 0.0;
    if (ff == 0.0 && fb == 0.0) continue;
    const uint32_t m = (1u << h->lo[p]) | (1u << h->hi[p]);
    ((m & ~lmask) ? remote : local).push_back(FlipTerm{m, 0u, ff, fb, 0.0});
  }
  h->h_terms.clear();
  h->h_terms.insert(h->h_terms.end(), fields.begin(), fields.end());
  h->h_terms.insert(h->h_terms.end(), local.begin(), local.end());
  h->h_terms.insert(h->h_terms.end(), remote.begin(), remote.end());
  h->nterms = static_cast<int>(h->h_terms.size());
  h->terms_nfield = static_cast<int>(fields.size());
  h->terms_nlocal = static_cast<int>(local.size());
  h->terms_nremote = static_cast<int>(remote.size());
  h->terms_tb = tb;
  h->terms_scale = scale;
  h->terms_valid = true;
  const size_t bytes = (h->h_terms.size() + 1) * sizeof(FlipTerm);
  if (bytes > h->terms_cap) {
    if (h->d_terms) SPX_HIP(hipFree(h->d_terms));
    SPX_HIP(hipMalloc(&h->d_terms, bytes));
    h->terms_cap = bytes;
  }
  if (!h->h_terms.empty())
    SPX_HIP(hipMemcpyAsync(h->d_terms, h->h_terms.data(), h->h_terms.size() * sizeof(FlipTerm),
                           hipMemcpyHostToDevice, c->stream));
  static_assert(sizeof(FlipTerm) == 32, "flip term layout");
  return SPX_OK;
}

static int launch_k(spx_ham* h, int kind, const KArgs& a, unsigned nvec) {
  spx_ctx* c = h->ctx;
  const size_t n = static_cast<size_t>(1) << h->nspins;
  const dim3 grid(static_cast<unsigned>(n >> a.tb), nvec);
  const size_t lds = sizeof(double2) << a.tb;
  switch (kind) {
    case K_APPLY: k_horder<K_APPLY><<<grid, kBlock, lds, c->stream>>>(a); break;
    case K_ACCUM: k_horder<K_ACCUM><<<grid, kBlock, lds, c->stream>>>(a); break;
    case K_ORDER1: k_horder<K_ORDER1><<<grid, kBlock, lds, c->stream>>>(a); break;
    case K_ORDER2: k_horder<K_ORDER2><<<grid, kBlock, lds, c->stream>>>(a); break;
    default: k_horder<K_ORDERN><<<grid, kBlock, lds, c->stream>>>(a); break;
  }
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

static KArgs base_args(spx_ham* h, double sdiag) {
  KArgs a{};
  a.diag = h->d_diag;
  a.terms = h->d_terms;
  a.nfield = h->terms_nfield;
  a.nlocal = h->terms_nlocal;
  a.nremote = h->terms_nremote;
  a.tb = h->terms_tb;
  a.sdiag = sdiag;
  return a;
}

int launch_apply(spx_ham* h, double scale, const double2* in, double2* out, bool accumulate) {
  SPX_TRY(ham_prepare(h, scale));
  KArgs a = base_args(h, scale);
  a.in[0] = in;
  a.out[0] = out;
  a.psi[0] = nullptr;
  return launch_k(h, accumulate ? K_ACCUM : K_APPLY, a, 1);
}

int launch_order(spx_ham* h, int kind, const OrderLaunch& v, double a0, double c1r, double c1i,
                 double cr, double ci) {
  KArgs a = base_args(h, h->terms_scale);
  for (int i = 0; i < v.nvec; ++i) {
    a.in[i] = v.in[i];
    a.out[i] = v.out[i];
    a.psi[i] = v.psi[i];
  }
  a.a0 = a0;
  a.c1r = c1r;
  a.c1i = c1i;
  a.cr = cr;
  a.ci = ci;
  const int k = kind == 1 ? K_ORDER1 : kind == 2 ? K_ORDER2 : K_ORDERN;
  return launch_k(h, k, a, static_cast<unsigned>(v.nvec));
}

int launch_fill(spx_ctx* c, double2* y, size_t n) {
  k_fill<<<blocks_for(n), kBlock, 0, c->stream>>>(y, n);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_scale(spx_ctx* c, double2* y, double f, size_t n) {
  k_scale<<<blocks_for(n), kBlock, 0, c->stream>>>(y, f, n);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_negate(spx_ctx* c, double2* y, size_t n) {
  k_negate<<<blocks_for(n), kBlock, 0, c->stream>>>(y, n);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_add_scaled(spx_ctx* c, double2* y, double cr, double ci, const double2* x, size_t n) {
  k_add_scaled<<<blocks_for(n), kBlock, 0, c->stream>>>(y, cr, ci, x, n);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_axpby(spx_ctx* c, double2* y, double a, const double2* x, double br, double bi, size_t n) {
  k_axpby<<<blocks_for(n), kBlock, 0, c->stream>>>(y, a, x, br, bi, n);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_sigma(spx_ctx* c, int axis, double coef, const double2* in, size_t spin, double2* out,
                 size_t n, bool accumulate) {
  const size_t half = n / 2;
  k_sigma<<<blocks_for(half), kBlock, 0, c->stream>>>(axis, coef, in, static_cast<uint32_t>(spin),
                                                      out, half, accumulate ? 1 : 0);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_total_spin(spx_ctx* c, int axis, const double2* in, double2* out, size_t L) {
  const size_t n = static_cast<size_t>(1) << L;
  k_total_spin<<<blocks_for(n), kBlock, 0, c->stream>>>(axis, in, out, static_cast<uint32_t>(L));
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_rotate(spx_ctx* c, double2* s, size_t n, uint32_t spin, const double* u8) {
  Rot u;
  for (int i = 0; i < 8; ++i) u.m[i] = u8[i];
  const size_t half = n / 2;
  k_rotate<<<blocks_for(half), kBlock, 0, c->stream>>>(s, half, spin, u);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_two_spin(spx_ctx* c, double2* s, size_t n, uint32_t lo, uint32_t hi, const double* u32) {
  Gate4 u;
  for (int i = 0; i < 32; ++i) u.m[i] = u32[i];
  const size_t quarter = n / 4;
  k_two_spin<<<blocks_for(quarter), kBlock, 0, c->stream>>>(s, quarter, lo, hi, u);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int reduce_norm2(spx_ctx* c, const double2* x, size_t n, double* out) {
  return reduce(c, OpNorm{x}, n, out, nullptr);
}

int reduce_inner(spx_ctx* c, const double2* self, const double2* other, size_t n, double* re,
                 double* im) {
  return reduce(c, OpInner{self, other}, n, re, im);
}

int reduce_corr(spx_ctx* c, int axis, const double2* state, const double2* ref, size_t L, double* re,
                double* im) {
  return reduce(c, OpCorr{axis, static_cast<uint32_t>(L), state, ref}, static_cast<size_t>(1) << L,
                re, im);
}

int reduce_mag(spx_ctx* c, int axis, uint32_t spin, const double2* state, size_t L, double* out) {
  return reduce(c, OpMag{axis, spin, state}, static_cast<size_t>(1) << L, out, nullptr);
}

int reduce_energy(spx_ham* h, const double2* state, double* out) {
  SPX_TRY(ham_prepare(h, 1.0));
  OpEnergy op{state, h->d_diag, h->d_terms, h->terms_nfield,
              h->terms_nlocal + h->terms_nremote};
  return reduce(h->ctx, op, static_cast<size_t>(1) << h->nspins, out, nullptr);
}

}  // namespace spx
