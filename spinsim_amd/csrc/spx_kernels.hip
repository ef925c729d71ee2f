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
#include <utility>

#include "spx_common.hpp"

namespace spx {

namespace {

constexpr int kBlock = 256;

struct KArgs {
  const double2* in[kMaxMany];
  double2* out[kMaxMany];
  double2* psi[kMaxMany];
  const double* diag;
  // Flip-term table, grouped by where the partner amplitude lives relative to
  // a tile of 2^tb consecutive amplitudes:
  //   [0, nlocal)                       both bits (or the field bit) < tb: own tile in LDS
  //   [grp[r], grp[r+1]) for r < g      one bit is tb + r: neighbour tile x ^ (1 << r),
  //                                     mask = the local bit (0 for a field on tb + r)
  //   [grp[g], grp[g] + nboth)          both bits >= tb: mask in tile-index space
  const FlipTerm* terms;
  int nlocal, g, nboth;
  int grp[kMaxSpins + 1];
  int tb;
  int xcd_remap;
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

// One H application (or fused Chebyshev order) over a tile of 2^tb
// amplitudes; each thread owns AMPS amplitudes j = tid + r * blockDim.x.
template <int KIND, int AMPS>
__global__ __launch_bounds__(1024) void k_horder(KArgs a) {
  extern __shared__ double2 smem[];
  const uint32_t T = 1u << a.tb;
  const uint32_t B = blockDim.x;
  double2* own = smem;
  double2* nb = smem + T;
  const int v = blockIdx.y;
  const double2* __restrict__ in = a.in[v];

  // Tile index. Workgroups are dealt to the XCDs round-robin; with the remap
  // the three most significant tile bits select the XCD, so every pair
  // among the other bits finds its partner tile on the same XCD's L2.
  uint32_t x = blockIdx.x;
  if (a.xcd_remap && a.g >= 3) x = ((x & 7u) << (a.g - 3)) | (x >> 3);
  const size_t base = static_cast<size_t>(x) << a.tb;

  double2 acc[AMPS], xv[AMPS], pre[AMPS];
  uint32_t jj[AMPS];
#pragma unroll
  for (int r = 0; r < AMPS; ++r) {
    jj[r] = threadIdx.x + r * B;
    xv[r] = in[base + jj[r]];
  }
  if (a.g > 0) {
    const size_t nbase = static_cast<size_t>(x ^ 1u) << a.tb;
#pragma unroll
    for (int r = 0; r < AMPS; ++r) pre[r] = in[nbase + jj[r]];
  }
#pragma unroll
  for (int r = 0; r < AMPS; ++r) {
    own[jj[r]] = xv[r];
    const double f = a.sdiag * a.diag[base + jj[r]];
    acc[r] = make_double2(f * xv[r].x, f * xv[r].y);
  }
  __syncthreads();

  const FlipTerm* __restrict__ tt = a.terms;
  for (int t = 0; t < a.nlocal; ++t) {
    const FlipTerm ft = tt[t];
    if (ft.kind) {
#pragma unroll
      for (int r = 0; r < AMPS; ++r)
        field_term(acc[r], (jj[r] & ft.mask) != 0, ft.anti, ft.aligned, own[jj[r] ^ ft.mask]);
    } else {
#pragma unroll
      for (int r = 0; r < AMPS; ++r) {
        const double c = (__popc(jj[r] & ft.mask) == 1) ? ft.anti : ft.aligned;
        if (c != 0.0) cmad(acc[r], c, own[jj[r] ^ ft.mask]);
      }
    }
  }

  for (int rb = 0; rb < a.g; ++rb) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < AMPS; ++r) nb[jj[r]] = pre[r];
    __syncthreads();
    if (rb + 1 < a.g) {
      const size_t nbase = static_cast<size_t>(x ^ (1u << (rb + 1))) << a.tb;
#pragma unroll
      for (int r = 0; r < AMPS; ++r) pre[r] = in[nbase + jj[r]];
    }
    const uint32_t xb = (x >> rb) & 1u;
    for (int t = a.grp[rb]; t < a.grp[rb + 1]; ++t) {
      const FlipTerm ft = tt[t];
      if (ft.kind) {
#pragma unroll
        for (int r = 0; r < AMPS; ++r) field_term(acc[r], xb != 0, ft.anti, ft.aligned, nb[jj[r]]);
      } else {
#pragma unroll
        for (int r = 0; r < AMPS; ++r) {
          const bool anti = ((jj[r] & ft.mask) != 0) != (xb != 0);
          const double c = anti ? ft.anti : ft.aligned;
          if (c != 0.0) cmad(acc[r], c, nb[jj[r] ^ ft.mask]);
        }
      }
    }
  }

  const int b0 = a.grp[a.g];
  for (int t = 0; t < a.nboth; ++t) {
    const FlipTerm ft = tt[b0 + t];
    const double c = (__popc(x & ft.mask) == 1) ? ft.anti : ft.aligned;  // uniform per tile
    if (c != 0.0) {
      const size_t pbase = static_cast<size_t>(x ^ ft.mask) << a.tb;
#pragma unroll
      for (int r = 0; r < AMPS; ++r) cmad(acc[r], c, in[pbase + jj[r]]);
    }
  }

#pragma unroll
  for (int r = 0; r < AMPS; ++r) {
    const size_t k = base + jj[r];
    const double2 ac = acc[r];
    if constexpr (KIND == K_APPLY) {
      a.out[v][k] = ac;
    } else if constexpr (KIND == K_ACCUM) {
      double2 o = a.out[v][k];
      o.x += ac.x;
      o.y += ac.y;
      a.out[v][k] = o;
    } else if constexpr (KIND == K_ORDER1) {
      // T_1 = Hhat T_0: the general recurrence's 2 Hhat T_0 halved.
      a.out[v][k] = make_double2(0.5 * ac.x, 0.5 * ac.y);
    } else if constexpr (KIND == K_ORDER2) {
      // psi still holds T_0; psi = a0 T_0 + c1 T_1 + c2 T_2.
      const double2 t0 = a.psi[v][k];
      const double2 t1 = xv[r];
      const double2 t2 = make_double2(ac.x - t0.x, ac.y - t0.y);
      a.out[v][k] = t2;
      double2 p = make_double2(a.a0 * t0.x, a.a0 * t0.y);
      p.x += a.c1r * t1.x - a.c1i * t1.y;
      p.y += a.c1r * t1.y + a.c1i * t1.x;
      p.x += a.cr * t2.x - a.ci * t2.y;
      p.y += a.cr * t2.y + a.ci * t2.x;
      a.psi[v][k] = p;
    } else {
      const double2 tm2 = a.out[v][k];
      const double2 tn = make_double2(ac.x - tm2.x, ac.y - tm2.y);
      a.out[v][k] = tn;
      double2 p = a.psi[v][k];
      p.x += a.cr * tn.x - a.ci * tn.y;
      p.y += a.cr * tn.y + a.ci * tn.x;
      a.psi[v][k] = p;
    }
  }
}

// ------------------------------------------------------------ order kernel v3
// One amplitude per thread and a tile of T = blockDim.x <= 1024 amplitudes,
// so index bits 0-5 are lane bits and bits 6..tb-1 are the same for a whole
// wavefront. Every term whose activity (flip-flop: the two spins
// anti-aligned) does not involve a lane bit is decided once per wavefront
// (scalar branch) instead of per lane. The own tile and all g neighbour
// tiles (tile ^ one remote bit) are loaded together into LDS behind one
// barrier; pairs with both bits remote read their partner tile from global
// memory in batches of eight loads in flight.
constexpr int kMaxG3 = 12;

struct K3Args {
  const double2* in[kMaxMany];
  double2* out[kMaxMany];
  double2* psi[kMaxMany];
  const double* diag;
  const FlipTerm* terms;
  // segment ends: [0] local fields, [1] wave-uniform local pairs, [2] per-lane
  // local pairs, then per remote bit r: [3+3r] field on tb+r, [4+3r]
  // wave-uniform (local bit >= 6), [5+3r] per-lane; [3+3g] pairs with both
  // bits remote (tile-index masks)
  int off[3 + 3 * kMaxG3 + 1];
  int tb, g;
  int nterms;     // table length (padded), staged into LDS after the tiles
  int xcd_remap;
  int flipboth;   // any nonzero flip-both (jx != jy) weight
  double sdiag;
  double a0, c1r, c1i, cr, ci;
};

constexpr int kChunk = 8;

template <int KIND, int G, bool FLIPBOTH>
__global__ __launch_bounds__(1024) void k_order3(K3Args a) {
  extern __shared__ double2 smem[];
  const uint32_t T = blockDim.x;
  const uint32_t j = threadIdx.x;
  const uint32_t wj = __builtin_amdgcn_readfirstlane(j) & ~63u;
  const int v = blockIdx.y;
  const double2* __restrict__ in = a.in[v];
  uint32_t x = blockIdx.x;
  if (a.xcd_remap && G >= 3) x = ((x & 7u) << (G - 3)) | (x >> 3);
  const size_t base = static_cast<size_t>(x) << a.tb;
  const size_t k = base + j;

  // every global operand of this thread, issued up front
  const double2 xv = in[k];
  double2 nbv[G > 0 ? G : 1];
#pragma unroll
  for (int r = 0; r < G; ++r) nbv[r] = in[(static_cast<size_t>(x ^ (1u << r)) << a.tb) + j];
  const double dg = a.diag[k];
  double2 prev = make_double2(0.0, 0.0), psiv = make_double2(0.0, 0.0);
  if constexpr (KIND == K_ACCUM || KIND == K_ORDERN) prev = a.out[v][k];
  if constexpr (KIND == K_ORDER2 || KIND == K_ORDERN) psiv = a.psi[v][k];

  smem[j] = xv;
#pragma unroll
  for (int r = 0; r < G; ++r) smem[(1 + r) * T + j] = nbv[r];
  // The flip-term table goes to LDS too: read from there, its entries
  // do not serialise against the LDS gathers the way scalar loads
  // (which share the LGKM counter and may return out of order) would.
  FlipTerm* lterms = reinterpret_cast<FlipTerm*>(smem + (1 + G) * T);
  for (int i = static_cast<int>(j); i < 2 * a.nterms; i += static_cast<int>(T))
    reinterpret_cast<double2*>(lterms)[i] = reinterpret_cast<const double2*>(a.terms)[i];
  __syncthreads();

  const double2* own = smem;
  const FlipTerm* tt = lterms;
  const double f = a.sdiag * dg;
  double2 acc = make_double2(f * xv.x, f * xv.y);

  // Every segment is padded by the host to a multiple of kChunk with inert
  // terms (mask 0, zero weights), so each chunk is straight-line code: its
  // table entries, then its LDS gathers, then its FMAs, with the memory
  // operations of the whole chunk in flight together. Without flip-both
  // weights (the dipolar case) an inactive lane simply sits out the gather
  // and the FMA under the exec mask.
  for (int t0 = 0; t0 < a.off[0]; t0 += kChunk) {
    double2 p[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) p[u] = own[j ^ tt[t0 + u].mask];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const FlipTerm ft = tt[t0 + u];
      field_term(acc, (j & ft.mask) != 0, ft.anti, ft.aligned, p[u]);
    }
  }
  // wave-uniform local pairs: scalar activity test
  for (int t0 = a.off[0]; t0 < a.off[1]; t0 += kChunk) {
    double c[kChunk];
    double2 p[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const FlipTerm ft = tt[t0 + u];
      c[u] = (__popc(wj & ft.mask) == 1) ? ft.anti : ft.aligned;
      if (c[u] != 0.0) p[u] = own[j ^ ft.mask];
    }
#pragma unroll
    for (int u = 0; u < kChunk; ++u)
      if (c[u] != 0.0) cmad(acc, c[u], p[u]);
  }
  // per-lane local pairs
  for (int t0 = a.off[1]; t0 < a.off[2]; t0 += kChunk) {
    double2 p[kChunk];
    bool on[kChunk];
    if (FLIPBOTH) {
      double c[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const FlipTerm ft = tt[t0 + u];
        c[u] = (__popc(j & ft.mask) == 1) ? ft.anti : ft.aligned;
        p[u] = own[j ^ ft.mask];
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u) cmad(acc, c[u], p[u]);
    } else {
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const uint32_t m = tt[t0 + u].mask;
        on[u] = __popc(j & m) == 1;
        if (on[u]) p[u] = own[j ^ m];
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u)
        if (on[u]) cmad(acc, tt[t0 + u].anti, p[u]);
    }
  }
  for (int r = 0; r < G; ++r) {
    const double2* nb = smem + (1 + r) * T;
    const bool xb = ((x >> r) & 1u) != 0;
    const int s0 = a.off[2 + 3 * r], s1 = a.off[3 + 3 * r], s2 = a.off[4 + 3 * r],
              s3 = a.off[5 + 3 * r];
    for (int t = s0; t < s1; ++t) {
      const FlipTerm ft = tt[t];
      field_term(acc, xb, ft.anti, ft.aligned, nb[j]);
    }
    for (int t0 = s1; t0 < s2; t0 += kChunk) {
      double c[kChunk];
      double2 p[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const FlipTerm ft = tt[t0 + u];
        c[u] = (((wj & ft.mask) != 0) != xb) ? ft.anti : ft.aligned;
        if (ft.mask == 0) c[u] = 0.0;
        if (c[u] != 0.0) p[u] = nb[j ^ ft.mask];
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u)
        if (c[u] != 0.0) cmad(acc, c[u], p[u]);
    }
    for (int t0 = s2; t0 < s3; t0 += kChunk) {
      double2 p[kChunk];
      bool on[kChunk];
      if (FLIPBOTH) {
        double c[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          const FlipTerm ft = tt[t0 + u];
          c[u] = (((j & ft.mask) != 0) != xb) ? ft.anti : ft.aligned;
          if (ft.mask == 0) c[u] = 0.0;
          p[u] = nb[j ^ ft.mask];
        }
#pragma unroll
        for (int u = 0; u < kChunk; ++u) cmad(acc, c[u], p[u]);
      } else {
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          const uint32_t m = tt[t0 + u].mask;
          on[u] = m != 0 && (((j & m) != 0) != xb);
          if (on[u]) p[u] = nb[j ^ m];
        }
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
          if (on[u]) cmad(acc, tt[t0 + u].anti, p[u]);
      }
    }
  }
  // pairs with both bits remote: the partner tile from global memory,
  // activity decided per tile
  const int e0 = a.off[2 + 3 * G], e1 = a.off[3 + 3 * G];
  for (int t0 = e0; t0 < e1; t0 += kChunk) {
    double c[kChunk];
    double2 p[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const FlipTerm ft = tt[t0 + u];
      c[u] = (__popc(x & ft.mask) == 1) ? ft.anti : ft.aligned;
      if (ft.mask == 0) c[u] = 0.0;
      if (c[u] != 0.0) p[u] = in[(static_cast<size_t>(x ^ ft.mask) << a.tb) + j];
    }
#pragma unroll
    for (int u = 0; u < kChunk; ++u)
      if (c[u] != 0.0) cmad(acc, c[u], p[u]);
  }

  if constexpr (KIND == K_APPLY) {
    a.out[v][k] = acc;
  } else if constexpr (KIND == K_ACCUM) {
    a.out[v][k] = make_double2(prev.x + acc.x, prev.y + acc.y);
  } else if constexpr (KIND == K_ORDER1) {
    a.out[v][k] = make_double2(0.5 * acc.x, 0.5 * acc.y);
  } else if constexpr (KIND == K_ORDER2) {
    const double2 t0 = psiv, t1 = xv;
    const double2 t2 = make_double2(acc.x - t0.x, acc.y - t0.y);
    a.out[v][k] = t2;
    double2 q = make_double2(a.a0 * t0.x, a.a0 * t0.y);
    q.x += a.c1r * t1.x - a.c1i * t1.y;
    q.y += a.c1r * t1.y + a.c1i * t1.x;
    q.x += a.cr * t2.x - a.ci * t2.y;
    q.y += a.cr * t2.y + a.ci * t2.x;
    a.psi[v][k] = q;
  } else {
    const double2 tn = make_double2(acc.x - prev.x, acc.y - prev.y);
    a.out[v][k] = tn;
    psiv.x += a.cr * tn.x - a.ci * tn.y;
    psiv.y += a.cr * tn.y + a.ci * tn.x;
    a.psi[v][k] = psiv;
  }
}

// ------------------------------------------------------ dense-pair kernel
// The hot configuration: every pair (a < b) may carry a flip-flop weight
// (all-to-all dipolar couplings), no flip-both weight and no transverse
// field. The pair structure is then fixed by L alone, so it is compiled in:
// each pair's partner offset, its tile class (own tile / neighbour tile /
// partner tile in global memory) and whether its activity test is per lane
// or per wavefront are constants; only the L(L-1)/2 weights are read, eight
// per scalar load. One amplitude per thread, tiles of 2^TB amplitudes.
// Pair q of (0,1),(0,2),...,(0,L-1),(1,2),...: its two spins.
constexpr int pair_a(int L, int q) {
  int x = 0;
  while (q >= L - 1 - x) {
    q -= L - 1 - x;
    ++x;
  }
  return x;
}
constexpr int pair_b(int L, int q) {
  int x = 0;
  while (q >= L - 1 - x) {
    q -= L - 1 - x;
    ++x;
  }
  return x + 1 + q;
}

struct DenseLane {
  uint32_t j, wj, x;
  const double2* smem;
  const double2* in;
};

// Gather of pair Q for this lane: partner from the own tile, a neighbour
// tile (both in LDS) or the partner tile in global memory.
template <int L, int TB, int Q>
__device__ __forceinline__ void dense_gather(const DenseLane& d, double c, bool& on, double2& p) {
  constexpr int NP = L * (L - 1) / 2;
  if constexpr (Q < NP) {
    constexpr int A = pair_a(L, Q), B = pair_b(L, Q);
    constexpr int T = 1 << TB;
    const uint32_t bA = A < 6 ? (d.j >> A) & 1u : A < TB ? (d.wj >> A) & 1u : (d.x >> (A - TB)) & 1u;
    const uint32_t bB = B < 6 ? (d.j >> B) & 1u : B < TB ? (d.wj >> B) & 1u : (d.x >> (B - TB)) & 1u;
    on = (bA != bB) && c != 0.0;
    if (on) {
      if constexpr (B < TB) {
        p = d.smem[d.j ^ ((1u << A) | (1u << B))];
      } else if constexpr (A < TB) {
        p = d.smem[(1 + B - TB) * T + (d.j ^ (1u << A))];
      } else {
        p = d.in[(static_cast<size_t>(d.x ^ ((1u << (A - TB)) | (1u << (B - TB)))) << TB) + d.j];
      }
    }
  } else {
    on = false;
  }
}

template <int L, int TB, int CH, int... U>
__device__ __forceinline__ void dense_chunk(std::integer_sequence<int, U...>, const DenseLane& d,
                                            const double* __restrict__ coef, double2& acc) {
  constexpr int NP = L * (L - 1) / 2;
  __builtin_amdgcn_sched_barrier(0);
  double c[sizeof...(U)];
  double2 p[sizeof...(U)];
  bool on[sizeof...(U)];
  ((c[U] = (CH * 8 + U < NP) ? coef[CH * 8 + U] : 0.0), ...);
  __builtin_amdgcn_sched_barrier(0);
  (dense_gather<L, TB, CH * 8 + U>(d, c[U], on[U], p[U]), ...);
  ((on[U] ? cmad(acc, c[U], p[U]) : void()), ...);
}

// Pairs are split into those whose partner is in LDS (first NL in pair
// order) and those with both bits remote (the last C(L-TB, 2)), whose partner
// tiles come from global memory. Each remote chunk is issued four LDS chunks
// before it is consumed, so its L2/MALL latency hides behind LDS work.
struct RemoteSlot {
  double c[8];
  double2 p[8];
  bool on[8];
};

template <int L, int TB>
struct DenseShape {
  static constexpr int NP = L * (L - 1) / 2;
  static constexpr int NR = (L - TB) * (L - TB - 1) / 2;
  static constexpr int NL = NP - NR;
  static constexpr int NCL = (NL + 7) / 8;
  static constexpr int NCR = (NR + 7) / 8;
};

template <int L, int TB, int CH, int... U>
__device__ __forceinline__ void lds_chunk(std::integer_sequence<int, U...>, const DenseLane& d,
                                          const double* __restrict__ coef, double2& acc) {
  constexpr int NL = DenseShape<L, TB>::NL;
  __builtin_amdgcn_sched_barrier(0);
  double c[8];
  double2 p[8];
  bool on[8];
  ((c[U] = (CH * 8 + U < NL) ? coef[CH * 8 + U] : 0.0), ...);
  __builtin_amdgcn_sched_barrier(0);
  ((CH * 8 + U < NL ? dense_gather<L, TB, CH * 8 + U>(d, c[U], on[U], p[U]) : void(on[U] = false)), ...);
  ((on[U] ? cmad(acc, c[U], p[U]) : void()), ...);
}

template <int L, int TB, int R, int... U>
__device__ __forceinline__ void remote_issue(std::integer_sequence<int, U...>, const DenseLane& d,
                                             const double* __restrict__ coef, RemoteSlot& s) {
  constexpr int NL = DenseShape<L, TB>::NL, NP = DenseShape<L, TB>::NP;
  __builtin_amdgcn_sched_barrier(0);
  ((s.c[U] = (NL + R * 8 + U < NP) ? coef[NL + R * 8 + U] : 0.0), ...);
  __builtin_amdgcn_sched_barrier(0);
  (dense_gather<L, TB, NL + R * 8 + U>(d, s.c[U], s.on[U], s.p[U]), ...);
}

template <int... U>
__device__ __forceinline__ void remote_consume(std::integer_sequence<int, U...>, const RemoteSlot& s,
                                               double2& acc) {
  __builtin_amdgcn_sched_barrier(0);
  ((s.on[U] ? cmad(acc, s.c[U], s.p[U]) : void()), ...);
}

template <int L, int TB, int CH>
__device__ __forceinline__ void dense_step(const DenseLane& d, const double* __restrict__ coef,
                                           double2& acc, RemoteSlot& slot) {
  using S = DenseShape<L, TB>;
  constexpr auto U8 = std::make_integer_sequence<int, 8>{};
  if constexpr (CH % 4 == 0 && CH / 4 < S::NCR) remote_issue<L, TB, CH / 4>(U8, d, coef, slot);
  lds_chunk<L, TB, CH>(U8, d, coef, acc);
  if constexpr (CH % 4 == 3 && CH / 4 < S::NCR) remote_consume(U8, slot, acc);
}

// remote chunks whose issue or consume fell past the last LDS chunk
template <int L, int TB, int R>
__device__ __forceinline__ void dense_tail(const DenseLane& d, const double* __restrict__ coef,
                                           double2& acc, RemoteSlot& slot) {
  using S = DenseShape<L, TB>;
  constexpr auto U8 = std::make_integer_sequence<int, 8>{};
  if constexpr (R * 4 + 3 >= S::NCL) {
    if constexpr (R * 4 >= S::NCL) remote_issue<L, TB, R>(U8, d, coef, slot);
    remote_consume(U8, slot, acc);
  }
}

template <int L, int TB, int... CH, int... R>
__device__ __forceinline__ void dense_all(std::integer_sequence<int, CH...>, std::integer_sequence<int, R...>,
                                          const DenseLane& d, const double* __restrict__ coef, double2& acc) {
  RemoteSlot slot;
  (dense_step<L, TB, CH>(d, coef, acc, slot), ...);
  (dense_tail<L, TB, R>(d, coef, acc, slot), ...);
}

// Per vector (blockIdx.y): a launch may carry the states of several
// Hamiltonians of one size.
struct KDArgs {
  const double2* in[kMaxMany];
  double2* out[kMaxMany];
  double2* psi[kMaxMany];
  const double* diag[kMaxMany];
  const double* coef[kMaxMany];   // [L(L-1)/2] flip-flop weights, pair order (0,1),(0,2),...
  int xcd_remap;
  double sdiag[kMaxMany];
  double a0[kMaxMany], c1r[kMaxMany], c1i[kMaxMany], cr[kMaxMany], ci[kMaxMany];
};

template <int KIND, int L, int TB>
__global__ __launch_bounds__(1 << TB) void k_dense(KDArgs a) {
  constexpr int T = 1 << TB;
  constexpr int G = L - TB;
  extern __shared__ double2 smem[];
  const uint32_t j = threadIdx.x;
  const uint32_t wj = __builtin_amdgcn_readfirstlane(j) & ~63u;
  const int v = blockIdx.y;
  const double2* __restrict__ in = a.in[v];
  uint32_t x = blockIdx.x;
  if (a.xcd_remap && G >= 3) x = ((x & 7u) << (G - 3)) | (x >> 3);
  const size_t base = static_cast<size_t>(x) << TB;
  const size_t k = base + j;

  const double2 xv = in[k];
  double2 nbv[G > 0 ? G : 1];
#pragma unroll
  for (int r = 0; r < G; ++r) nbv[r] = in[(static_cast<size_t>(x ^ (1u << r)) << TB) + j];
  const double dg = a.diag[v][k];
  double2 prev = make_double2(0.0, 0.0), psiv = make_double2(0.0, 0.0);
  if constexpr (KIND == K_ACCUM || KIND == K_ORDERN) prev = a.out[v][k];
  if constexpr (KIND == K_ORDER2 || KIND == K_ORDERN) psiv = a.psi[v][k];
  smem[j] = xv;
#pragma unroll
  for (int r = 0; r < G; ++r) smem[(1 + r) * T + j] = nbv[r];
  __syncthreads();

  const double f = a.sdiag[v] * dg;
  double2 acc = make_double2(f * xv.x, f * xv.y);
  const DenseLane d{j, wj, x, smem, in};
  dense_all<L, TB>(std::make_integer_sequence<int, DenseShape<L, TB>::NCL>{},
                   std::make_integer_sequence<int, DenseShape<L, TB>::NCR>{}, d, a.coef[v], acc);

  if constexpr (KIND == K_APPLY) {
    a.out[v][k] = acc;
  } else if constexpr (KIND == K_ACCUM) {
    a.out[v][k] = make_double2(prev.x + acc.x, prev.y + acc.y);
  } else if constexpr (KIND == K_ORDER1) {
    a.out[v][k] = make_double2(0.5 * acc.x, 0.5 * acc.y);
  } else if constexpr (KIND == K_ORDER2) {
    const double2 t0 = psiv, t1 = xv;
    const double2 t2 = make_double2(acc.x - t0.x, acc.y - t0.y);
    a.out[v][k] = t2;
    double2 qv = make_double2(a.a0[v] * t0.x, a.a0[v] * t0.y);
    qv.x += a.c1r[v] * t1.x - a.c1i[v] * t1.y;
    qv.y += a.c1r[v] * t1.y + a.c1i[v] * t1.x;
    qv.x += a.cr[v] * t2.x - a.ci[v] * t2.y;
    qv.y += a.cr[v] * t2.y + a.ci[v] * t2.x;
    a.psi[v][k] = qv;
  } else {
    const double2 tn = make_double2(acc.x - prev.x, acc.y - prev.y);
    a.out[v][k] = tn;
    psiv.x += a.cr[v] * tn.x - a.ci[v] * tn.y;
    psiv.y += a.cr[v] * tn.y + a.ci[v] * tn.x;
    a.psi[v][k] = psiv;
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
#pragma clang fp contract(off)
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

// addScaled (src/core/StateVector.cpp:21-32). The BLAS-1 kernels keep the
// reference's separate multiply and add (no FMA), so they are bitwise its own.
__global__ __launch_bounds__(kBlock) void k_add_scaled(double2* __restrict__ y, double cr, double ci,
                                                       const double2* __restrict__ x, size_t n) {
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
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

// The rotation of one pair (l, h) of a spin: the arithmetic of k_rotate.
__device__ __forceinline__ void rotate_pair(double2& a, double2& b, const Rot& u) {
#pragma clang fp contract(off)
  const double llr = u.m[0], lli = u.m[1], lhr = u.m[2], lhi = u.m[3];
  const double hlr = u.m[4], hli = u.m[5], hhr = u.m[6], hhi = u.m[7];
  const double2 x = a, y = b;
  a = make_double2((llr * x.x - lli * x.y) + (lhr * y.x - lhi * y.y),
                   (llr * x.y + lli * x.x) + (lhr * y.y + lhi * y.x));
  b = make_double2((hlr * x.x - hli * x.y) + (hhr * y.x - hhi * y.y),
                   (hlr * x.y + hli * x.x) + (hhr * y.y + hhi * y.x));
}

// The same rotation on every target spin below tb (mask: target bits), in
// ascending order, on a tile of 2^tb contiguous amplitudes in LDS: one pass
// over the state instead of one per target. Each pair sees exactly the
// arithmetic and the order of per-target k_rotate launches.
__global__ __launch_bounds__(256) void k_rotate_lo(double2* __restrict__ s, uint32_t tb, uint32_t mask, Rot u) {
  __shared__ double2 t[1024];
  const uint32_t T = 1u << tb;
  const size_t base = static_cast<size_t>(blockIdx.x) << tb;
  for (uint32_t i = threadIdx.x; i < T; i += 256) t[i] = s[base + i];
  __syncthreads();
  for (uint32_t b = 0; b < tb; ++b) {
    if (!((mask >> b) & 1u)) continue;   // uniform
    for (uint32_t q = threadIdx.x; q < T / 2; q += 256) {
      const uint32_t l = static_cast<uint32_t>(insert_zero(q, b)), h = l | (1u << b);
      rotate_pair(t[l], t[h], u);
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < T; i += 256) s[base + i] = t[i];
}

// The target spins from tb up (mask: target bits above tb, shifted down by
// tb; g = L - tb of them at most, 2^g <= 1024 rows), in ascending order: a
// workgroup takes 1024 / 2^g consecutive low indices (columns) and all 2^g
// rows through LDS.
__global__ __launch_bounds__(256) void k_rotate_hi(double2* __restrict__ s, uint32_t tb, uint32_t g, uint32_t mask,
                                                   Rot u) {
  __shared__ double2 t[1024];
  const uint32_t R = 1u << g, C = 1024u >> g;
  const size_t c0 = static_cast<size_t>(blockIdx.x) * C;
  for (uint32_t i = threadIdx.x; i < R * C; i += 256) {
    const uint32_t r = i / C, j = i % C;
    t[i] = s[(static_cast<size_t>(r) << tb) + c0 + j];
  }
  __syncthreads();
  for (uint32_t b = 0; b < g; ++b) {
    if (!((mask >> b) & 1u)) continue;   // uniform
    for (uint32_t q = threadIdx.x; q < R * C / 2; q += 256) {
      const uint32_t rl = static_cast<uint32_t>(insert_zero(q / C, b)), j = q % C;
      rotate_pair(t[rl * C + j], t[(rl | (1u << b)) * C + j], u);
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < R * C; i += 256) {
    const uint32_t r = i / C, j = i % C;
    s[(static_cast<size_t>(r) << tb) + c0 + j] = t[i];
  }
}

struct Gate4 {
  double m[32];
};

// TwoSpinGate::apply (src/core/Gate.cpp:90-121).
__global__ __launch_bounds__(kBlock) void k_two_spin(double2* __restrict__ s, size_t quarter,
                                                     uint32_t lo, uint32_t hi, Gate4 u) {
#pragma clang fp contract(off)
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

// The fixed-order reduction of op over n terms into dst[0..1] on the
// device (stream-ordered, no host wait).
template <class Op>
int reduce_dev(spx_ctx* c, Op op, size_t n, double* dst) {
  size_t nb = (n + kReduceThreads - 1) / kReduceThreads;
  if (nb > static_cast<size_t>(kReduceBlocks)) nb = kReduceBlocks;
  if (nb == 0) nb = 1;
  double2* partials = reinterpret_cast<double2*>(c->d_partials);
  k_reduce1<Op><<<static_cast<unsigned>(nb), kReduceThreads, 0, c->stream>>>(op, n, partials);
  k_reduce2<<<1, kReduceThreads, 0, c->stream>>>(partials, static_cast<int>(nb), dst);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <class Op>
int reduce(spx_ctx* c, Op op, size_t n, double* re, double* im) {
  SPX_TRY(reduce_dev(c, op, n, c->d_result));
  SPX_HIP(hipMemcpyAsync(c->h_result, c->d_result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  SPX_HIP(hipStreamSynchronize(c->stream));
  if (re) *re = c->h_result[0];
  if (im) *im = c->h_result[1];
  return SPX_OK;
}

}  // namespace

// --------------------------------------------------------------- launchers

static int env_int(const char* name, int fallback) {
  const char* e = std::getenv(name);
  return e ? std::atoi(e) : fallback;
}

// Tile bits (2^tb amplitudes staged in LDS per workgroup), threads per
// workgroup and the XCD-aware tile order: the Hamiltonian's schedule when
// set, else SPX_TILE_BITS / SPX_BLOCK / SPX_XCD_REMAP (sweeps), else 10 /
// 256 / off.
int tile_bits_for(const spx_ham* h) {
  static const int forced = env_int("SPX_TILE_BITS", 0);
  int tb = h->sched.tb > 0 ? h->sched.tb : forced > 0 ? forced : 10;
  if (tb > 12) tb = 12;
  if (tb < 6) tb = 6;
  if (static_cast<size_t>(tb) > h->nspins) tb = static_cast<int>(h->nspins);
  return tb;
}

int block_threads(const spx_ham* h) {
  static const int env = env_int("SPX_BLOCK", 256);
  const int b = h->sched.block > 0 ? h->sched.block : env;
  return b >= 64 && b <= 1024 ? b : 256;
}

int xcd_remap(const spx_ham* h) {
  static const int env = env_int("SPX_XCD_REMAP", 0);
  return h->sched.xcd >= 0 ? h->sched.xcd : env;
}

// The two specialised paths and their weights: the all-pairs gather kernel
// (dense = 1; L = 13..18, flip-flop couplings only: no transverse field and
// no flip-both weight at this scale) and the Hadamard transforms (dense = 2;
// L = 13..20, any couplings and fields). SPX_DENSE=0 (or a schedule with
// dense = 0) switches both off; without either, default_dense_path chooses.
int dense_tile_bits(const spx_ham* h) {
  static const int env = env_int("SPX_DENSE_TB", 10);
  const int tb = h->sched.dense_tb > 0 ? h->sched.dense_tb : env;
  return tb == 9 ? 9 : 10;
}

// Measured per launch (dipolar couplings, all states of the launch, µs,
// gathers / Hadamard): L = 15, width 8: 12.5 / 18.8; L = 16, width 2:
// 11.1 / 13.7, width 4: 17.3 / 22.2, width 8: 28.7 / 22.6; L = 17, width 2:
// 15.9 / 19.3, width 4: 37.7 / 30.4, width 8: 55.7 / 36.5; L = 18, width 1:
// 17.4 / 23.1, width 2: 35.1 / 24.4; from L = 19 only the Hadamard path
// exists. At L = 13 the all-pairs kernel replaces the table kernels, whose
// sixteen-state launches fill only 128 workgroups there: 64-echo CPMG
// ensembles of 128 realisations ran 9.6x faster (13.5 -> 130 per second).
// Sixteen states per launch (lockstep ensemble groups, 64-echo CPMG),
// realisations per second, gathers / Hadamard: L = 14, 82.7 / 61.8;
// L = 15, 29.2 / 40.1.
int default_dense_path(size_t L, int width, bool flipflop_only) {
  if (!flipflop_only) return hadamard_eligible(L) ? 2 : 0;
  if (L >= 19) return 2;
  if (L == 18) return width >= 2 ? 2 : 1;
  if (L == 17) return width >= 4 ? 2 : 1;
  if (L == 16) return width >= 8 ? 2 : 1;
  if (L == 15) return width >= 16 ? 2 : 1;
  return 1;
}

// Whether the Hamiltonian has flip-flop couplings only (no transverse field,
// jx = jy on every pair): the case in which the default path depends on the
// launch width. A heuristic on the raw values; the path itself is settled in
// prepare_dense after the skip rule.
bool ham_flipflop_only(const spx_ham* h) {
  for (size_t s = 0; s < h->nspins; ++s)
    if (h->hx[s] != 0.0 || h->hy[s] != 0.0) return false;
  for (size_t p = 0; p < h->lo.size(); ++p)
    if (h->jx[p] != h->jy[p]) return false;
  return true;
}

void set_launch_width(spx_ham* h, int width) {
  if (h->width == width) return;
  const bool ff = ham_flipflop_only(h);
  const bool changes = default_dense_path(h->nspins, h->width, ff) != default_dense_path(h->nspins, width, ff);
  h->width = width;
  if (changes) h->terms_valid = false;
}

static int prepare_dense(spx_ham* h, double scale) {
  static const int env = env_int("SPX_DENSE", -1);
  const size_t L = h->nspins;
  h->dense_ok = false;
  h->wht_ok = false;
  if (L < 2) return SPX_OK;
  // weights at this scale after the skip rule (as the flip table takes them)
  const size_t np = L * (L - 1) / 2;
  std::vector<double> anti(np, 0.0), aligned(np, 0.0), cx(L, 0.0), cy(L, 0.0);
  bool flipflop_only = true;
  for (size_t s = 0; s < L; ++s) {
    const double x = scale * 0.5 * h->hx[s], y = scale * 0.5 * h->hy[s];
    if (!(std::abs(x) > kNegligible || std::abs(y) > kNegligible)) continue;
    cx[s] = x;
    cy[s] = y;
    flipflop_only = false;
  }
  for (size_t p = 0; p < h->lo.size(); ++p) {
    const double quarter = scale * 0.25;
    double ff = quarter * (h->jx[p] + h->jy[p]);
    double fb = quarter * (h->jx[p] - h->jy[p]);
    if (!(std::abs(ff) > kNegligible)) ff = 0.0;
    if (!(std::abs(fb) > kNegligible)) fb = 0.0;
    if (fb != 0.0) flipflop_only = false;
    const size_t lo = h->lo[p], hi = h->hi[p];
    const size_t q = lo * L - lo * (lo + 1) / 2 + (hi - lo - 1);
    anti[q] = ff;
    aligned[q] = fb;
  }
  const int path = h->sched.dense >= 0 ? h->sched.dense
                   : env >= 0          ? env
                                       : default_dense_path(L, h->width, flipflop_only);
  if (path == 2 && hadamard_eligible(L)) {
    // the diagonal's terms after ham_prepare's skip rule
    std::vector<double> dfield(L, 0.0), dpair(np, 0.0);
    for (size_t s = 0; s < L; ++s) {
      const double cz = 0.5 * h->hz[s];
      if (std::abs(cz) > kNegligible) dfield[s] = cz;
    }
    for (size_t p = 0; p < h->lo.size(); ++p) {
      const double cz = 0.25 * h->jz[p];
      const size_t lo = h->lo[p], hi = h->hi[p];
      if (std::abs(cz) > kNegligible) dpair[lo * L - lo * (lo + 1) / 2 + (hi - lo - 1)] = cz;
    }
    SPX_TRY(hadamard_prepare(h, anti, aligned, cx, cy, dfield, dpair));
    h->wht_ok = true;
    return SPX_OK;
  }
  if (path != 1 || !flipflop_only || L < 13 || L > 18) return SPX_OK;
  if (!h->d_coef) SPX_HIP(hipMalloc(&h->d_coef, 32 * 31 / 2 * sizeof(double)));
  SPX_TRY(copy_to_device(h->ctx, h->d_coef, anti.data(), np * sizeof(double)));
  h->dense_ok = true;
  return SPX_OK;
}

// The cached sigma_z diagonal (rebuildDiagonal, src/core/Hamiltonian.cpp:54-80),
// rebuilt when hz or the couplings changed since it was built. Only the
// gather and table kernels read it (the Hadamard path forms the diagonal
// from small factor tables), so it is built on demand.
int ensure_diag(spx_ham* h) {
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
    SPX_TRY(copy_to_device(c, h->d_tab, tab.data(), tab.size() * sizeof(double)));
    k_build_diag<<<blocks_for(n), kBlock, 0, c->stream>>>(h->d_diag, n, h->d_tab, nf, np);
    SPX_HIP(hipGetLastError());
    h->diag_valid = true;
    h->diag_hz = h->hz;
    h->diag_jz = h->jz;
    h->diag_lo = h->lo;
    h->diag_hi = h->hi;
  }
  return SPX_OK;
}

static int prepare_terms(spx_ham* h, double scale);

int ham_prepare(spx_ham* h, double scale) {
  SPX_TRY(prepare_terms(h, scale));
  if (!h->wht_ok) SPX_TRY(ensure_diag(h));
  return SPX_OK;
}

// The launch tables of the order kernels at `scale` (field, pair and
// diagonal weights after the skip rule), rebuilt when any field or coupling
// or the scale changed.
static int prepare_terms(spx_ham* h, double scale) {
  spx_ctx* c = h->ctx;
  const size_t L = h->nspins;
  if (h->terms_valid && h->terms_scale == scale) return SPX_OK;
  SPX_TRY(prepare_dense(h, scale));
  if (h->wht_ok) {
    // the Hadamard path reads its factor tables only: no flip table
    h->terms_scale = scale;
    h->terms_valid = true;
    return SPX_OK;
  }

  const int tb = tile_bits_for(h);
  const int g = static_cast<int>(L) - tb;
  // the padded table is at most L fields + L(L-1)/2 pairs plus kChunk per segment
  const size_t table_max = sizeof(FlipTerm) * (L + L * (L - 1) / 2 + kChunk * (4 + 2 * (g > 0 ? g : 0)));
  if (h->sched.table != 2 && tb <= 10 && g <= kMaxG3 &&
      static_cast<size_t>(1 + (g > 0 ? g : 0)) * (sizeof(double2) << tb) + table_max <= 160 * 1024) {
    // v3 layout (see K3Args::off)
    std::vector<FlipTerm> seg[3];
    std::vector<std::vector<FlipTerm>> rseg(3 * (g > 0 ? g : 0));
    std::vector<FlipTerm> both;
    for (size_t s = 0; s < L; ++s) {
      const double cx = scale * 0.5 * h->hx[s];
      const double cy = scale * 0.5 * h->hy[s];
      if (!(std::abs(cx) > kNegligible || std::abs(cy) > kNegligible)) continue;
      if (static_cast<int>(s) < tb) seg[0].push_back(FlipTerm{1u << s, 1u, cx, cy, 0.0});
      else rseg[3 * (s - tb)].push_back(FlipTerm{0u, 1u, cx, cy, 0.0});
    }
    for (size_t p = 0; p < h->lo.size(); ++p) {
      const double quarter = scale * 0.25;
      double ff = quarter * (h->jx[p] + h->jy[p]);
      double fb = quarter * (h->jx[p] - h->jy[p]);
      if (!(std::abs(ff) > kNegligible)) ff = 0.0;
      if (!(std::abs(fb) > kNegligible)) fb = 0.0;
      if (ff == 0.0 && fb == 0.0) continue;
      const int lo = static_cast<int>(h->lo[p]), hi = static_cast<int>(h->hi[p]);
      if (hi < tb) {
        seg[lo >= 6 ? 1 : 2].push_back(FlipTerm{(1u << lo) | (1u << hi), 0u, ff, fb, 0.0});
      } else if (lo < tb) {
        rseg[3 * (hi - tb) + (lo >= 6 ? 1 : 2)].push_back(FlipTerm{1u << lo, 0u, ff, fb, 0.0});
      } else {
        both.push_back(FlipTerm{(1u << (lo - tb)) | (1u << (hi - tb)), 0u, ff, fb, 0.0});
      }
    }
    // pad every chunked segment to a multiple of kChunk with inert terms
    // (mask 0, zero weights); the field segments of remote bits are walked
    // term by term and stay unpadded
    h->terms_flipboth = 0;
    auto append = [&](const std::vector<FlipTerm>& sg, bool pad) {
      for (const FlipTerm& t : sg)
        if (t.kind == 0 && t.aligned != 0.0) h->terms_flipboth = 1;
      const size_t start = h->h_terms.size();
      h->h_terms.insert(h->h_terms.end(), sg.begin(), sg.end());
      if (pad)  // the chunk loops step from the segment's own start
        while ((h->h_terms.size() - start) % kChunk != 0)
          h->h_terms.push_back(FlipTerm{0u, 0u, 0.0, 0.0, 0.0});
      h->terms_off.push_back(static_cast<int>(h->h_terms.size()));
    };
    h->h_terms.clear();
    h->terms_off.clear();
    for (auto& sg : seg) append(sg, true);
    for (size_t i = 0; i < rseg.size(); ++i) append(rseg[i], i % 3 != 0);
    append(both, true);
    h->terms_variant = 3;
    h->nterms = static_cast<int>(h->h_terms.size());
    h->terms_g = g > 0 ? g : 0;
    h->terms_tb = tb;
    h->terms_scale = scale;
    h->terms_valid = true;
  } else {
  std::vector<FlipTerm> local;
  std::vector<std::vector<FlipTerm>> group(g > 0 ? g : 0);
  std::vector<FlipTerm> both;
  // local fields first, then pairs, each in the reference's order
  for (size_t s = 0; s < L; ++s) {
    const double cx = scale * 0.5 * h->hx[s];
    const double cy = scale * 0.5 * h->hy[s];
    if (!(std::abs(cx) > kNegligible || std::abs(cy) > kNegligible)) continue;
    if (static_cast<int>(s) < tb) local.push_back(FlipTerm{1u << s, 1u, cx, cy, 0.0});
    else group[s - tb].push_back(FlipTerm{0u, 1u, cx, cy, 0.0});
  }
  for (size_t p = 0; p < h->lo.size(); ++p) {
    const double quarter = scale * 0.25;
    double ff = quarter * (h->jx[p] + h->jy[p]);
    double fb = quarter * (h->jx[p] - h->jy[p]);
    if (!(std::abs(ff) > kNegligible)) ff = 0.0;
    if (!(std::abs(fb) > kNegligible)) fb = 0.0;
    if (ff == 0.0 && fb == 0.0) continue;
    const int lo = static_cast<int>(h->lo[p]), hi = static_cast<int>(h->hi[p]);
    if (hi < tb) local.push_back(FlipTerm{(1u << lo) | (1u << hi), 0u, ff, fb, 0.0});
    else if (lo < tb) group[hi - tb].push_back(FlipTerm{1u << lo, 0u, ff, fb, 0.0});
    else both.push_back(FlipTerm{(1u << (lo - tb)) | (1u << (hi - tb)), 0u, ff, fb, 0.0});
  }
  h->h_terms.clear();
  h->h_terms.insert(h->h_terms.end(), local.begin(), local.end());
  h->terms_grp.assign(1, static_cast<int>(h->h_terms.size()));
  for (int r = 0; r < g; ++r) {
    h->h_terms.insert(h->h_terms.end(), group[r].begin(), group[r].end());
    h->terms_grp.push_back(static_cast<int>(h->h_terms.size()));
  }
  h->h_terms.insert(h->h_terms.end(), both.begin(), both.end());
  h->nterms = static_cast<int>(h->h_terms.size());
  h->terms_nlocal = static_cast<int>(local.size());
  h->terms_nboth = static_cast<int>(both.size());
  h->terms_g = g > 0 ? g : 0;
  h->terms_tb = tb;
  h->terms_scale = scale;
  h->terms_valid = true;
  h->terms_variant = 2;
  }
  const size_t bytes = (h->h_terms.size() + 1) * sizeof(FlipTerm);
  if (bytes > h->terms_cap) {
    if (h->d_terms) SPX_HIP(hipFree(h->d_terms));
    SPX_HIP(hipMalloc(&h->d_terms, bytes));
    h->terms_cap = bytes;
  }
  if (!h->h_terms.empty()) {
    SPX_TRY(copy_to_device(c, h->d_terms, h->h_terms.data(), h->h_terms.size() * sizeof(FlipTerm)));
  }
  static_assert(sizeof(FlipTerm) == 32, "flip term layout");
  return SPX_OK;
}

template <int KIND>
static int launch_kind(spx_ctx* c, const KArgs& a, dim3 grid, unsigned block, int amps, size_t lds) {
  switch (amps) {
    case 1: k_horder<KIND, 1><<<grid, block, lds, c->stream>>>(a); break;
    case 2: k_horder<KIND, 2><<<grid, block, lds, c->stream>>>(a); break;
    case 4: k_horder<KIND, 4><<<grid, block, lds, c->stream>>>(a); break;
    case 8: k_horder<KIND, 8><<<grid, block, lds, c->stream>>>(a); break;
    default: k_horder<KIND, 16><<<grid, block, lds, c->stream>>>(a); break;
  }
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <int KIND, int G>
static void launch3g(spx_ctx* c, const K3Args& a, dim3 grid, unsigned block, size_t lds) {
  if (a.flipboth) k_order3<KIND, G, true><<<grid, block, lds, c->stream>>>(a);
  else k_order3<KIND, G, false><<<grid, block, lds, c->stream>>>(a);
}

template <int KIND>
static int launch3(spx_ctx* c, const K3Args& a, dim3 grid, unsigned block, size_t lds) {
  switch (a.g) {
    case 0: launch3g<KIND, 0>(c, a, grid, block, lds); break;
    case 1: launch3g<KIND, 1>(c, a, grid, block, lds); break;
    case 2: launch3g<KIND, 2>(c, a, grid, block, lds); break;
    case 3: launch3g<KIND, 3>(c, a, grid, block, lds); break;
    case 4: launch3g<KIND, 4>(c, a, grid, block, lds); break;
    case 5: launch3g<KIND, 5>(c, a, grid, block, lds); break;
    case 6: launch3g<KIND, 6>(c, a, grid, block, lds); break;
    case 7: launch3g<KIND, 7>(c, a, grid, block, lds); break;
    case 8: launch3g<KIND, 8>(c, a, grid, block, lds); break;
    case 9: launch3g<KIND, 9>(c, a, grid, block, lds); break;
    case 10: launch3g<KIND, 10>(c, a, grid, block, lds); break;
    case 11: launch3g<KIND, 11>(c, a, grid, block, lds); break;
    default: launch3g<KIND, 12>(c, a, grid, block, lds); break;
  }
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

static int launch_k3(spx_ham* h, int kind, const KArgs& b, unsigned nvec) {
  K3Args a{};
  for (int i = 0; i < kMaxMany; ++i) {
    a.in[i] = b.in[i];
    a.out[i] = b.out[i];
    a.psi[i] = b.psi[i];
  }
  a.diag = b.diag;
  a.terms = h->d_terms;
  for (size_t i = 0; i < h->terms_off.size(); ++i) a.off[i] = h->terms_off[i];
  a.tb = h->terms_tb;
  a.g = h->terms_g;
  a.nterms = h->nterms;
  a.flipboth = h->terms_flipboth;
  a.xcd_remap = b.xcd_remap;
  a.sdiag = b.sdiag;
  a.a0 = b.a0;
  a.c1r = b.c1r;
  a.c1i = b.c1i;
  a.cr = b.cr;
  a.ci = b.ci;
  spx_ctx* c = h->ctx;
  const size_t n = static_cast<size_t>(1) << h->nspins;
  const unsigned T = 1u << a.tb;
  const dim3 grid(static_cast<unsigned>(n >> a.tb), nvec);
  const size_t lds = sizeof(double2) * T * (1 + a.g) + sizeof(FlipTerm) * a.nterms;
  switch (kind) {
    case K_APPLY: return launch3<K_APPLY>(c, a, grid, T, lds);
    case K_ACCUM: return launch3<K_ACCUM>(c, a, grid, T, lds);
    case K_ORDER1: return launch3<K_ORDER1>(c, a, grid, T, lds);
    case K_ORDER2: return launch3<K_ORDER2>(c, a, grid, T, lds);
    default: return launch3<K_ORDERN>(c, a, grid, T, lds);
  }
}

template <int KIND, int L, int TB>
static void launch_dense_lt(spx_ctx* c, const KDArgs& a, unsigned nvec) {
  const dim3 grid(1u << (L - TB), nvec);
  const size_t lds = sizeof(double2) * (static_cast<size_t>(1 + L - TB) << TB);
  k_dense<KIND, L, TB><<<grid, 1u << TB, lds, c->stream>>>(a);
}

// Tile of 2^10 amplitudes (one workgroup of 1024 threads per CU) or 2^9
// (two of 512); SPX_DENSE_TB selects.
template <int KIND, int L>
static void launch_dense_l(spx_ctx* c, const KDArgs& a, unsigned nvec, int tb) {
  if (tb == 9) launch_dense_lt<KIND, L, 9>(c, a, nvec);
  else launch_dense_lt<KIND, L, 10>(c, a, nvec);
}

template <int KIND>
static int launch_dense_kind(spx_ham* h, const KDArgs& a, unsigned nvec) {
  spx_ctx* c = h->ctx;
  const int tb = dense_tile_bits(h);
  switch (h->nspins) {
    case 13: launch_dense_l<KIND, 13>(c, a, nvec, tb); break;
    case 14: launch_dense_l<KIND, 14>(c, a, nvec, tb); break;
    case 15: launch_dense_l<KIND, 15>(c, a, nvec, tb); break;
    case 16: launch_dense_l<KIND, 16>(c, a, nvec, tb); break;
    case 17: launch_dense_l<KIND, 17>(c, a, nvec, tb); break;
    default: launch_dense_l<KIND, 18>(c, a, nvec, tb); break;
  }
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

static int launch_dense_args(spx_ham* h, int kind, const KDArgs& a, unsigned nvec) {
  switch (kind) {
    case K_ORDER1: return launch_dense_kind<K_ORDER1>(h, a, nvec);
    case K_ORDER2: return launch_dense_kind<K_ORDER2>(h, a, nvec);
    case K_ORDERN: return launch_dense_kind<K_ORDERN>(h, a, nvec);
    case K_ACCUM: return launch_dense_kind<K_ACCUM>(h, a, nvec);
    default: return launch_dense_kind<K_APPLY>(h, a, nvec);
  }
}

// One Hamiltonian for every vector of the launch.
static int launch_dense(spx_ham* h, int kind, const KArgs& b, unsigned nvec) {
  KDArgs a{};
  for (int i = 0; i < kMaxMany; ++i) {
    a.in[i] = b.in[i];
    a.out[i] = b.out[i];
    a.psi[i] = b.psi[i];
    a.diag[i] = b.diag;
    a.coef[i] = h->d_coef;
    a.sdiag[i] = b.sdiag;
    a.a0[i] = b.a0;
    a.c1r[i] = b.c1r;
    a.c1i[i] = b.c1i;
    a.cr[i] = b.cr;
    a.ci[i] = b.ci;
  }
  a.xcd_remap = b.xcd_remap;
  return launch_dense_args(h, kind, a, nvec);
}

static int launch_wht(spx_ham* h, int kind, const KArgs& b, unsigned nvec) {
  HadamardLaunch a{};
  for (int i = 0; i < kMaxMany; ++i) {
    a.in[i] = b.in[i];
    a.out[i] = b.out[i];
    a.psi[i] = b.psi[i];
    a.ham[i] = h;
    a.slot[i] = i;
    a.sdiag[i] = b.sdiag;
    a.w[i] = OrderWeights{b.a0, b.c1r, b.c1i, b.cr, b.ci};
  }
  return launch_hadamard(kind, a, nvec);
}

static int launch_k(spx_ham* h, int kind, const KArgs& a, unsigned nvec) {
  if (h->wht_ok) return launch_wht(h, kind, a, nvec);
  if (h->dense_ok) return launch_dense(h, kind, a, nvec);
  if (h->terms_variant == 3) return launch_k3(h, kind, a, nvec);
  spx_ctx* c = h->ctx;
  const size_t n = static_cast<size_t>(1) << h->nspins;
  const uint32_t T = 1u << a.tb;
  unsigned block = static_cast<unsigned>(block_threads(h));
  if (block > T) block = T;
  int amps = static_cast<int>(T / block);
  if (amps > 16) {
    amps = 16;
    block = T / 16;
  }
  const dim3 grid(static_cast<unsigned>(n >> a.tb), nvec);
  const size_t lds = sizeof(double2) * (a.g > 0 ? 2 * T : T);
  switch (kind) {
    case K_APPLY: return launch_kind<K_APPLY>(c, a, grid, block, amps, lds);
    case K_ACCUM: return launch_kind<K_ACCUM>(c, a, grid, block, amps, lds);
    case K_ORDER1: return launch_kind<K_ORDER1>(c, a, grid, block, amps, lds);
    case K_ORDER2: return launch_kind<K_ORDER2>(c, a, grid, block, amps, lds);
    default: return launch_kind<K_ORDERN>(c, a, grid, block, amps, lds);
  }
}

static KArgs base_args(spx_ham* h, double sdiag) {
  KArgs a{};
  a.diag = h->d_diag;
  a.terms = h->d_terms;
  a.nlocal = h->terms_nlocal;
  a.g = h->terms_g;
  a.nboth = h->terms_nboth;
  if (h->terms_variant == 2)
    for (int r = 0; r <= h->terms_g; ++r) a.grp[r] = h->terms_grp[r];
  a.tb = h->terms_tb;
  a.xcd_remap = xcd_remap(h);
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
  if (h->wht_ok) {
    HadamardLaunch b{};
    for (int i = 0; i < v.nvec; ++i) {
      b.in[i] = v.in[i];
      b.out[i] = v.out[i];
      b.psi[i] = v.psi[i];
      b.ham[i] = h;
      b.slot[i] = i;
      b.sdiag[i] = h->terms_scale;
      b.w[i] = OrderWeights{a0, c1r, c1i, cr, ci};
      b.next[i] = v.uv_next[i];
    }
    b.ready = v.uv_ready;
    return launch_hadamard(k, b, static_cast<unsigned>(v.nvec));
  }
  return launch_k(h, k, a, static_cast<unsigned>(v.nvec));
}

bool order_set_compatible(spx_ham* const* hs, size_t n) {
  if (n == 0 || !hs[0] || !hs[0]->terms_valid) return false;
  const spx_ham* h = hs[0];
  for (size_t i = 1; i < n; ++i) {
    const spx_ham* g = hs[i];
    if (!g || !g->terms_valid || g->nspins != h->nspins || g->ctx != h->ctx) return false;
    if (h->wht_ok) {
      if (!g->wht_ok) return false;
    } else if (h->dense_ok) {
      if (!g->dense_ok || dense_tile_bits(g) != dense_tile_bits(h) || xcd_remap(g) != xcd_remap(h)) return false;
    } else if (g != h) {
      return false;
    }
  }
  return true;
}

int launch_order_set(spx_ham* const* hs, int kind, const OrderLaunch& v, const OrderWeights* w) {
  const unsigned nvec = static_cast<unsigned>(v.nvec);
  if (nvec < 1 || nvec > static_cast<unsigned>(kMaxMany) || !order_set_compatible(hs, nvec))
    return fail(SPX_EINVAL, "these Hamiltonians cannot share an order launch");
  spx_ham* h = hs[0];
  const int k = kind == 1 ? K_ORDER1 : kind == 2 ? K_ORDER2 : K_ORDERN;
  if (h->wht_ok) {
    HadamardLaunch a{};
    for (unsigned i = 0; i < nvec; ++i) {
      a.in[i] = v.in[i];
      a.out[i] = v.out[i];
      a.psi[i] = v.psi[i];
      a.ham[i] = hs[i];
      int slot = 0;
      for (unsigned j = 0; j < i; ++j) slot += hs[j] == hs[i] ? 1 : 0;
      a.slot[i] = slot;
      a.sdiag[i] = hs[i]->terms_scale;
      a.w[i] = w[i];
      a.next[i] = v.uv_next[i];
    }
    a.ready = v.uv_ready;
    return launch_hadamard(k, a, nvec);
  }
  if (h->dense_ok) {
    KDArgs a{};
    for (unsigned i = 0; i < nvec; ++i) {
      if (w[i].defer != 0) return fail(SPX_EINVAL, "batched accumulator updates need the Hadamard path");
      a.in[i] = v.in[i];
      a.out[i] = v.out[i];
      a.psi[i] = v.psi[i];
      a.diag[i] = hs[i]->d_diag;
      a.coef[i] = hs[i]->d_coef;
      a.sdiag[i] = hs[i]->terms_scale;
      a.a0[i] = w[i].a0;
      a.c1r[i] = w[i].c1r;
      a.c1i[i] = w[i].c1i;
      a.cr[i] = w[i].cr;
      a.ci[i] = w[i].ci;
    }
    a.xcd_remap = xcd_remap(h);
    return launch_dense_args(h, k, a, nvec);
  }
  for (unsigned i = 0; i < nvec; ++i)
    if (w[i].a0 != w[0].a0 || w[i].c1r != w[0].c1r || w[i].c1i != w[0].c1i || w[i].cr != w[0].cr ||
        w[i].ci != w[0].ci || w[i].defer != 0)
      return fail(SPX_EINVAL, "the table kernels take one set of weights per launch");
  return launch_order(h, kind, v, w[0].a0, w[0].c1r, w[0].c1i, w[0].cr, w[0].ci);
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

// SpinRotation::apply over a target list (src/core/Gate.cpp:60-85): targets
// in ascending order without repeats, L <= 20, run as one LDS pass over the
// low 10 bits and one over the rest; any other list, one launch per target
// in list order.
int launch_rotate_targets(spx_ctx* c, double2* s, size_t L, const uint32_t* targets, size_t nt,
                          const double* u8) {
  bool ascending = L <= 20;
  for (size_t i = 1; i < nt && ascending; ++i) ascending = targets[i] > targets[i - 1];
  if (!ascending || nt < 2) {
    for (size_t i = 0; i < nt; ++i) SPX_TRY(launch_rotate(c, s, size_t(1) << L, targets[i], u8));
    return SPX_OK;
  }
  Rot u;
  for (int i = 0; i < 8; ++i) u.m[i] = u8[i];
  const uint32_t tb = L < 10 ? static_cast<uint32_t>(L) : 10u;
  uint32_t lo = 0, hi = 0;
  for (size_t i = 0; i < nt; ++i) {
    if (targets[i] < tb) lo |= 1u << targets[i];
    else hi |= 1u << (targets[i] - tb);
  }
  if (lo) {
    k_rotate_lo<<<static_cast<unsigned>((size_t(1) << L) >> tb), 256, 0, c->stream>>>(s, tb, lo, u);
    SPX_HIP(hipGetLastError());
  }
  if (hi) {
    const uint32_t g = static_cast<uint32_t>(L) - tb;
    const uint32_t C = 1024u >> g;
    k_rotate_hi<<<static_cast<unsigned>((size_t(1) << tb) / C), 256, 0, c->stream>>>(s, tb, g, hi, u);
    SPX_HIP(hipGetLastError());
  }
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

// The same sums left on the device at dst (two doubles each: re, im; the
// norm's sum of squares, not its square root), for callers that read many
// at once.
int reduce_norm2_dev(spx_ctx* c, const double2* x, size_t n, double* dst) {
  return reduce_dev(c, OpNorm{x}, n, dst);
}

int reduce_corr_dev(spx_ctx* c, int axis, const double2* state, const double2* ref, size_t L, double* dst) {
  return reduce_dev(c, OpCorr{axis, static_cast<uint32_t>(L), state, ref}, static_cast<size_t>(1) << L, dst);
}

int reduce_mag_dev(spx_ctx* c, int axis, uint32_t spin, const double2* state, size_t L, double* dst) {
  return reduce_dev(c, OpMag{axis, spin, state}, static_cast<size_t>(1) << L, dst);
}

// Energy::measure (src/core/Observable.cpp:112-120): H psi into the
// Hamiltonian's scratch vector, then Re <H psi|psi> by the fixed-order
// reduction.
int reduce_energy(spx_ham* h, const double2* state, double* out) {
  const size_t n = static_cast<size_t>(1) << h->nspins;
  if (!h->d_scratch) SPX_HIP(hipMalloc(&h->d_scratch, n * sizeof(double2)));
  SPX_TRY(launch_apply(h, 1.0, state, h->d_scratch, false));
  return reduce(h->ctx, OpInner{state, h->d_scratch}, n, out, nullptr);
}

int reduce_energy_dev(spx_ham* h, const double2* state, double* dst) {
  const size_t n = static_cast<size_t>(1) << h->nspins;
  if (!h->d_scratch) SPX_HIP(hipMalloc(&h->d_scratch, n * sizeof(double2)));
  SPX_TRY(launch_apply(h, 1.0, state, h->d_scratch, false));
  return reduce_dev(h->ctx, OpInner{state, h->d_scratch}, n, dst);
}

}  // namespace spx
