// The off-diagonal part of H through Walsh-Hadamard transforms (gfx950, FP64).
//
// In the Pauli basis of the reference's kernels (spin s = bit s, set bit =
// up), the coupled pair (i, j) contributes anti (|ud> <-> |du>) and aligned
// (|dd> <-> |uu>) weights (scale (jx + jy) / 4 and scale (jx - jy) / 4) and a
// transverse field cx sigma_x + cy sigma_y (scale hx / 2, scale hy / 2),
// applied term by term in src/core/Hamiltonian.cpp:95-127 and
// src/core/SpinOperators.cpp:211-291. With X, Y the standard Pauli matrices
// in the bit basis (the reference's sigma_y is -Y there),
//
//   anti (s+s- + s-s+) + aligned (s+s+ + s-s-)
//       = (anti + aligned)/2 X_i X_j + (anti - aligned)/2 Y_i Y_j,
//
// and Y = S X S^+ with S = diag(i^popcount(k)). X strings are diagonal in the
// Hadamard basis (W X_i W = 2^L Z_i, W the unnormalised Walsh-Hadamard
// transform on all L bits), so the whole off-diagonal part is
//
//   H_off psi = W A W psi + S W B W (S^+ psi),
//   A(s) = [sum_q (anti + aligned)/2 z_i z_j + sum_s cx_s z_s] / 2^L,
//   B(s) = [sum_q (anti - aligned)/2 z_i z_j - sum_s cy_s z_s] / 2^L,
//
// z = +1 for bit 0, -1 for bit 1; 2^-L is exact. For flip-flop couplings
// alone (aligned = 0, no transverse field) A = B.
//
// This costs O(L) butterflies per amplitude instead of one partner gather
// per term (L(L-1)/2 pairs for all-to-all couplings, plus L fields during
// pulses), and every butterfly is unmasked. The 2^L index is split into a
// row part (the low TB = 10 bits, a contiguous tile) and a column part (the
// high G = L - TB bits); W = W_high W_low, so one Chebyshev order is three
// passes:
//   1. rows:    U = W_low psi, V = W_low (S^+ psi)            per tile, in LDS
//   2. columns: U = W_high A W_high U, V = W_high B W_high V  8 columns per workgroup
//   3. rows:    H psi = sdiag diag psi + W_low U + S W_low V, fused with the
//               same Chebyshev / accumulate epilogues as the gather kernels.
// In a Chebyshev expansion the input of order n+1 is the output T_n of order
// n, so pass 3 also runs pass 1 on T_n while it is in registers: an order
// after the first costs two launches.
// A and B are not stored per amplitude: with k = (x, c) (x the G column-pass
// bits, c the TB row bits), A(k) = A_lo(c) + A_hi(x) + sum_j z_j(x) A_j(c),
// where A_lo / A_hi collect the pairs and fields inside c / inside x and A_j
// the pairs across (likewise B); the tables hold 2^TB + 2^G + G 2^TB values
// each and stay in L2. The diagonal is not read per amplitude either:
// diag(x, c) = D_lo(c) + D_hi(x) + sum_{i < TB} z_i(c) K_i(x), with
// K_i(x) = sum_j (pair weight of i and TB + j) z_j(x), from a table of
// 2^TB + 2^G (1 + TB) values; pass 3 reads D_hi(x) and K(x) once per tile.
// The transforms are exact sums and differences, the phases i^k exact swaps
// and sign changes; results agree with the gather kernels to rounding (a few
// ulps of the amplitude scale).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <utility>

#include "spx_common.hpp"

namespace spx {
namespace {

constexpr int kTB = 10;                // row (tile) bits
constexpr int kCols = 8;               // columns per workgroup of the column pass
constexpr int kRowThreads = 1 << (kTB - 2);   // two vectors x 2^(TB-3) threads

// Everything per vector (blockIdx.y): a launch may carry the states of
// several Hamiltonians of one size.
struct WhtArgs {
  const double2* in[kMaxMany];
  double2* out[kMaxMany];
  double2* psi[kMaxMany];
  double2* u[kMaxMany];         // scratch: W A W psi (row transforms between passes)
  double2* w[kMaxMany];         // scratch: W B W S^+ psi
  const double* dz[kMaxMany];   // diagonal factors: D_lo[2^TB], then per x: D_hi, K_0..K_{TB-1}
  const double* e[kMaxMany];    // A factors: A_lo[2^TB], A_hi[2^G], A_j[G][2^TB]
  const double* eb[kMaxMany];   // B factors, same layout
  double sdiag[kMaxMany];
  double a0[kMaxMany], c1r[kMaxMany], c1i[kMaxMany], cr[kMaxMany], ci[kMaxMany];
  double pr[kMaxMany], pi[kMaxMany], qr[kMaxMany], qi[kMaxMany];
  int defer[kMaxMany];          // accumulator batching (OrderWeights::defer)
  int next[kMaxMany];           // pass 3 also leaves the row transforms of its output
  int xcd;                      // each vector's workgroups on one XCD (vector count a multiple of 8)
};

// Vector-per-XCD placement. The dispatcher hands the workgroups of these
// launches (linear id w, x fastest) to XCD w mod 8 in turn, so when the
// vector count is a multiple of 8 the item of workgroup w is taken as vector
// (w mod 8) + 8 floor(w / 8 / per), item (w / 8) mod per, where per is a
// vector's workgroup count: all of a vector's workgroups, in every pass, run
// on one XCD, whose L2 keeps what one pass writes for the next. Off, or for
// any other placement by the dispatcher, every item still runs exactly once.
__device__ __forceinline__ void xcd_item(int on, uint32_t per, uint32_t& v, uint32_t& k) {
  if (!on) {
    v = blockIdx.y;
    k = blockIdx.x;
    return;
  }
  const uint32_t w = blockIdx.x + gridDim.x * blockIdx.y, q = w >> 3;
  v = (w & 7u) + 8u * (q / per);
  k = q % per;
}

// q += (cr + i ci) t, the one form every accumulator update takes
__device__ __forceinline__ void caxpy(double2& q, double cr, double ci, double2 t) {
  q.x = fma(-ci, t.y, fma(cr, t.x, q.x));
  q.y = fma(ci, t.x, fma(cr, t.y, q.y));
}

enum { K_APPLY = 0, K_ACCUM = 1, K_ORDER1 = 2, K_ORDER2 = 3, K_ORDERN = 4 };

// v i^e (e taken mod 4): a swap and sign changes, exact.
__device__ __forceinline__ double2 mul_ipow(double2 v, uint32_t e) {
  const bool odd = e & 1u, neg = e & 2u;
  double re = odd ? -v.y : v.x, im = odd ? v.x : v.y;
  return make_double2(neg ? -re : re, neg ? -im : im);
}
// S^+ and S at amplitude k: i^-popcount(k) and i^popcount(k)
__device__ __forceinline__ uint32_t sdag(uint32_t k) { return 0u - static_cast<uint32_t>(__popc(k)); }
__device__ __forceinline__ uint32_t sphase(uint32_t k) { return static_cast<uint32_t>(__popc(k)); }

// LDS slot of element i of a row tile: bits 0-2 XOR bits 3-5, so that in
// every round eight consecutive lanes touch eight distinct 16-byte slots of
// one 128-byte bank line.
__device__ __forceinline__ uint32_t rslot(uint32_t i) { return i ^ ((i >> 3) & 7u); }

// Global loads of a pass are all issued into registers before the first
// dependent LDS write or global store: the compiler cannot prove that the
// state pointers do not alias, and would otherwise serialise the memory
// round trips (load, wait, store, next load).

// Butterflies of one round on the thread's eight values v[r] (r = its three
// register bits): stages s in [S0, S1).
template <int S0, int S1>
__device__ __forceinline__ void butterflies(double2 (&v)[8]) {
#pragma unroll
  for (int s = S0; s < S1; ++s) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (r & (1 << s)) continue;
      const double2 a = v[r], b = v[r | (1 << s)];
      v[r] = make_double2(a.x + b.x, a.y + b.y);
      v[r | (1 << s)] = make_double2(a.x - b.x, a.y - b.y);
    }
  }
}

// Row tiles with index bits 8 and 9 in registers (the four values a thread
// holds, k = t + 256 m): those two butterfly stages run on the registers and
// only bits 0..7 go through LDS, in three rounds over four 256-value batches.
__device__ __forceinline__ void bfly_m(double2 (&x)[4]) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      if (m & (1 << s)) continue;
      const double2 a = x[m], b = x[m | (1 << s)];
      x[m] = make_double2(a.x + b.x, a.y + b.y);
      x[m | (1 << s)] = make_double2(a.x - b.x, a.y - b.y);
    }
}
// One LDS round over index bits [LO, LO + 3) of a 256-value batch: the thread
// (batch, sub) takes eight values into registers and runs stages S0..2 of
// them (a stage below S0 was done by the round before).
template <int LO, int S0>
__device__ __forceinline__ void low8_round(double2* __restrict__ S, uint32_t batch, uint32_t sub) {
  const uint32_t base = (batch << 8) | ((sub >> LO) << (LO + 3)) | (sub & ((1u << LO) - 1u));
  double2 v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = S[rslot(base | (static_cast<uint32_t>(r) << LO))];
  butterflies<S0, 3>(v);
#pragma unroll
  for (int r = 0; r < 8; ++r) S[rslot(base | (static_cast<uint32_t>(r) << LO))] = v[r];
  __syncthreads();
}
// u = thread index within the half (128 threads, 8 values each)
__device__ __forceinline__ void wht_low8(double2* S, uint32_t u) {
  const uint32_t batch = u >> 5, sub = u & 31u;
  low8_round<0, 0>(S, batch, sub);
  low8_round<3, 0>(S, batch, sub);
  low8_round<5, 1>(S, batch, sub);
}

// Pass 1: one workgroup per row tile x (2^TB consecutive amplitudes) and
// vector; threads [0, 128) transform psi, [128, 256) transform S^+ psi.
__global__ __launch_bounds__(kRowThreads) void k_wht_rows_fwd(WhtArgs a) {
  constexpr uint32_t T = 1u << kTB, NT = kRowThreads, H = NT / 2, M = T / NT;
  extern __shared__ double2 smem[];
  const uint32_t t = threadIdx.x;
  uint32_t vv, tile;
  xcd_item(a.xcd, gridDim.x, vv, tile);
  const int v = static_cast<int>(vv);
  const size_t base = static_cast<size_t>(tile) << kTB;
  const double2* __restrict__ in = a.in[v];
  double2 x[M];
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) x[m] = in[base + t + m * NT];
  static_assert(M == 4 && NT == 256, "bits 8 and 9 of the row tile are the register index m");
  double2 yu[M], yv[M];
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    const uint32_t i = t + m * NT;
    yu[m] = x[m];
    yv[m] = mul_ipow(x[m], sdag(static_cast<uint32_t>(base + i)));
  }
  bfly_m(yu);
  bfly_m(yv);
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    smem[rslot(t + m * NT)] = yu[m];
    smem[T + rslot(t + m * NT)] = yv[m];
  }
  __syncthreads();
  const uint32_t half = t >= H ? 1u : 0u;
  wht_low8(smem + half * T, t - half * H);
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    yu[m] = smem[rslot(t + m * NT)];
    yv[m] = smem[T + rslot(t + m * NT)];
  }
  double2* __restrict__ U = a.u[v];
  double2* __restrict__ V = a.w[v];
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    U[base + t + m * NT] = yu[m];
    V[base + t + m * NT] = yv[m];
  }
}

// Index of value r of a thread with free bits u in a round whose register
// bits are [LO, LO + 3).
template <int LO>
__device__ __forceinline__ uint32_t round_index(uint32_t u, int r) {
  return ((u >> LO) << (LO + 3)) | (u & ((1u << LO) - 1u)) | (static_cast<uint32_t>(r) << LO);
}

// Column-pass round k, register bits [lo, lo + 3): the forward transform
// takes bits 0, 3, 6, ... (the last round aligned to the top, its lower
// stages already done); the inverse starts in the forward's last layout with
// all three stages and walks down to bit 0.
constexpr int fwd_lo(int G, int k) { return 3 * k + 3 <= G ? 3 * k : G - 3; }
constexpr int fwd_s0(int G, int k) { return 3 * k - fwd_lo(G, k); }
constexpr int inv_rounds(int G) { return 1 + (G - 3 + 2) / 3; }
constexpr int inv_top(int G, int k) { return k == 0 ? G : G - 3 - 3 * (k - 1); }   // bits [lo, top)
constexpr int inv_lo(int G, int k) { return k == 0 ? G - 3 : (inv_top(G, k) - 3 > 0 ? inv_top(G, k) - 3 : 0); }

template <int LO>
__device__ __forceinline__ void col_store(double2* S, uint32_t u, uint32_t c, const double2 (&v)[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) S[(round_index<LO>(u, r) << 3) | c] = v[r];
}
template <int LO>
__device__ __forceinline__ void col_load(const double2* S, uint32_t u, uint32_t c, double2 (&v)[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = S[(round_index<LO>(u, r) << 3) | c];
}

// forward rounds 1 .. NF-1 (round 0 runs on the values loaded from memory)
template <int G, int K>
__device__ __forceinline__ void col_forward(double2* S, uint32_t u, uint32_t c, double2 (&v)[8]) {
  constexpr int NF = (G + 2) / 3;
  if constexpr (K < NF) {
    col_store<fwd_lo(G, K - 1)>(S, u, c, v);
    __syncthreads();
    col_load<fwd_lo(G, K)>(S, u, c, v);
    butterflies<fwd_s0(G, K), 3>(v);
    col_forward<G, K + 1>(S, u, c, v);
  }
}
// inverse rounds 1 .. NI-1 (round 0 runs in registers after the product with A or B)
template <int G, int K>
__device__ __forceinline__ void col_inverse(double2* S, uint32_t u, uint32_t c, double2 (&v)[8]) {
  if constexpr (K < inv_rounds(G)) {
    col_store<inv_lo(G, K - 1)>(S, u, c, v);
    __syncthreads();
    col_load<inv_lo(G, K)>(S, u, c, v);
    butterflies<0, inv_top(G, K) - inv_lo(G, K)>(v);
    col_inverse<G, K + 1>(S, u, c, v);
  }
}

// Pass 2: one workgroup per group of 8 columns (low indices) and per scratch
// vector; 2^G threads (column c = t mod 8, free row bits u = t / 8) hold 8
// columns x 2^G rows and transform the rows in place: the first round on the
// values as loaded (register bits 0-2: 128-byte row segments), then through
// LDS; the product with A (U) or B (V) and the first inverse round in registers; the last
// inverse round (register bits 0-2 again) straight back to memory.
// Up to G = 8 the 32 KiB of LDS per workgroup allow five workgroups per CU;
// the register budget is held to match (at most 102 VGPRs), measured 2%
// faster at L = 18 than the compiler's own 107 VGPRs and four per CU.
template <int G>
__global__ __launch_bounds__(1 << G) __attribute__((amdgpu_waves_per_eu(G <= 8 ? 5 : 1))) void k_wht_cols(WhtArgs a) {
  static_assert(inv_lo(G, inv_rounds(G) - 1) == 0, "the inverse ends at bit 0");
  constexpr uint32_t T = 1u << kTB, NT = 1u << G;
  constexpr int LT = G - 3;   // register bits of the forward's last round
  extern __shared__ double2 smem[];
  const uint32_t t = threadIdx.x, c = t & 7u, u = t >> 3;
  // items: (vector, V half, column group); blockIdx.y = 2 vector + half
  uint32_t vv, item;
  if (a.xcd) {
    xcd_item(1, 2u * gridDim.x, vv, item);
  } else {
    vv = blockIdx.y >> 1;
    item = (blockIdx.y & 1u) * gridDim.x + blockIdx.x;
  }
  const int v = static_cast<int>(vv);
  const bool vhalf = item >= gridDim.x;
  double2* __restrict__ X = vhalf ? a.w[v] : a.u[v];
  const uint32_t col = (vhalf ? item - gridDim.x : item) * kCols + c;
  double2 x[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = X[(static_cast<size_t>(round_index<0>(u, r)) << kTB) | col];
  // A (or B) at the elements of the forward's last layout: rows u | r << LT
  const double* __restrict__ tab = vhalf ? a.eb[v] : a.e[v];
  double e[8];
  {
    double part = tab[col];
#pragma unroll
    for (int j = 0; j < LT; ++j) part += ((u >> j) & 1u) ? -tab[T + NT + j * T + col] : tab[T + NT + j * T + col];
    double top[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) top[j] = tab[T + NT + (LT + j) * T + col];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      double s = part + tab[T + round_index<LT>(u, r)];
#pragma unroll
      for (int j = 0; j < 3; ++j) s += ((r >> j) & 1) ? -top[j] : top[j];
      e[r] = s;
    }
  }
  butterflies<0, 3>(x);
  col_forward<G, 1>(smem, u, c, x);
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = make_double2(e[r] * x[r].x, e[r] * x[r].y);
  butterflies<0, 3>(x);
  col_inverse<G, 1>(smem, u, c, x);
#pragma unroll
  for (int r = 0; r < 8; ++r) X[(static_cast<size_t>(round_index<0>(u, r)) << kTB) | col] = x[r];
  (void)NT;
}

// Pass 3: the row transforms back, the diagonal term and the epilogue of
// the gather kernels (apply / accumulate / first, second and general
// Chebyshev order, src/core/Propagator.cpp:58-72). The epilogue's operands
// are loaded before the transform so that their latency overlaps it.
template <int KIND>
__global__ __launch_bounds__(kRowThreads) void k_wht_rows_bwd(WhtArgs a) {
  constexpr uint32_t T = 1u << kTB, NT = kRowThreads, H = NT / 2, M = T / NT;
  extern __shared__ double2 smem[];
  const uint32_t t = threadIdx.x;
  uint32_t vv, tile;
  xcd_item(a.xcd, gridDim.x, vv, tile);
  const int v = static_cast<int>(vv);
  const size_t base = static_cast<size_t>(tile) << kTB;
  // read here, rewritten in place at the end when the next order is chained
  const double2* U = a.u[v];
  const double2* V = a.w[v];
  const double2* __restrict__ in = a.in[v];
  double2* __restrict__ out = a.out[v];
  double2* __restrict__ psi = a.psi[v];
  const double* __restrict__ dlo = a.dz[v];
  const double* __restrict__ drow = a.dz[v] + T + static_cast<size_t>(tile) * (kTB + 1);
  const double sdiag = a.sdiag[v], a0 = a.a0[v], c1r = a.c1r[v], c1i = a.c1i[v], cr = a.cr[v], ci = a.ci[v];
  const int defer = KIND == K_ORDERN ? a.defer[v] : 0;   // uniform over the workgroup
  double2 xu[M], xw[M], xv[M], prev[M], pv[M];
  double dg[M];
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    xu[m] = U[base + t + m * NT];
    xw[m] = V[base + t + m * NT];
  }
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    const size_t k = base + t + m * NT;
    xv[m] = in[k];
    dg[m] = dlo[t + m * NT];
    if constexpr (KIND == K_ACCUM || KIND == K_ORDERN) prev[m] = out[k];
    if constexpr (KIND == K_ORDER2 || KIND == K_ORDERN)
      if (defer != 1) pv[m] = psi[k];
  }
  static_assert(M == 4 && NT == 256, "bits 8 and 9 of the row tile are the register index m");
  bfly_m(xu);
  bfly_m(xw);
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    smem[rslot(t + m * NT)] = xu[m];
    smem[T + rslot(t + m * NT)] = xw[m];
  }
  __syncthreads();
  const uint32_t half = t >= H ? 1u : 0u;
  wht_low8(smem + half * T, t - half * H);
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    xu[m] = smem[rslot(t + m * NT)];
    xw[m] = smem[T + rslot(t + m * NT)];
  }
  {
    // diag at c = t + m NT: the bits of t are shared by the thread's M
    // elements, the bits above come from m
    static_assert(M == 4 && NT == 256, "row tile split assumed below");
    double kr[kTB];
#pragma unroll
    for (int i = 0; i < kTB; ++i) kr[i] = drow[1 + i];
    double own = drow[0];
#pragma unroll
    for (int i = 0; i < 8; ++i) own += ((t >> i) & 1u) ? -kr[i] : kr[i];
#pragma unroll
    for (uint32_t m = 0; m < M; ++m) {
      const double up = ((m & 1u) ? -kr[8] : kr[8]) + ((m & 2u) ? -kr[9] : kr[9]);
      dg[m] = dg[m] + (own + up);
    }
  }
  double2 o[M];
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    const size_t k = base + t + m * NT;
    const double2 sv = mul_ipow(xw[m], sphase(static_cast<uint32_t>(k)));
    const double f = sdiag * dg[m];
    double2 acc;
    acc.x = fma(f, xv[m].x, xu[m].x + sv.x);
    acc.y = fma(f, xv[m].y, xu[m].y + sv.y);
    if constexpr (KIND == K_APPLY) {
      o[m] = acc;
    } else if constexpr (KIND == K_ACCUM) {
      o[m] = make_double2(prev[m].x + acc.x, prev[m].y + acc.y);
    } else if constexpr (KIND == K_ORDER1) {
      o[m] = make_double2(0.5 * acc.x, 0.5 * acc.y);
    } else if constexpr (KIND == K_ORDER2) {
      const double2 t0 = pv[m], t1 = xv[m];
      const double2 t2 = make_double2(acc.x - t0.x, acc.y - t0.y);
      o[m] = t2;
      double2 q = make_double2(a0 * t0.x, a0 * t0.y);
      caxpy(q, c1r, c1i, t1);
      caxpy(q, cr, ci, t2);
      pv[m] = q;
    } else {
      const double2 tn = make_double2(acc.x - prev[m].x, acc.y - prev[m].y);
      o[m] = tn;
      if (defer == 3) caxpy(pv[m], a.qr[v], a.qi[v], prev[m]);
      if (defer >= 2) caxpy(pv[m], a.pr[v], a.pi[v], xv[m]);
      if (defer != 1) caxpy(pv[m], cr, ci, tn);
    }
  }
  // T_n of a vector's last order (next != 1) is read by no later launch
  const bool keep = !(KIND == K_ORDER2 || KIND == K_ORDERN) || a.next[v] == 1;
#pragma unroll
  for (uint32_t m = 0; m < M; ++m) {
    const size_t k = base + t + m * NT;
    if (keep) out[k] = o[m];
    if constexpr (KIND == K_ORDER2 || KIND == K_ORDERN)
      if (defer != 1) psi[k] = pv[m];
  }
  if constexpr (KIND == K_ORDER1 || KIND == K_ORDER2 || KIND == K_ORDERN) {
    if (a.next[v]) {  // uniform over the workgroup
      // pass 1 of the next order on T_n, or (next = 2, a vector's last order)
      // of the next step's first order on the finished accumulator: every
      // thread has read its transform values out of LDS before the barrier
      __syncthreads();
#pragma unroll
      for (uint32_t m = 0; m < M; ++m) {
        const uint32_t i = t + m * NT;
        double2 src = o[m];
        if constexpr (KIND == K_ORDER2 || KIND == K_ORDERN)
          if (a.next[v] == 2) src = pv[m];
        xu[m] = src;
        xw[m] = mul_ipow(src, sdag(static_cast<uint32_t>(base + i)));
      }
      bfly_m(xu);
      bfly_m(xw);
#pragma unroll
      for (uint32_t m = 0; m < M; ++m) {
        smem[rslot(t + m * NT)] = xu[m];
        smem[T + rslot(t + m * NT)] = xw[m];
      }
      __syncthreads();
      wht_low8(smem + half * T, t - half * H);
#pragma unroll
      for (uint32_t m = 0; m < M; ++m) {
        xu[m] = smem[rslot(t + m * NT)];
        xw[m] = smem[T + rslot(t + m * NT)];
      }
      double2* __restrict__ Un = a.u[v];
      double2* __restrict__ Vn = a.w[v];
#pragma unroll
      for (uint32_t m = 0; m < M; ++m) {
        Un[base + t + m * NT] = xu[m];
        Vn[base + t + m * NT] = xw[m];
      }
    }
  }
}

// Pass 2 at G = 3 (L = 13): a column's eight values are the whole column
// transform, so one thread takes one column, runs the forward stages, the
// product with A (U) or B (V) and the inverse stages in registers, and
// consecutive threads take consecutive columns (whole 1 KiB rows per
// wavefront). The same arithmetic as k_wht_cols<3>, whose workgroups would
// be eight threads.
__global__ __launch_bounds__(256) void k_wht_cols3(WhtArgs a) {
  constexpr uint32_t T = 1u << kTB, NT = 8;
  const int v = static_cast<int>(blockIdx.y >> 1);
  const bool vhalf = blockIdx.y & 1u;
  double2* __restrict__ X = vhalf ? a.w[v] : a.u[v];
  const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  double2 x[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = X[(static_cast<size_t>(r) << kTB) | col];
  const double* __restrict__ tab = vhalf ? a.eb[v] : a.e[v];
  double e[8];
  {
    const double part = tab[col];
    double top[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) top[j] = tab[T + NT + j * T + col];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      double s = part + tab[T + r];
#pragma unroll
      for (int j = 0; j < 3; ++j) s += ((r >> j) & 1) ? -top[j] : top[j];
      e[r] = s;
    }
  }
  butterflies<0, 3>(x);
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = make_double2(e[r] * x[r].x, e[r] * x[r].y);
  butterflies<0, 3>(x);
#pragma unroll
  for (int r = 0; r < 8; ++r) X[(static_cast<size_t>(r) << kTB) | col] = x[r];
}

template <int G>
void launch_cols(hipStream_t st, const WhtArgs& a, unsigned nvec) {
  const dim3 grid((1u << kTB) / kCols, 2 * nvec);
  k_wht_cols<G><<<grid, 1u << G, sizeof(double2) * (kCols << G), st>>>(a);
}

template <int KIND>
void launch_bwd(hipStream_t st, const WhtArgs& a, dim3 grid) {
  k_wht_rows_bwd<KIND><<<grid, kRowThreads, sizeof(double2) * (2u << kTB), st>>>(a);
}

}  // namespace

// Largest size whose launches place each vector on one XCD by default.
constexpr size_t kWhtXcdMaxSpins = 17;

bool hadamard_eligible(size_t L) { return L >= kHadamardMinSpins && L <= kHadamardMaxSpins; }

// The factors of A and B (see the file comment) from the weights at the
// apply scale: pair q (order (0,1),(0,2),...) anti / aligned, spin s cx / cy.
// Sums over the unscaled combinations, then the exact scale 2^-L.
int hadamard_prepare(spx_ham* h, const std::vector<double>& anti, const std::vector<double>& aligned,
                     const std::vector<double>& cx, const std::vector<double>& cy,
                     const std::vector<double>& dfield, const std::vector<double>& dpair) {
  const size_t L = h->nspins;
  if (!hadamard_eligible(L)) return fail(SPX_EINVAL, "Hadamard path needs 13 <= L <= 20");
  const size_t T = size_t(1) << kTB, G = L - kTB, R = size_t(1) << G, TS = T + R + G * T;
  // tables already built for these weights: point at them
  std::vector<double> key;
  key.reserve(2 * anti.size() + 2 * L + L + dpair.size());
  for (const auto* v : {&anti, &aligned, &cx, &cy, &dfield, &dpair}) key.insert(key.end(), v->begin(), v->end());
  for (auto& t : h->wht_cache)
    if (t.key == key) {
      t.used = ++h->wht_clock;
      h->d_wht_e = t.dev;
      h->wht_tab = TS;
      return hadamard_reserve(h, 1);
    }
  const int sc = -static_cast<int>(L);
  auto q_of = [&](size_t lo, size_t hi) { return lo * L - lo * (lo + 1) / 2 + (hi - lo - 1); };
  auto z = [](size_t k, size_t bit) { return ((k >> bit) & 1u) ? -1.0 : 1.0; };
  const size_t DS = T + R * (kTB + 1);
  std::vector<double> tab(2 * TS + DS, 0.0);
  for (int which = 0; which < 2; ++which) {
    // which 0: A (X strings), 1: B (Y strings)
    auto pair = [&](size_t lo, size_t hi) {
      const size_t q = q_of(lo, hi);
      return which == 0 ? 0.5 * (anti[q] + aligned[q]) : 0.5 * (anti[q] - aligned[q]);
    };
    auto field = [&](size_t s) { return which == 0 ? cx[s] : -cy[s]; };
    double* t = tab.data() + which * TS;
    for (size_t c = 0; c < T; ++c) {
      double e = 0.0;
      for (size_t i = 0; i < static_cast<size_t>(kTB); ++i)
        for (size_t k = i + 1; k < static_cast<size_t>(kTB); ++k) e += pair(i, k) * z(c, i) * z(c, k);
      for (size_t s = 0; s < static_cast<size_t>(kTB); ++s) e += field(s) * z(c, s);
      t[c] = std::ldexp(e, sc);
    }
    for (size_t x = 0; x < R; ++x) {
      double e = 0.0;
      for (size_t i = 0; i < G; ++i)
        for (size_t k = i + 1; k < G; ++k) e += pair(kTB + i, kTB + k) * z(x, i) * z(x, k);
      for (size_t j = 0; j < G; ++j) e += field(kTB + j) * z(x, j);
      t[T + x] = std::ldexp(e, sc);
    }
    for (size_t j = 0; j < G; ++j)
      for (size_t c = 0; c < T; ++c) {
        double e = 0.0;
        for (size_t i = 0; i < static_cast<size_t>(kTB); ++i) e += pair(i, kTB + j) * z(c, i);
        t[T + R + j * T + c] = std::ldexp(e, sc);
      }
  }
  {
    // the diagonal (rebuildDiagonal's terms: field c at bit s adds -c z_s,
    // pair weight c adds c z_i z_j), unscaled
    double* d = tab.data() + 2 * TS;
    for (size_t c = 0; c < T; ++c) {
      double e = 0.0;
      for (size_t s = 0; s < static_cast<size_t>(kTB); ++s) e -= dfield[s] * z(c, s);
      for (size_t i = 0; i < static_cast<size_t>(kTB); ++i)
        for (size_t k = i + 1; k < static_cast<size_t>(kTB); ++k) e += dpair[q_of(i, k)] * z(c, i) * z(c, k);
      d[c] = e;
    }
    for (size_t x = 0; x < R; ++x) {
      double* row = d + T + x * (kTB + 1);
      double e = 0.0;
      for (size_t j = 0; j < G; ++j) e -= dfield[kTB + j] * z(x, j);
      for (size_t i = 0; i < G; ++i)
        for (size_t k = i + 1; k < G; ++k) e += dpair[q_of(kTB + i, kTB + k)] * z(x, i) * z(x, k);
      row[0] = e;
      for (size_t i = 0; i < static_cast<size_t>(kTB); ++i) {
        double kx = 0.0;
        for (size_t j = 0; j < G; ++j) kx += dpair[q_of(i, kTB + j)] * z(x, j);
        row[1 + i] = kx;
      }
    }
  }
  // A cache slot: a new one up to kWhtCache, else the least recently used.
  // The upload is ordered on the context's stream after every launch that
  // may still read the slot's old tables (a second stream of a split set is
  // joined back into it before the set evolve returns), so no host wait is
  // needed; only the pinned source is reused after its previous upload.
  constexpr size_t kWhtCache = 8;
  spx_ctx* c = h->ctx;
  spx_ham::WhtTables* slot = nullptr;
  if (h->wht_cache.size() < kWhtCache) {
    h->wht_cache.emplace_back();
    slot = &h->wht_cache.back();
  } else {
    slot = &h->wht_cache[0];
    for (auto& t : h->wht_cache)
      if (t.used < slot->used) slot = &t;
  }
  if (slot->count < tab.size()) {
    if (slot->copied) SPX_HIP(hipEventSynchronize(slot->copied));
    if (slot->dev) SPX_HIP(hipFree(slot->dev));
    if (slot->host) SPX_HIP(hipHostFree(slot->host));
    slot->dev = nullptr;
    slot->host = nullptr;
    slot->count = 0;
    SPX_HIP(hipMalloc(&slot->dev, tab.size() * sizeof(double)));
    SPX_HIP(hipHostMalloc(&slot->host, tab.size() * sizeof(double), hipHostMallocDefault));
    slot->count = tab.size();
  }
  if (!slot->copied) SPX_HIP(hipEventCreateWithFlags(&slot->copied, hipEventDisableTiming));
  SPX_HIP(hipEventSynchronize(slot->copied));   // the pinned source is free again
  std::copy(tab.begin(), tab.end(), slot->host);
  SPX_HIP(hipMemcpyAsync(slot->dev, slot->host, tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  SPX_HIP(hipEventRecord(slot->copied, c->stream));
  slot->key = std::move(key);
  slot->used = ++h->wht_clock;
  h->d_wht_e = slot->dev;
  h->wht_tab = TS;
  return hadamard_reserve(h, 1);
}

int hadamard_reserve(spx_ham* h, int slots) {
  if (slots < 1 || slots > kMaxMany) return fail(SPX_EINVAL, "scratch slots out of range");
  if (slots <= h->wht_slots) return SPX_OK;
  const size_t n = static_cast<size_t>(1) << h->nspins;
  if (h->d_wht) {
    SPX_HIP(hipStreamSynchronize(h->ctx->stream));  // launches in flight may still use it
    SPX_HIP(hipFree(h->d_wht));
    h->d_wht = nullptr;
    h->wht_slots = 0;
  }
  SPX_HIP(hipMalloc(&h->d_wht, 2 * static_cast<size_t>(slots) * n * sizeof(double2)));
  h->wht_slots = slots;
  return SPX_OK;
}

int launch_hadamard(int kind, const HadamardLaunch& b, unsigned nvec) {
  if (nvec < 1 || nvec > static_cast<unsigned>(kMaxMany)) return fail(SPX_EINVAL, "Hadamard launch out of range");
  const spx_ham* h0 = b.ham[0];
  const size_t L = h0->nspins, n = static_cast<size_t>(1) << L;
  if (!hadamard_eligible(L)) return fail(SPX_EINVAL, "Hadamard launch out of range");
  WhtArgs a{};
  {
    // placement of each vector on one XCD: SPX_WHT_XCD = 0 off, 1 on, else
    // on up to kWhtXcdMaxSpins
    static const int forced = [] {
      const char* e = std::getenv("SPX_WHT_XCD");
      return e ? std::atoi(e) : -1;
    }();
    const bool want = forced >= 0 ? forced != 0 : L <= kWhtXcdMaxSpins;
    a.xcd = want && nvec % 8 == 0 ? 1 : 0;
  }
  for (unsigned i = 0; i < nvec; ++i) {
    const spx_ham* h = b.ham[i];
    if (!h || h->nspins != L || !h->wht_ok || !h->d_wht || b.slot[i] < 0 || b.slot[i] >= h->wht_slots)
      return fail(SPX_EINVAL, "Hadamard launch mixes sizes, or its Hamiltonians lack prepared scratch");
    a.in[i] = b.in[i];
    a.out[i] = b.out[i];
    a.psi[i] = b.psi[i];
    a.u[i] = h->d_wht + (2 * static_cast<size_t>(b.slot[i])) * n;
    a.w[i] = h->d_wht + (2 * static_cast<size_t>(b.slot[i]) + 1) * n;
    a.dz[i] = h->d_wht_e + 2 * h->wht_tab;
    a.e[i] = h->d_wht_e;
    a.eb[i] = h->d_wht_e + h->wht_tab;
    a.sdiag[i] = b.sdiag[i];
    a.a0[i] = b.w[i].a0;
    a.c1r[i] = b.w[i].c1r;
    a.c1i[i] = b.w[i].c1i;
    a.cr[i] = b.w[i].cr;
    a.ci[i] = b.w[i].ci;
    a.pr[i] = b.w[i].pr;
    a.pi[i] = b.w[i].pi;
    a.qr[i] = b.w[i].qr;
    a.qi[i] = b.w[i].qi;
    a.defer[i] = b.w[i].defer;
    if (a.defer[i] < 0 || a.defer[i] > 3) return fail(SPX_EINVAL, "accumulator batching code out of range");
    a.next[i] = b.next[i];
    if (a.next[i] < 0 || a.next[i] > 2 || (a.next[i] == 2 && kind != K_ORDER2 && kind != K_ORDERN))
      return fail(SPX_EINVAL, "chained transforms of an accumulator need a Chebyshev order of 2 or more");
  }
  hipStream_t st = h0->ctx->stream;
  const dim3 rows(static_cast<unsigned>(n >> kTB), nvec);
  if (!b.ready) k_wht_rows_fwd<<<rows, kRowThreads, sizeof(double2) * (2u << kTB), st>>>(a);
  switch (L - kTB) {
    case 3: k_wht_cols3<<<dim3((1u << kTB) / 256, 2 * nvec), 256, 0, st>>>(a); break;
    case 4: launch_cols<4>(st, a, nvec); break;
    case 5: launch_cols<5>(st, a, nvec); break;
    case 6: launch_cols<6>(st, a, nvec); break;
    case 7: launch_cols<7>(st, a, nvec); break;
    case 8: launch_cols<8>(st, a, nvec); break;
    case 9: launch_cols<9>(st, a, nvec); break;
    default: launch_cols<10>(st, a, nvec); break;
  }
  switch (kind) {
    case K_APPLY: launch_bwd<K_APPLY>(st, a, rows); break;
    case K_ACCUM: launch_bwd<K_ACCUM>(st, a, rows); break;
    case K_ORDER1: launch_bwd<K_ORDER1>(st, a, rows); break;
    case K_ORDER2: launch_bwd<K_ORDER2>(st, a, rows); break;
    default: launch_bwd<K_ORDERN>(st, a, rows); break;
  }
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

}  // namespace spx
