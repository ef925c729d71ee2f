// Resident ensemble runner: one persistent launch runs many realisations of
// one pulse program (LocalRunner::run + Sequence::run, the reference's
// src/ensemble/EnsembleRunner.cpp:27-88 and src/sequence/Sequence.cpp:35-106).
//
// A workgroup pulls realisation indices from a queue (ordered by estimated
// work, largest first, on the host) and runs the whole program of one
// realisation with its state and reference vectors in registers; a batch of
// one size runs the walker instantiated for that L, a batch of several sizes
// one kernel that runs each realisation with the walker of its own L: 512
// threads, thread t owning amplitudes k = t + 512 i. Up to L = 10 each
// Chebyshev order stages T_{n-1} of the state and of the reference side by
// side in LDS for the partner gathers, under one pair of barriers. At L = 11
// and 12 the off-diagonal part goes through Walsh-Hadamard transforms
// instead (about 2.3x less LDS traffic than the gathers there), each vector
// held by 8 amplitudes per lane. An evolve step's pair and field weights and
// the pair masks are staged in LDS once per step. The weights are per
// realisation and per evolve step. All reductions are fixed-order block
// trees, so every realisation's numbers are independent of which workgroup
// ran it and of timing.
#include <utility>

#include "spx_batch.hpp"

namespace spx {
namespace {

constexpr int kBT = 512;  // threads per workgroup

struct RArgs {
  const DevOp* ops;
  int nops;
  const DevPulse* pulses;
  const DevEvolve* evo;
  const DevReal* reals;
  const int* order;
  int nreal;
  const double* w;
  const double* psi0;
  double* out;
  int width;
  const int* obs_codes;
  const int* obs_spins;
  int nobs;
  int ref_axis;
  int* queue;
};

constexpr int pa(int L, int q) {
  int x = 0;
  while (q >= L - 1 - x) {
    q -= L - 1 - x;
    ++x;
  }
  return x;
}
constexpr int pb(int L, int q) {
  int x = 0;
  while (q >= L - 1 - x) {
    q -= L - 1 - x;
    ++x;
  }
  return x + 1 + q;
}

__device__ __forceinline__ void cmadd(double2& acc, double c, double2 p) {
  acc.x = fma(c, p.x, acc.x);
  acc.y = fma(c, p.y, acc.y);
}

// Fixed-order block sum of two doubles (all 512 threads call it).
__device__ __forceinline__ double2 block_sum(double* red, double a, double b) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = a;
  red[kBT + t] = b;
  __syncthreads();
  for (int s = kBT / 2; s > 0; s >>= 1) {
    if (t < s) {
      red[t] += red[t + s];
      red[kBT + t] += red[kBT + t + s];
    }
    __syncthreads();
  }
  const double2 r = make_double2(red[0], red[kBT]);
  __syncthreads();
  return r;
}

template <int L>
struct Res {
  static constexpr int N = 1 << L;
  static constexpr int A = N > kBT ? N / kBT : 1;    // amplitudes per thread
  static constexpr int RB = N > kBT ? 9 : L;         // first register bit
  static constexpr int NP = L * (L - 1) / 2;
};

// LDS amplitudes of a realisation of L spins: up to L = 10 the state's and
// the reference's T_{n-1} side by side (for the gathers); at L = 11 and 12
// two transpose buffers of 2^12 amplitudes (for the transforms).
template <int L>
constexpr int kLdsAmps = L >= 11 ? 8192 : L == 10 ? 4096 : 2 * Res<L>::N;

// bit `b` of amplitude k = t + 512 i
template <int L>
__device__ __forceinline__ uint32_t kbit(uint32_t t, int i, int b) {
  return b < Res<L>::RB ? (t >> b) & 1u : (static_cast<uint32_t>(i) >> (b - Res<L>::RB)) & 1u;
}

// sigma_z part of the Pauli-basis operator at the own amplitudes, terms in
// the order of rebuildDiagonal (src/core/Hamiltonian.cpp:54-80).
template <int L>
__device__ void build_diag(double* dg, uint32_t t, const double* __restrict__ zw) {
  constexpr int A = Res<L>::A;
#pragma unroll
  for (int i = 0; i < A; ++i) dg[i] = 0.0;
  for (int s = 0; s < L; ++s) {
    const double c = zw[s];
    if (c == 0.0) continue;
#pragma unroll
    for (int i = 0; i < A; ++i) dg[i] += kbit<L>(t, i, s) ? c : -c;
  }
  int q = 0;
  for (int a = 0; a < L; ++a)
    for (int b = a + 1; b < L; ++b, ++q) {
      const double c = zw[L + q];
      if (c == 0.0) continue;
#pragma unroll
      for (int i = 0; i < A; ++i) dg[i] += (kbit<L>(t, i, a) == kbit<L>(t, i, b)) ? c : -c;
    }
}

template <int L>
__device__ __forceinline__ void stage(double2* lds, const double2* v, uint32_t t) {
  constexpr int A = Res<L>::A;
  __syncthreads();
  if (t < static_cast<uint32_t>(Res<L>::N))
#pragma unroll
    for (int i = 0; i < A; ++i) lds[t + kBT * i] = v[i];
  __syncthreads();
}

// Pair and field terms of H x at the own amplitudes k = t + 512 i for an
// evolve step whose weights sit in LDS: lw[q] = (anti, aligned) of pair q,
// lw[NP + s] = (cx, cy) of spin s, at the scale, and lm[q] = the pair's bit
// mask. The whole vector is staged in LDS, so partner k ^ m is read from there
// for every pair; a pair term is added for every amplitude, with a zero
// weight where the reference skips it. The pair loop is a runtime loop, so
// the partner reads of a few consecutive pairs are in flight together
// without the register cost of unrolling all L(L-1)/2 of them.
template <int L>
__device__ __forceinline__ void apply_h_lds(const double2* lds, const double2* xv, double2* acc, uint32_t t,
                                            const double* dg, double sdiag, const double2* lw,
                                            const uint32_t* lm, bool fields) {
  constexpr int A = Res<L>::A, NP = Res<L>::NP;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    const double f = sdiag * dg[i];
    acc[i] = make_double2(f * xv[i].x, f * xv[i].y);
  }
  if (fields) {
    for (int s = 0; s < L; ++s) {
      const double2 c = lw[NP + s];
      if (c.x == 0.0 && c.y == 0.0) continue;
#pragma unroll
      for (int i = 0; i < A; ++i) {
        const uint32_t k = t + kBT * i;
        const double2 p = lds[k ^ (1u << s)];
        const double sy = ((k >> s) & 1u) ? -c.y : c.y;
        acc[i].x = fma(c.x, p.x, acc[i].x);
        acc[i].x = fma(-sy, p.y, acc[i].x);
        acc[i].y = fma(c.x, p.y, acc[i].y);
        acc[i].y = fma(sy, p.x, acc[i].y);
      }
    }
  }
#pragma unroll 5
  for (int q = 0; q < NP; ++q) {
    const uint32_t m = lm[q];
    const double2 wq = lw[q];
#pragma unroll
    for (int i = 0; i < A; ++i) {
      const uint32_t k = t + kBT * i;
      const bool on = __popc(k & m) == 1;
      cmadd(acc[i], on ? wq.x : wq.y, lds[k ^ m]);
    }
  }
}

// The weights of one evolve step (or of the energy's Hamiltonian) into LDS
// for apply_h_lds; the leading barrier lets earlier readers finish, and the
// first staging barrier after it makes them visible.
template <int L>
__device__ __forceinline__ void stage_weights(double2* lw, uint32_t* lm, const double* __restrict__ w,
                                              const DevEvolve& e, uint32_t t) {
  constexpr int NP = Res<L>::NP;
  __syncthreads();
  if (t < static_cast<uint32_t>(NP)) {
    lw[t] = make_double2(w[e.pair_off + 2 * t], w[e.pair_off + 2 * t + 1]);
    lm[t] = (1u << pa(L, static_cast<int>(t))) | (1u << pb(L, static_cast<int>(t)));
  } else if (t < static_cast<uint32_t>(NP + L)) {
    const uint32_t sp = t - NP;
    lw[t] = e.has_field ? make_double2(w[e.field_off + 2 * sp], w[e.field_off + 2 * sp + 1])
                        : make_double2(0.0, 0.0);
  }
}

// nsteps Chebyshev steps (Propagator::applyOneStep / evolve,
// src/core/Propagator.cpp:46-100) of the state and the reference together
// (L <= 10): one staging barrier pair per order for both (T_{n-1} of the
// state in lds[0, N), of the reference in lds[N, 2N)), the step's weights
// read from LDS; T_1 = acc / 2, T_n = acc - T_{n-2}, psi += c_n T_n.
// Returns false on a zero norm.
template <int L>
__device__ bool evolve_both(double2* st, double2* rf, double2* lds, double* red, double2* lw, uint32_t* lm,
                            uint32_t t,
                            const double* dg, const double* __restrict__ w, const DevEvolve& e,
                            const DevOp& op) {
  constexpr int A = Res<L>::A, N = Res<L>::N;
  const bool active = t < static_cast<uint32_t>(N);
  stage_weights<L>(lw, lm, w, e, t);
  const double* a = w + e.coef_off;
  const double scale = -2.0 / e.bound;
  const bool fields = e.has_field != 0;
  double2 sprev[A], scur[A], rprev[A], rcur[A], acc[A];
  for (int step = 0; step < op.nsteps; ++step) {
#pragma unroll
    for (int i = 0; i < A; ++i) {
      scur[i] = st[i];
      st[i] = make_double2(st[i].x * a[0], st[i].y * a[0]);
      rcur[i] = rf[i];
      rf[i] = make_double2(rf[i].x * a[0], rf[i].y * a[0]);
    }
    double wre = op.mode == 0 ? 0.0 : -1.0, wim = op.mode == 0 ? -1.0 : 0.0;
    for (int n = 1; n <= e.order; ++n) {
      __syncthreads();
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          lds[t + kBT * i] = scur[i];
          lds[N + t + kBT * i] = rcur[i];
        }
      __syncthreads();
      const double cr = 2.0 * a[n] * wre, ci = 2.0 * a[n] * wim;
      if (active) {
        apply_h_lds<L>(lds, scur, acc, t, dg, scale, lw, lm, fields);
#pragma unroll
        for (int i = 0; i < A; ++i) {
          double2 tn;
          if (n == 1) tn = make_double2(0.5 * acc[i].x, 0.5 * acc[i].y);
          else tn = make_double2(acc[i].x - sprev[i].x, acc[i].y - sprev[i].y);
          st[i].x += cr * tn.x - ci * tn.y;
          st[i].y += cr * tn.y + ci * tn.x;
          sprev[i] = scur[i];
          scur[i] = tn;
        }
        apply_h_lds<L>(lds + N, rcur, acc, t, dg, scale, lw, lm, fields);
#pragma unroll
        for (int i = 0; i < A; ++i) {
          double2 tn;
          if (n == 1) tn = make_double2(0.5 * acc[i].x, 0.5 * acc[i].y);
          else tn = make_double2(acc[i].x - rprev[i].x, acc[i].y - rprev[i].y);
          rf[i].x += cr * tn.x - ci * tn.y;
          rf[i].y += cr * tn.y + ci * tn.x;
          rprev[i] = rcur[i];
          rcur[i] = tn;
        }
      }
      if (op.mode == 0) {
        const double pr = wre;
        wre = wim;
        wim = -pr;
      } else {
        wre = -wre;
        wim = -wim;
      }
    }
    if (op.autonorm) {
      double ss = 0.0, sr = 0.0;
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          ss += st[i].x * st[i].x + st[i].y * st[i].y;
          sr += rf[i].x * rf[i].x + rf[i].y * rf[i].y;
        }
      const double2 nn = block_sum(red, ss, sr);
      const double ns = sqrt(nn.x), nr = sqrt(nn.y);
      if (ns == 0.0 || nr == 0.0) return false;
      const double is = 1.0 / ns, ir = 1.0 / nr;
#pragma unroll
      for (int i = 0; i < A; ++i) {
        st[i] = make_double2(st[i].x * is, st[i].y * is);
        rf[i] = make_double2(rf[i].x * ir, rf[i].y * ir);
      }
    }
  }
  return true;
}

// ---------------------------------------------- L = 11, 12: Walsh-Hadamard apply
// The off-diagonal part through the identity of spx_hadamard.hip,
//   H_off x = W A W x + S W B W S^+ x,
// with W the unnormalised Walsh-Hadamard transform on all L bits, A and B
// its diagonals from the pair and field weights (times 2^-L) and
// S = i^popcount. The partner gathers (L(L-1)/2 reads of 16 B per amplitude
// and order: 1056 B at L = 12, 880 B at L = 11) become four L-bit
// transforms of 8 values per lane: register stages and three LDS transposes
// each way (384 B of LDS traffic per amplitude). A vector is held by
// 2^LB lanes (LB = L - 3): the whole workgroup at L = 12 (state and
// reference take turns), half of it at L = 11 (the state in threads
// 0..255, the reference in 256..511, both through the same code and
// barriers). Layouts of the index c for lane u, register j:
//   N  c = u | j << LB       (k = u + 2^LB j)
//   L1 c = j | u << 3        (registers = bits 0-2)
//   L2 registers = bits 3-5, L3 registers = bits 6-8 (the transform domain;
//   at L = 11 register bit 2 there is index bit 8, already transformed)
// and the LDS slot c ^ ((c >> 3) & 7) on bits 0-2, so that each layout's
// eight consecutive lanes hit eight distinct 16-byte slots.
namespace wres {
constexpr unsigned kA = 8;
template <int LB>
__host__ __device__ constexpr unsigned c_n(unsigned u, unsigned j) { return u | (j << LB); }
__host__ __device__ constexpr unsigned c_l1(unsigned u, unsigned j) { return j | (u << 3); }
__host__ __device__ constexpr unsigned c_l2(unsigned u, unsigned j) { return (u & 7u) | (j << 3) | ((u >> 3) << 6); }
__host__ __device__ constexpr unsigned c_l3(unsigned u, unsigned j) { return (u & 63u) | (j << 6) | ((u >> 6) << 9); }
__host__ __device__ constexpr unsigned slot(unsigned c) { return (c & ~7u) | ((c ^ (c >> 3)) & 7u); }
// per-lane bases plus compile-time offsets (immediate LDS offsets)
__host__ __device__ constexpr unsigned b_l1(unsigned u, unsigned k) { return (u << 3) + ((k ^ u) & 7u); }
__host__ __device__ constexpr unsigned b_l2(unsigned u, unsigned k) { return ((u >> 3) << 6) + ((u ^ k) & 7u); }
__host__ __device__ constexpr unsigned b_l3(unsigned u) { return (u & 56u) + ((u >> 6) << 9) + ((u ^ (u >> 3)) & 7u); }
template <int LB>
constexpr bool slots_ok() {
  for (unsigned u = 0; u < (1u << LB); ++u)
    for (unsigned j = 0; j < kA; ++j) {
      if (slot(c_n<LB>(u, j)) != slot(u) + (j << LB)) return false;
      if (slot(c_l1(u, j)) != b_l1(u, j)) return false;
      if (slot(c_l2(u, j)) != b_l2(u, j) + (j << 3)) return false;
      if (slot(c_l3(u, j)) != b_l3(u) + (j << 6)) return false;
    }
  return true;
}
static_assert(slots_ok<9>() && slots_ok<8>(), "slot bases");
enum { N = 0, L1 = 1, L2 = 2, L3 = 3 };

template <int LB, int LAY>
__device__ __forceinline__ unsigned at(unsigned u, unsigned j) {
  if constexpr (LAY == N) return slot(u) + (j << LB);
  else if constexpr (LAY == L1) return b_l1(u, j);
  else if constexpr (LAY == L2) return b_l2(u, j) + (j << 3);
  else return b_l3(u) + (j << 6);
}

template <int B0, int B1>
__device__ __forceinline__ void bfly(double2 (&v)[kA]) {
#pragma unroll
  for (int st = B0; st < B1; ++st)
#pragma unroll
    for (int j = 0; j < static_cast<int>(kA); ++j) {
      if (j & (1 << st)) continue;
      const double2 a = v[j], b = v[j | (1 << st)];
      v[j] = make_double2(a.x + b.x, a.y + b.y);
      v[j | (1 << st)] = make_double2(a.x - b.x, a.y - b.y);
    }
}

// One transpose through LDS buffer S. Consecutive transposes alternate
// between two buffers, so one barrier each suffices: a thread reaches the
// next transpose's barrier only after its reads of this one, and the
// transpose after that (the only one to write S again) starts after that
// barrier. The lane index reaches the slot arithmetic through an empty asm,
// so that the bases are recomputed at each transpose instead of held in
// registers.
// Threads that hold no vector (active false, whole wavefronts) take part in
// the barrier only.
template <int LB, int FROM, int TO>
__device__ __forceinline__ void xpose(double2* S, uint32_t u, double2 (&v)[kA], bool active) {
  unsigned uv = u;
  asm volatile("" : "+v"(uv));
  if (active)
#pragma unroll
    for (unsigned j = 0; j < kA; ++j) S[at<LB, FROM>(uv, j)] = v[j];
  __syncthreads();
  if (active)
#pragma unroll
    for (unsigned j = 0; j < kA; ++j) v[j] = S[at<LB, TO>(uv, j)];
}

// register stages of the transform domain: bits 6-8, or 6-7 when bit 8 is
// among the natural layout's register bits (LB = 8)
template <int LB>
constexpr int kL3Stages = LB == 8 ? 2 : 3;

// The forward and inverse transforms use the buffers S0, S1, S0 and S1,
// S0, S1: every transpose alternates with the one before (a transform pair
// is six transposes, so the next pair starts again with S0).
// NST: register stages of the natural layout (2 when its register bit 2 is
// not an index bit of the transform, as for two L = 10 vectors held as one
// 11-bit vector).
template <int LB, int NST = 3>
__device__ __forceinline__ void fwd(double2* S0, double2* S1, uint32_t u, double2 (&v)[kA], bool active = true) {
  if (active) bfly<0, NST>(v);
  xpose<LB, N, L1>(S0, u, v, active);
  if (active) bfly<0, 3>(v);
  xpose<LB, L1, L2>(S1, u, v, active);
  if (active) bfly<0, 3>(v);
  xpose<LB, L2, L3>(S0, u, v, active);
  if (active) bfly<0, kL3Stages<LB>>(v);
}
template <int LB, int NST = 3>
__device__ __forceinline__ void inv(double2* S0, double2* S1, uint32_t u, double2 (&v)[kA], bool active = true) {
  if (active) bfly<0, kL3Stages<LB>>(v);
  xpose<LB, L3, L2>(S1, u, v, active);
  if (active) bfly<0, 3>(v);
  xpose<LB, L2, L1>(S0, u, v, active);
  if (active) bfly<0, 3>(v);
  xpose<LB, L1, N>(S1, u, v, active);
  if (active) bfly<0, NST>(v);
}

// v i^e (e mod 4): a swap and sign changes, exact
__device__ __forceinline__ double2 ipow(double2 v, unsigned e) {
  const bool odd = e & 1u, neg = e & 2u;
  const double re = odd ? -v.y : v.x, im = odd ? v.x : v.y;
  return make_double2(neg ? -re : re, neg ? -im : im);
}
}  // namespace wres

// Chebyshev steps with the Walsh-Hadamard apply for the vector held by this
// thread (lane u of 2^(L-3), 8 amplitudes k = u + 2^(L-3) j; its transposes
// alternate between the buffers S0 and S1 of 2^L amplitudes): the same recursion and accumulator updates (T_n = acc - T_{n-2}
// is formed as scale diag T_{n-1} - T_{n-2} plus the two transform terms).
// The diagonal at the own amplitudes and A and B at the transform-domain
// ones are formed once per evolve op from the step's weights; the step's
// weights must be staged in lw, lm (stage_weights) and visible.
// LS < L (LS = 10, L = 11): the lane holds two LS-spin vectors as one
// L-bit vector, index bit LS telling them apart (register bit 2 of the
// natural layout), and only threads 0..255 hold it; the transforms skip that
// bit, the tables never see it (their masks cover spins 0..LS-1), and
// S = i^popcount gains a constant i on the second vector, which cancels in
// S W B W S^+.
template <int L, int LS = L>
__device__ bool evolve_wht(double2 (&v)[wres::kA], double2* S0, double2* S1, double* red, const double2* lw,
                           const uint32_t* lm,
                           uint32_t u, const double* __restrict__ w, const DevEvolve& e, const DevOp& op) {
  constexpr int LB = L - 3, A = static_cast<int>(wres::kA), NP = Res<LS>::NP;
  constexpr bool kTwo = LS < L;
  constexpr int NST = kTwo ? 2 : 3;
  static_assert((L == 11 || L == 12) && (LS == L || (L == 11 && LS == 10)),
                "the Walsh apply is laid out for L = 11 and 12 (two L = 10 vectors as one of 11 bits)");
  const bool active = !kTwo || threadIdx.x < (1u << LB);
  // runtime loops over pairs and spins outside, the eight amplitudes inside
  // (unrolled, so that the arrays stay in registers)
  double ea[A], eb[A], dg[A];
  uint32_t sw[A], kn[A];
#pragma unroll
  for (int j = 0; j < A; ++j) {
    sw[j] = wres::c_l3(u, static_cast<unsigned>(j));
    kn[j] = wres::c_n<LB>(u, static_cast<unsigned>(j));
    ea[j] = 0.0;
    eb[j] = 0.0;
    dg[j] = 0.0;
  }
  for (int q = 0; q < NP; ++q) {
    const double2 c = lw[q];
    const uint32_t m = lm[q];
    const double pa = 0.5 * (c.x + c.y), pb = 0.5 * (c.x - c.y);
#pragma unroll
    for (int j = 0; j < A; ++j) {
      const double z = (__popc(sw[j] & m) & 1) ? -1.0 : 1.0;
      ea[j] = fma(pa, z, ea[j]);
      eb[j] = fma(pb, z, eb[j]);
    }
  }
  for (int sp = 0; sp < LS; ++sp) {
    const double2 c = lw[NP + sp];
#pragma unroll
    for (int j = 0; j < A; ++j) {
      const double z = ((sw[j] >> sp) & 1u) ? -1.0 : 1.0;
      ea[j] = fma(c.x, z, ea[j]);
      eb[j] = fma(-c.y, z, eb[j]);
    }
  }
  // the sigma_z diagonal in rebuildDiagonal's order (as build_diag)
  const double* zw = w + e.zdiag_off;
  for (int sp = 0; sp < LS; ++sp) {
    const double c = zw[sp];
    if (c == 0.0) continue;
#pragma unroll
    for (int j = 0; j < A; ++j) dg[j] += ((kn[j] >> sp) & 1u) ? c : -c;
  }
  for (int qa = 0, q = 0; qa < LS; ++qa)
    for (int qb = qa + 1; qb < LS; ++qb, ++q) {
      const double c = zw[LS + q];
      if (c == 0.0) continue;
#pragma unroll
      for (int j = 0; j < A; ++j) dg[j] += (((kn[j] >> qa) & 1u) == ((kn[j] >> qb) & 1u)) ? c : -c;
    }
#pragma unroll
  for (int j = 0; j < A; ++j) {
    ea[j] = ldexp(ea[j], -LS);
    eb[j] = ldexp(eb[j], -LS);
  }
  const double* a = w + e.coef_off;
  const double scale = -2.0 / e.bound;
  const unsigned pcu = static_cast<unsigned>(__popc(u));
  double2 tprev[A], tcur[A], x[A];
  for (int step = 0; step < op.nsteps; ++step) {
    if (active)
#pragma unroll
      for (int i = 0; i < A; ++i) {
        tcur[i] = v[i];
        v[i] = make_double2(v[i].x * a[0], v[i].y * a[0]);
      }
    double wre = op.mode == 0 ? 0.0 : -1.0, wim = op.mode == 0 ? -1.0 : 0.0;
    for (int n = 1; n <= e.order; ++n) {
      // W A W T_{n-1}
      if (active) {
#pragma unroll
        for (int i = 0; i < A; ++i) {
          const double f = scale * dg[i];
          const double2 b = n >= 2 ? tprev[i] : make_double2(0.0, 0.0);
          tprev[i] = make_double2(fma(f, tcur[i].x, -b.x), fma(f, tcur[i].y, -b.y));
        }
#pragma unroll
        for (int i = 0; i < A; ++i) x[i] = tcur[i];
      }
      wres::fwd<LB, NST>(S0, S1, u, x, active);
      if (active)
#pragma unroll
        for (int j = 0; j < A; ++j) x[j] = make_double2(ea[j] * x[j].x, ea[j] * x[j].y);
      wres::inv<LB, NST>(S0, S1, u, x, active);
      // S W B W S^+ T_{n-1}
      if (active) {
#pragma unroll
        for (int i = 0; i < A; ++i) tprev[i] = make_double2(tprev[i].x + x[i].x, tprev[i].y + x[i].y);
#pragma unroll
        for (int i = 0; i < A; ++i) x[i] = wres::ipow(tcur[i], 0u - (pcu + static_cast<unsigned>(__popc(i))));
      }
      wres::fwd<LB, NST>(S0, S1, u, x, active);
      if (active)
#pragma unroll
        for (int j = 0; j < A; ++j) x[j] = make_double2(eb[j] * x[j].x, eb[j] * x[j].y);
      wres::inv<LB, NST>(S0, S1, u, x, active);
      const double cr = 2.0 * a[n] * wre, ci = 2.0 * a[n] * wim;
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          const double2 sv = wres::ipow(x[i], pcu + static_cast<unsigned>(__popc(i)));
          double2 tn = make_double2(tprev[i].x + sv.x, tprev[i].y + sv.y);
          if (n == 1) tn = make_double2(0.5 * tn.x, 0.5 * tn.y);
          v[i].x += cr * tn.x - ci * tn.y;
          v[i].y += cr * tn.y + ci * tn.x;
          tprev[i] = tcur[i];
          tcur[i] = tn;
        }
      if (op.mode == 0) {
        const double pr = wre;
        wre = wim;
        wim = -pr;
      } else {
        wre = -wre;
        wim = -wim;
      }
    }
    if (op.autonorm) {
      // each vector's own norm: at L = 11 the halves hold different vectors,
      // with two L = 10 vectors registers 0..3 and 4..7
      double s0 = 0.0, s1 = 0.0;
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          const double q = v[i].x * v[i].x + v[i].y * v[i].y;
          if (kTwo && i >= A / 2) s1 += q;
          else s0 += q;
        }
      const bool upper = !kTwo && LB == 8 && threadIdx.x >= (1u << LB);
      const double2 nn = block_sum(red, upper ? 0.0 : s0, upper ? s0 : s1);
      if (nn.x == 0.0 || ((kTwo || LB == 8) && nn.y == 0.0)) return false;
      const double iv0 = 1.0 / sqrt(upper ? nn.y : nn.x), iv1 = 1.0 / sqrt(nn.y);
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          const double iv = kTwo && i >= A / 2 ? iv1 : iv0;
          v[i] = make_double2(v[i].x * iv, v[i].y * iv);
        }
    }
  }
  return true;
}

// Both vectors of an evolve op through evolve_wht. L = 12: the state, then
// the reference, each held by all 512 threads (their walker layout is the
// transform's natural one). L = 11: the walker's layout (4 amplitudes of
// each vector per thread) is moved through LDS to 8 amplitudes of one vector
// per thread and back.
template <int L>
__device__ bool evolve_pair_wht(double2* st, double2* rf, double2* lds, double* red, double2* lw, uint32_t* lm,
                                uint32_t t, const double* __restrict__ w, const DevEvolve& e, const DevOp& op) {
  stage_weights<L>(lw, lm, w, e, t);
  __syncthreads();
  if constexpr (L == 12) {
    double2 v[wres::kA];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = st[i];
    if (!evolve_wht<12>(v, lds, lds + 4096, red, lw, lm, t, w, e, op)) return false;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      st[i] = v[i];
      v[i] = rf[i];
    }
    if (!evolve_wht<12>(v, lds, lds + 4096, red, lw, lm, t, w, e, op)) return false;
#pragma unroll
    for (int i = 0; i < 8; ++i) rf[i] = v[i];
    return true;
  } else if constexpr (L == 10) {
    // state and reference as one 11-bit vector (index bit 10: the
    // reference) in threads 0..255, 8 amplitudes each; threads 256..511 join
    // the barriers only
    constexpr int N = Res<L>::N, A = Res<L>::A;   // 1024, 2
    const uint32_t u = t & 255u;
    const bool holder = t < 256u;
#pragma unroll
    for (int i = 0; i < A; ++i) {
      lds[t + kBT * i] = st[i];
      lds[N + t + kBT * i] = rf[i];
    }
    __syncthreads();
    double2 v[wres::kA];
    if (holder)
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = lds[u + 256 * j];
    __syncthreads();
    const bool ok = evolve_wht<11, 10>(v, lds, lds + 2 * N, red, lw, lm, u, w, e, op);
    if (holder)
#pragma unroll
      for (int j = 0; j < 8; ++j) lds[u + 256 * j] = v[j];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < A; ++i) {
      st[i] = lds[t + kBT * i];
      rf[i] = lds[N + t + kBT * i];
    }
    __syncthreads();
    return ok;
  } else {
    constexpr int N = Res<L>::N, A = Res<L>::A;   // 2048, 4
    static_assert(L == 11 && A == 4, "half-workgroup vectors are laid out for L = 11");
    const uint32_t h = t >> 8, u = t & 255u;
#pragma unroll
    for (int i = 0; i < A; ++i) {
      lds[t + kBT * i] = st[i];
      lds[N + t + kBT * i] = rf[i];
    }
    __syncthreads();
    double2 v[wres::kA];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = lds[h * N + u + 256 * j];
    __syncthreads();
    const bool ok = evolve_wht<11>(v, lds + h * N, lds + 2 * N + h * N, red, lw, lm, u, w, e, op);
#pragma unroll
    for (int j = 0; j < 8; ++j) lds[h * N + u + 256 * j] = v[j];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < A; ++i) {
      st[i] = lds[t + kBT * i];
      rf[i] = lds[N + t + kBT * i];
    }
    __syncthreads();
    return ok;
  }
}

// SpinRotation::apply (src/core/Gate.cpp:60-85), targets in ascending order.
template <int L>
__device__ void rotate_vec(double2* v, double2* lds, uint32_t t, const DevPulse& p) {
  constexpr int A = Res<L>::A;
  const bool active = t < static_cast<uint32_t>(Res<L>::N);
  const double llr = p.u[0], lli = p.u[1], lhr = p.u[2], lhi = p.u[3];
  const double hlr = p.u[4], hli = p.u[5], hhr = p.u[6], hhi = p.u[7];
  for (int s = 0; s < L; ++s) {
    if (!((p.targets >> s) & 1u)) continue;
    stage<L>(lds, v, t);
    if (!active) continue;
    double2 nv[A];
#pragma unroll
    for (int i = 0; i < A; ++i) {
      const double2 q = lds[(t + kBT * i) ^ (1u << s)];
      const double2 own = v[i];
      if (kbit<L>(t, i, s) == 0) {  // own is the down (lo) amplitude
        nv[i] = make_double2((llr * own.x - lli * own.y) + (lhr * q.x - lhi * q.y),
                             (llr * own.y + lli * own.x) + (lhr * q.y + lhi * q.x));
      } else {
        nv[i] = make_double2((hlr * q.x - hli * q.y) + (hhr * own.x - hhi * own.y),
                             (hlr * q.y + hli * q.x) + (hhr * own.y + hhi * own.x));
      }
    }
#pragma unroll
    for (int i = 0; i < A; ++i) v[i] = nv[i];
  }
}

// sum_s sigma_axis(s) psi at own amplitude i (psi staged in LDS).
template <int L>
__device__ __forceinline__ double2 total_spin(const double2* lds, const double2* psi, uint32_t t, int i,
                                              int axis) {
  double2 w = make_double2(0.0, 0.0);
  for (int s = 0; s < L; ++s) {
    const bool up = kbit<L>(t, i, s) != 0;
    if (axis == 2) {
      w.x += up ? psi[i].x : -psi[i].x;
      w.y += up ? psi[i].y : -psi[i].y;
      continue;
    }
    const double2 p = lds[(t + kBT * i) ^ (1u << s)];
    if (axis == 0) {
      w.x += p.x;
      w.y += p.y;
    } else if (up) {
      w.x += p.y;
      w.y -= p.x;
    } else {
      w.x -= p.y;
      w.y += p.x;
    }
  }
  return w;
}

template <int L>
__device__ void measure(double2* st, double2* rf, double2* lds, double* red, double2* lw, uint32_t* lm,
                        uint32_t t, double* dg, const RArgs& g, const DevReal& R, int row) {
  constexpr int A = Res<L>::A;
  const bool active = t < static_cast<uint32_t>(Res<L>::N);
  double* out = g.out + R.out_off + static_cast<long long>(row) * g.width;
  int at = 0;
  for (int o = 0; o < g.nobs; ++o) {
    const int code = g.obs_codes[o];
    if (code <= 2) {  // corr: <ref| sum sigma |state>
      stage<L>(lds, st, t);
      double sr = 0.0, si = 0.0;
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          const double2 w = total_spin<L>(lds, st, t, i, code);
          const double2 r = rf[i];
          sr += r.x * w.x + r.y * w.y;
          si += r.x * w.y - r.y * w.x;
        }
      const double2 s = block_sum(red, sr, si);
      if (t == 0) {
        out[at] = s.x;
        out[at + 1] = s.y;
      }
      at += 2;
    } else if (code <= 5) {  // mag: Re <state| sigma(spin) |state>
      const int axis = code - 3, spin = g.obs_spins[o];
      stage<L>(lds, st, t);
      double sr = 0.0;
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          const double2 x = st[i];
          const bool up = kbit<L>(t, i, spin) != 0;
          double2 w;
          if (axis == 2) {
            w = up ? x : make_double2(-x.x, -x.y);
          } else {
            const double2 p = lds[(t + kBT * i) ^ (1u << spin)];
            w = axis == 0 ? p : (up ? make_double2(p.y, -p.x) : make_double2(-p.y, p.x));
          }
          sr += x.x * w.x + x.y * w.y;
        }
      const double2 s = block_sum(red, sr, 0.0);
      if (t == 0) out[at] = s.x;
      at += 1;
    } else if (code == 6) {  // norms
      double a = 0.0, b = 0.0;
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) {
          a += st[i].x * st[i].x + st[i].y * st[i].y;
          b += rf[i].x * rf[i].x + rf[i].y * rf[i].y;
        }
      const double2 s = block_sum(red, a, b);
      if (t == 0) {
        out[at] = sqrt(s.x);
        out[at + 1] = sqrt(s.y);
      }
      at += 2;
    } else {  // energy: Re <H state|state> with the base fields
      const DevEvolve& me = g.evo[R.measure_off];
      build_diag<L>(dg, t, g.w + me.zdiag_off);
      stage_weights<L>(lw, lm, g.w, me, t);
      stage<L>(lds, st, t);
      double2 hx[A];
      double sr = 0.0;
      if (active) {
        apply_h_lds<L>(lds, st, hx, t, dg, 1.0, lw, lm, me.has_field != 0);
#pragma unroll
        for (int i = 0; i < A; ++i) sr += hx[i].x * st[i].x + hx[i].y * st[i].y;
      }
      const double2 s = block_sum(red, sr, 0.0);
      if (t == 0) out[at] = s.x;
      at += 1;
    }
  }
}

template <int L>
__device__ void run_real(const RArgs& g, const DevReal& R, double2* lds, double* red, double2* lw, uint32_t* lm) {
  constexpr int A = Res<L>::A;
  const uint32_t t = threadIdx.x;
  const bool active = t < static_cast<uint32_t>(Res<L>::N);
  double2 st[A], rf[A];
  double dg[A];
  if (active)
#pragma unroll
    for (int i = 0; i < A; ++i) {
      const long long k = t + kBT * i;
      st[i] = make_double2(g.psi0[R.psi_off + 2 * k], g.psi0[R.psi_off + 2 * k + 1]);
    }
  // reference state A|psi> = sum sigma_axis psi (prepareReferenceState)
  stage<L>(lds, st, t);
  if (active)
#pragma unroll
    for (int i = 0; i < A; ++i) rf[i] = total_spin<L>(lds, st, t, i, g.ref_axis);

  int stack_pc[16], stack_left[16];
  int depth = 0, row = 0;
  for (int pc = 0; pc < g.nops; ++pc) {
    const DevOp op = g.ops[pc];
    if (op.kind == OP_EVOLVE) {
      const DevEvolve e = g.evo[R.evolve_off + op.arg];
      if (e.order == 0 || op.nsteps == 0) continue;
      bool ok;
      if constexpr (L >= 10) {
        ok = evolve_pair_wht<L>(st, rf, lds, red, lw, lm, t, g.w, e, op);
      } else {
        build_diag<L>(dg, t, g.w + e.zdiag_off);
        ok = evolve_both<L>(st, rf, lds, red, lw, lm, t, dg, g.w, e, op);
      }
      if (!ok) {
        if (t == 0) atomicExch(g.queue + 1, 1);  // zero norm under autonormalisation
        return;
      }
    } else if (op.kind == OP_PULSE) {
      const DevPulse p = g.pulses[op.arg];
      rotate_vec<L>(st, lds, t, p);
      rotate_vec<L>(rf, lds, t, p);
    } else if (op.kind == OP_MEASURE) {
      measure<L>(st, rf, lds, red, lw, lm, t, dg, g, R, row);
      ++row;
    } else if (op.kind == OP_REPEAT) {
      if (op.arg == 0) {
        pc = op.match;
      } else {
        stack_pc[depth] = pc;
        stack_left[depth] = op.arg;
        ++depth;
      }
    } else if (op.kind == OP_END) {
      if (--stack_left[depth - 1] > 0) {
        pc = stack_pc[depth - 1];
      } else {
        --depth;
      }
    }
  }
}

// The LDS of a realisation of L spins: 2^L amplitudes (two vectors' worth
// where both evolve together), the reduction scratch, the evolve step's
// weights and pair masks.
template <int L>
__device__ void run_sized(const RArgs& g, const DevReal& R, double2* smem) {
  double2* lds = smem;
  double* red = reinterpret_cast<double*>(smem + kLdsAmps<L>);               // 2 x kBT
  double2* lw = reinterpret_cast<double2*>(red + 2 * kBT);                   // NP + L weights
  uint32_t* lm = reinterpret_cast<uint32_t*>(lw + Res<L>::NP + L);           // NP pair masks
  run_real<L>(g, R, lds, red, lw, lm);
}

template <int... Ls>
__device__ void run_any(std::integer_sequence<int, Ls...>, int L, const RArgs& g, const DevReal& R,
                        double2* smem) {
  ((L == Ls + 1 ? run_sized<Ls + 1>(g, R, smem) : void()), ...);
}

// One persistent kernel per L: its registers are sized for that L alone.
template <int L>
__global__ __launch_bounds__(kBT) void k_resident(RArgs g) {
  extern __shared__ double2 smem[];
  __shared__ int next;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(g.queue, 1);
    __syncthreads();
    const int slot = next;
    if (slot >= g.nreal) return;
    run_sized<L>(g, g.reals[g.order[slot]], smem);
  }
}

// Every size in one persistent kernel: a workgroup pulls the next
// realisation of one queue (the host orders it by estimated work, largest
// first, across sizes) and runs the walker of its L. Its registers are those
// of the largest walker; from L = 8 up every walker already needs more than
// a second workgroup per CU would leave it, so the larger sizes lose nothing.
__global__ __launch_bounds__(kBT) void k_resident_all(RArgs g) {
  extern __shared__ double2 smem[];
  __shared__ int next;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(g.queue, 1);
    __syncthreads();
    const int slot = next;
    if (slot >= g.nreal) return;
    const DevReal& R = g.reals[g.order[slot]];
    run_any(std::make_integer_sequence<int, 12>{}, R.nspins, g, R, smem);
  }
}

}  // namespace

size_t resident_lds_bytes(int L) {
  const size_t amps = L >= 11 ? 8192 : L == 10 ? 4096 : static_cast<size_t>(2) << L;   // kLdsAmps<L>
  return sizeof(double2) * amps + sizeof(double) * 2 * kBT +
         sizeof(double2) * static_cast<size_t>(L * (L - 1) / 2 + L) +
         sizeof(uint32_t) * static_cast<size_t>(L * (L - 1) / 2);
}

int launch_resident_all(spx_ctx* c, const DevOp* ops, int nops, const DevPulse* pulses, const DevEvolve* evo,
                        const DevReal* reals, const int* order, int nreal, const double* weights,
                        const double* psi0, double* out, int width, const int* obs_codes, const int* obs_spins,
                        int nobs, int reference_axis, int* queue, int nblocks, size_t lds, hipStream_t stream) {
  (void)c;
  RArgs g{ops, nops, pulses, evo, reals, order, nreal, weights, psi0, out, width,
          obs_codes, obs_spins, nobs, reference_axis, queue};
  k_resident_all<<<nblocks, kBT, lds, stream>>>(g);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int launch_resident(spx_ctx* c, int L, const DevOp* ops, int nops, const DevPulse* pulses,
                    const DevEvolve* evo, const DevReal* reals, const int* order, int nreal,
                    const double* weights, const double* psi0, double* out, int width,
                    const int* obs_codes, const int* obs_spins, int nobs, int reference_axis,
                    int* queue, int nblocks, hipStream_t stream) {
  RArgs g{ops, nops, pulses, evo, reals, order, nreal, weights, psi0, out, width,
          obs_codes, obs_spins, nobs, reference_axis, queue};
  const size_t lds = resident_lds_bytes(L);
  switch (L) {
    case 1: k_resident<1><<<nblocks, kBT, lds, stream>>>(g); break;
    case 2: k_resident<2><<<nblocks, kBT, lds, stream>>>(g); break;
    case 3: k_resident<3><<<nblocks, kBT, lds, stream>>>(g); break;
    case 4: k_resident<4><<<nblocks, kBT, lds, stream>>>(g); break;
    case 5: k_resident<5><<<nblocks, kBT, lds, stream>>>(g); break;
    case 6: k_resident<6><<<nblocks, kBT, lds, stream>>>(g); break;
    case 7: k_resident<7><<<nblocks, kBT, lds, stream>>>(g); break;
    case 8: k_resident<8><<<nblocks, kBT, lds, stream>>>(g); break;
    case 9: k_resident<9><<<nblocks, kBT, lds, stream>>>(g); break;
    case 10: k_resident<10><<<nblocks, kBT, lds, stream>>>(g); break;
    case 11: k_resident<11><<<nblocks, kBT, lds, stream>>>(g); break;
    case 12: k_resident<12><<<nblocks, kBT, lds, stream>>>(g); break;
    default: return fail(SPX_EINVAL, "resident runner holds at most 12 spins");
  }
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

}  // namespace spx
