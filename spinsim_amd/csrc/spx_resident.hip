// Resident ensemble runner: one persistent launch runs many realisations of
// one pulse program (LocalRunner::run + Sequence::run, the reference's
// src/ensemble/EnsembleRunner.cpp:27-88 and src/sequence/Sequence.cpp:35-106).
//
// A workgroup pulls realisation indices from a queue (ordered longest
// first on the host) and runs the whole program of one realisation with its
// state and reference vectors in registers: 512 threads, thread t owning
// amplitudes k = t + 512 i. Each Chebyshev order stages T_{n-1} in LDS for
// the partner gathers (at most 2^12 x 16 B = 64 KiB); partners that differ
// only in register bits come straight from registers. The pair structure is
// compiled per L (8..12), the weights are per realisation and per evolve
// step. All reductions are fixed-order block trees, so every realisation's
// numbers are independent of which workgroup ran it and of timing.
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

// bit `b` of amplitude k = t + 512 i
template <int L>
__device__ __forceinline__ uint32_t kbit(uint32_t t, int i, int b) {
  return b < Res<L>::RB ? (t >> b) & 1u : (static_cast<uint32_t>(i) >> (b - Res<L>::RB)) & 1u;
}

// Partner k ^ m of the own amplitude i: from registers when m touches only
// register bits, from LDS otherwise.
template <int L>
__device__ __forceinline__ double2 partner(const double2* lds, const double2* v, uint32_t t, int i,
                                           uint32_t m) {
  constexpr int RB = Res<L>::RB;
  const uint32_t ml = m & ((1u << RB) - 1u);
  const uint32_t mr = m >> RB;
  if (ml == 0) return v[i ^ static_cast<int>(mr)];
  return lds[(t ^ ml) + kBT * (static_cast<uint32_t>(i) ^ mr)];
}

// acc_i = (H x)_k at the own amplitudes, H given by its weights at the
// scale already folded in (diag unscaled, multiplied by `sdiag`).
template <int L, int... Q>
__device__ __forceinline__ void apply_pairs(std::integer_sequence<int, Q...>, const double2* lds,
                                            const double2* xv, double2* acc, uint32_t t,
                                            const double* __restrict__ pw) {
  constexpr int A = Res<L>::A;
  auto one = [&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int a = pa(L, q), b = pb(L, q);
    constexpr uint32_t m = (1u << a) | (1u << b);
    const double anti = pw[2 * q], aligned = pw[2 * q + 1];
#pragma unroll
    for (int i = 0; i < A; ++i) {
      const bool on = kbit<L>(t, i, a) != kbit<L>(t, i, b);
      const double c = on ? anti : aligned;
      if (c != 0.0) cmadd(acc[i], c, partner<L>(lds, xv, t, i, m));
    }
  };
  (one(std::integral_constant<int, Q>{}), ...);
}

template <int L>
__device__ void apply_h(const double2* lds, const double2* xv, double2* acc, uint32_t t,
                        const double* dg, double sdiag, const double* __restrict__ w,
                        const DevEvolve& e) {
  constexpr int A = Res<L>::A;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    const double f = sdiag * dg[i];
    acc[i] = make_double2(f * xv[i].x, f * xv[i].y);
  }
  if (e.has_field) {
    const double* fw = w + e.field_off;
    for (int s = 0; s < L; ++s) {
      const double cx = fw[2 * s], cy = fw[2 * s + 1];
      if (cx == 0.0 && cy == 0.0) continue;
#pragma unroll
      for (int i = 0; i < A; ++i) {
        const double2 p = partner<L>(lds, xv, t, i, 1u << s);
        const double sy = kbit<L>(t, i, s) ? -cy : cy;
        acc[i].x = fma(cx, p.x, acc[i].x);
        acc[i].x = fma(-sy, p.y, acc[i].x);
        acc[i].y = fma(cx, p.y, acc[i].y);
        acc[i].y = fma(sy, p.x, acc[i].y);
      }
    }
  }
  apply_pairs<L>(std::make_integer_sequence<int, Res<L>::NP>{}, lds, xv, acc, t, w + e.pair_off);
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

// nsteps Chebyshev steps of one vector (Propagator::applyOneStep /
// evolve, src/core/Propagator.cpp:46-100). Returns false on a zero norm.
template <int L>
__device__ bool evolve_vec(double2* v, double2* lds, double* red, uint32_t t, const double* dg,
                           const double* __restrict__ w, const DevEvolve& e, const DevOp& op) {
  constexpr int A = Res<L>::A;
  const bool active = t < static_cast<uint32_t>(Res<L>::N);
  const double* a = w + e.coef_off;
  const double scale = -2.0 / e.bound;
  double2 tprev[A], tcur[A], acc[A];
  for (int step = 0; step < op.nsteps; ++step) {
#pragma unroll
    for (int i = 0; i < A; ++i) {
      tcur[i] = v[i];
      v[i] = make_double2(v[i].x * a[0], v[i].y * a[0]);
    }
    double wre = op.mode == 0 ? 0.0 : -1.0, wim = op.mode == 0 ? -1.0 : 0.0;
    for (int n = 1; n <= e.order; ++n) {
      stage<L>(lds, tcur, t);
      if (active) apply_h<L>(lds, tcur, acc, t, dg, scale, w, e);
      const double cr = 2.0 * a[n] * wre, ci = 2.0 * a[n] * wim;
#pragma unroll
      for (int i = 0; i < A; ++i) {
        double2 tn;
        if (n == 1) tn = make_double2(0.5 * acc[i].x, 0.5 * acc[i].y);
        else tn = make_double2(acc[i].x - tprev[i].x, acc[i].y - tprev[i].y);
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
      double s = 0.0;
      if (active)
#pragma unroll
        for (int i = 0; i < A; ++i) s += v[i].x * v[i].x + v[i].y * v[i].y;
      const double nrm = sqrt(block_sum(red, s, 0.0).x);
      if (nrm == 0.0) return false;
      const double inv = 1.0 / nrm;
#pragma unroll
      for (int i = 0; i < A; ++i) v[i] = make_double2(v[i].x * inv, v[i].y * inv);
    }
  }
  return true;
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
      const double2 q = partner<L>(lds, v, t, i, 1u << s);
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
    const double2 p = partner<L>(lds, psi, t, i, 1u << s);
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
__device__ void measure(double2* st, double2* rf, double2* lds, double* red, uint32_t t, double* dg,
                        const RArgs& g, const DevReal& R, int row) {
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
            const double2 p = partner<L>(lds, st, t, i, 1u << spin);
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
      stage<L>(lds, st, t);
      double2 hx[A];
      double sr = 0.0;
      if (active) {
        apply_h<L>(lds, st, hx, t, dg, 1.0, g.w, me);
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
__device__ void run_real(const RArgs& g, const DevReal& R, double2* lds, double* red) {
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
      build_diag<L>(dg, t, g.w + e.zdiag_off);
      if (!evolve_vec<L>(st, lds, red, t, dg, g.w, e, op) ||
          !evolve_vec<L>(rf, lds, red, t, dg, g.w, e, op)) {
        if (t == 0) atomicExch(g.queue + 1, 1);  // zero norm under autonormalisation
        return;
      }
    } else if (op.kind == OP_PULSE) {
      const DevPulse p = g.pulses[op.arg];
      rotate_vec<L>(st, lds, t, p);
      rotate_vec<L>(rf, lds, t, p);
    } else if (op.kind == OP_MEASURE) {
      measure<L>(st, rf, lds, red, t, dg, g, R, row);
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

// One persistent kernel per L: its registers are sized for that L alone.
template <int L>
__global__ __launch_bounds__(kBT) void k_resident(RArgs g) {
  extern __shared__ double2 smem[];
  double2* lds = smem;                                           // 2^L amplitudes
  double* red = reinterpret_cast<double*>(smem + Res<L>::N);    // 2 x kBT
  __shared__ int next;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(g.queue, 1);
    __syncthreads();
    const int slot = next;
    if (slot >= g.nreal) return;
    run_real<L>(g, g.reals[g.order[slot]], lds, red);
  }
}

}  // namespace

int launch_resident(spx_ctx* c, int L, const DevOp* ops, int nops, const DevPulse* pulses,
                    const DevEvolve* evo, const DevReal* reals, const int* order, int nreal,
                    const double* weights, const double* psi0, double* out, int width,
                    const int* obs_codes, const int* obs_spins, int nobs, int reference_axis,
                    int* queue, int nblocks) {
  RArgs g{ops, nops, pulses, evo, reals, order, nreal, weights, psi0, out, width,
          obs_codes, obs_spins, nobs, reference_axis, queue};
  const size_t lds = sizeof(double2) * (static_cast<size_t>(1) << L) + sizeof(double) * 2 * kBT;
  switch (L) {
    case 1: k_resident<1><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 2: k_resident<2><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 3: k_resident<3><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 4: k_resident<4><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 5: k_resident<5><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 6: k_resident<6><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 7: k_resident<7><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 8: k_resident<8><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 9: k_resident<9><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 10: k_resident<10><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 11: k_resident<11><<<nblocks, kBT, lds, c->stream>>>(g); break;
    case 12: k_resident<12><<<nblocks, kBT, lds, c->stream>>>(g); break;
    default: return fail(SPX_EINVAL, "resident runner holds at most 12 spins");
  }
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

}  // namespace spx
