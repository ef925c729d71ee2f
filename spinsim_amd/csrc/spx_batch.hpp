// Pulse programs and the batched ensemble runner (internal).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "spx_common.hpp"
#include "spx_ensemble.hpp"

namespace spx {

enum OpKind { OP_EVOLVE = 0, OP_PULSE = 1, OP_MEASURE = 2, OP_REPEAT = 3, OP_END = 4 };

// One step of the flattened Sequence AST (include/spinsim/sequence/Sequence.hpp:19-57).
struct ProgOp {
  int kind = OP_MEASURE;
  double tau = 0.0;
  size_t nsteps = 1;
  int mode = 0;
  bool autonorm = false;
  std::vector<double> fields;     // [3L] override, empty = keep base fields
  bool uniform = false;           // override is ufield on every spin
  double ufield[3] = {0, 0, 0};
  double axis[3] = {0, 0, 1};
  double angle = 0.0;
  std::vector<uint32_t> targets;  // empty = every spin
  size_t count = 1;               // repeat count
  size_t match = 0;               // REPEAT <-> END partner index
};

size_t program_measurements(const struct ::spx_program& p, size_t b, size_t e);

// The field table an evolve step uses for a realisation of L spins: its
// own override (per spin or uniform), else the realisation's base fields.
inline const std::vector<double>& step_fields(const ProgOp& op, const std::vector<double>& base,
                                              size_t L, std::vector<double>& tmp) {
  if (op.uniform) {
    tmp.resize(3 * L);
    for (size_t s = 0; s < L; ++s)
      for (int c = 0; c < 3; ++c) tmp[3 * s + c] = op.ufield[c];
    return tmp;
  }
  return op.fields.empty() ? base : op.fields;
}

struct Realisation {
  size_t nspins = 0;
  std::vector<double> fields;     // [3L]
  std::vector<uint32_t> lo, hi;
  std::vector<double> jx, jy, jz;
  std::vector<double> psi0;       // [2 * 2^L]; empty when drawn on use (below)
  // Realisations added by (seed, index) keep only their couplings and fields
  // and draw their initial state when a run needs it (a state of L = 16 is
  // 1 MiB; a plain Monte-Carlo arm holds tens of thousands).
  bool drawn_state = false;
  DipolarSpec spec{};
  uint64_t index = 0;
  // The initial state into `out` (2 * 2^L doubles).
  int initial_state(double* out) const;
};

}  // namespace spx

struct spx_program {
  std::vector<spx::ProgOp> ops;
  std::vector<size_t> open;
};

struct spx_batch {
  spx_ctx* ctx = nullptr;
  spx_program program;
  double epsilon = 1e-7;
  int reference_axis = 0;
  std::vector<int> obs_codes, obs_spins;
  size_t width = 0;
  std::vector<spx::Realisation> items;
};

namespace spx {

// Device program of the resident runner: the flattened Sequence with every
// evolve step numbered (its per-realisation expansion and weights live in
// the realisation tables) and every pulse's 2x2 matrix precomputed.
struct DevOp {
  int kind;        // OP_EVOLVE / OP_PULSE / OP_MEASURE / OP_REPEAT / OP_END
  int arg;         // evolve: step number; pulse: pulse number; repeat: count
  int match;       // REPEAT <-> END partner
  int nsteps;
  int mode;
  int autonorm;
};

struct DevPulse {
  double u[8];          // ll, lh, hl, hh (re, im)
  uint32_t targets;     // bit mask of target spins
  int pad;
};

// One evolve step of one realisation: expansion and Hamiltonian weights at
// scale -2/E (flip-flop and flip-both per pair in pair order, transverse
// field per spin), plus the sigma_z terms for the diagonal.
struct DevEvolve {
  double bound;
  int order;            // K; 0 = identity
  int coef_off;         // K + 1 expansion coefficients
  int pair_off;         // 2 * L(L-1)/2 weights (anti, aligned) at scale
  int field_off;        // 2 * L weights (cx, cy) at scale
  int zdiag_off;        // L + L(L-1)/2 sigma_z weights (0.5 hz, 0.25 jz), unscaled
  int has_field;
  int pad;
};

struct DevReal {
  int nspins;
  int evolve_off;       // index of its first DevEvolve (one per evolve step)
  int measure_off;      // pair/field/zdiag weights at scale 1 (energy)
  long long psi_off;    // doubles into the initial-state buffer
  long long out_off;    // doubles into the output buffer
};

// Runs every realisation of `b`; times [points], values [n][points][width].
int batch_run_resident(spx_batch* b, double* times, double* values);
int batch_run_streaming(spx_batch* b, size_t first, size_t count, double* times, double* values);
int launch_resident(spx_ctx* c, int L, const DevOp* ops, int nops, const DevPulse* pulses,
                    const DevEvolve* evo, const DevReal* reals, const int* order, int nreal,
                    const double* weights, const double* psi0, double* out, int width,
                    const int* obs_codes, const int* obs_spins, int nobs, int reference_axis,
                    int* queue, int nblocks, hipStream_t stream);
// Realisations of several sizes (L <= 12) in one persistent launch;
// `lds` covers the largest size's layout (resident_lds_bytes).
int launch_resident_all(spx_ctx* c, const DevOp* ops, int nops, const DevPulse* pulses, const DevEvolve* evo,
                        const DevReal* reals, const int* order, int nreal, const double* weights,
                        const double* psi0, double* out, int width, const int* obs_codes, const int* obs_spins,
                        int nobs, int reference_axis, int* queue, int nblocks, size_t lds, hipStream_t stream);
size_t resident_lds_bytes(int L);
}  // namespace spx
