// Pulse programs and the batched ensemble runner (internal).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "spx_common.hpp"

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
  double axis[3] = {0, 0, 1};
  double angle = 0.0;
  std::vector<uint32_t> targets;  // empty = every spin
  size_t count = 1;               // repeat count
  size_t match = 0;               // REPEAT <-> END partner index
};

size_t program_measurements(const struct ::spx_program& p, size_t b, size_t e);

struct Realisation {
  size_t nspins = 0;
  std::vector<double> fields;     // [3L]
  std::vector<uint32_t> lo, hi;
  std::vector<double> jx, jy, jz;
  std::vector<double> psi0;       // [2 * 2^L]
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
// Runs every realisation of `b`; times [points], values [n][points][width].
int batch_run_resident(spx_batch* b, double* times, double* values);
int batch_run_streaming(spx_batch* b, size_t first, size_t count, double* times, double* values);
}  // namespace spx
