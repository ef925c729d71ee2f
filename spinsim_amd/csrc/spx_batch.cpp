// Batched ensemble runner: the realisation loop of LocalRunner::run
// (src/ensemble/EnsembleRunner.cpp:27-88) with each realisation's program
// walked as Sequence::run does (src/sequence/Sequence.cpp:35-106).
#include <cmath>
#include <cstring>
#include <vector>

#include "spx_batch.hpp"

using namespace spx;

namespace spx {

size_t program_measurements(const spx_program& p, size_t b, size_t e) {
  size_t total = 0;
  for (size_t i = b; i < e; ++i) {
    const ProgOp& op = p.ops[i];
    if (op.kind == OP_MEASURE) {
      ++total;
    } else if (op.kind == OP_REPEAT) {
      total += op.count * program_measurements(p, i + 1, op.match);
      i = op.match;
    }
  }
  return total;
}

namespace {

size_t obs_width(const std::vector<int>& codes) {
  size_t w = 0;
  for (int c : codes) w += (c <= 2 || c == 6) ? 2 : 1;
  return w;
}

struct Walker {
  spx_batch* b;
  const Realisation* r;
  spx_ham* h;
  spx_state* state;
  spx_state* ref;
  spx_prop* prop;
  double time = 0.0;
  double* times;
  double* values;
  size_t row = 0;

  int measure() {
    double* out = values + row * b->width;
    size_t at = 0;
    for (size_t o = 0; o < b->obs_codes.size(); ++o) {
      const int code = b->obs_codes[o];
      if (code <= 2) {
        SPX_TRY(spx_measure_corr(code, state, ref, &out[at], &out[at + 1]));
        at += 2;
      } else if (code <= 5) {
        SPX_TRY(spx_measure_mag(code - 3, static_cast<size_t>(b->obs_spins[o]), state, &out[at]));
        at += 1;
      } else if (code == 6) {
        SPX_TRY(spx_state_norm(state, &out[at]));
        SPX_TRY(spx_state_norm(ref, &out[at + 1]));
        at += 2;
      } else {
        SPX_TRY(spx_measure_energy(h, state, &out[at]));
        at += 1;
      }
    }
    if (times) times[row] = time;
    ++row;
    return SPX_OK;
  }

  int run(size_t beg, size_t end) {
    const std::vector<ProgOp>& ops = b->program.ops;
    const size_t L = r->nspins;
    for (size_t i = beg; i < end; ++i) {
      const ProgOp& op = ops[i];
      if (op.kind == OP_EVOLVE) {
        std::vector<double> saved;
        if (!op.fields.empty()) {
          if (op.fields.size() != 3 * L)
            return fail(SPX_EINVAL, "step field table does not match the number of spins");
          saved.resize(3 * L);
          for (size_t s = 0; s < L; ++s) {
            SPX_TRY(spx_ham_get_field(h, s, &saved[3 * s]));
            SPX_TRY(spx_ham_set_field(h, s, op.fields[3 * s], op.fields[3 * s + 1], op.fields[3 * s + 2]));
          }
        }
        spx_state* both[2] = {state, ref};
        int rc = spx_prop_evolve_many(prop, h, op.mode, both, 2, op.tau, op.nsteps, op.autonorm);
        if (!saved.empty())
          for (size_t s = 0; s < L; ++s)
            spx_ham_set_field(h, s, saved[3 * s], saved[3 * s + 1], saved[3 * s + 2]);
        SPX_TRY(rc);
        time += op.tau * static_cast<double>(op.nsteps);
      } else if (op.kind == OP_PULSE) {
        std::vector<uint32_t> all;
        const uint32_t* t = op.targets.data();
        size_t nt = op.targets.size();
        if (nt == 0) {
          for (size_t s = 0; s < L; ++s) all.push_back(static_cast<uint32_t>(s));
          t = all.data();
          nt = all.size();
        }
        SPX_TRY(spx_rotate(state, op.axis[0], op.axis[1], op.axis[2], op.angle, t, nt));
        SPX_TRY(spx_rotate(ref, op.axis[0], op.axis[1], op.axis[2], op.angle, t, nt));
      } else if (op.kind == OP_MEASURE) {
        SPX_TRY(measure());
      } else if (op.kind == OP_REPEAT) {
        for (size_t k = 0; k < op.count; ++k) SPX_TRY(run(i + 1, op.match));
        i = op.match;
      }
    }
    return SPX_OK;
  }
};

}  // namespace

// Per-realisation device objects, one kernel sequence per program step.
int batch_run_streaming(spx_batch* b, size_t first, size_t count, double* times, double* values) {
  const size_t points = program_measurements(b->program, 0, b->program.ops.size());
  for (size_t i = first; i < first + count; ++i) {
    const Realisation& r = b->items[i];
    spx_ham* h = nullptr;
    spx_state* state = nullptr;
    spx_state* ref = nullptr;
    spx_prop* prop = nullptr;
    int rc = spx_ham_create(b->ctx, r.nspins, &h);
    if (rc == SPX_OK) rc = spx_state_create(b->ctx, r.nspins, &state);
    if (rc == SPX_OK) rc = spx_state_create(b->ctx, r.nspins, &ref);
    if (rc == SPX_OK) rc = spx_prop_create(b->ctx, r.nspins, b->epsilon, &prop);
    for (size_t s = 0; rc == SPX_OK && s < r.nspins; ++s)
      rc = spx_ham_set_field(h, s, r.fields[3 * s], r.fields[3 * s + 1], r.fields[3 * s + 2]);
    for (size_t p = 0; rc == SPX_OK && p < r.lo.size(); ++p)
      rc = spx_ham_set_coupling(h, r.lo[p], r.hi[p], r.jx[p], r.jy[p], r.jz[p]);
    if (rc == SPX_OK) rc = spx_state_upload(state, r.psi0.data());
    if (rc == SPX_OK) rc = spx_apply_total_spin(b->reference_axis, state, ref);
    if (rc == SPX_OK) {
      Walker w{b, &r, h, state, ref, prop, 0.0, (i == first) ? times : nullptr,
               values + (i - first) * points * b->width};
      rc = w.run(0, b->program.ops.size());
    }
    spx_prop_destroy(prop);
    spx_state_destroy(ref);
    spx_state_destroy(state);
    spx_ham_destroy(h);
    SPX_TRY(rc);
  }
  return SPX_OK;
}

}  // namespace spx

extern "C" {

int spx_batch_create(spx_ctx* ctx, const spx_program* program, double epsilon, int reference_axis,
                     const int* obs_codes, const int* obs_spins, size_t nobs, spx_batch** out) {
  if (!ctx || !program || !out) return fail(SPX_EINVAL, "null argument");
  if (!program->open.empty()) return fail(SPX_EINVAL, "program has an unclosed repeat");
  if (!(epsilon > 0.0)) return fail(SPX_EINVAL, "epsilon must be positive");
  if (reference_axis < 0 || reference_axis > 2) return fail(SPX_EINVAL, "axis must be one of x, y, z");
  for (size_t o = 0; o < nobs; ++o)
    if (obs_codes[o] < 0 || obs_codes[o] > 7) return fail(SPX_EINVAL, "unknown observable code");
  auto* b = new spx_batch();
  b->ctx = ctx;
  b->program = *program;
  b->epsilon = epsilon;
  b->reference_axis = reference_axis;
  b->obs_codes.assign(obs_codes, obs_codes + nobs);
  b->obs_spins.assign(obs_spins, obs_spins + nobs);
  b->width = obs_width(b->obs_codes);
  *out = b;
  return SPX_OK;
}

int spx_batch_destroy(spx_batch* b) {
  delete b;
  return SPX_OK;
}

int spx_batch_add(spx_batch* b, size_t nspins, const double* fields3L, size_t npairs, const uint32_t* lo,
                  const uint32_t* hi, const double* jx, const double* jy, const double* jz,
                  const double* psi0) {
  if (nspins == 0 || nspins > kMaxSpins) return fail(SPX_ERANGE, "number of spins must be in [1, kMaxSpins]");
  for (size_t o = 0; o < b->obs_codes.size(); ++o)
    if (b->obs_codes[o] >= 3 && b->obs_codes[o] <= 5 &&
        (b->obs_spins[o] < 0 || static_cast<size_t>(b->obs_spins[o]) >= nspins))
      return fail(SPX_ERANGE, "observable names a spin outside the system");
  for (const ProgOp& op : b->program.ops) {
    if (op.kind == OP_EVOLVE && !op.fields.empty() && op.fields.size() != 3 * nspins)
      return fail(SPX_EINVAL, "step field table does not match the number of spins");
    if (op.kind == OP_PULSE)
      for (uint32_t t : op.targets)
        if (t >= nspins) return fail(SPX_ERANGE, "rotation names a spin outside the system");
  }
  Realisation r;
  r.nspins = nspins;
  r.fields.assign(3 * nspins, 0.0);
  if (fields3L) r.fields.assign(fields3L, fields3L + 3 * nspins);
  for (size_t p = 0; p < npairs; ++p) {
    if (lo[p] == hi[p]) return fail(SPX_EINVAL, "a spin cannot couple to itself");
    if (lo[p] >= nspins || hi[p] >= nspins) return fail(SPX_ERANGE, "coupled spin index out of range");
  }
  r.lo.assign(lo, lo + npairs);
  r.hi.assign(hi, hi + npairs);
  r.jx.assign(jx, jx + npairs);
  r.jy.assign(jy, jy + npairs);
  r.jz.assign(jz, jz + npairs);
  r.psi0.assign(psi0, psi0 + (static_cast<size_t>(2) << nspins));
  b->items.push_back(std::move(r));
  return SPX_OK;
}

int spx_batch_width(const spx_batch* b, size_t* width) {
  *width = b->width;
  return SPX_OK;
}

int spx_batch_clear(spx_batch* b) {
  b->items.clear();
  return SPX_OK;
}

int spx_batch_run(spx_batch* b, double* times, double* values) {
  if (b->items.empty()) return fail(SPX_EINVAL, "an ensemble needs at least one realisation");
  return batch_run_streaming(b, 0, b->items.size(), times, values);
}

}  // extern "C"
