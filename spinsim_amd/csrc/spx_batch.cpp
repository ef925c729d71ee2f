// Batched ensemble runner: the realisation loop of LocalRunner::run
// (src/ensemble/EnsembleRunner.cpp:27-88) with each realisation's program
// walked as Sequence::run does (src/sequence/Sequence.cpp:35-106).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "spx_batch.hpp"
#include "spx_ensemble.hpp"

using namespace spx;

namespace spx {

int Realisation::initial_state(double* out) const {
  if (!drawn_state) {
    std::copy(psi0.begin(), psi0.end(), out);
    return SPX_OK;
  }
  const size_t np = nspins * (nspins - 1) / 2;
  std::vector<uint32_t> l(np + 1), h(np + 1);
  std::vector<double> z(np + 1), f(nspins);
  const int rc = draw_realisation(spec, index, l.data(), h.data(), z.data(), f.data(), out);
  if (rc == SPX_EDOMAIN) return fail(rc, "cannot normalize a zero state");
  if (rc != SPX_OK) return fail(rc, "could not place the spins at the minimum separation");
  return SPX_OK;
}

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

// Realisations that share a program walk it in lockstep as a group: every
// evolve of the program is one evolve_set over each member's state and
// reference (under that member's Hamiltonian), so their Chebyshev orders
// share launches; pulses and measurements are per member. Each member sees
// exactly the arithmetic of walking it alone.
struct Member {
  const Realisation* r;
  spx_ham* h;
  spx_state* state;
  spx_state* ref;
  double* times;   // null unless this member records the time column
  double* values;
  size_t row = 0;
};

// Per measured value, four doubles of raw sums on the device: (re, im) of
// the observable's fixed-order reduction, and for norm the second vector's.
constexpr size_t kRawPerObs = 4;

struct Walker {
  spx_batch* b;
  std::vector<Member>& m;
  spx_prop* prop;
  double* d_raw;        // [member][row][observable][kRawPerObs]
  size_t points;
  double time = 0.0;

  // The observables' reductions are left on the device (no host wait per
  // measurement); finish() reads them all once the group is done.
  int measure(size_t k) {
    Member& x = m[k];
    const size_t nobs = b->obs_codes.size();
    double* raw = d_raw + (k * points + x.row) * nobs * kRawPerObs;
    const size_t L = x.state->nspins;
    spx_ctx* c = x.state->ctx;
    for (size_t o = 0; o < nobs; ++o) {
      const int code = b->obs_codes[o];
      double* dst = raw + o * kRawPerObs;
      if (code <= 2) {
        SPX_TRY(reduce_corr_dev(c, code, x.state->d, x.ref->d, L, dst));
      } else if (code <= 5) {
        SPX_TRY(reduce_mag_dev(c, code - 3, static_cast<uint32_t>(b->obs_spins[o]), x.state->d, L, dst));
      } else if (code == 6) {
        SPX_TRY(reduce_norm2_dev(c, x.state->d, x.state->n, dst));
        SPX_TRY(reduce_norm2_dev(c, x.ref->d, x.ref->n, dst + 2));
      } else {
        SPX_TRY(reduce_energy_dev(x.h, x.state->d, dst));
      }
    }
    if (x.times) x.times[x.row] = time;
    ++x.row;
    return SPX_OK;
  }

  // Rows from the raw sums, as the single-state entry points finish them:
  // corr (re, im), mag re, norm sqrt of each sum, energy re.
  int finish(const std::vector<double>& raw) {
    const size_t nobs = b->obs_codes.size();
    for (size_t k = 0; k < m.size(); ++k)
      for (size_t row = 0; row < m[k].row; ++row) {
        double* out = m[k].values + row * b->width;
        const double* r = raw.data() + (k * points + row) * nobs * kRawPerObs;
        size_t at = 0;
        for (size_t o = 0; o < nobs; ++o, r += kRawPerObs) {
          const int code = b->obs_codes[o];
          if (code <= 2) {
            out[at++] = r[0];
            out[at++] = r[1];
          } else if (code == 6) {
            out[at++] = std::sqrt(r[0]);
            out[at++] = std::sqrt(r[2]);
          } else {
            out[at++] = r[0];
          }
        }
      }
    return SPX_OK;
  }

  int evolve(const ProgOp& op) {
    const size_t g = m.size();
    std::vector<std::vector<double>> saved(g);
    int rc = SPX_OK;
    for (size_t k = 0; k < g && rc == SPX_OK && (!op.fields.empty() || op.uniform); ++k) {
      const size_t L = m[k].r->nspins;
      std::vector<double> tmp;
      const std::vector<double>& f = step_fields(op, m[k].r->fields, L, tmp);
      if (f.size() != 3 * L) {
        rc = fail(SPX_EINVAL, "step field table does not match the number of spins");
        break;
      }
      saved[k].resize(3 * L);
      for (size_t s = 0; s < L && rc == SPX_OK; ++s) rc = spx_ham_get_field(m[k].h, s, &saved[k][3 * s]);
      for (size_t s = 0; s < L && rc == SPX_OK; ++s) rc = spx_ham_set_field(m[k].h, s, f[3 * s], f[3 * s + 1], f[3 * s + 2]);
    }
    if (rc == SPX_OK) {
      std::vector<spx_ham*> hs;
      std::vector<spx_state*> ss;
      for (Member& x : m) {
        hs.push_back(x.h);
        hs.push_back(x.h);
        ss.push_back(x.state);
        ss.push_back(x.ref);
      }
      rc = spx_prop_evolve_set(prop, hs.data(), op.mode, ss.data(), ss.size(), op.tau, op.nsteps, op.autonorm);
    }
    for (size_t k = 0; k < g; ++k)
      for (size_t s = 0; 3 * s < saved[k].size(); ++s)
        spx_ham_set_field(m[k].h, s, saved[k][3 * s], saved[k][3 * s + 1], saved[k][3 * s + 2]);
    SPX_TRY(rc);
    time += op.tau * static_cast<double>(op.nsteps);
    return SPX_OK;
  }

  int run(size_t beg, size_t end) {
    const std::vector<ProgOp>& ops = b->program.ops;
    const size_t L = m.front().r->nspins;
    for (size_t i = beg; i < end; ++i) {
      const ProgOp& op = ops[i];
      if (op.kind == OP_EVOLVE) {
        SPX_TRY(evolve(op));
      } else if (op.kind == OP_PULSE) {
        std::vector<uint32_t> all;
        const uint32_t* t = op.targets.data();
        size_t nt = op.targets.size();
        if (nt == 0) {
          for (size_t s = 0; s < L; ++s) all.push_back(static_cast<uint32_t>(s));
          t = all.data();
          nt = all.size();
        }
        for (Member& x : m) {
          SPX_TRY(spx_rotate(x.state, op.axis[0], op.axis[1], op.axis[2], op.angle, t, nt));
          SPX_TRY(spx_rotate(x.ref, op.axis[0], op.axis[1], op.axis[2], op.angle, t, nt));
        }
      } else if (op.kind == OP_MEASURE) {
        for (size_t k = 0; k < m.size(); ++k) SPX_TRY(measure(k));
      } else if (op.kind == OP_REPEAT) {
        for (size_t k = 0; k < op.count; ++k) SPX_TRY(run(i + 1, op.match));
        i = op.match;
      }
    }
    return SPX_OK;
  }
};

}  // namespace

// Realisations per lockstep group: a state and a reference each, so one
// group fills an order launch.
constexpr size_t kGroup = kMaxMany / 2;

// Device objects of one worker, reused across its groups (allocation and
// release synchronise the device, which would serialise the workers);
// rebuilt only when L changes.
struct WorkerObjects {
  spx_ctx* c = nullptr;
  size_t L = 0;
  spx_ham* h[kGroup] = {};
  spx_state* state[kGroup] = {};
  spx_state* ref[kGroup] = {};
  spx_prop* prop = nullptr;
  double* psi_pin[kGroup] = {};   // pinned staging of the members' initial states
  double* d_raw = nullptr;        // the group's raw measurement sums
  size_t raw_cap = 0;             // doubles
  std::vector<double> raw;

  void release() {
    if (c) (void)hipStreamSynchronize(c->stream);
    for (double*& p : psi_pin) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
    }
    if (d_raw) (void)hipFree(d_raw);
    d_raw = nullptr;
    raw_cap = 0;
    spx_prop_destroy(prop);
    prop = nullptr;
    for (size_t k = 0; k < kGroup; ++k) {
      spx_state_destroy(ref[k]);
      spx_state_destroy(state[k]);
      spx_ham_destroy(h[k]);
      ref[k] = state[k] = nullptr;
      h[k] = nullptr;
    }
    L = 0;
  }
  ~WorkerObjects() { release(); }

  int prepare(size_t nspins, double epsilon) {
    if (L == nspins) {
      // same size: clear the Hamiltonians' fields and couplings (their device
      // caches are keyed on them and rebuild on first use)
      for (spx_ham* x : h) {
        std::fill(x->hx.begin(), x->hx.end(), 0.0);
        std::fill(x->hy.begin(), x->hy.end(), 0.0);
        std::fill(x->hz.begin(), x->hz.end(), 0.0);
        x->lo.clear();
        x->hi.clear();
        x->jx.clear();
        x->jy.clear();
        x->jz.clear();
        x->terms_valid = false;
        x->sched_set = false;
      }
      return SPX_OK;
    }
    release();
    for (size_t k = 0; k < kGroup; ++k) {
      SPX_TRY(spx_ham_create(c, nspins, &h[k]));
      SPX_TRY(spx_state_create(c, nspins, &state[k]));
      SPX_TRY(spx_state_create(c, nspins, &ref[k]));
    }
    SPX_TRY(spx_prop_create(c, nspins, epsilon, &prop));
    // the workers' own streams already overlap their groups; a second stream
    // per worker measured slower (L = 14, 16 realisations: 559 -> 448 /s)
    SPX_TRY(spx_prop_set_streams(prop, 1));
    for (double*& p : psi_pin)
      SPX_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), (size_t(2) << nspins) * sizeof(double),
                            hipHostMallocDefault));
    L = nspins;
    return SPX_OK;
  }

  int reserve_raw(size_t doubles) {
    if (doubles <= raw_cap) return SPX_OK;
    if (c) SPX_HIP(hipStreamSynchronize(c->stream));
    if (d_raw) SPX_HIP(hipFree(d_raw));
    d_raw = nullptr;
    raw_cap = 0;
    SPX_HIP(hipMalloc(&d_raw, doubles * sizeof(double)));
    raw_cap = doubles;
    return SPX_OK;
  }
};

// One group (realisations idx[0..n), all of one size) on a worker's
// objects; realisation i writes its rows at values_of(i), and the one whose
// index is `timed` also the time column.
template <typename Rows>
static int run_group(spx_batch* b, WorkerObjects& o, const size_t* idx, size_t n, size_t timed, double* times,
                     Rows values_of) {
  SPX_TRY(o.prepare(b->items[idx[0]].nspins, b->epsilon));
  const size_t points = program_measurements(b->program, 0, b->program.ops.size());
  SPX_TRY(o.reserve_raw(std::max<size_t>(1, n * points * b->obs_codes.size() * kRawPerObs)));
  std::vector<Member> m;
  for (size_t k = 0; k < n; ++k) {
    const Realisation& r = b->items[idx[k]];
    for (size_t s = 0; s < r.nspins; ++s)
      SPX_TRY(spx_ham_set_field(o.h[k], s, r.fields[3 * s], r.fields[3 * s + 1], r.fields[3 * s + 2]));
    for (size_t p = 0; p < r.lo.size(); ++p)
      SPX_TRY(spx_ham_set_coupling(o.h[k], r.lo[p], r.hi[p], r.jx[p], r.jy[p], r.jz[p]));
    // the pinned staging of member k was last read by the previous group's
    // upload, which the end of that group waited for
    SPX_TRY(r.initial_state(o.psi_pin[k]));
    SPX_HIP(hipMemcpyAsync(o.state[k]->d, o.psi_pin[k], o.state[k]->n * sizeof(double2), hipMemcpyHostToDevice,
                           o.c->stream));
    SPX_TRY(spx_apply_total_spin(b->reference_axis, o.state[k], o.ref[k]));
    m.push_back(Member{&r, o.h[k], o.state[k], o.ref[k], idx[k] == timed ? times : nullptr, values_of(idx[k])});
  }
  Walker w{b, m, o.prop, o.d_raw, points};
  SPX_TRY(w.run(0, b->program.ops.size()));
  const size_t nraw = n * points * b->obs_codes.size() * kRawPerObs;
  o.raw.resize(std::max<size_t>(1, nraw));
  SPX_TRY(copy_to_host(o.c, o.raw.data(), o.d_raw, nraw * sizeof(double)));
  return w.finish(o.raw);
}

// Realisations in lockstep groups of up to kGroup of one size (by size, in
// index order within a size), the groups one after another on the batch's
// context or, with several workers, group g on worker g mod W, each worker
// a host thread with its own context (stream), as LocalRunner's pool runs
// jobs concurrently. Every realisation sees the same arithmetic on its own
// rows, so the output depends neither on the grouping nor on the workers.
int batch_run_streaming(spx_batch* b, size_t first, size_t count, double* times, double* values) {
  const size_t points = program_measurements(b->program, 0, b->program.ops.size());
  static const int env_workers = [] {
    const char* e = std::getenv("SPX_BATCH_WORKERS");
    return e ? std::atoi(e) : 8;
  }();
  static const size_t group_max = [] {
    const char* e = std::getenv("SPX_BATCH_GROUP");
    const int g = e ? std::atoi(e) : static_cast<int>(kGroup);
    return static_cast<size_t>(std::max(1, std::min(g, static_cast<int>(kGroup))));
  }();
  std::vector<size_t> idx(count);
  for (size_t i = 0; i < count; ++i) idx[i] = first + i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t x, size_t y) { return b->items[x].nspins < b->items[y].nspins; });
  std::vector<std::pair<size_t, size_t>> groups;   // [begin, end) into idx
  for (size_t g = 0; g < count;) {
    size_t e = g + 1;
    while (e < count && e - g < group_max && b->items[idx[e]].nspins == b->items[idx[g]].nspins) ++e;
    groups.emplace_back(g, e);
    g = e;
  }
  auto values_of = [&](size_t i) { return values + (i - first) * points * b->width; };
  const size_t workers =
      std::min<size_t>(groups.size(), static_cast<size_t>(std::max(1, std::min(env_workers, 8))));
  if (workers <= 1) {
    WorkerObjects o;
    o.c = b->ctx;
    for (const auto& [g, e] : groups) SPX_TRY(run_group(b, o, &idx[g], e - g, first, times, values_of));
    return SPX_OK;
  }
  std::vector<int> rcs(workers, SPX_OK);
  std::vector<std::string> msgs(workers);
  std::vector<std::thread> pool;
  for (size_t w = 0; w < workers; ++w)
    pool.emplace_back([&, w] {
      spx_ctx* c = nullptr;
      int rc = spx_ctx_create(b->ctx->device, &c);
      {
        WorkerObjects o;
        o.c = c;
        for (size_t gi = w; rc == SPX_OK && gi < groups.size(); gi += workers)
          rc = run_group(b, o, &idx[groups[gi].first], groups[gi].second - groups[gi].first, first, times,
                         values_of);
        if (rc != SPX_OK) msgs[w] = spx_last_error();
      }
      if (c) spx_ctx_destroy(c);
      rcs[w] = rc;
    });
  for (auto& t : pool) t.join();
  for (size_t w = 0; w < workers; ++w)
    if (rcs[w] != SPX_OK) return fail(rcs[w], msgs[w]);
  return SPX_OK;
}

namespace {

// Spectral bound of one field table + coupling set (spectralBound,
// src/core/Hamiltonian.cpp:82-93), same summation order.
double bound_of(const std::vector<double>& f, const Realisation& r) {
  double b = 0.0;
  for (size_t s = 0; s < r.nspins; ++s)
    b += 0.5 * (std::abs(f[3 * s]) + std::abs(f[3 * s + 1]) + std::abs(f[3 * s + 2]));
  for (size_t p = 0; p < r.lo.size(); ++p)
    b += 0.25 * (std::abs(r.jx[p]) + std::abs(r.jy[p]) + std::abs(r.jz[p]));
  return b;
}

// Hamiltonian weights of one field table at `scale`, appended to `w`:
// pairs (anti, aligned) in pair order, fields (cx, cy), sigma_z terms.
void append_weights(std::vector<double>& w, const std::vector<double>& f, const Realisation& r,
                    double scale, DevEvolve& e) {
  const size_t L = r.nspins, np = L * (L - 1) / 2;
  e.pair_off = static_cast<int>(w.size());
  w.resize(w.size() + 2 * np, 0.0);
  for (size_t p = 0; p < r.lo.size(); ++p) {
    const double quarter = scale * 0.25;
    const double ff = quarter * (r.jx[p] + r.jy[p]);
    const double fb = quarter * (r.jx[p] - r.jy[p]);
    const size_t lo = r.lo[p], hi = r.hi[p];
    const size_t q = lo * L - lo * (lo + 1) / 2 + (hi - lo - 1);
    w[e.pair_off + 2 * q] = std::abs(ff) > kNegligible ? ff : 0.0;
    w[e.pair_off + 2 * q + 1] = std::abs(fb) > kNegligible ? fb : 0.0;
  }
  e.field_off = static_cast<int>(w.size());
  e.has_field = 0;
  for (size_t s = 0; s < L; ++s) {
    const double cx = scale * 0.5 * f[3 * s];
    const double cy = scale * 0.5 * f[3 * s + 1];
    const bool keep = std::abs(cx) > kNegligible || std::abs(cy) > kNegligible;
    w.push_back(keep ? cx : 0.0);
    w.push_back(keep ? cy : 0.0);
    e.has_field |= keep ? 1 : 0;
  }
  e.zdiag_off = static_cast<int>(w.size());
  for (size_t s = 0; s < L; ++s) {
    const double c = 0.5 * f[3 * s + 2];
    w.push_back(std::abs(c) <= kNegligible ? 0.0 : c);
  }
  std::vector<double> zz(np, 0.0);
  for (size_t p = 0; p < r.lo.size(); ++p) {
    const size_t lo = r.lo[p], hi = r.hi[p];
    const double c = 0.25 * r.jz[p];
    zz[lo * L - lo * (lo + 1) / 2 + (hi - lo - 1)] = std::abs(c) <= kNegligible ? 0.0 : c;
  }
  w.insert(w.end(), zz.begin(), zz.end());
}

template <class T>
int upload(spx_ctx* c, T** d, const std::vector<T>& h) {
  const size_t bytes = (h.size() ? h.size() : 1) * sizeof(T);
  hipError_t e = hipMalloc(d, bytes);
  if (e != hipSuccess) return fail(SPX_ENOMEM, "hipMalloc for the batch tables failed");
  return copy_to_device(c, *d, h.data(), h.size() * sizeof(T));
}

bool resident_ok(const spx_batch* b) {
  for (const Realisation& r : b->items)
    if (r.nspins > 12) return false;
  int depth = 0, maxd = 0;
  for (const ProgOp& op : b->program.ops) {
    if (op.kind == OP_REPEAT) maxd = std::max(maxd, ++depth);
    if (op.kind == OP_END) --depth;
  }
  return maxd <= 16;
}

}  // namespace

// All realisations in one persistent launch (L <= 12).
int batch_run_resident(spx_batch* b, double* times, double* values) {
  spx_ctx* c = b->ctx;
  const std::vector<ProgOp>& prog = b->program.ops;
  const size_t n = b->items.size();
  const size_t points = program_measurements(b->program, 0, prog.size());

  std::vector<DevOp> ops;
  std::vector<DevPulse> pulses;
  std::vector<const ProgOp*> evolves;
  for (const ProgOp& op : prog) {
    DevOp d{op.kind, 0, static_cast<int>(op.match), static_cast<int>(op.nsteps), op.mode,
            op.autonorm ? 1 : 0};
    if (op.kind == OP_EVOLVE) {
      d.arg = static_cast<int>(evolves.size());
      evolves.push_back(&op);
    } else if (op.kind == OP_PULSE) {
      // Axis3::normalized + Rotation2x2::make (src/core/Gate.cpp:20-51)
      const double len = std::sqrt(op.axis[0] * op.axis[0] + op.axis[1] * op.axis[1] + op.axis[2] * op.axis[2]);
      const double nx = op.axis[0] / len, ny = op.axis[1] / len, nz = op.axis[2] / len;
      const double cs = std::cos(0.5 * op.angle), sn = std::sin(0.5 * op.angle);
      DevPulse p{{cs, sn * nz, sn * ny, -sn * nx, -sn * ny, -sn * nx, cs, -sn * nz}, 0u, 0};
      if (op.targets.empty()) p.targets = 0xffffffffu;
      for (uint32_t tg : op.targets) p.targets |= 1u << tg;
      d.arg = static_cast<int>(pulses.size());
      pulses.push_back(p);
    } else if (op.kind == OP_REPEAT) {
      d.arg = static_cast<int>(op.count);
    }
    ops.push_back(d);
  }

  // per-realisation expansions and weights, built on host threads: each
  // realisation's tables with offsets relative to its own block of weights
  // (an evolve op whose fields, mode and tau equal an earlier one's shares
  // that one's tables), then concatenated in index order, so the tables do
  // not depend on the threads
  struct Tables {
    std::vector<DevEvolve> evo;   // one per evolve op, then the measure weights
    std::vector<double> w;
    double cost = 0.0;
  };
  std::vector<Tables> tables(n);
  {
    std::atomic<size_t> next{0};
    std::atomic<int> status{SPX_OK};
    std::string msg;
    auto build = [&](size_t i) -> int {
      const Realisation& r = b->items[i];
      Tables& T = tables[i];
      std::vector<std::vector<double>> seen_f;   // fields of the distinct evolve ops so far
      std::vector<const ProgOp*> seen_op;
      std::vector<size_t> seen_at;
      for (const ProgOp* op : evolves) {
        std::vector<double> tmp;
        const std::vector<double>& f = step_fields(*op, r.fields, r.nspins, tmp);
        size_t k = 0;
        for (; k < seen_op.size(); ++k)
          if (seen_op[k]->mode == op->mode && seen_op[k]->tau == op->tau && seen_f[k] == f) break;
        if (k < seen_op.size()) {
          const DevEvolve e = T.evo[seen_at[k]];
          T.cost += static_cast<double>(e.order) * static_cast<double>(op->nsteps);
          T.evo.push_back(e);
          continue;
        }
        DevEvolve e{};
        const double sb = bound_of(f, r);
        std::vector<double> coeffs;
        double eb = 1.0;
        SPX_TRY(cheb_build(op->mode, sb, op->tau, b->epsilon, 6000, coeffs, eb));
        e.bound = eb;
        e.order = static_cast<int>(coeffs.size()) - 1;
        e.coef_off = static_cast<int>(T.w.size());
        T.w.insert(T.w.end(), coeffs.begin(), coeffs.end());
        append_weights(T.w, f, r, e.order > 0 ? -2.0 / eb : 1.0, e);
        T.cost += static_cast<double>(e.order) * static_cast<double>(op->nsteps);
        seen_f.push_back(f);
        seen_op.push_back(op);
        seen_at.push_back(T.evo.size());
        T.evo.push_back(e);
      }
      DevEvolve me{};
      append_weights(T.w, r.fields, r, 1.0, me);
      T.evo.push_back(me);
      return SPX_OK;
    };
    auto work = [&] {
      for (;;) {
        const size_t i = next.fetch_add(1);
        if (i >= n || status.load() != SPX_OK) return;
        const int rc = build(i);
        if (rc != SPX_OK) {
          int ok = SPX_OK;
          if (status.compare_exchange_strong(ok, rc)) msg = spx_last_error();
          return;
        }
      }
    };
    const size_t threads = std::min<size_t>(std::max<size_t>(1, n / 64),
                                            std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (status.load() != SPX_OK) return fail(status.load(), msg);
  }
  std::vector<DevEvolve> evo;
  std::vector<DevReal> reals(n);
  std::vector<double> w, psi;
  std::vector<double> cost(n, 0.0);
  long long psi_off = 0;
  for (size_t i = 0; i < n; ++i) {
    const Realisation& r = b->items[i];
    DevReal& R = reals[i];
    const Tables& T = tables[i];
    if (w.size() + T.w.size() > static_cast<size_t>(0x7fffffff)) return fail(SPX_ENOMEM, "batch tables too large");
    const int base = static_cast<int>(w.size());
    R.nspins = static_cast<int>(r.nspins);
    R.evolve_off = static_cast<int>(evo.size());
    for (DevEvolve e : T.evo) {
      e.coef_off += base;
      e.pair_off += base;
      e.field_off += base;
      e.zdiag_off += base;
      evo.push_back(e);
    }
    R.measure_off = static_cast<int>(evo.size()) - 1;
    w.insert(w.end(), T.w.begin(), T.w.end());
    cost[i] = T.cost;
    R.psi_off = psi_off;
    psi_off += static_cast<long long>(static_cast<size_t>(2) << r.nspins);
    R.out_off = static_cast<long long>(i * points * b->width);
    cost[i] *= static_cast<double>(1u << r.nspins);
  }
  tables.clear();
  // initial states into one buffer, drawn or copied on host threads
  psi.resize(static_cast<size_t>(psi_off));
  {
    std::atomic<size_t> next{0};
    std::atomic<int> status{SPX_OK};
    std::string msg;
    auto work = [&] {
      for (;;) {
        const size_t i = next.fetch_add(1);
        if (i >= n || status.load() != SPX_OK) return;
        const int rc = b->items[i].initial_state(psi.data() + reals[i].psi_off);
        if (rc != SPX_OK) {
          int ok = SPX_OK;
          if (status.compare_exchange_strong(ok, rc)) msg = spx_last_error();
          return;
        }
      }
    };
    const size_t threads = std::min<size_t>(std::max<size_t>(1, n / 64),
                                            std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (status.load() != SPX_OK) return fail(status.load(), msg);
  }
  if (w.size() > static_cast<size_t>(0x7fffffff)) return fail(SPX_ENOMEM, "batch tables too large");
  // One queue, largest estimated work first (orders x steps x amplitudes
  // x pair and field terms), so that the last realisations to start are
  // the short ones.
  bool mixed = false;
  for (size_t i = 1; i < n && !mixed; ++i) mixed = reals[i].nspins != reals[0].nspins;
  std::vector<int> order(n);
  for (size_t i = 0; i < n; ++i) {
    order[i] = static_cast<int>(i);
    const double L = reals[i].nspins;
    cost[i] *= L * (L - 1.0) / 2.0 + L;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });

  DevOp* d_ops = nullptr;
  DevPulse* d_pulses = nullptr;
  DevEvolve* d_evo = nullptr;
  DevReal* d_reals = nullptr;
  int *d_order = nullptr, *d_codes = nullptr, *d_spins = nullptr, *d_queue = nullptr;
  double *d_w = nullptr, *d_psi = nullptr, *d_out = nullptr;
  std::vector<int> queue0(2, 0);   // queue counter, zero-norm flag
  int rc = SPX_OK;
  const size_t nout = n * points * b->width;
  if (rc == SPX_OK) rc = upload(c, &d_ops, ops);
  if (rc == SPX_OK) rc = upload(c, &d_pulses, pulses);
  if (rc == SPX_OK) rc = upload(c, &d_evo, evo);
  if (rc == SPX_OK) rc = upload(c, &d_reals, reals);
  if (rc == SPX_OK) rc = upload(c, &d_order, order);
  if (rc == SPX_OK) rc = upload(c, &d_codes, b->obs_codes);
  if (rc == SPX_OK) rc = upload(c, &d_spins, b->obs_spins);
  if (rc == SPX_OK) rc = upload(c, &d_queue, queue0);
  if (rc == SPX_OK) rc = upload(c, &d_w, w);
  if (rc == SPX_OK) rc = upload(c, &d_psi, psi);
  if (rc == SPX_OK) rc = upload(c, &d_out, std::vector<double>(nout ? nout : 1, 0.0));
  // One persistent launch: the walker of the realisations' one size, or,
  // for several sizes, the kernel that runs each realisation with the walker
  // of its own L (a size-by-size launch sequence left the largest size's
  // last realisations running alone; on 4096 realisations at L = 8..12, one
  // launch for all sizes was 3.5% faster than concurrent launches per size).
  const int nreal = static_cast<int>(n);
  const int nblocks = std::min(nreal, 2 * c->arch_cus);
  if (rc == SPX_OK && mixed) {
    size_t lds = 0;
    for (size_t i = 0; i < n; ++i) lds = std::max(lds, resident_lds_bytes(reals[i].nspins));
    rc = launch_resident_all(c, d_ops, static_cast<int>(ops.size()), d_pulses, d_evo, d_reals, d_order, nreal, d_w,
                             d_psi, d_out, static_cast<int>(b->width), d_codes, d_spins,
                             static_cast<int>(b->obs_codes.size()), b->reference_axis, d_queue, nblocks, lds,
                             c->stream);
  } else if (rc == SPX_OK) {
    rc = launch_resident(c, reals[0].nspins, d_ops, static_cast<int>(ops.size()), d_pulses, d_evo, d_reals, d_order,
                         nreal, d_w, d_psi, d_out, static_cast<int>(b->width), d_codes, d_spins,
                         static_cast<int>(b->obs_codes.size()), b->reference_axis, d_queue, nblocks, c->stream);
  }
  std::vector<int> flags(queue0.size(), 0);
  if (rc == SPX_OK) rc = copy_to_host(c, values, d_out, nout * sizeof(double));
  if (rc == SPX_OK) rc = copy_to_host(c, flags.data(), d_queue, flags.size() * sizeof(int));
  // the launch has finished (or none was made) before the buffers go
  (void)hipStreamSynchronize(c->stream);
  for (void* p : {static_cast<void*>(d_ops), static_cast<void*>(d_pulses), static_cast<void*>(d_evo),
                  static_cast<void*>(d_reals), static_cast<void*>(d_order), static_cast<void*>(d_codes),
                  static_cast<void*>(d_spins), static_cast<void*>(d_queue), static_cast<void*>(d_w),
                  static_cast<void*>(d_psi), static_cast<void*>(d_out)})
    if (p) (void)hipFree(p);
  SPX_TRY(rc);
  if (flags[1]) return fail(SPX_EDOMAIN, "cannot normalize a zero state");
  // times: Sequence::run's clock (src/sequence/Sequence.cpp:47)
  if (times) {
    double now = 0.0;
    size_t row = 0;
    std::vector<std::pair<size_t, size_t>> stack;
    for (size_t pc = 0; pc < prog.size(); ++pc) {
      const ProgOp& op = prog[pc];
      if (op.kind == OP_EVOLVE) {
        now += op.tau * static_cast<double>(op.nsteps);
      } else if (op.kind == OP_MEASURE) {
        times[row++] = now;
      } else if (op.kind == OP_REPEAT) {
        if (op.count == 0) pc = op.match;
        else stack.emplace_back(pc, op.count);
      } else if (op.kind == OP_END) {
        if (--stack.back().second > 0) pc = stack.back().first;
        else stack.pop_back();
      }
    }
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

int spx_batch_add_dipolar(spx_batch* b, uint64_t seed, size_t first, size_t count, size_t nspins, double scale,
                          double min_separation, int angular, double disorder) {
  if (!b) return fail(SPX_EINVAL, "null batch");
  const DipolarSpec spec{seed, nspins, scale, min_separation, angular != 0, disorder};
  SPX_TRY(check_spec(spec));
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
  const size_t np = nspins * (nspins - 1) / 2;
  std::vector<Realisation> drawn(count);
  std::atomic<size_t> next{0};
  std::atomic<int> status{SPX_OK};
  auto work = [&] {
    std::vector<double> hz(nspins);
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= count || status.load() != SPX_OK) return;
      Realisation& r = drawn[i];
      r.nspins = nspins;
      r.lo.resize(np);
      r.hi.resize(np);
      r.jz.resize(np);
      const int rc = draw_dipolar(spec, first + i, r.lo.data(), r.hi.data(), r.jz.data(), hz.data());
      if (rc != SPX_OK) {
        status.store(rc);
        return;
      }
      r.drawn_state = true;
      r.spec = spec;
      r.index = first + i;
      r.jx.resize(np);
      r.jy.resize(np);
      for (size_t p = 0; p < np; ++p) r.jx[p] = r.jy[p] = -0.5 * r.jz[p];
      r.fields.assign(3 * nspins, 0.0);
      for (size_t q = 0; q < nspins; ++q) r.fields[3 * q + 2] = hz[q];
    }
  };
  const size_t threads = std::min<size_t>(
      std::max<size_t>(1, count / 4), std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency())));
  std::vector<std::thread> pool;
  for (size_t t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (status.load() != SPX_OK) return fail(status.load(), "could not place the spins at the minimum separation");
  for (Realisation& r : drawn) b->items.push_back(std::move(r));
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
  static const int forced = [] {
    const char* e = std::getenv("SPX_BATCH_STREAMING");
    return e ? std::atoi(e) : 0;
  }();
  if (!forced && resident_ok(b)) return batch_run_resident(b, times, values);
  return batch_run_streaming(b, 0, b->items.size(), times, values);
}

}  // extern "C"
