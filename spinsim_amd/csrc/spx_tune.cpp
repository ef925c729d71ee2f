// Autotuner of the order-kernel schedule (no reference counterpart).
//
// For one Hamiltonian at one step bound it times every eligible schedule —
// the Hadamard-sandwich path, the all-pairs kernel at tile bits 10 and 9,
// the table kernels at tile bits
// 6..10, each with and without the XCD-aware tile order, and the
// per-amplitude-group kernel at two block sizes — on the general Chebyshev
// order (T_n = 2Ĥ T_{n-1} - T_{n-2}, psi += c T_n) and keeps the fastest.
// Small tiles (6, 7 bits) are tried up to L = 16, where few large tiles
// would leave most compute units idle.
// Winners are cached per problem shape (L, whether the flip terms are
// flip-flop pairs only, number of flip terms, states per launch), so the
// realisations of an ensemble pay
// for one tuning. Schedules change the order of floating-point additions, not
// the terms, so results agree to rounding across schedules.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "spx_common.hpp"

namespace spx {
namespace {

using Key = std::tuple<size_t, bool, size_t, int>;

std::mutex g_mu;
std::map<Key, Schedule> g_cache;
int g_mode = 0;

// The shape a schedule depends on, at the scale of one step: L, whether the
// flip terms are flip-flop pairs only (the all-pairs and Hadamard paths), the
// flip-term count, and how many states share each launch.
Key shape_of(const spx_ham* h, double scale, int nvec) {
  bool fields = false, flipboth = false;
  size_t terms = 0;
  for (size_t s = 0; s < h->nspins; ++s) {
    const bool on = std::abs(scale * 0.5 * h->hx[s]) > kNegligible || std::abs(scale * 0.5 * h->hy[s]) > kNegligible;
    fields |= on;
    terms += on ? 1 : 0;
  }
  for (size_t p = 0; p < h->lo.size(); ++p) {
    const double ff = scale * 0.25 * (h->jx[p] + h->jy[p]);
    const double fb = scale * 0.25 * (h->jx[p] - h->jy[p]);
    flipboth |= std::abs(fb) > kNegligible;
    terms += (std::abs(ff) > kNegligible || std::abs(fb) > kNegligible) ? 1 : 0;
  }
  return Key{h->nspins, !fields && !flipboth, terms, nvec};
}

std::string describe(const Schedule& s, const spx_ham* h) {
  if (h->wht_ok) return "hadamard";
  if (h->dense_ok) return "all-pairs tb=" + std::to_string(dense_tile_bits(h)) + " xcd=" + std::to_string(xcd_remap(h));
  const std::string kind = h->terms_variant == 3 ? "table" : "group";
  std::string d = kind + " tb=" + std::to_string(h->terms_tb) + " xcd=" + std::to_string(xcd_remap(h));
  if (h->terms_variant != 3) d += " block=" + std::to_string(block_threads(h));
  (void)s;
  return d;
}

std::vector<Schedule> candidates(const spx_ham* h, bool flipflop) {
  std::vector<Schedule> c;
  const int L = static_cast<int>(h->nspins);
  if (hadamard_eligible(h->nspins)) {
    Schedule s;
    s.dense = 2;
    c.push_back(s);
  }
  if (flipflop && L >= 13 && L <= 18)
    for (int tb : {10, 9})
      for (int xcd : {0, 1}) {
        Schedule s;
        s.dense = 1;
        s.dense_tb = tb;
        s.xcd = xcd;
        c.push_back(s);
      }
  for (int tb : {10, 9, 8, 7, 6}) {
    if (tb > L || (tb < 8 && L > 16)) continue;
    for (int xcd : {0, 1}) {
      if (xcd == 1 && L - tb < 3) continue;
      Schedule s;
      s.dense = 0;
      s.tb = tb;
      s.xcd = xcd;
      c.push_back(s);
    }
  }
  for (int tb : {10, 9}) {
    if (tb > L) continue;
    for (int block : {256, 1024}) {
      Schedule s;
      s.dense = 0;
      s.tb = tb;
      s.table = 2;
      s.block = block;
      c.push_back(s);
    }
  }
  return c;
}

struct Scratch {
  double2* v[3 * kMaxMany] = {};
  ~Scratch() {
    for (double2* p : v)
      if (p) (void)hipFree(p);
  }
};

// nvec states share every timed launch (the width of a batched evolve).
int tune(spx_ham* h, double bound, int nvec, std::string* report) {
  if (!(bound > 0.0)) return fail(SPX_EDOMAIN, "spectral bound must be positive");
  if (nvec < 1 || nvec > kMaxMany) return fail(SPX_EINVAL, "states per launch must be in [1, 16]");
  const double scale = -2.0 / bound;
  const Key key = shape_of(h, scale, nvec);
  spx_ctx* c = h->ctx;
  const size_t n = static_cast<size_t>(1) << h->nspins;
  Scratch sc;
  for (int i = 0; i < 3 * nvec; ++i) {
    SPX_HIP(hipMalloc(&sc.v[i], n * sizeof(double2)));
    SPX_HIP(hipMemsetAsync(sc.v[i], 0, n * sizeof(double2), c->stream));
  }
  hipEvent_t e0, e1;
  SPX_HIP(hipEventCreate(&e0));
  SPX_HIP(hipEventCreate(&e1));
  const int reps = n >= (size_t(1) << 16) ? 12 : 40;
  double best = 1e300;
  Schedule best_s;
  std::string log;
  int rc = SPX_OK;
  for (const Schedule& s : candidates(h, std::get<1>(key))) {
    h->sched = s;
    h->terms_valid = false;
    if ((rc = ham_prepare(h, scale)) != SPX_OK) break;
    if ((s.dense == 1 && !h->dense_ok) || (s.dense == 2 && !h->wht_ok)) continue;
    if (h->wht_ok && (rc = hadamard_reserve(h, nvec)) != SPX_OK) break;
    OrderLaunch v{};
    for (int i = 0; i < nvec; ++i) {
      v.in[i] = sc.v[3 * i];
      v.out[i] = sc.v[3 * i + 1];
      v.psi[i] = sc.v[3 * i + 2];
    }
    v.nvec = nvec;
    // the steady state of an expansion: each order hands its row transforms
    // to the next (Hadamard path; the other kernels ignore the flags)
    for (int i = 0; i < nvec; ++i) v.uv_next[i] = 1;
    for (int i = 0; i < 2 && rc == SPX_OK; ++i) {
      rc = launch_order(h, 3, v, 0.0, 0.0, 0.0, 1e-3, 0.0);
      v.uv_ready = 1;
    }
    if (rc == SPX_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(SPX_EDEVICE, "hipEventRecord failed");
    for (int i = 0; i < reps && rc == SPX_OK; ++i) rc = launch_order(h, 3, v, 0.0, 0.0, 0.0, 1e-3, 0.0);
    if (rc == SPX_OK && hipEventRecord(e1, c->stream) != hipSuccess) rc = fail(SPX_EDEVICE, "hipEventRecord failed");
    float ms = 0.f;
    if (rc == SPX_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = fail(SPX_EDEVICE, "timing the candidate schedule failed");
    if (rc != SPX_OK) break;
    const double us = 1e3 * ms / reps;
    log += describe(s, h) + ": " + std::to_string(us).substr(0, 7) + " us; ";
    if (us < best) {
      best = us;
      best_s = s;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != SPX_OK) return rc;
  h->sched = best_s;
  h->sched_set = true;
  h->terms_valid = false;
  SPX_TRY(ham_prepare(h, scale));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best_s;
  }
  if (report) *report = "chosen " + describe(best_s, h) + " (" + std::to_string(best).substr(0, 7) + " us/order); " + log;
  return SPX_OK;
}

}  // namespace

int autotune_before_evolve(spx_ham* h, double bound, int nvec) {
  int mode;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    mode = g_mode;
  }
  if (mode == 0 || h->sched_set) return SPX_OK;
  const Key key = shape_of(h, -2.0 / bound, nvec);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      h->sched = it->second;
      h->sched_set = true;
      h->terms_valid = false;
      return SPX_OK;
    }
  }
  return tune(h, bound, nvec, nullptr);
}

}  // namespace spx

using namespace spx;

extern "C" {

int spx_autotune_mode(int mode) {
  if (mode != 0 && mode != 1) return fail(SPX_EINVAL, "autotune mode must be 0 or 1");
  std::lock_guard<std::mutex> lk(g_mu);
  g_mode = mode;
  return SPX_OK;
}

int spx_ham_autotune(spx_ham* h, double bound, char* report, size_t len) {
  return spx_ham_autotune_batch(h, bound, 1, report, len);
}

int spx_ham_autotune_batch(spx_ham* h, double bound, int nvec, char* report, size_t len) {
  if (!h) return fail(SPX_EINVAL, "null Hamiltonian");
  std::string r;
  SPX_TRY(tune(h, bound, nvec, &r));
  if (report && len > 0) {
    const size_t k = std::min(len - 1, r.size());
    std::copy(r.begin(), r.begin() + static_cast<long>(k), report);
    report[k] = '\0';
  }
  return SPX_OK;
}

int spx_ham_set_schedule(spx_ham* h, int tile_bits, int block, int xcd, int dense, int dense_tile_bits) {
  if (!h) return fail(SPX_EINVAL, "null Hamiltonian");
  if (tile_bits != 0 && (tile_bits < 6 || tile_bits > 12)) return fail(SPX_EINVAL, "tile bits must be 0 or 6..12");
  if (block != 0 && block != 256 && block != 512 && block != 1024) return fail(SPX_EINVAL, "block must be 0, 256, 512 or 1024");
  if (dense_tile_bits != 0 && dense_tile_bits != 9 && dense_tile_bits != 10)
    return fail(SPX_EINVAL, "all-pairs tile bits must be 0, 9 or 10");
  Schedule s;
  s.tb = tile_bits;
  s.block = block;
  s.xcd = xcd < 0 ? -1 : (xcd ? 1 : 0);
  if (dense > 2) return fail(SPX_EINVAL, "dense must be -1, 0, 1 (pair gathers) or 2 (Hadamard transforms)");
  s.dense = dense < 0 ? -1 : dense;
  s.dense_tb = dense_tile_bits;
  h->sched = s;
  h->sched_set = true;
  h->terms_valid = false;
  return SPX_OK;
}

}  // extern "C"
