// The extern "C" surface of libspx (include/spx.h): contexts, device state
// vectors, Hamiltonians, the Chebyshev propagator driver, gates,
// observables, pulse programs and the batched ensemble runner.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "spx_batch.hpp"
#include "spx_common.hpp"

namespace spx {
const char* last_error();
void last_convergence(double* b, double* t, double* e, size_t* o);
}  // namespace spx

using namespace spx;

namespace {


int check_state(const spx_state* s) {
  if (!s || !s->d) return fail(SPX_EINVAL, "null state");
  return SPX_OK;
}

int same_space(const spx_state* a, const spx_state* b) {
  SPX_TRY(check_state(a));
  SPX_TRY(check_state(b));
  if (a->n != b->n) return fail(SPX_EINVAL, "state vectors span different spaces");
  return SPX_OK;
}

int set_device(spx_ctx* c) {
  SPX_HIP(hipSetDevice(c->device));
  return SPX_OK;
}

int reserve_bounce(spx_ctx* c, size_t bytes) {
  const size_t want = bytes < kBounceBytes ? bytes : kBounceBytes;
  if (want <= c->bounce_cap) return SPX_OK;
  SPX_HIP(hipStreamSynchronize(c->stream));
  if (c->bounce) SPX_HIP(hipHostFree(c->bounce));
  c->bounce = nullptr;
  c->bounce_cap = 0;
  SPX_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->bounce), want, hipHostMallocDefault));
  c->bounce_cap = want;
  return SPX_OK;
}

}  // namespace

namespace spx {

int copy_to_device(spx_ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return SPX_OK;
  SPX_TRY(reserve_bounce(c, bytes));
  auto* d = static_cast<unsigned char*>(dst);
  const auto* h = static_cast<const unsigned char*>(src);
  for (size_t off = 0; off < bytes; off += c->bounce_cap) {
    const size_t n = bytes - off < c->bounce_cap ? bytes - off : c->bounce_cap;
    // the previous piece's copy has completed: the bounce buffer is free
    std::memcpy(c->bounce, h + off, n);
    SPX_HIP(hipMemcpyAsync(d + off, c->bounce, n, hipMemcpyHostToDevice, c->stream));
    SPX_HIP(hipStreamSynchronize(c->stream));
  }
  return SPX_OK;
}

int copy_to_host(spx_ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return SPX_OK;
  SPX_TRY(reserve_bounce(c, bytes));
  auto* h = static_cast<unsigned char*>(dst);
  const auto* d = static_cast<const unsigned char*>(src);
  for (size_t off = 0; off < bytes; off += c->bounce_cap) {
    const size_t n = bytes - off < c->bounce_cap ? bytes - off : c->bounce_cap;
    SPX_HIP(hipMemcpyAsync(c->bounce, d + off, n, hipMemcpyDeviceToHost, c->stream));
    SPX_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(h + off, c->bounce, n);
  }
  return SPX_OK;
}

}  // namespace spx

extern "C" {

const char* spx_last_error(void) { return spx::last_error(); }

int spx_last_convergence(double* bound, double* tau, double* epsilon, size_t* order) {
  spx::last_convergence(bound, tau, epsilon, order);
  return SPX_OK;
}

// ------------------------------------------------------------------ context

int spx_device_count(int* count) {
  int n = 0;
  SPX_HIP(hipGetDeviceCount(&n));
  *count = n;
  return SPX_OK;
}

int spx_ctx_create(int device, spx_ctx** out) {
  if (!out) return fail(SPX_EINVAL, "null output");
  int n = 0;
  SPX_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(SPX_ERANGE, "device index out of range");
  SPX_HIP(hipSetDevice(device));
  auto* c = new spx_ctx();
  c->device = device;
  SPX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  SPX_HIP(hipMalloc(&c->d_partials, 2 * kReduceBlocks * sizeof(double)));
  SPX_HIP(hipMalloc(&c->d_result, 16 * sizeof(double)));
  SPX_HIP(hipHostMalloc(&c->h_result, 16 * sizeof(double), hipHostMallocDefault));
  for (int i = 0; i < kMaxEvents; ++i) SPX_HIP(hipEventCreate(&c->events[i]));
  hipDeviceProp_t prop;
  SPX_HIP(hipGetDeviceProperties(&prop, device));
  c->arch_cus = prop.multiProcessorCount;
  *out = c;
  return SPX_OK;
}

int spx_ctx_destroy(spx_ctx* c) {
  if (!c) return SPX_OK;
  // teardown is best effort: a failure here leaves nothing to recover
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  for (int i = 0; i < kMaxEvents; ++i)
    if (c->events[i]) (void)hipEventDestroy(c->events[i]);
  (void)hipFree(c->d_partials);
  (void)hipFree(c->d_result);
  (void)hipHostFree(c->h_result);
  if (c->bounce) (void)hipHostFree(c->bounce);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return SPX_OK;
}

int spx_ctx_synchronize(spx_ctx* c) {
  SPX_HIP(hipStreamSynchronize(c->stream));
  return SPX_OK;
}

int spx_ctx_stream(spx_ctx* c, void** stream) {
  *stream = static_cast<void*>(c->stream);
  return SPX_OK;
}

int spx_ctx_event_record(spx_ctx* c, int slot) {
  if (slot < 0 || slot >= kMaxEvents) return fail(SPX_ERANGE, "event slot out of range");
  SPX_HIP(hipEventRecord(c->events[slot], c->stream));
  return SPX_OK;
}

int spx_ctx_event_elapsed(spx_ctx* c, int s0, int s1, float* ms) {
  if (s0 < 0 || s0 >= kMaxEvents || s1 < 0 || s1 >= kMaxEvents)
    return fail(SPX_ERANGE, "event slot out of range");
  SPX_HIP(hipEventSynchronize(c->events[s1]));
  SPX_HIP(hipEventElapsedTime(ms, c->events[s0], c->events[s1]));
  return SPX_OK;
}

// -------------------------------------------------------------- StateVector

int spx_state_create(spx_ctx* c, size_t nspins, spx_state** out) {
  if (!c || !out) return fail(SPX_EINVAL, "null argument");
  if (nspins > kMaxSpins) return fail(SPX_ERANGE, "number of spins above kMaxSpins");
  SPX_TRY(set_device(c));
  auto* s = new spx_state();
  s->ctx = c;
  s->nspins = nspins;
  s->n = static_cast<size_t>(1) << nspins;
  hipError_t e = hipMalloc(&s->d, s->n * sizeof(double2));
  if (e != hipSuccess) {
    delete s;
    return fail(SPX_ENOMEM, "hipMalloc of a state vector failed");
  }
  int rc = launch_fill(c, s->d, s->n);
  if (rc != SPX_OK) {
    (void)hipFree(s->d);
    delete s;
    return rc;
  }
  *out = s;
  return SPX_OK;
}

// Destroy functions do not touch the owning context (it may already be
// gone); hipFree waits for outstanding work on the buffer itself.
int spx_state_destroy(spx_state* s) {
  if (!s) return SPX_OK;
  (void)hipFree(s->d);
  delete s;
  return SPX_OK;
}

int spx_state_nspins(const spx_state* s, size_t* nspins) {
  SPX_TRY(check_state(s));
  *nspins = s->nspins;
  return SPX_OK;
}

int spx_state_upload(spx_state* s, const double* host) {
  SPX_TRY(check_state(s));
  if (!host) return fail(SPX_EINVAL, "null host buffer");
  return copy_to_device(s->ctx, s->d, host, s->n * sizeof(double2));
}

int spx_state_download(const spx_state* s, double* host) {
  SPX_TRY(check_state(s));
  if (!host) return fail(SPX_EINVAL, "null host buffer");
  return copy_to_host(s->ctx, host, s->d, s->n * sizeof(double2));
}

int spx_state_copy(spx_state* dst, const spx_state* src) {
  SPX_TRY(same_space(dst, src));
  if (dst == src) return SPX_OK;
  SPX_HIP(hipMemcpyAsync(dst->d, src->d, src->n * sizeof(double2), hipMemcpyDeviceToDevice,
                         dst->ctx->stream));
  return SPX_OK;
}

int spx_state_set_zero(spx_state* s) {
  SPX_TRY(check_state(s));
  return launch_fill(s->ctx, s->d, s->n);
}

int spx_state_set_basis(spx_state* s, size_t index) {
  SPX_TRY(check_state(s));
  if (index >= s->n) return fail(SPX_ERANGE, "basis state index outside the 2^L state space");
  SPX_TRY(launch_fill(s->ctx, s->d, s->n));
  static const double one[2] = {1.0, 0.0};
  return copy_to_device(s->ctx, s->d + index, one, sizeof(double2));
}

int spx_state_scale(spx_state* s, double f) {
  SPX_TRY(check_state(s));
  return launch_scale(s->ctx, s->d, f, s->n);
}

int spx_state_negate(spx_state* s) {
  SPX_TRY(check_state(s));
  return launch_negate(s->ctx, s->d, s->n);
}

int spx_state_add_scaled(spx_state* dst, double cr, double ci, const spx_state* src) {
  SPX_TRY(same_space(dst, src));
  return launch_add_scaled(dst->ctx, dst->d, cr, ci, src->d, dst->n);
}

int spx_state_norm(const spx_state* s, double* norm) {
  SPX_TRY(check_state(s));
  double sum = 0.0;
  SPX_TRY(reduce_norm2(s->ctx, s->d, s->n, &sum));
  *norm = std::sqrt(sum);
  return SPX_OK;
}

int spx_state_normalize(spx_state* s, double* norm_before) {
  double nrm = 0.0;
  SPX_TRY(spx_state_norm(s, &nrm));
  if (norm_before) *norm_before = nrm;
  if (nrm == 0.0) return fail(SPX_EDOMAIN, "cannot normalize a zero state");
  return launch_scale(s->ctx, s->d, 1.0 / nrm, s->n);
}

int spx_state_inner(const spx_state* self, const spx_state* other, double* re, double* im) {
  SPX_TRY(same_space(self, other));
  return reduce_inner(self->ctx, self->d, other->d, self->n, re, im);
}

// ------------------------------------------------------- Bessel / Chebyshev

int spx_bessel(int kind, double x, size_t count, double* values, size_t* nonzero) {
  if (!values && count) return fail(SPX_EINVAL, "null output");
  return spx::bessel(kind, x, count, values, nonzero);
}

int spx_cheb_build(int mode, double bound, double tau, double epsilon, size_t max_order,
                   double* coeffs, size_t capacity, size_t* order, double* effective_bound) {
  std::vector<double> c;
  double eb = 1.0;
  SPX_TRY(cheb_build(mode, bound, tau, epsilon, max_order, c, eb));
  if (c.size() > capacity) return fail(SPX_EINVAL, "coefficient buffer too small");
  std::memcpy(coeffs, c.data(), c.size() * sizeof(double));
  *order = c.size() - 1;
  *effective_bound = eb;
  return SPX_OK;
}

// -------------------------------------------------------------- Hamiltonian

int spx_ham_create(spx_ctx* c, size_t nspins, spx_ham** out) {
  if (!c || !out) return fail(SPX_EINVAL, "null argument");
  if (nspins == 0 || nspins > kMaxSpins)
    return fail(SPX_ERANGE, "number of spins must be in [1, kMaxSpins]");
  auto* h = new spx_ham();
  h->ctx = c;
  h->nspins = nspins;
  h->hx.assign(nspins, 0.0);
  h->hy.assign(nspins, 0.0);
  h->hz.assign(nspins, 0.0);
  *out = h;
  return SPX_OK;
}

int spx_ham_destroy(spx_ham* h) {
  if (!h) return SPX_OK;
  if (h->d_diag) (void)hipFree(h->d_diag);
  if (h->d_terms) (void)hipFree(h->d_terms);
  if (h->d_tab) (void)hipFree(h->d_tab);
  if (h->d_scratch) (void)hipFree(h->d_scratch);
  if (h->d_coef) (void)hipFree(h->d_coef);
  for (auto& t : h->wht_cache) {
    if (t.copied) (void)hipEventSynchronize(t.copied);
    if (t.copied) (void)hipEventDestroy(t.copied);
    if (t.dev) (void)hipFree(t.dev);
    if (t.host) (void)hipHostFree(t.host);
  }
  if (h->d_wht) (void)hipFree(h->d_wht);
  delete h;
  return SPX_OK;
}

int spx_ham_set_field(spx_ham* h, size_t spin, double x, double y, double z) {
  if (spin >= h->nspins) return fail(SPX_ERANGE, "spin index out of range");
  // the Hadamard path's factor tables hold hz, the other paths' tables the
  // transverse fields
  if (h->hx[spin] != x || h->hy[spin] != y || h->hz[spin] != z) h->terms_valid = false;
  h->hx[spin] = x;
  h->hy[spin] = y;
  h->hz[spin] = z;
  return SPX_OK;
}

int spx_ham_get_field(const spx_ham* h, size_t spin, double* xyz) {
  if (spin >= h->nspins) return fail(SPX_ERANGE, "spin index out of range");
  xyz[0] = h->hx[spin];
  xyz[1] = h->hy[spin];
  xyz[2] = h->hz[spin];
  return SPX_OK;
}

int spx_ham_set_coupling(spx_ham* h, size_t a, size_t b, double jx, double jy, double jz) {
  if (a == b) return fail(SPX_EINVAL, "a spin cannot couple to itself");
  const size_t lo = a < b ? a : b, hi = a < b ? b : a;
  if (hi >= h->nspins) return fail(SPX_ERANGE, "coupled spin index out of range");
  h->terms_valid = false;
  for (size_t p = 0; p < h->lo.size(); ++p)
    if (h->lo[p] == lo && h->hi[p] == hi) {
      h->jx[p] = jx;
      h->jy[p] = jy;
      h->jz[p] = jz;
      return SPX_OK;
    }
  h->lo.push_back(static_cast<uint32_t>(lo));
  h->hi.push_back(static_cast<uint32_t>(hi));
  h->jx.push_back(jx);
  h->jy.push_back(jy);
  h->jz.push_back(jz);
  return SPX_OK;
}

int spx_ham_num_couplings(const spx_ham* h, size_t* n) {
  *n = h->lo.size();
  return SPX_OK;
}

int spx_ham_get_coupling(const spx_ham* h, size_t i, uint32_t* lo, uint32_t* hi, double* jxyz) {
  if (i >= h->lo.size()) return fail(SPX_ERANGE, "coupling index out of range");
  *lo = h->lo[i];
  *hi = h->hi[i];
  jxyz[0] = h->jx[i];
  jxyz[1] = h->jy[i];
  jxyz[2] = h->jz[i];
  return SPX_OK;
}

// spectralBound (src/core/Hamiltonian.cpp:82-93), same summation order.
int spx_ham_spectral_bound(const spx_ham* h, double* bound) {
  double b = 0.0;
  for (size_t s = 0; s < h->nspins; ++s)
    b += 0.5 * (std::abs(h->hx[s]) + std::abs(h->hy[s]) + std::abs(h->hz[s]));
  for (size_t p = 0; p < h->lo.size(); ++p)
    b += 0.25 * (std::abs(h->jx[p]) + std::abs(h->jy[p]) + std::abs(h->jz[p]));
  *bound = b;
  return SPX_OK;
}

int spx_ham_apply(spx_ham* h, int kind, double bound, const spx_state* in, spx_state* out) {
  SPX_TRY(same_space(in, out));
  if (in->nspins != h->nspins) return fail(SPX_EINVAL, "state and Hamiltonian disagree on spin count");
  if (in == out) return fail(SPX_EINVAL, "kernels require distinct input and output");
  SPX_TRY(set_device(h->ctx));
  if (kind == 0) return launch_apply(h, 1.0, in->d, out->d, false);
  if (!(bound > 0.0)) return fail(SPX_EDOMAIN, "spectral bound must be positive");
  return launch_apply(h, -2.0 / bound, in->d, out->d, kind == 2);
}

int spx_ham_diagonal(spx_ham* h, double* host) {
  SPX_TRY(set_device(h->ctx));
  SPX_TRY(ensure_diag(h));
  const size_t n = static_cast<size_t>(1) << h->nspins;
  return copy_to_host(h->ctx, host, h->d_diag, n * sizeof(double));
}

int spx_ham_schedule(const spx_ham* h, char* buf, size_t len) {
  if (!h || !buf || len == 0) return fail(SPX_EINVAL, "null argument");
  std::string name;
  if (!h->terms_valid) name = "unprepared";
  else if (h->wht_ok)
    name = "k_wht_cols<" + std::to_string(h->nspins - 10) + "> + k_wht_rows_bwd<K_ORDERN>";
  else if (h->dense_ok)
    name = "k_dense<K_ORDERN, " + std::to_string(h->nspins) + ", " + std::to_string(dense_tile_bits(h)) + ">";
  else if (h->terms_variant == 3)
    name = "k_order3<K_ORDERN, G=" + std::to_string(h->terms_g) + ">, tile bits " + std::to_string(h->terms_tb);
  else
    name = "k_horder<K_ORDERN>, tile bits " + std::to_string(h->terms_tb);
  const size_t n = std::min(len - 1, name.size());
  std::memcpy(buf, name.data(), n);
  buf[n] = '\0';
  return SPX_OK;
}

// --------------------------------------------------------------- Propagator

int spx_prop_set_graphs(spx_prop* p, int enable) {
  if (!p) return fail(SPX_EINVAL, "null propagator");
  p->graphs = enable ? 1 : 0;
  return SPX_OK;
}

int spx_prop_set_streams(spx_prop* p, int streams) {
  if (!p) return fail(SPX_EINVAL, "null propagator");
  if (streams != 1 && streams != 2) return fail(SPX_EINVAL, "streams must be 1 or 2");
  p->streams = streams;
  return SPX_OK;
}

int spx_prop_create(spx_ctx* c, size_t nspins, double epsilon, spx_prop** out) {
  if (!c || !out) return fail(SPX_EINVAL, "null argument");
  if (!(epsilon > 0.0)) return fail(SPX_EINVAL, "epsilon must be positive");
  if (nspins > kMaxSpins) return fail(SPX_ERANGE, "number of spins above kMaxSpins");
  auto* p = new spx_prop();
  p->ctx = c;
  p->nspins = nspins;
  p->epsilon = epsilon;
  *out = p;
  return SPX_OK;
}

int spx_prop_destroy(spx_prop* p) {
  if (!p) return SPX_OK;
  for (auto& d : p->d_scratch)
    if (d) (void)hipFree(d);
  delete p;
  return SPX_OK;
}

int spx_prop_last_order(const spx_prop* p, size_t* order) {
  *order = p->last_order;
  return SPX_OK;
}

int spx_prop_epsilon(const spx_prop* p, double* epsilon) {
  *epsilon = p->epsilon;
  return SPX_OK;
}

namespace {

int ensure_scratch(spx_prop* p, size_t nvec) {
  const size_t n = static_cast<size_t>(1) << p->nspins;
  for (size_t i = 0; i < 2 * nvec; ++i)
    if (!p->d_scratch[i]) {
      hipError_t e = hipMalloc(&p->d_scratch[i], n * sizeof(double2));
      if (e != hipSuccess) return fail(SPX_ENOMEM, "hipMalloc of propagator scratch failed");
    }
  return SPX_OK;
}

// applyOneStep (src/core/Propagator.cpp:46-73) for up to kMaxMany states:
// T_1 into B; T_2 into A with psi = a0 T_0 + c1 T_1 + c2 T_2 (psi still holds
// T_0 until then, so no copy of it is made); from there T_n overwrites
// T_{n-2} in place and psi += c_n T_n in the same pass.
constexpr size_t kGraphMinSteps = 4;

int one_step_set(spx_prop* p, spx_ham* const* hs, int mode, spx_state* const* states, size_t nvec,
                 const std::vector<double>* const* a, bool chain_in = false, bool chain_out = false,
                 size_t slot0 = 0);

int one_step(spx_prop* p, spx_ham* h, int mode, spx_state* const* states, size_t nvec,
             const std::vector<double>& a) {
  spx_ham* hs[kMaxMany];
  const std::vector<double>* cs[kMaxMany];
  for (size_t i = 0; i < nvec; ++i) {
    hs[i] = h;
    cs[i] = &a;
  }
  return one_step_set(p, hs, mode, states, nvec, cs);
}

int evolve_impl(spx_prop* p, spx_ham* h, int mode, spx_state* const* states, size_t nstates,
                const std::vector<double>* prebuilt, double prebuilt_bound, double tau,
                size_t nsteps, int autonormalize) {
  if (!p || !h || !states) return fail(SPX_EINVAL, "null argument");
  // the order launches go to the Hamiltonian's stream and the first-order
  // update to the propagator's: both, and every state, must share one
  // context or the launches of one step would run on unordered streams
  if (h->ctx != p->ctx) return fail(SPX_EINVAL, "Hamiltonian belongs to another context");
  if (nstates == 0) return SPX_OK;
  for (size_t i = 0; i < nstates; ++i) {
    SPX_TRY(check_state(states[i]));
    if (states[i]->ctx != p->ctx) return fail(SPX_EINVAL, "state belongs to another context");
    if (states[i]->nspins != h->nspins)
      return fail(SPX_EINVAL, "state and Hamiltonian disagree on spin count");
    if (states[i]->nspins != p->nspins)
      return fail(SPX_EINVAL, "state does not match the propagator's size");
    for (size_t j = 0; j < i; ++j)
      if (states[j] == states[i]) return fail(SPX_EINVAL, "a state appears twice");
  }
  if (mode != 0 && mode != 1) return fail(SPX_EINVAL, "unknown time mode");
  if (nsteps == 0) return SPX_OK;

  std::vector<double> coeffs;
  double bound = 1.0;
  if (prebuilt) {
    coeffs = *prebuilt;
    bound = prebuilt_bound;
    if (coeffs.empty()) return fail(SPX_EINVAL, "empty expansion");
  } else {
    double sb = 0.0;
    spx_ham_spectral_bound(h, &sb);
    SPX_TRY(cheb_build(mode, sb, tau, p->epsilon, 6000, coeffs, bound));
    p->last_order = coeffs.size() - 1;
  }
  if (coeffs.size() == 1) return SPX_OK;  // identity: psi untouched
  if (!(bound > 0.0)) return fail(SPX_EDOMAIN, "spectral bound must be positive");
  SPX_TRY(set_device(p->ctx));
  SPX_TRY(ensure_scratch(p, nstates < kMaxMany ? nstates : kMaxMany));
  const int width = static_cast<int>(nstates < kMaxMany ? nstates : kMaxMany);
  SPX_TRY(autotune_before_evolve(h, bound, width));
  set_launch_width(h, width);
  SPX_TRY(ham_prepare(h, -2.0 / bound));
  if (h->wht_ok) SPX_TRY(hadamard_reserve(h, static_cast<int>(nstates < kMaxMany ? nstates : kMaxMany)));
  // Without autonormalisation every step is the same sequence of launches
  // (same expansion, same buffers): record it once as a HIP graph and replay
  // it, so the host issues one launch per step instead of one per order.
  if (!autonormalize && p->graphs && nsteps >= kGraphMinSteps) {
    hipStream_t st = p->ctx->stream;
    SPX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = SPX_OK;
    for (size_t first = 0; first < nstates && rc == SPX_OK; first += kMaxMany) {
      const size_t cnt = nstates - first < static_cast<size_t>(kMaxMany) ? nstates - first : kMaxMany;
      rc = one_step(p, h, mode, states + first, cnt, coeffs);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (rc != SPX_OK || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc != SPX_OK ? rc : fail(SPX_EDEVICE, std::string("capturing a step: ") + hipGetErrorString(ec));
    }
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (size_t step = 0; step < nsteps && e == hipSuccess; ++step) e = hipGraphLaunch(exec, st);
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(SPX_EDEVICE, std::string("replaying a step: ") + hipGetErrorString(e));
    return SPX_OK;
  }
  for (size_t step = 0; step < nsteps; ++step) {
    for (size_t first = 0; first < nstates; first += kMaxMany) {
      const size_t cnt = nstates - first < static_cast<size_t>(kMaxMany) ? nstates - first : kMaxMany;
      SPX_TRY(one_step(p, h, mode, states + first, cnt, coeffs));
    }
    if (autonormalize)
      for (size_t i = 0; i < nstates; ++i) SPX_TRY(spx_state_normalize(states[i], nullptr));
  }
  return SPX_OK;
}

// Accumulator batching (OrderWeights::defer) for general order n of an
// expansion of order N on the Hadamard path: the terms of orders 3, 4, 5 are
// added at order 5, those of 6, 7, 8 at 8, ..., and whatever waits at N.
int defer_code(size_t n, size_t N) {
  const size_t k = (n - 3) % 3;
  if (n == N) return k == 0 ? 0 : static_cast<int>(k) + 1;
  return k == 2 ? 3 : 1;
}

// One step of a set of states, state i under hs[i] with its own expansion
// *a[i] (order a[i]->size() - 1 >= 1), sorted longest expansion first so that
// the states still expanding at order n are a prefix of the launch. Every
// state sees exactly the arithmetic of one_step on its own.
// chain_in: the Hamiltonians' Hadamard scratch already holds the row
// transforms of every state (left by the previous step's chain_out), so the
// first order skips its forward row pass. chain_out: each state's last order
// leaves the row transforms of its finished state there. Both need the
// Hadamard path and every expansion of order 2 or more; the arithmetic is
// that of the unchained step. State i uses the propagator's scratch pair
// slot0 + i.
int one_step_set(spx_prop* p, spx_ham* const* hs, int mode, spx_state* const* states, size_t nvec,
                 const std::vector<double>* const* a, bool chain_in, bool chain_out, size_t slot0) {
  double2* A[kMaxMany];
  double2* B[kMaxMany];
  for (size_t i = 0; i < nvec; ++i) {
    A[i] = p->d_scratch[2 * (slot0 + i)];
    B[i] = p->d_scratch[2 * (slot0 + i) + 1];
  }
  double wre = mode == 0 ? 0.0 : -1.0, wim = mode == 0 ? -1.0 : 0.0;
  auto advance = [&] {
    if (mode == 0) {
      const double pr = wre;
      wre = wim;
      wim = -pr;
    } else {
      wre = -wre;
      wim = -wim;
    }
  };
  auto order_of = [&](size_t i) { return a[i]->size() - 1; };
  OrderLaunch v{};
  OrderWeights w[kMaxMany] = {};
  v.nvec = static_cast<int>(nvec);
  for (size_t i = 0; i < nvec; ++i) {
    v.in[i] = states[i]->d;
    v.out[i] = B[i];
    v.psi[i] = states[i]->d;
    v.uv_next[i] = order_of(i) >= 2;   // each order hands its output's row transforms on
  }
  v.uv_ready = chain_in ? 1 : 0;
  SPX_TRY(launch_order_set(hs, 1, v, w));
  double c1r[kMaxMany], c1i[kMaxMany];
  for (size_t i = 0; i < nvec; ++i) {
    c1r[i] = 2.0 * (*a[i])[1] * wre;
    c1i[i] = 2.0 * (*a[i])[1] * wim;
  }
  advance();
  size_t act = nvec;
  while (act > 0 && order_of(act - 1) == 1) {
    --act;
    SPX_TRY(launch_axpby(p->ctx, states[act]->d, (*a[act])[0], B[act], c1r[act], c1i[act], states[act]->n));
  }
  if (act == 0) return SPX_OK;
  v.nvec = static_cast<int>(act);
  v.uv_ready = 1;
  for (size_t i = 0; i < act; ++i) {
    v.in[i] = B[i];
    v.out[i] = A[i];
    v.uv_next[i] = order_of(i) >= 3 ? 1 : (chain_out ? 2 : 0);
    w[i] = OrderWeights{(*a[i])[0], c1r[i], c1i[i], 2.0 * (*a[i])[2] * wre, 2.0 * (*a[i])[2] * wim};
  }
  SPX_TRY(launch_order_set(hs, 2, v, w));
  // weights of orders n-1 and n-2 for batched accumulator updates
  double p1r[kMaxMany], p1i[kMaxMany], p2r[kMaxMany] = {}, p2i[kMaxMany] = {};
  for (size_t i = 0; i < act; ++i) {
    p1r[i] = w[i].cr;
    p1i[i] = w[i].ci;
  }
  const bool batch_psi = hs[0]->wht_ok;   // every Hamiltonian of the set then is
  advance();
  const size_t top = order_of(0);
  for (size_t n = 3; n <= top; ++n) {
    while (act > 0 && order_of(act - 1) < n) --act;
    v.nvec = static_cast<int>(act);
    for (size_t i = 0; i < act; ++i) {
      const bool odd = (n % 2) == 1;
      v.in[i] = odd ? A[i] : B[i];
      v.out[i] = odd ? B[i] : A[i];
      v.uv_next[i] = order_of(i) > n ? 1 : (chain_out ? 2 : 0);
      w[i] = OrderWeights{0.0, 0.0, 0.0, 2.0 * (*a[i])[n] * wre, 2.0 * (*a[i])[n] * wim};
      w[i].pr = p1r[i];
      w[i].pi = p1i[i];
      w[i].qr = p2r[i];
      w[i].qi = p2i[i];
      w[i].defer = batch_psi ? defer_code(n, order_of(i)) : 0;
      p2r[i] = p1r[i];
      p2i[i] = p1i[i];
      p1r[i] = w[i].cr;
      p1i[i] = w[i].ci;
    }
    SPX_TRY(launch_order_set(hs, 3, v, w));
    advance();
  }
  return SPX_OK;
}

}  // namespace

int spx_prop_evolve(spx_prop* p, spx_ham* h, int mode, spx_state* psi, double tau, size_t nsteps,
                    int autonormalize) {
  spx_state* one[1] = {psi};
  return evolve_impl(p, h, mode, one, 1, nullptr, 0.0, tau, nsteps, autonormalize);
}

int spx_prop_evolve_many(spx_prop* p, spx_ham* h, int mode, spx_state* const* states, size_t nstates,
                         double tau, size_t nsteps, int autonormalize) {
  return evolve_impl(p, h, mode, states, nstates, nullptr, 0.0, tau, nsteps, autonormalize);
}

namespace {

// Redirects the launches a context issues to another of its streams while in
// scope. A context is driven by one host thread at a time, so no other
// caller sees the change.
class OnStream {
 public:
  OnStream(spx_ctx* c, hipStream_t s) : c_(c), saved_(c->stream) { c->stream = s; }
  ~OnStream() { c_->stream = saved_; }
  OnStream(const OnStream&) = delete;
  OnStream& operator=(const OnStream&) = delete;

 private:
  spx_ctx* c_;
  hipStream_t saved_;
};

int ensure_side(spx_ctx* c) {
  if (!c->side) SPX_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  if (!c->fork) SPX_HIP(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  if (!c->join) SPX_HIP(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
  return SPX_OK;
}

// Tuning, launch width and device tables of the Hamiltonians for launches of
// `width` states.
int prepare_set(spx_ham* const* hams, const std::vector<size_t>& idx, const std::vector<double>& bounds,
                int width) {
  for (size_t i : idx) {
    SPX_TRY(autotune_before_evolve(hams[i], bounds[i], width));
    set_launch_width(hams[i], width);
    SPX_TRY(ham_prepare(hams[i], -2.0 / bounds[i]));
  }
  return SPX_OK;
}

// The set evolve over given expansions (coeffs[i], bounds[i] per state).
//
// Two streams: an order over a set is two dependent launches (column pass,
// then row pass), and each ends in a tail where its last workgroups drain
// while the rest of the chip idles. Sets on the Hadamard path therefore run
// on the context's stream and a second stream, so that one part's tails
// overlap the other part's passes:
//   * 4..kMaxMany states that split into two parts of at least two states
//     with no Hamiltonian in both (distinct realisations, or each
//     realisation's state and reference side by side) run as those two
//     halves (at L = 18, eight states: 10% less time per step);
//   * more than kMaxMany states, with no Hamiltonian in two launch groups,
//     run their groups alternately on the two streams, in groups of half the
//     set up to 2 kMaxMany states (at L = 18, sixteen states: 10%).
// Each group is what the set's own launches would do to its states (bitwise
// the same results); groups on different streams share no buffer (a
// Hamiltonian's transform scratch serves its own states; the propagator's
// scratch pairs are split by position, or by stream). The second stream
// starts after the work already on the context's stream and is joined back
// into it before return.
int evolve_set_impl(spx_prop* p, spx_ham* const* hams, int mode, spx_state* const* states, size_t count,
                    const std::vector<std::vector<double>>& coeffs, const std::vector<double>& bounds,
                    size_t nsteps, int autonormalize) {
  size_t top = 0;
  for (size_t i = 0; i < count; ++i) top = std::max(top, coeffs[i].size() - 1);
  p->last_order = top;
  SPX_TRY(set_device(p->ctx));
  SPX_TRY(ensure_scratch(p, count < kMaxMany ? count : kMaxMany));
  // longest expansion first (ties by position), identities (order 0) dropped
  std::vector<size_t> idx;
  for (size_t i = 0; i < count; ++i)
    if (coeffs[i].size() > 1) {
      if (!(bounds[i] > 0.0)) return fail(SPX_EDOMAIN, "spectral bound must be positive");
      idx.push_back(i);
    }
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t x, size_t y) { return coeffs[x].size() > coeffs[y].size(); });
  const size_t n = idx.size();
  // split point for two streams: the most even one that leaves no
  // Hamiltonian on both sides (0: none)
  auto separates = [&](size_t h) {
    for (size_t x = 0; x < h; ++x)
      for (size_t y = h; y < n; ++y)
        if (hams[idx[x]] == hams[idx[y]]) return false;
    return true;
  };
  const size_t kM = static_cast<size_t>(kMaxMany);
  const bool two = p->streams == 2 && !autonormalize;
  size_t half = 0;
  if (two && n >= 4 && n <= kM)
    for (size_t h = 2; h + 2 <= n; ++h)
      if (separates(h) && (half == 0 || std::max(h, n - h) < std::max(half, n - half))) half = h;
  const bool alternate = two && n > kM;   // groups alternate between the streams
  // The path is chosen at the one-stream width first, so that a schedule
  // the tuner picks there (it sticks to the Hamiltonian) is the one-stream
  // choice; the narrower launches of a split follow only if every state
  // stays on the Hadamard path at their width.
  int width = static_cast<int>(n < kM ? n : kM);
  SPX_TRY(prepare_set(hams, idx, bounds, width));
  bool split = half != 0 || alternate;
  for (size_t i : idx) split = split && hams[i]->wht_ok;
  // launch width of the split: the larger half, or half the set (at most
  // kMaxMany) when groups alternate
  const int narrow = static_cast<int>(half ? std::max(half, n - half) : n <= 2 * kM ? (n + 1) / 2 : kM);
  // without a tuned schedule, a width at which the flip-flop default turns
  // to the gathers would only rebuild every Hamiltonian's tables twice per
  // call and fall back: such sets keep one stream
  for (size_t i : idx)
    split = split && (hams[i]->sched_set ||
                      default_dense_path(hams[i]->nspins, narrow, ham_flipflop_only(hams[i])) == 2);
  if (split) {
    if (narrow != width) {
      SPX_TRY(prepare_set(hams, idx, bounds, narrow));
      bool stays = true;
      for (size_t i : idx) stays = stays && hams[i]->wht_ok;
      if (stays) {
        width = narrow;
      } else {
        split = false;
        SPX_TRY(prepare_set(hams, idx, bounds, width));
      }
    }
  }
  for (size_t i : idx)
    if (hams[i]->wht_ok) {
      int copies = 0;   // one scratch slot per appearance, at most one launch's worth
      for (size_t j : idx) copies += hams[j] == hams[i] ? 1 : 0;
      SPX_TRY(hadamard_reserve(hams[i], copies < kMaxMany ? copies : kMaxMany));
    }
  // groups of up to `width` consecutive states whose Hamiltonians can share
  // launches, within each half when split; the third member says which
  // stream a group's launches go to
  struct Group {
    size_t g, e;
    bool side;
  };
  std::vector<Group> groups;
  auto add_groups = [&](size_t beg, size_t end, bool side) {
    for (size_t g = beg; g < end;) {
      size_t e = g + 1;
      while (e < end && e - g < static_cast<size_t>(width)) {
        spx_ham* hs[kMaxMany];
        for (size_t k = g; k <= e; ++k) hs[k - g] = hams[idx[k]];
        if (!order_set_compatible(hs, e - g + 1)) break;
        ++e;
      }
      groups.push_back({g, e, side});
      g = e;
    }
  };
  if (split && half) {
    add_groups(0, half, false);
    add_groups(half, n, true);
  } else {
    add_groups(0, n, false);
  }
  // On the Hadamard path, with no normalisation between steps and every
  // expansion of order 2 or more, each step's last orders leave the row
  // transforms of the finished states for the next step's first order, which
  // then skips its forward row pass. This needs each Hamiltonian's scratch
  // slots to serve one launch group: no Hamiltonian in two groups.
  bool shared = false;   // a Hamiltonian in two launch groups
  for (size_t a = 0; a < groups.size() && !shared; ++a)
    for (size_t b = 0; b < a && !shared; ++b)
      for (size_t x = groups[a].g; x < groups[a].e && !shared; ++x)
        for (size_t y = groups[b].g; y < groups[b].e && !shared; ++y) shared = hams[idx[x]] == hams[idx[y]];
  if (split && alternate) {
    split = groups.size() >= 2 && !shared;
    for (size_t gi = 0; gi < groups.size(); ++gi) groups[gi].side = split && gi % 2 == 1;
    if (split) SPX_TRY(ensure_scratch(p, 2 * kM));
  }
  bool chain = !autonormalize && nsteps > 1 && !shared;
  for (size_t i : idx) chain = chain && hams[i]->wht_ok && coeffs[i].size() > 2;
  if (split) {
    SPX_TRY(ensure_side(p->ctx));
    SPX_HIP(hipEventRecord(p->ctx->fork, p->ctx->stream));
    SPX_HIP(hipStreamWaitEvent(p->ctx->side, p->ctx->fork, 0));
  }
  int rc = SPX_OK;
  for (size_t step = 0; step < nsteps && rc == SPX_OK; ++step) {
    for (const Group& grp : groups) {
      if (rc != SPX_OK) break;
      const size_t g = grp.g, e = grp.e;
      spx_ham* hs[kMaxMany];
      spx_state* ss[kMaxMany];
      const std::vector<double>* cs[kMaxMany];
      for (size_t k = g; k < e; ++k) {
        hs[k - g] = hams[idx[k]];
        ss[k - g] = states[idx[k]];
        cs[k - g] = &coeffs[idx[k]];
      }
      OnStream on(p->ctx, grp.side ? p->ctx->side : p->ctx->stream);
      // scratch pairs: by position within a split set, by stream when
      // groups alternate
      const size_t slot0 = !split ? 0 : alternate ? (grp.side ? kM : 0) : g;
      rc = one_step_set(p, hs, mode, ss, e - g, cs, chain && step > 0, chain && step + 1 < nsteps, slot0);
    }
    if (autonormalize)
      for (size_t i = 0; i < count && rc == SPX_OK; ++i) rc = spx_state_normalize(states[i], nullptr);
  }
  if (split) {
    // joined after a failed launch too, so that later work on the context's
    // stream stays ordered after everything issued here
    hipError_t e = hipEventRecord(p->ctx->join, p->ctx->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(p->ctx->stream, p->ctx->join, 0);
    if (rc == SPX_OK && e != hipSuccess)
      return fail(SPX_EDEVICE, std::string("joining the second stream: ") + hipGetErrorString(e));
  }
  return rc;
}

int check_set(spx_prop* p, spx_ham* const* hams, int mode, spx_state* const* states, size_t count) {
  if (!p || !hams || !states) return fail(SPX_EINVAL, "null argument");
  for (size_t i = 0; i < count; ++i) {
    if (!hams[i]) return fail(SPX_EINVAL, "null Hamiltonian");
    SPX_TRY(check_state(states[i]));
    if (states[i]->nspins != hams[i]->nspins) return fail(SPX_EINVAL, "state and Hamiltonian disagree on spin count");
    if (states[i]->nspins != p->nspins) return fail(SPX_EINVAL, "state does not match the propagator's size");
    if (hams[i]->ctx != p->ctx) return fail(SPX_EINVAL, "Hamiltonian belongs to another context");
    // the launches and the join of a split set are ordered on the
    // propagator's context; a state read through another context's stream
    // would not wait for them
    if (states[i]->ctx != p->ctx) return fail(SPX_EINVAL, "state belongs to another context");
    for (size_t j = 0; j < i; ++j)
      if (states[j] == states[i]) return fail(SPX_EINVAL, "a state appears twice");
  }
  if (mode != 0 && mode != 1) return fail(SPX_EINVAL, "unknown time mode");
  return SPX_OK;
}

}  // namespace

// LocalRunner::run (src/ensemble/EnsembleRunner.cpp:27-88) evolves each
// realisation under its own Hamiltonian on a pool of threads; here up to
// kMaxMany of them share every order launch (the Hadamard and all-pairs
// paths take per-state Hamiltonians), each with its own spectral bound and
// expansion order. Results are bitwise those of spx_prop_evolve per state.
int spx_prop_evolve_set(spx_prop* p, spx_ham* const* hams, int mode, spx_state* const* states, size_t count,
                        double tau, size_t nsteps, int autonormalize) {
  SPX_TRY(check_set(p, hams, mode, states, count));
  if (count == 0 || nsteps == 0) return SPX_OK;
  std::vector<std::vector<double>> coeffs(count);
  std::vector<double> bounds(count);
  for (size_t i = 0; i < count; ++i) {
    double sb = 0.0;
    spx_ham_spectral_bound(hams[i], &sb);
    SPX_TRY(cheb_build(mode, sb, tau, p->epsilon, 6000, coeffs[i], bounds[i]));
  }
  return evolve_set_impl(p, hams, mode, states, count, coeffs, bounds, nsteps, autonormalize);
}

// The same over prebuilt expansions (spx_prop_steps for a set): state i takes
// coeffs[i][0..orders[i]] at bound bounds[i].
int spx_prop_steps_set(spx_prop* p, spx_ham* const* hams, int mode, spx_state* const* states, size_t count,
                       const double* const* coeffs, const size_t* orders, const double* bounds, size_t nsteps,
                       int autonormalize) {
  SPX_TRY(check_set(p, hams, mode, states, count));
  if (count == 0 || nsteps == 0) return SPX_OK;
  if (!coeffs || !orders || !bounds) return fail(SPX_EINVAL, "null expansion");
  std::vector<std::vector<double>> c(count);
  std::vector<double> b(bounds, bounds + count);
  for (size_t i = 0; i < count; ++i) {
    if (!coeffs[i]) return fail(SPX_EINVAL, "null expansion");
    c[i].assign(coeffs[i], coeffs[i] + orders[i] + 1);
  }
  return evolve_set_impl(p, hams, mode, states, count, c, b, nsteps, autonormalize);
}

int spx_prop_steps(spx_prop* p, spx_ham* h, int mode, spx_state* psi, const double* coeffs,
                   size_t order, double bound, size_t nsteps, int autonormalize) {
  if (!coeffs) return fail(SPX_EINVAL, "null coefficients");
  std::vector<double> c(coeffs, coeffs + order + 1);
  spx_state* one[1] = {psi};
  return evolve_impl(p, h, mode, one, 1, &c, bound, 0.0, nsteps, autonormalize);
}

// -------------------------------------------------------------------- Gates

int spx_rotate(spx_state* s, double ax, double ay, double az, double angle, const uint32_t* targets,
               size_t ntargets) {
  SPX_TRY(check_state(s));
  // Axis3::normalized and Rotation2x2::make (src/core/Gate.cpp:20-51)
  const double length = std::sqrt(ax * ax + ay * ay + az * az);
  if (!(length > 0.0)) return fail(SPX_EINVAL, "rotation axis must have non-zero length");
  if (ntargets == 0 || !targets) return fail(SPX_EINVAL, "a rotation must name at least one spin");
  const double nx = ax / length, ny = ay / length, nz = az / length;
  const double c = std::cos(0.5 * angle), sn = std::sin(0.5 * angle);
  const double u[8] = {c, sn * nz, sn * ny, -sn * nx, -sn * ny, -sn * nx, c, -sn * nz};
  for (size_t t = 0; t < ntargets; ++t)
    if (targets[t] >= s->nspins) return fail(SPX_ERANGE, "rotation names a spin outside the system");
  return launch_rotate_targets(s->ctx, s->d, s->nspins, targets, ntargets, u);
}

int spx_two_spin_gate(spx_state* s, const double* u16, size_t a, size_t b) {
  SPX_TRY(check_state(s));
  if (a == b) return fail(SPX_EINVAL, "a spin cannot couple to itself");
  const size_t lo = a < b ? a : b, hi = a < b ? b : a;
  if (hi >= s->nspins) return fail(SPX_ERANGE, "gate names a spin outside the system");
  return launch_two_spin(s->ctx, s->d, s->n, static_cast<uint32_t>(lo), static_cast<uint32_t>(hi), u16);
}

// -------------------------------------------------------------- Observables

int spx_apply_sigma(int axis, double c, const spx_state* in, size_t spin, spx_state* out,
                    int accumulate) {
  SPX_TRY(same_space(in, out));
  if (in == out) return fail(SPX_EINVAL, "kernels require distinct input and output");
  if (axis < 0 || axis > 2) return fail(SPX_EINVAL, "axis must be one of x, y, z");
  if (spin >= in->nspins) return fail(SPX_ERANGE, "spin index out of range");
  return launch_sigma(in->ctx, axis, c, in->d, spin, out->d, in->n, accumulate != 0);
}

int spx_apply_total_spin(int axis, const spx_state* in, spx_state* out) {
  SPX_TRY(same_space(in, out));
  if (in == out) return fail(SPX_EINVAL, "kernels require distinct input and output");
  if (axis < 0 || axis > 2) return fail(SPX_EINVAL, "axis must be one of x, y, z");
  return launch_total_spin(in->ctx, axis, in->d, out->d, in->nspins);
}

int spx_measure_corr(int axis, const spx_state* state, const spx_state* ref, double* re, double* im) {
  SPX_TRY(same_space(state, ref));
  if (axis < 0 || axis > 2) return fail(SPX_EINVAL, "axis must be one of x, y, z");
  return reduce_corr(state->ctx, axis, state->d, ref->d, state->nspins, re, im);
}

int spx_measure_mag(int axis, size_t spin, const spx_state* state, double* value) {
  SPX_TRY(check_state(state));
  if (axis < 0 || axis > 2) return fail(SPX_EINVAL, "axis must be one of x, y, z");
  if (spin >= state->nspins) return fail(SPX_ERANGE, "observable names a spin outside the system");
  return reduce_mag(state->ctx, axis, static_cast<uint32_t>(spin), state->d, state->nspins, value);
}

int spx_measure_energy(spx_ham* h, const spx_state* state, double* value) {
  SPX_TRY(check_state(state));
  if (state->nspins != h->nspins) return fail(SPX_EINVAL, "state and Hamiltonian disagree on spin count");
  return reduce_energy(h, state->d, value);
}

// ------------------------------------------------------------------ Program

int spx_program_create(spx_program** out) {
  *out = new spx_program();
  return SPX_OK;
}

int spx_program_destroy(spx_program* p) {
  delete p;
  return SPX_OK;
}

int spx_program_evolve(spx_program* p, double tau, size_t nsteps, int mode, int autonormalize,
                       const double* fields3L, size_t nspins) {
  if (mode != 0 && mode != 1) return fail(SPX_EINVAL, "unknown time mode");
  ProgOp op;
  op.kind = OP_EVOLVE;
  op.tau = tau;
  op.nsteps = nsteps;
  op.mode = mode;
  op.autonorm = autonormalize != 0;
  if (fields3L) op.fields.assign(fields3L, fields3L + 3 * nspins);
  p->ops.push_back(op);
  return SPX_OK;
}

int spx_program_evolve_uniform(spx_program* p, double tau, size_t nsteps, int mode, int autonormalize,
                               double hx, double hy, double hz) {
  if (mode != 0 && mode != 1) return fail(SPX_EINVAL, "unknown time mode");
  ProgOp op;
  op.kind = OP_EVOLVE;
  op.tau = tau;
  op.nsteps = nsteps;
  op.mode = mode;
  op.autonorm = autonormalize != 0;
  op.uniform = true;
  op.ufield[0] = hx;
  op.ufield[1] = hy;
  op.ufield[2] = hz;
  p->ops.push_back(op);
  return SPX_OK;
}

int spx_program_pulse(spx_program* p, double ax, double ay, double az, double angle,
                      const uint32_t* targets, size_t ntargets) {
  const double length = std::sqrt(ax * ax + ay * ay + az * az);
  if (!(length > 0.0)) return fail(SPX_EINVAL, "rotation axis must have non-zero length");
  ProgOp op;
  op.kind = OP_PULSE;
  op.axis[0] = ax;
  op.axis[1] = ay;
  op.axis[2] = az;
  op.angle = angle;
  if (targets && ntargets) op.targets.assign(targets, targets + ntargets);
  p->ops.push_back(op);
  return SPX_OK;
}

int spx_program_measure(spx_program* p) {
  ProgOp op;
  op.kind = OP_MEASURE;
  p->ops.push_back(op);
  return SPX_OK;
}

int spx_program_repeat(spx_program* p, size_t count) {
  if (p->open.size() >= 32) return fail(SPX_EINVAL, "repeat nesting deeper than 32");
  ProgOp op;
  op.kind = OP_REPEAT;
  op.count = count;
  p->open.push_back(p->ops.size());
  p->ops.push_back(op);
  return SPX_OK;
}

int spx_program_end(spx_program* p) {
  if (p->open.empty()) return fail(SPX_EINVAL, "end without an open repeat");
  const size_t b = p->open.back();
  p->open.pop_back();
  ProgOp op;
  op.kind = OP_END;
  op.match = b;
  p->ops[b].match = p->ops.size();
  p->ops.push_back(op);
  return SPX_OK;
}

int spx_program_measurements(const spx_program* p, size_t* count) {
  if (!p->open.empty()) return fail(SPX_EINVAL, "program has an unclosed repeat");
  *count = program_measurements(*p, 0, p->ops.size());
  return SPX_OK;
}

}  // extern "C"
