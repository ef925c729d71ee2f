// Internal definitions shared by the libspx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "spx.h"

namespace spx {

// Coefficients at or below this magnitude are treated as absent, as in the
// reference's accumulate (src/core/Hamiltonian.cpp:12-16).
constexpr double kNegligible = 1.0e-14;
constexpr size_t kMaxSpins = 30;          // include/spinsim/core/Bits.hpp:13
constexpr int kMaxEvents = 16;
constexpr int kMaxMany = 16;              // states per batched evolve launch
constexpr int kReduceBlocks = 1024;       // fixed reduction fan-in (determinism)
constexpr int kReduceThreads = 256;

// ---------------------------------------------------------------- errors
int fail(int code, const std::string& msg);
int ok();
void set_convergence(double bound, double tau, double eps, size_t order);

#define SPX_HIP(expr)                                                      \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess)                                                  \
      return ::spx::fail(SPX_EDEVICE, std::string(#expr ": ") +            \
                                          hipGetErrorString(e_));          \
  } while (0)

#define SPX_TRY(expr)              \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != SPX_OK) return rc_; \
  } while (0)

// --------------------------------------------------------- device tables
// One flip term of the Pauli-basis Hamiltonian, pre-multiplied by the apply
// scale: for a coupled pair (mask = bit lo | bit hi) `anti` is the
// flip-flop weight (|ud> <-> |du>, scale*(jx+jy)/4) and `aligned` the
// flip-both weight (|dd> <-> |uu>, scale*(jx-jy)/4); for a transverse field
// (mask = one bit) `anti` = scale*hx/2 and `aligned` = scale*hy/2 and the
// term is (cx -/+ i cy) on the partner depending on the spin's bit.
struct FlipTerm {
  uint32_t mask;
  uint32_t kind;  // 0 = pair, 1 = field
  double anti;
  double aligned;
  double pad;
};

// Launch schedule of the order kernels for one Hamiltonian; zero or -1
// fields keep the library default (environment overridable).
struct Schedule {
  int tb = 0;        // tile bits of the table kernels
  int block = 0;     // threads per workgroup of the per-amplitude-group kernel
  int xcd = -1;      // XCD-aware tile order
  int dense = -1;    // all-pairs kernel allowed
  int dense_tb = 0;  // its tile bits (9 or 10); dense = 2 selects the Hadamard path
  int table = 0;     // 0: best eligible, 2: force the per-amplitude-group kernel
};

// ----------------------------------------------------------------- objects
}  // namespace spx

struct spx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double* d_partials = nullptr;   // [2 * kReduceBlocks]
  double* d_result = nullptr;     // [16]
  double* h_result = nullptr;     // pinned [16]
  hipEvent_t events[spx::kMaxEvents] = {};
  int arch_cus = 0;
  // second stream of a set evolve split in two (created on first use), and
  // the events that order it after, and join it back into, `stream`
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // pinned bounce buffer of copy_to_device / copy_to_host (grown on use)
  unsigned char* bounce = nullptr;
  size_t bounce_cap = 0;
};

struct spx_state {
  spx_ctx* ctx = nullptr;
  size_t nspins = 0;
  size_t n = 0;           // 2^L amplitudes
  double2* d = nullptr;   // device, interleaved complex
};

struct spx_ham {
  spx_ctx* ctx = nullptr;
  size_t nspins = 0;
  std::vector<double> hx, hy, hz;
  std::vector<uint32_t> lo, hi;
  std::vector<double> jx, jy, jz;
  // device sigma_z diagonal (Pauli units) and the (hz, jz) it was built from
  double* d_diag = nullptr;
  bool diag_valid = false;
  std::vector<double> diag_hz, diag_jz;
  std::vector<uint32_t> diag_lo, diag_hi;
  // device flip table for the most recent scale
  spx::FlipTerm* d_terms = nullptr;
  size_t terms_cap = 0;
  int nterms = 0;
  int terms_nlocal = 0, terms_nboth = 0, terms_g = 0, terms_tb = 0;
  std::vector<int> terms_grp;   // v2 group offsets
  std::vector<int> terms_off;   // v3 segment ends
  int terms_variant = 0;
  int terms_flipboth = 0;
  bool dense_ok = false;         // specialised all-pairs kernel usable
  double* d_coef = nullptr;      // its weights
  bool wht_ok = false;           // Hadamard-sandwich flip-flop path selected
  double* d_wht_e = nullptr;     // its Hadamard-basis diagonals A then B, factored (spx_hadamard.hip);
                                 // points into the table cache below
  size_t wht_tab = 0;            // doubles per factor table (B starts there)
  // Factor tables of the field/coupling/scale sets this Hamiltonian has
  // seen (a pulse program alternates between a few), so that returning to
  // one costs no host work and no upload. Owned here; freed on destroy.
  struct WhtTables {
    std::vector<double> key;     // the weights the tables were built from
    double* host = nullptr;      // pinned copy of the tables (the upload's source)
    double* dev = nullptr;
    size_t count = 0;            // doubles in the tables
    hipEvent_t copied = nullptr; // after the upload from `host`
    uint64_t used = 0;           // last-use stamp
  };
  std::vector<WhtTables> wht_cache;
  uint64_t wht_clock = 0;
  double2* d_wht = nullptr;      // its transform scratch: wht_slots pairs of 2^L vectors
  int wht_slots = 0;
  double terms_scale = 0.0;
  bool terms_valid = false;
  std::vector<spx::FlipTerm> h_terms;
  double* d_tab = nullptr;   // diagonal build table
  double2* d_scratch = nullptr;  // H psi for the energy observable
  size_t tab_cap = 0;
  int width = 1;                 // states sharing each order launch of the latest evolve
  spx::Schedule sched;           // launch schedule (autotuned or default)
  bool sched_set = false;        // chosen by the autotuner (or its cache)
};

struct spx_prop {
  spx_ctx* ctx = nullptr;
  size_t nspins = 0;
  double epsilon = 1e-7;
  size_t last_order = 0;
  // T_{n-1}, T_{n-2} pairs: slots [0, kMaxMany) for the launches on the
  // context's stream, [kMaxMany, 2 kMaxMany) for a second stream's groups
  double2* d_scratch[4 * spx::kMaxMany] = {};
  int graphs = 0;   // record one step as a HIP graph and replay it
  int streams = 2;  // 2: a set evolve may run as two halves on two streams
};

namespace spx {

// ---------------------------------------------------------- host helpers
int bessel(int kind, double x, size_t count, double* out, size_t* nonzero);
int cheb_build(int mode, double bound, double tau, double eps, size_t max_order,
               std::vector<double>& coeffs, double& eff_bound);

// ------------------------------------------------------ device launchers
struct OrderLaunch {
  const double2* in[kMaxMany];   // operand of H
  double2* out[kMaxMany];        // result / T_n (in-place over T_{n-2})
  double2* psi[kMaxMany];        // accumulator
  int nvec;
  // Hadamard path only (ignored elsewhere): the previous launch of these
  // vectors already left the row transforms of `in` in the Hamiltonians'
  // scratch (uv_ready), and vector i's launch leaves those of `out` for the
  // next one (uv_next[i] = 1); consecutive Chebyshev orders chain this way.
  // uv_next[i] = 2 on a vector's last order leaves those of its finished
  // accumulator instead, the input of the next step's first order.
  int uv_ready;
  int uv_next[kMaxMany];
};

int ham_prepare(spx_ham* h, double scale);   // diagonal + flip table at scale

// The off-diagonal part of H through Walsh-Hadamard transforms
// (spx_hadamard.hip): any couplings and fields, three passes per application
// (two per Chebyshev order after the first), L in range.
constexpr size_t kHadamardMinSpins = 13, kHadamardMaxSpins = 20;
// Chebyshev weights of one vector of an order launch (see launch_order).
// `defer` (general orders on the Hadamard path only, 0 elsewhere) batches
// the accumulator updates: 0 psi += c T_n; 1 leave psi (the term waits);
// 2 psi += p T_{n-1} + c T_n; 3 psi += q T_{n-2} + p T_{n-1} + c T_n, in
// that order, so psi is read and written once per three orders.
struct OrderWeights {
  double a0, c1r, c1i, cr, ci;
  double pr = 0.0, pi = 0.0, qr = 0.0, qi = 0.0;
  int defer = 0;
};
// Vector i of a launch is propagated under ham[i] (prepared at its own
// scale sdiag[i]) and uses scratch pair slot[i] of that Hamiltonian.
struct HadamardLaunch {
  const double2* in[kMaxMany];
  double2* out[kMaxMany];
  double2* psi[kMaxMany];
  const spx_ham* ham[kMaxMany];
  int slot[kMaxMany];
  double sdiag[kMaxMany];
  OrderWeights w[kMaxMany];
  int ready;             // see OrderLaunch::uv_ready
  int next[kMaxMany];    // see OrderLaunch::uv_next
};
bool hadamard_eligible(size_t L);
// Weights at the apply scale after the skip rule: pair q (order (0,1),(0,2),...)
// anti / aligned, spin s transverse cx / cy (both zero for a skipped field);
// the diagonal's terms unscaled, as rebuildDiagonal takes them: dfield[s] =
// hz_s / 2 and dpair[q] = jz_q / 4, zero where skipped.
int hadamard_prepare(spx_ham* h, const std::vector<double>& anti, const std::vector<double>& aligned,
                     const std::vector<double>& cx, const std::vector<double>& cy,
                     const std::vector<double>& dfield, const std::vector<double>& dpair);
int launch_hadamard(int kind, const HadamardLaunch& b, unsigned nvec);
// Scratch for `slots` vectors of one launch under h (call before any launch
// or graph capture that needs them; launches check and fail otherwise).
int hadamard_reserve(spx_ham* h, int slots);
int launch_apply(spx_ham* h, double scale, const double2* in, double2* out, bool accumulate);
// kind: 1 = first order (out = 0.5 acc), 2 = second order (fused psi
// a0 T0 + c1 T1 + c2 T2), 3 = general order (T = acc - out; psi += c T)
int launch_order(spx_ham* h, int kind, const OrderLaunch& v, double a0, double c1r, double c1i,
                 double cr, double ci);
// The same over a set of states, vector i under hs[i] (prepared at its own
// scale) with weights w[i]. The Hadamard and all-pairs paths take distinct
// Hamiltonians in one launch; order_set_compatible says whether these
// prepared Hamiltonians can share launches (else they are stepped one by one).
bool order_set_compatible(spx_ham* const* hs, size_t n);
int launch_order_set(spx_ham* const* hs, int kind, const OrderLaunch& v, const OrderWeights* w);
int launch_axpby(spx_ctx* c, double2* y, double a, const double2* x, double br, double bi, size_t n);
int launch_fill(spx_ctx* c, double2* y, size_t n);
int launch_scale(spx_ctx* c, double2* y, double f, size_t n);
int launch_negate(spx_ctx* c, double2* y, size_t n);
int launch_add_scaled(spx_ctx* c, double2* y, double cr, double ci, const double2* x, size_t n);
int launch_sigma(spx_ctx* c, int axis, double coef, const double2* in, size_t spin, double2* out,
                 size_t n, bool accumulate);
int launch_total_spin(spx_ctx* c, int axis, const double2* in, double2* out, size_t L);
int launch_rotate(spx_ctx* c, double2* s, size_t n, uint32_t spin, const double* u8);
int launch_two_spin(spx_ctx* c, double2* s, size_t n, uint32_t lo, uint32_t hi, const double* u32);

// deterministic reductions; results land in host doubles
int reduce_norm2(spx_ctx* c, const double2* x, size_t n, double* out);
int reduce_inner(spx_ctx* c, const double2* self, const double2* other, size_t n, double* re,
                 double* im);
int reduce_corr(spx_ctx* c, int axis, const double2* state, const double2* ref, size_t L,
                double* re, double* im);
int reduce_mag(spx_ctx* c, int axis, uint32_t spin, const double2* state, size_t L, double* out);
int reduce_energy(spx_ham* h, const double2* state, double* out);
int reduce_norm2_dev(spx_ctx* c, const double2* x, size_t n, double* dst);
int reduce_corr_dev(spx_ctx* c, int axis, const double2* state, const double2* ref, size_t L, double* dst);
int reduce_mag_dev(spx_ctx* c, int axis, uint32_t spin, const double2* state, size_t L, double* dst);
int reduce_energy_dev(spx_ham* h, const double2* state, double* dst);
int launch_rotate_targets(spx_ctx* c, double2* s, size_t L, const uint32_t* targets, size_t nt,
                          const double* u8);

// tunables: the Hamiltonian's schedule, else the environment (for sweeps),
// else the defaults
int tile_bits_for(const spx_ham* h);
int block_threads(const spx_ham* h);
int xcd_remap(const spx_ham* h);
int dense_tile_bits(const spx_ham* h);
// The path used when neither the schedule nor SPX_DENSE names one (0 table
// kernels, 1 all-pairs gathers, 2 Hadamard transforms), by L, launch width
// and whether the couplings are flip-flop only; set_launch_width records an
// evolve's width and invalidates the prepared tables when that changes the
// choice.
int default_dense_path(size_t L, int width, bool flipflop_only);
bool ham_flipflop_only(const spx_ham* h);
int ensure_diag(spx_ham* h);
// Host <-> device copies of pageable host memory through the context's
// pinned bounce buffer (at most kBounceBytes, allocated on first use), in
// pieces, each ordered on the context's stream; both return when the copy
// is complete. An asynchronous copy from or to pageable memory makes the
// runtime copy on the CPU into its own staging memory; these never ask it
// to. The context's stream is synchronised, so the caller's earlier work on
// it has finished too.
constexpr size_t kBounceBytes = size_t(4) << 20;
int copy_to_device(spx_ctx* c, void* dst, const void* src, size_t bytes);
int copy_to_host(spx_ctx* c, void* dst, const void* src, size_t bytes);
void set_launch_width(spx_ham* h, int width);
// the autotuner's process-wide mode and its hook in the propagator
int autotune_before_evolve(spx_ham* h, double bound, int nvec);

}  // namespace spx
