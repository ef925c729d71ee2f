/*
 * spx.h -- C ABI of the MI355X-native spinsim propagator library (libspx.so).
 *
 * The reference (BekturMur/Spin_Systems) has no FFI: its boundary is the C++
 * class API of namespace spinsim. This header is the thin C layer beneath a
 * drop-in re-implementation of that API (spinsim_amd/host, namespace
 * spinsim): each entry point below names the reference member it backs.
 * Plain pointers and sizes only; every function returns an spx_status.
 *
 * Storage: a state of L spins is 2^L complex FP64 amplitudes, interleaved
 * (re, im), amplitude k at [2k], [2k+1]; spin s owns bit s of k, set bit = up
 * (include/spinsim/core/StateVector.hpp:53-60, include/spinsim/core/Bits.hpp:15-27).
 * Device vectors live in HBM and are handed across as opaque handles; host
 * buffers cross only through spx_state_upload / spx_state_download.
 *
 * Threading: one spx_ctx per host thread (each owns a HIP stream and its
 * reduction scratch). Objects created from a context must be used from the
 * thread that owns it. Calls on one context run in stream order.
 */
#ifndef SPX_H
#define SPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The C++ layer rethrows them as the reference's exception
 * types (include/spinsim/core of the reference): */
typedef enum {
  SPX_OK = 0,
  SPX_EINVAL = 1,     /* std::invalid_argument */
  SPX_ERANGE = 2,     /* std::out_of_range */
  SPX_EDOMAIN = 3,    /* std::domain_error (bound <= 0, zero norm) */
  SPX_ECONVERGE = 4,  /* ChebyshevConvergenceError (include/spinsim/core/Chebyshev.hpp:21-36) */
  SPX_EOVERFLOW = 5,  /* std::overflow_error (besselI, x > 709) */
  SPX_ERUNTIME = 6,   /* std::runtime_error */
  SPX_EDEVICE = 7,    /* HIP runtime failure */
  SPX_ENOMEM = 8      /* std::bad_alloc / hipErrorOutOfMemory */
} spx_status;

typedef struct spx_ctx spx_ctx;
typedef struct spx_state spx_state;
typedef struct spx_ham spx_ham;
typedef struct spx_prop spx_prop;
typedef struct spx_program spx_program;
typedef struct spx_batch spx_batch;

/* Message of the last failure on the calling thread ("" if none). */
const char* spx_last_error(void);
/* Context of the last SPX_ECONVERGE on this thread: bound, tau, epsilon,
 * attempted order (ChebyshevConvergenceError accessors, Chebyshev.hpp:27-31). */
int spx_last_convergence(double* bound, double* tau, double* epsilon, size_t* order);

/* ------------------------------------------------------------ context */
int spx_device_count(int* count);
int spx_ctx_create(int device, spx_ctx** out);
int spx_ctx_destroy(spx_ctx* ctx);
int spx_ctx_synchronize(spx_ctx* ctx);
/* The context's hipStream_t, for interop and for timing on that stream. */
int spx_ctx_stream(spx_ctx* ctx, void** stream);
/* Timing on the context's stream: record event `slot` (0..15); elapsed ms
 * between two recorded slots (synchronises on the later one). */
int spx_ctx_event_record(spx_ctx* ctx, int slot);
int spx_ctx_event_elapsed(spx_ctx* ctx, int slot0, int slot1, float* ms);

/* ------------------------------------------------- StateVector (a1)
 * include/spinsim/core/StateVector.hpp:82-140, src/core/StateVector.cpp:10-75 */
int spx_state_create(spx_ctx* ctx, size_t nspins, spx_state** out); /* StateVector(L), zeroed */
int spx_state_destroy(spx_state* s);
int spx_state_nspins(const spx_state* s, size_t* nspins);
int spx_state_upload(spx_state* s, const double* host);          /* 2*2^L doubles */
int spx_state_download(const spx_state* s, double* host);
int spx_state_copy(spx_state* dst, const spx_state* src);         /* operator= */
int spx_state_set_zero(spx_state* s);                             /* setZero */
int spx_state_set_basis(spx_state* s, size_t index);              /* setBasisState */
int spx_state_scale(spx_state* s, double factor);                 /* scale */
int spx_state_negate(spx_state* s);                               /* negate */
int spx_state_add_scaled(spx_state* dst, double cr, double ci, const spx_state* src); /* addScaled */
int spx_state_norm(const spx_state* s, double* norm);             /* norm */
int spx_state_normalize(spx_state* s, double* norm_before);       /* normalize */
int spx_state_inner(const spx_state* self, const spx_state* other, double* re, double* im); /* innerProduct: <other|self> */

/* ------------------------------------------ Bessel / Chebyshev (a6, a7)
 * Host FP64. src/core/Bessel.cpp:45-149, src/core/Chebyshev.cpp:54-107.
 * kind 0 = J, 1 = I. mode 0 = TimeMode::Real, 1 = TimeMode::Imaginary. */
int spx_bessel(int kind, double x, size_t count, double* values, size_t* nonzero);
int spx_cheb_build(int mode, double bound, double tau, double epsilon, size_t max_order,
                   double* coeffs, size_t capacity, size_t* order, double* effective_bound);

/* ---------------------------------------------- Hamiltonian (a2-a5)
 * include/spinsim/core/Hamiltonian.hpp:34-85, src/core/Hamiltonian.cpp:20-147 */
int spx_ham_create(spx_ctx* ctx, size_t nspins, spx_ham** out);   /* Hamiltonian(L) */
int spx_ham_destroy(spx_ham* h);
int spx_ham_set_field(spx_ham* h, size_t spin, double x, double y, double z);  /* setField */
int spx_ham_get_field(const spx_ham* h, size_t spin, double* xyz);             /* field */
int spx_ham_set_coupling(spx_ham* h, size_t a, size_t b, double jx, double jy, double jz); /* setCoupling */
int spx_ham_num_couplings(const spx_ham* h, size_t* n);                        /* couplings().size() */
int spx_ham_get_coupling(const spx_ham* h, size_t i, uint32_t* lo, uint32_t* hi, double* jxyz);
int spx_ham_spectral_bound(const spx_ham* h, double* bound);                   /* spectralBound */
/* kind 0: out = H in (applyPhysical); 1: out = -2H/bound in
 * (applyTwiceNormalized); 2: out += -2H/bound in (accumulateTwiceNormalized). */
int spx_ham_apply(spx_ham* h, int kind, double bound, const spx_state* in, spx_state* out);
/* The cached sigma_z diagonal (rebuildDiagonal), downloaded to host [2^L]. */
int spx_ham_diagonal(spx_ham* h, double* host);
/* Name of the order kernel the most recently prepared scale dispatches to
 * (no reference counterpart: a schedule report for tuning and benchmarks). */
int spx_ham_schedule(const spx_ham* h, char* buf, size_t len);

/* ------------------------------------------------- Autotuner (a16)
 * No reference counterpart. Times every eligible launch schedule of the
 * order kernels (the Hadamard-sandwich path, the all-pairs kernel at tile
 * bits 10/9, table kernels at tile bits 6..10 (6 and 7 up to L = 16), with
 * and without the XCD-aware tile order, the general kernel at two block
 * sizes) on this Hamiltonian at the step bound `bound`, keeps the fastest
 * and caches it per problem shape (L, whether the couplings are flip-flop
 * only, flip-term count) for the process; `report` receives the timings. */
int spx_ham_autotune(spx_ham* h, double bound, char* report, size_t len);
/* The same with `nvec` states (1..16) sharing every timed launch, the width of
 * a batched evolve (spx_prop_evolve_many / spx_prop_evolve_set); winners are
 * cached per width as well. */
int spx_ham_autotune_batch(spx_ham* h, double bound, int nvec, char* report, size_t len);
/* 0: default schedules (the default); 1: every evolve on a Hamiltonian
 * without a schedule adopts the cached one of its shape and launch width,
 * tuning it first if absent. Schedules reorder floating-point additions only. */
int spx_autotune_mode(int mode);
/* Fix a schedule by hand; 0 / -1 fields keep the defaults. `dense`: -1
 * default, 0 table kernels only, 1 the all-pairs gather kernel (L = 13..18,
 * flip-flop couplings only: no transverse field and jx = jy, else the table
 * kernels), 2 the Walsh-Hadamard path (L = 13..20, any couplings and
 * fields). */
int spx_ham_set_schedule(spx_ham* h, int tile_bits, int block, int xcd, int dense, int dense_tile_bits);

/* ------------------------------------------------ Propagator (a8)
 * include/spinsim/core/Propagator.hpp:21-55, src/core/Propagator.cpp:38-100 */
int spx_prop_create(spx_ctx* ctx, size_t nspins, double epsilon, spx_prop** out);
int spx_prop_destroy(spx_prop* p);
/* Propagator::evolve on one state. */
int spx_prop_evolve(spx_prop* p, spx_ham* h, int mode, spx_state* psi, double tau,
                    size_t nsteps, int autonormalize);
/* The same step applied to `nstates` states under one Hamiltonian (the
 * state and its reference in Sequence::runEvolve, src/sequence/Sequence.cpp:35-48),
 * sharing one expansion and one diagonal read per order. */
int spx_prop_evolve_many(spx_prop* p, spx_ham* h, int mode, spx_state* const* states,
                         size_t nstates, double tau, size_t nsteps, int autonormalize);
/* State i evolved under hams[i] (a realisation each, as in LocalRunner::run,
 * src/ensemble/EnsembleRunner.cpp:27-88, which evolves them one per thread):
 * each with its own spectral bound and expansion order, up to 16 states per
 * order launch where their Hamiltonians take the Hadamard or all-pairs path
 * together (else one by one). Every state's result is bitwise that of
 * spx_prop_evolve on it alone; last_order reports the longest expansion. */
int spx_prop_evolve_set(spx_prop* p, spx_ham* const* hams, int mode, spx_state* const* states,
                        size_t count, double tau, size_t nsteps, int autonormalize);
/* spx_prop_evolve_set over prebuilt expansions (spx_prop_steps for a set):
 * state i takes coeffs[i][0..orders[i]] at spectral bound bounds[i]. */
int spx_prop_steps_set(spx_prop* p, spx_ham* const* hams, int mode, spx_state* const* states,
                       size_t count, const double* const* coeffs, const size_t* orders,
                       const double* bounds, size_t nsteps, int autonormalize);
int spx_prop_last_order(const spx_prop* p, size_t* order);                      /* lastOrder */
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops). spx_prop_evolve on
 * one state with no Propagator object and no device allocation, no host
 * synchronisation and no copy: the Chebyshev scratch vectors and the
 * diagonal live in a static device array for at most 16 spins (2.5 MiB per
 * device, held until the process ends), and the Hamiltonian's fields and
 * couplings are read from its host side. L > 16 is rejected with
 * SPX_ERANGE; there is no autonormalisation. Every call on a device shares
 * that device's array: calls that overlap on one device race (calls on one
 * context's stream do not overlap). `order` (may be NULL) receives the
 * expansion order. Results agree with spx_prop_evolve to rounding. */
int spx_prop_evolve_noalloc(spx_ham* h, int mode, spx_state* psi, double tau, size_t nsteps, double epsilon,
                            size_t* order);
/* No reference counterpart: with enable = 1, evolve calls of at least 4
 * steps without autonormalisation record one step as a HIP graph and replay
 * it; with enable = 0 (the default) every order is launched from the host.
 * Both run the same kernels with the same arguments (bitwise the same
 * result); on this part the host launches already keep the device busy. */
int spx_prop_set_graphs(spx_prop* p, int enable);
/* No reference counterpart: with streams = 2 (the default), a set evolve
 * (spx_prop_evolve_set, spx_prop_steps_set) on the Hadamard path, without
 * autonormalisation, uses the context's stream and a second stream of the
 * context, so that launches overlap: 4..16 states that split into two parts
 * of two or more states with no Hamiltonian in both run as those halves;
 * more than 16 states with no Hamiltonian in two launch groups alternate
 * their groups between the streams. The second stream is joined back before
 * return.
 * streams = 1 issues every launch on the context's stream. Results are
 * bitwise the same either way. */
int spx_prop_set_streams(spx_prop* p, int streams);
int spx_prop_epsilon(const spx_prop* p, double* epsilon);                       /* epsilon */
/* Hot loop with a prebuilt expansion: `nsteps` applyOneStep calls of
 * `order` fused Chebyshev orders (coeffs[0..order]) at `bound`. */
int spx_prop_steps(spx_prop* p, spx_ham* h, int mode, spx_state* psi, const double* coeffs,
                   size_t order, double bound, size_t nsteps, int autonormalize);

/* --------------------------------------------------- Gates (a9)
 * include/spinsim/core/Gate.hpp:37-93, src/core/Gate.cpp:20-121 */
/* SpinRotation(axis, angle, targets).apply: exp(-i angle/2 n.sigma) on each target. */
int spx_rotate(spx_state* s, double ax, double ay, double az, double angle,
               const uint32_t* targets, size_t ntargets);
/* TwoSpinGate(matrix, pair).apply: u16 is 16 complex row-major (32 doubles),
 * basis (down-down, low-up, high-up, up-up). */
int spx_two_spin_gate(spx_state* s, const double* u16, size_t a, size_t b);

/* ----------------------------------------- Kernels and observables (a5, a10)
 * include/spinsim/core/SpinOperators.hpp:22-45, src/core/Observable.cpp:13-120.
 * axis 0 = x, 1 = y, 2 = z. */
/* accumulate=1: out += c sigma_axis(spin) in (axpySigma*);
 * accumulate=0: out = sigma_axis(spin) in (applySigma*). */
int spx_apply_sigma(int axis, double c, const spx_state* in, size_t spin, spx_state* out,
                    int accumulate);
/* out = sum_i sigma_axis(i) in: prepareReferenceState (src/ensemble/Ensemble.cpp:47-58). */
int spx_apply_total_spin(int axis, const spx_state* in, spx_state* out);
/* corr:<axis> = <ref| sum_i sigma_axis(i) |state>: TotalSpinCorrelator::measure. */
int spx_measure_corr(int axis, const spx_state* state, const spx_state* ref, double* re, double* im);
/* mag:<spin>:<axis> = Re <state|sigma_axis(spin)|state>: SingleSpinMagnetization. */
int spx_measure_mag(int axis, size_t spin, const spx_state* state, double* value);
/* energy = Re <state|H|state>: Energy::measure. */
int spx_measure_energy(spx_ham* h, const spx_state* state, double* value);

/* ------------------------------------------- Pulse program (a11)
 * The Sequence AST (include/spinsim/sequence/Sequence.hpp:19-104) flattened
 * for the device runner: Repeat is a begin/end bracket. */
int spx_program_create(spx_program** out);
int spx_program_destroy(spx_program* p);
/* Evolve{tau, nsteps, mode, autonormalize, fields}; fields3L NULL = keep base. */
int spx_program_evolve(spx_program* p, double tau, size_t nsteps, int mode, int autonormalize,
                       const double* fields3L, size_t nspins);
/* Evolve step whose field table is (hx, hy, hz) on every spin, whatever a
 * realisation's L: an evolve step's `hx`/`hy`/`hz` keys
 * (docs/configuration.md of the reference, `evolve` table), with the
 * unnamed components passed as their base values. */
int spx_program_evolve_uniform(spx_program* p, double tau, size_t nsteps, int mode,
                               int autonormalize, double hx, double hy, double hz);
/* Pulse{SpinRotation(axis, angle, targets)}; ntargets 0 = every spin. */
int spx_program_pulse(spx_program* p, double ax, double ay, double az, double angle,
                      const uint32_t* targets, size_t ntargets);
int spx_program_measure(spx_program* p);                        /* Measure{} */
int spx_program_repeat(spx_program* p, size_t count);            /* Repeat{count, body...} */
int spx_program_end(spx_program* p);                             /* ...end of body */
int spx_program_measurements(const spx_program* p, size_t* count); /* measurementCount */

/* ------------------------------------ Batched ensemble runner (a13-a15)
 * Many realisations of one experiment in one persistent launch, each
 * realisation's state and reference resident in LDS (L <= 12). Realisation r
 * is fully described by the host: its L, fields, couplings and initial
 * state vector (drawn on the host from Rng::forRealization(seed, index) in
 * runRealization order, src/io/Experiment.cpp:10-55). Observables: codes
 * 0..2 corr:x/y/z (2 values), 3..5 mag:x/y/z of obs_spin (1 value),
 * 6 norm (2 values), 7 energy (1 value). Output rows are written per
 * realisation, [r][point][value], and folded on the host in index order. */
int spx_batch_create(spx_ctx* ctx, const spx_program* program, double epsilon,
                     int reference_axis, const int* obs_codes, const int* obs_spins,
                     size_t nobs, spx_batch** out);
int spx_batch_destroy(spx_batch* b);
/* Adds realisation `slot` (0-based, dense). psi0 is the normalised initial
 * state (2*2^L doubles); fields3L the base fields after disorder; couplings
 * as SoA. */
int spx_batch_add(spx_batch* b, size_t nspins, const double* fields3L, size_t npairs,
                  const uint32_t* lo, const uint32_t* hi, const double* jx, const double* jy,
                  const double* jz, const double* psi0);
/* Adds realisations first .. first+count-1 of the dipolar ensemble as
 * runRealization builds them (src/io/Experiment.cpp:10-55) for a run with
 * zero base fields, dipolar geometry, Gaussian Sz disorder and a random
 * initial state: the lattice, the disorder and the typical state drawn from
 * Rng::forRealization(seed, index) in that order (spx_dipolar_realisation,
 * spx_dipolar_state), on a pool of host threads; slots follow index order. */
int spx_batch_add_dipolar(spx_batch* b, uint64_t seed, size_t first, size_t count, size_t nspins,
                          double scale, double min_separation, int angular, double disorder);
/* Runs every added realisation; times [points], values [n][points][width]. */
int spx_batch_run(spx_batch* b, double* times, double* values);
int spx_batch_width(const spx_batch* b, size_t* width);
int spx_batch_clear(spx_batch* b);

/* ---------------------------------------------- Disorder draws (a16)
 * Host-side draws of realisation (seed, index) in runRealization's order
 * (src/io/Experiment.cpp:10-55): the 2D dipolar lattice of
 * generateDipolarLattice2D (src/ensemble/Ensemble.cpp:60-104; `angular`
 * multiplies by 1 - 3 cos^2 of the angle to the in-plane x axis), then
 * the Gaussian Sz disorder of applyFieldDisorder (Ensemble.cpp:106-112),
 * from Rng::forRealization (include/spinsim/ensemble/Rng.hpp:25-58).
 * Returns SPX_ECONVERGE when a spin cannot be placed at min_separation
 * (the reference loops forever there). */
int spx_dipolar_realisation(uint64_t seed, size_t index, size_t nspins, double scale,
                            double min_separation, int angular, double disorder, uint32_t* lo,
                            uint32_t* hi, double* jx, double* jy, double* jz, double* hz);
/* The typical initial state of realisation (seed, index): the draws of
 * spx_dipolar_realisation, then prepareInitialState's independent Gaussian
 * real and imaginary parts per amplitude on the same stream, normalised
 * (src/ensemble/Ensemble.cpp:20-45, src/io/Experiment.cpp:17-30);
 * psi: 2*2^L doubles. */
int spx_dipolar_state(uint64_t seed, size_t index, size_t nspins, double scale, double min_separation,
                      int angular, double disorder, double* psi);
/* Realisations first .. first+count-1 as dense symmetric J [count][L][L]
 * (the secular jz of each pair) and fields h [count][L], drawn on a pool
 * of host threads; the result does not depend on the pool size. */
int spx_dipolar_graphs(uint64_t seed, size_t first, size_t count, size_t nspins, double scale,
                       double min_separation, int angular, double disorder, double* J, double* h);

#ifdef __cplusplus
}
#endif
#endif /* SPX_H */
