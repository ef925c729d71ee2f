/*
 * CPU oracle for the spinsim Chebyshev propagator and its disorder-averaging
 * loop. TEST INFRASTRUCTURE ONLY: this is the checker that the GPU path is
 * compared against and the CPU baseline timed beside it. Nothing in the
 * shipped library links, loads or calls it.
 *
 * A plain-C restatement of the reference algorithm (BekturMur/Spin_Systems),
 * FP64, compiled with -ffp-contract=off like the reference
 * (CMakeLists.txt:16-22) so that the arithmetic order of every kernel is the
 * reference's own. Each function cites the reference file:line it follows.
 *
 * Storage convention (include/spinsim/core/StateVector.hpp:53-60):
 * amplitude k of a 2^L state occupies psi[2k] (re) and psi[2k+1] (im);
 * spin s owns bit s of k, a set bit is "up" (include/spinsim/core/Bits.hpp:15-27).
 *
 * Parity pinning (how each part is checked before it is trusted):
 *   rng        bit-exact vs the reference Rng.cpp compiled into oracle/_ref,
 *              the C++-standard mt19937_64 known answer, committed draws;
 *   bessel     tests/data/bessel_reference.txt (Cody rjbesl/ribesl) to 1e-14,
 *              recurrences and sum rules, bit-exact vs oracle/_ref;
 *   chebyshev  order rule vs closed behaviour (growth with tau, identity
 *              shortcut, convergence error), propagator vs dense expm;
 *   kernels    bit-exact vs reference SpinOperators.cpp / Hamiltonian.cpp in
 *              oracle/_ref, and vs dense Kronecker-product Pauli matrices;
 *   propagator dense eigendecomposition (numpy eigh) to 1e-10 at L<=6,
 *              time reversal, Rabi closed form, norm/energy conservation;
 *   gates      dense U = cos - i sin n.sigma, pi_x = -i sigma_x, norm;
 *   observables dense correlator, known-state magnetisation;
 *   ensemble   geometry and initial state: parity with the reference's
 *              Ensemble.cpp is pinned only through the pinned RNG plus the
 *              reference tests' properties (pair count, Jx=Jy=-Jz/2, Jz>0,
 *              reproducibility, t=0 correlator ~ L) -- the reference file
 *              needs <format>, absent here, so it cannot be compiled;
 *   accumulator textbook set {2,4,4,4,5,5,7,9}, Chan merge == one pass.
 */
#ifndef SPINSIM_ORACLE_H
#define SPINSIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: include/spinsim/ensemble/Rng.hpp:25-58, src/ensemble/Rng.cpp:8-23 */
typedef struct {
  uint64_t mt[312];
  size_t mti;
  double spare;
  int has_spare;
} so_rng;

uint64_t so_splitmix(uint64_t z);
void so_rng_seed(so_rng* r, uint64_t seed);                 /* Rng(seed) */
void so_rng_for_realization(so_rng* r, uint64_t base, uint64_t index);
uint64_t so_mt_next(so_rng* r);                             /* engine_() */
double so_rng_uniform(so_rng* r);
double so_rng_gaussian(so_rng* r);
void so_mt_raw(uint64_t seed, size_t n, uint64_t* out);     /* raw mt19937_64(seed) */
void so_rng_draws(uint64_t base, uint64_t index, size_t n, int gaussian, double* out);

/* ---- Bessel: src/core/Bessel.cpp:20-149. Return 0 ok, -1 invalid arg,
 * -2 overflow (I with x > 709), -3 normalisation failure. */
int so_bessel_j(double x, size_t count, double* out, size_t* nonzero);
int so_bessel_i(double x, size_t count, double* out, size_t* nonzero);

/* ---- Chebyshev: src/core/Chebyshev.cpp:24-107. mode 0 real, 1 imaginary.
 * coeffs must hold max_order + 4 doubles. Returns 0 ok, -1 invalid,
 * -2 overflow, -4 convergence error. */
int so_cheb_build(int mode, double bound, double tau, double eps, size_t max_order,
                  double* coeffs, size_t* order, double* eff_bound);

/* ---- StateVector: src/core/StateVector.cpp:10-75 (n = 2^L amplitudes) */
void so_state_zero(double* psi, size_t n);
void so_state_scale(double* psi, size_t n, double f);
void so_state_negate(double* psi, size_t n);
void so_state_add_scaled(double* dst, size_t n, double cr, double ci, const double* src);
double so_state_norm(const double* psi, size_t n);
int so_state_normalize(double* psi, size_t n, double* norm_out);
void so_state_inner(const double* self, const double* other, size_t n, double* re, double* im);

/* ---- kernels: src/core/SpinOperators.cpp:90-291 (Pauli units) */
void so_axpy_sigma(int axis, double c, const double* in, size_t L, size_t spin, double* out);
void so_apply_sigma(int axis, const double* in, size_t L, size_t spin, double* out);
void so_axpy_coupling(double jx, double jy, double jz, const double* in, size_t L,
                      size_t lo, size_t hi, double* out);
void so_axpy_diagonal(double c, const double* diag, const double* in, size_t L, double* out);
void so_axpy_transverse(double cx, double cy, const double* in, size_t L, size_t spin, double* out);
void so_axpy_flip_flop(double c, const double* in, size_t L, size_t lo, size_t hi, double* out);
void so_axpy_flip_both(double c, const double* in, size_t L, size_t lo, size_t hi, double* out);

/* ---- Hamiltonian: src/core/Hamiltonian.cpp:20-147 */
typedef struct {
  size_t L;
  size_t npairs;
  double* hx; double* hy; double* hz;          /* [L] */
  uint32_t* lo; uint32_t* hi;                  /* [npairs], lo < hi */
  double* jx; double* jy; double* jz;          /* [npairs] */
  double* diag;                                /* [2^L] cache */
  int diag_valid;
  size_t cap;
} so_ham;

so_ham* so_ham_new(size_t L);
void so_ham_free(so_ham* h);
int so_ham_set_field(so_ham* h, size_t spin, double x, double y, double z);
int so_ham_set_coupling(so_ham* h, size_t a, size_t b, double jx, double jy, double jz);
double so_ham_bound(const so_ham* h);
void so_ham_rebuild_diag(so_ham* h);
const double* so_ham_diag(so_ham* h);
void so_ham_accumulate(so_ham* h, double scale, const double* in, double* out);
void so_ham_apply_physical(so_ham* h, const double* in, double* out);
int so_ham_apply_twice_normalized(so_ham* h, double bound, const double* in, double* out);
int so_ham_accumulate_twice_normalized(so_ham* h, double bound, const double* in, double* out);

/* ---- Propagator: src/core/Propagator.cpp:14-100. Returns 0 ok or the
 * Chebyshev error code; *order_out = lastOrder(). */
int so_evolve(so_ham* h, int mode, double* psi, double tau, size_t nsteps,
              int autonormalize, double eps, size_t* order_out);
/* One fused Chebyshev step with a prebuilt expansion (applyOneStep). */
void so_apply_one_step(so_ham* h, const double* coeffs, size_t order, double bound,
                       int mode, double* psi, double* cur, double* prev);

/* ---- Gates: src/core/Gate.cpp:20-121 */
int so_rotate(double* psi, size_t L, double ax, double ay, double az, double angle,
              const uint32_t* targets, size_t ntargets);
int so_two_spin_gate(double* psi, size_t L, const double* u16_interleaved,
                     size_t lo, size_t hi);

/* ---- Observables: src/core/Observable.cpp:13-120 */
void so_apply_total_spin(int axis, const double* in, size_t L, double* out);
void so_corr(int axis, const double* state, const double* ref, size_t L, double* scratch,
             double* re, double* im);
double so_mag(int axis, size_t spin, const double* state, size_t L, double* scratch);
double so_energy(so_ham* h, const double* state, double* scratch);

/* ---- Ensemble: src/ensemble/Ensemble.cpp:20-112 */
int so_prepare_initial_state(double* psi, size_t L, int kind, so_rng* r, size_t basis);
void so_prepare_reference(const double* psi, size_t L, int axis, double* ref);
/* Writes npairs = L(L-1)/2 couplings in (j<i) generation order. angular=1 adds
 * the secular (1 - 3 cos^2 theta) factor with cos theta = dx/r (no counterpart
 * in the reference generator, which emits Jz = scale/r^3). */
int so_dipolar_lattice(size_t L, so_rng* r, double scale, double min_sep, int angular,
                       uint32_t* lo, uint32_t* hi, double* jx, double* jy, double* jz);
void so_field_disorder(so_ham* h, so_rng* r, double sigma);

/* ---- Sequence program (include/spinsim/sequence/Sequence.hpp:19-104),
 * flattened: repeat blocks are begin/end markers. */
enum { SO_OP_EVOLVE = 0, SO_OP_PULSE = 1, SO_OP_MEASURE = 2, SO_OP_REPEAT = 3, SO_OP_END = 4 };
typedef struct so_program so_program;
so_program* so_program_new(void);
void so_program_free(so_program* p);
int so_program_evolve(so_program* p, double tau, size_t nsteps, int mode, int autonorm,
                      const double* fields3L, size_t L);
int so_program_pulse(so_program* p, double ax, double ay, double az, double angle,
                     const uint32_t* targets, size_t ntargets);
int so_program_measure(so_program* p);
int so_program_repeat(so_program* p, size_t count);
int so_program_end(so_program* p);
size_t so_program_measurements(const so_program* p);

/* Observables: code 0..2 corr:x/y/z (2 values), 3..5 mag:<spin>:x/y/z
 * (1 value), 6 norm (2 values), 7 energy (1 value). */
enum { SO_OBS_CORR = 0, SO_OBS_MAG = 3, SO_OBS_NORM = 6, SO_OBS_ENERGY = 7 };

typedef struct {
  size_t L;
  const double* base_fields;   /* [3L] x,y,z per spin, or NULL for zero */
  int initial_state;           /* 0 random, 1 up, 2 down, 3 basis */
  size_t basis_index;
  uint64_t seed;
  int generate_geometry;
  double dipolar_scale;
  double min_separation;       /* reference default 0.05 */
  int angular;                 /* 0 reference generator, 1 (1-3cos^2) factor */
  double field_disorder;
  int reference_axis;          /* 0 x, 1 y, 2 z */
  double epsilon;
  const int* obs_codes;        /* [nobs] */
  const int* obs_spins;        /* [nobs] spin for mag, else ignored */
  size_t nobs;
  const so_program* program;
} so_config;

size_t so_config_width(const so_config* c);
/* runRealization (src/io/Experiment.cpp:10-55). times [points], values
 * [points*width]. Returns 0 or a negative error code. */
int so_run_realization(const so_config* c, uint64_t index, double* times, double* values);
/* LocalRunner::run (src/ensemble/EnsembleRunner.cpp:27-88) over indices
 * [0, n) with `threads` workers, per-realisation rows kept in `per` if
 * non-NULL ([n*points*width]) and folded in index order by Welford into
 * mean/sd/stderr [points*width]. */
int so_run_ensemble(const so_config* c, size_t n, size_t threads, double* times,
                    double* per, double* mean, double* sd, double* se);

/* ---- Accumulator: src/ensemble/Ensemble.cpp:114-166 */
typedef struct { size_t count, width; double* mean; double* m2; } so_acc;
so_acc* so_acc_new(size_t width);
void so_acc_free(so_acc* a);
int so_acc_add(so_acc* a, const double* sample);
int so_acc_merge(so_acc* a, const so_acc* b);
void so_acc_stats(const so_acc* a, double* mean, double* sd, double* se);

/* ---- timing helper for the CPU baseline: `orders` fused Chebyshev orders
 * (negate + accumulateTwiceNormalized + addScaled) on `threads` independent
 * states of one Hamiltonian; returns wall seconds. */
double so_time_orders(so_ham* h, size_t orders, size_t threads);

#ifdef __cplusplus
}
#endif
#endif
