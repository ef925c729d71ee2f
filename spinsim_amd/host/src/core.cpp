// Device plumbing, StateVector, Bessel/Chebyshev and Hamiltonian of the
// spinsim drop-in, over the libspx C ABI.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "spinsim/core/Bessel.hpp"
#include "spinsim/core/Chebyshev.hpp"
#include "spinsim/core/Device.hpp"
#include "spinsim/core/Hamiltonian.hpp"
#include "spinsim/core/StateVector.hpp"

namespace spinsim {

// ------------------------------------------------------------------ device

void check(int status) {
  if (status == SPX_OK) return;
  const std::string msg = spx_last_error();
  switch (status) {
    case SPX_EINVAL: throw std::invalid_argument(msg);
    case SPX_ERANGE: throw std::out_of_range(msg);
    case SPX_EDOMAIN: throw std::domain_error(msg);
    case SPX_ECONVERGE: {
      double b = 0, t = 0, e = 0;
      std::size_t o = 0;
      spx_last_convergence(&b, &t, &e, &o);
      throw ChebyshevConvergenceError(msg, b, t, e, o);
    }
    case SPX_EOVERFLOW: throw std::overflow_error(msg);
    case SPX_ENOMEM: throw std::bad_alloc();
    case SPX_EDEVICE: throw std::runtime_error("device failure: " + msg);
    default: throw std::runtime_error(msg);
  }
}

DeviceContext::DeviceContext(int device) { check(spx_ctx_create(device, &ctx_)); }

DeviceContext::~DeviceContext() {
  if (ctx_) spx_ctx_destroy(ctx_);
}

void DeviceContext::synchronize() const { check(spx_ctx_synchronize(ctx_)); }

std::shared_ptr<DeviceContext> DeviceContext::current() {
  thread_local std::shared_ptr<DeviceContext> ctx;
  if (!ctx) {
    const char* e = std::getenv("SPINSIM_DEVICE");
    ctx = std::make_shared<DeviceContext>(e ? std::atoi(e) : 0);
  }
  return ctx;
}

// -------------------------------------------------------------- StateVector

StateVector::StateVector(std::size_t nspins) : nspins_(nspins), ctx_(DeviceContext::current()) {
  if (nspins > kMaxSpins) throw std::out_of_range("number of spins above kMaxSpins");
  check(spx_state_create(ctx_->get(), nspins, &dev_));
  fresh_ = Fresh::Device;
}

StateVector::StateVector(const StateVector& other) : nspins_(other.nspins_) {
  if (!other.dev_) return;
  ctx_ = DeviceContext::current();
  check(spx_state_create(ctx_->get(), nspins_, &dev_));
  spx_state* src = other.device(false);
  if (other.ctx_ != ctx_) other.ctx_->synchronize();
  check(spx_state_copy(dev_, src));
  fresh_ = Fresh::Device;
}

StateVector& StateVector::operator=(const StateVector& other) {
  if (this == &other) return *this;
  StateVector tmp(other);
  swap(tmp);
  return *this;
}

StateVector::StateVector(StateVector&& other) noexcept { swap(other); }

StateVector& StateVector::operator=(StateVector&& other) noexcept {
  swap(other);
  return *this;
}

StateVector::~StateVector() {
  if (dev_) spx_state_destroy(dev_);
}

void StateVector::pull() const {
  if (!dev_) return;
  if (fresh_ == Fresh::Device || host_.size() != 2 * stateCount(nspins_)) {
    host_.resize(2 * stateCount(nspins_));
    check(spx_state_download(dev_, host_.data()));
    fresh_ = Fresh::Both;
  }
}

double* StateVector::data() {
  pull();
  fresh_ = Fresh::Host;
  return host_.data();
}

const double* StateVector::data() const {
  pull();
  return host_.data();
}

std::span<double> StateVector::amplitudes() { return {data(), 2 * size()}; }
std::span<const double> StateVector::amplitudes() const { return {data(), 2 * size()}; }

spx_state* StateVector::device(bool willWrite) const {
  if (!dev_) throw std::invalid_argument("empty state vector");
  if (fresh_ == Fresh::Host) {
    check(spx_state_upload(dev_, host_.data()));
    fresh_ = Fresh::Both;
  }
  if (willWrite) fresh_ = Fresh::Device;
  return dev_;
}

void StateVector::setZero() {
  if (!dev_) return;
  check(spx_state_set_zero(device(true)));
}

void StateVector::scale(double factor) {
  if (!dev_) return;
  check(spx_state_scale(device(true), factor));
}

void StateVector::negate() {
  if (!dev_) return;
  check(spx_state_negate(device(true)));
}

void StateVector::addScaled(double cr, double ci, const StateVector& other) {
  if (!dev_ && !other.dev_) return;
  check(spx_state_add_scaled(device(true), cr, ci, other.device(false)));
}

void StateVector::setBasisState(std::size_t index) {
  if (index >= size()) throw std::out_of_range("basis state index outside the 2^L state space");
  check(spx_state_set_basis(device(true), index));
}

double StateVector::norm() const {
  if (!dev_) return 0.0;
  double n = 0.0;
  check(spx_state_norm(device(false), &n));
  return n;
}

double StateVector::normalize() {
  if (!dev_) throw std::domain_error("cannot normalize a zero state");
  double n = 0.0;
  check(spx_state_normalize(device(true), &n));
  return n;
}

std::pair<double, double> StateVector::innerProduct(const StateVector& other) const {
  if (!dev_ && !other.dev_) return {0.0, 0.0};
  double re = 0.0, im = 0.0;
  check(spx_state_inner(device(false), other.device(false), &re, &im));
  return {re, im};
}

void StateVector::swap(StateVector& other) noexcept {
  std::swap(nspins_, other.nspins_);
  std::swap(ctx_, other.ctx_);
  std::swap(dev_, other.dev_);
  std::swap(host_, other.host_);
  std::swap(fresh_, other.fresh_);
}

// -------------------------------------------------------- Bessel/Chebyshev

namespace {

BesselSequence bessel(int kind, double x, std::size_t count) {
  if (count == 0) throw std::invalid_argument("Bessel sequence length must be positive");
  BesselSequence out;
  out.values.assign(count, 0.0);
  check(spx_bessel(kind, x, count, out.values.data(), &out.nonzero));
  return out;
}

}  // namespace

BesselSequence besselJ(double x, std::size_t count) { return bessel(0, x, count); }
BesselSequence besselI(double x, std::size_t count) { return bessel(1, x, count); }

ChebyshevExpansion ChebyshevExpansion::build(TimeMode mode, double spectralBound, double tau,
                                             double epsilon, std::size_t maxOrder) {
  ChebyshevExpansion out;
  std::vector<double> buf(maxOrder + 8);
  std::size_t order = 0;
  double eff = 1.0;
  check(spx_cheb_build(mode == TimeMode::Real ? 0 : 1, spectralBound, tau, epsilon, maxOrder,
                       buf.data(), buf.size(), &order, &eff));
  out.coefficients_.assign(buf.begin(), buf.begin() + static_cast<std::ptrdiff_t>(order + 1));
  out.bound_ = eff;
  return out;
}

// -------------------------------------------------------------- Hamiltonian

Hamiltonian::Hamiltonian(std::size_t nspins) : nspins_(nspins) {
  if (nspins == 0 || nspins > kMaxSpins)
    throw std::out_of_range("number of spins must be in [1, kMaxSpins]");
  fields_.assign(nspins, Field{});
  create();
}

void Hamiltonian::create() {
  ctx_ = DeviceContext::current();
  check(spx_ham_create(ctx_->get(), nspins_, &dev_));
  for (SpinIndex s = 0; s < nspins_; ++s)
    check(spx_ham_set_field(dev_, s, fields_[s].x, fields_[s].y, fields_[s].z));
  for (const Coupling& c : couplings_)
    check(spx_ham_set_coupling(dev_, c.pair.low(), c.pair.high(), c.jx, c.jy, c.jz));
}

Hamiltonian::Hamiltonian(const Hamiltonian& other)
    : nspins_(other.nspins_), fields_(other.fields_), couplings_(other.couplings_) {
  create();
}

Hamiltonian& Hamiltonian::operator=(const Hamiltonian& other) {
  if (this == &other) return *this;
  if (dev_) spx_ham_destroy(dev_);
  dev_ = nullptr;
  nspins_ = other.nspins_;
  fields_ = other.fields_;
  couplings_ = other.couplings_;
  create();
  return *this;
}

Hamiltonian::~Hamiltonian() {
  if (dev_) spx_ham_destroy(dev_);
}

void Hamiltonian::setField(SpinIndex spin, Field field) {
  if (spin >= nspins_) throw std::out_of_range("spin index out of range");
  check(spx_ham_set_field(dev_, spin, field.x, field.y, field.z));
  fields_[spin] = field;
}

const Field& Hamiltonian::field(SpinIndex spin) const {
  if (spin >= nspins_) throw std::out_of_range("spin index out of range");
  return fields_[spin];
}

void Hamiltonian::setCoupling(SpinPair pair, double jx, double jy, double jz) {
  if (pair.high() >= nspins_) throw std::out_of_range("coupled spin index out of range");
  check(spx_ham_set_coupling(dev_, pair.low(), pair.high(), jx, jy, jz));
  for (Coupling& c : couplings_)
    if (c.pair == pair) {
      c.jx = jx;
      c.jy = jy;
      c.jz = jz;
      return;
    }
  couplings_.push_back(Coupling{pair, jx, jy, jz});
}

double Hamiltonian::spectralBound() const noexcept {
  double b = 0.0;
  spx_ham_spectral_bound(dev_, &b);
  return b;
}

void Hamiltonian::applyPhysical(const StateVector& in, StateVector& out) const {
  check(spx_ham_apply(dev_, 0, 0.0, in.device(false), out.device(true)));
}

void Hamiltonian::applyTwiceNormalized(double bound, const StateVector& in, StateVector& out) const {
  if (!(bound > 0.0)) throw std::domain_error("spectral bound must be positive");
  check(spx_ham_apply(dev_, 1, bound, in.device(false), out.device(true)));
}

void Hamiltonian::accumulateTwiceNormalized(double bound, const StateVector& in,
                                            StateVector& out) const {
  if (!(bound > 0.0)) throw std::domain_error("spectral bound must be positive");
  check(spx_ham_apply(dev_, 2, bound, in.device(false), out.device(true)));
}

}  // namespace spinsim
