#pragma once
// Device plumbing shared by the spinsim classes: the per-thread libspx
// context and the mapping of spx status codes onto the exception types the
// reference throws (include/spinsim/core/*.hpp).

#include <memory>

#include "spx.h"

namespace spinsim {

/// A libspx context (one HIP stream); shared by every object created from it.
class DeviceContext {
 public:
  explicit DeviceContext(int device);
  ~DeviceContext();
  DeviceContext(const DeviceContext&) = delete;
  DeviceContext& operator=(const DeviceContext&) = delete;

  [[nodiscard]] spx_ctx* get() const noexcept { return ctx_; }
  void synchronize() const;

  /// The calling thread's context on the default device (SPINSIM_DEVICE, or
  /// 0), created on first use. Each host thread gets its own stream.
  static std::shared_ptr<DeviceContext> current();

 private:
  spx_ctx* ctx_ = nullptr;
};

/// Throws the reference's exception type for a non-OK spx status.
void check(int status);

}  // namespace spinsim
