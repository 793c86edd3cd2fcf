// Host-side helpers shared by the translation units of the C ABI: the
// per-thread error text, HIP error mapping, device selection and the
// derived collision parameters.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "lbm_collision.hpp"
#include "lbmh.h"

namespace lbmh {

// text of the calling thread's last error (defined in lbmh.hip)
extern thread_local std::string g_error;

inline int fail(int code, const std::string& msg)
{
  g_error = msg;
  return code;
}

inline int hip_fail(hipError_t e, const char* where)
{
  return fail(LBMH_E_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

#define LBMH_HIP(call)                                        \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return ::lbmh::hip_fail(e_, #call); \
  } while (0)

inline size_t storage_size(int s) { return s == LBMH_F64 ? 8 : (s == LBMH_F32 ? 4 : 2); }

inline uint32_t grid_for(uint32_t n, uint32_t block) { return (n + block - 1) / block; }

// RAII device buffer for temporaries.
struct DeviceBuffer {
  void* p = nullptr;
  ~DeviceBuffer()
  {
    if (p) (void)hipFree(p);
  }
};

// Make `dev` current for the scope of a call, restore the caller's device after.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev)
  {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard()
  {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

inline bool valid_bulk_params(int kind, const lbmh_bulk_params& p)
{
  if (!(p.omega > 0.0 && p.omega < 2.0)) return fail(LBMH_E_INVALID, "omega must lie in (0, 2)"), false;
  if ((kind == LBMH_TRT || kind == LBMH_FORCED_TRT) && !(p.magic > 0.0))
    return fail(LBMH_E_INVALID, "TRT magic parameter must be positive"), false;
  if (kind == LBMH_SMAGORINSKY && !(p.smago >= 0.0))
    return fail(LBMH_E_INVALID, "Smagorinsky constant must be non-negative"), false;
  return true;
}

// Parameters the collision kernels read, derived once on the host in FP64.
template <typename V>
BulkParams<V> derive_bulk_params(int q, const lbmh_bulk_params& p)
{
  BulkParams<V> b;
  const double omega = p.omega;
  const double tau = 1.0 / omega;
  b.omega = V(omega);
  b.omega2 = V(1.0 / (p.magic / (tau - 0.5) + 0.5));
  b.guo = V(1.0 - omega * 0.5);
  for (int d = 0; d < 3; ++d) {
    b.force[d] = V(p.force[d]);
    b.half_force[d] = V(p.force[d] * 0.5);
  }
  if (q == 19) {
    // constants of OpenLB's generated D3Q19 Smagorinsky kernel
    // (src/cse/dynamics/d7a4bdcd...cse.h:110)
    b.sm_a = V(3.00000046417339);
    b.sm_b = V(p.smago * p.smago / std::sqrt(2.0));
    b.sm_k = V(0.0277777691819762);
  } else {
    // collisionLES.h:74-80: tau_eff = tau/2 + sqrt(tau^2 + 18 sqrt2 C^2 |Pi|/rho)/2
    b.sm_a = V(0.5);
    b.sm_b = V(18.0 * std::sqrt(2.0) * p.smago * p.smago);
    b.sm_k = V(1.0);
  }
  b.tau = V(tau);
  b.half_tau = V(0.5 * tau);
  return b;
}

}  // namespace lbmh
