// Population storage: stored type S, arithmetic type V, the exact power-of-two
// scale of FP16 storage, and the addressing of a periodic-shift column.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace lbmh {

// Stored type S, arithmetic type V, and the power-of-two factor the stored
// deviation carries.  FP16 keeps fb_i * 2^12: the scaling is exact, it moves
// FP16's subnormal range (absolute precision 6e-8) below 1.5e-8 so that small
// deviations keep 11 significant bits, and |fb_i| <= 16 stays below FP16's
// maximum (f_i >= 0 bounds fb_i >= -1/3 from below).
template <typename S> struct StorageTraits;
template <> struct StorageTraits<double> {
  using V = double;
  static constexpr double scale = 1.0, inv_scale = 1.0;
};
template <> struct StorageTraits<float> {
  using V = float;
  static constexpr float scale = 1.0f, inv_scale = 1.0f;
};
template <> struct StorageTraits<_Float16> {
  using V = float;
  static constexpr float scale = 4096.0f, inv_scale = 1.0f / 4096.0f;
};

template <typename S, typename V>
__device__ __forceinline__ V load_pop(const S* p) { return V(*p) * V(StorageTraits<S>::inv_scale); }
template <typename S, typename V>
__device__ __forceinline__ void store_pop(S* p, V v) { *p = S(v * V(StorageTraits<S>::scale)); }

// Element k of a column through a 32-bit byte offset from the column base, so
// the base stays in scalar registers and each address costs one VGPR.
template <typename S>
__device__ __forceinline__ S* column_at(S* column, uint32_t k)
{
  using Byte = std::conditional_t<std::is_const_v<S>, const char, char>;
  return reinterpret_cast<S*>(reinterpret_cast<Byte*>(column) + uint32_t(k * uint32_t(sizeof(S))));
}

__device__ __forceinline__ uint32_t wrap_index(uint32_t n, uint32_t shift, uint32_t N)
{
  const uint32_t k = n + shift;
  return k >= N ? k - N : k;
}

}  // namespace lbmh
