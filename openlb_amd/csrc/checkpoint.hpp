// OpenLB checkpoints of one block: the byte stream that
// Serializable::save(std::uint8_t*) writes for a ConcreteBlockLattice
// (src/core/serializer.hh:106-110, src/io/serializerIO.hh:96-106) and
// Serializable::load(const std::uint8_t*) reads back.
//
// The block's serializable data is one entry per allocated field, in the
// order of a std::map keyed by the field array's typeid name
// (src/core/data.h:516-526, data.hh:52-62).  A field of D components is D
// columns (src/core/columnVector.h:155-176):
//   population column (CyclicColumn, src/core/platform/cpu/sisd/column.hh:
//   66-89 and cuda/column.hh:298-321): int64 shift, uint64 count, T[count]
//   other columns (Column, sisd/column.hh:38-63): uint64 count, T[count]
// with count the number of cells of the padded block and T OpenLB's value
// type (double or float).  The population data are stored as they lie in
// memory: logical cell n of column i is element (n + shift_i) mod count.
//
// Besides POPULATION, OpenLB allocates the fields the block's dynamics use:
//   descriptors::FORCE              forced bulk dynamics (3 components)
//   descriptors::VELOCITY           BounceBackVelocity walls
//   momenta::FixedDensity::RHO      BounceBack, BounceBackVelocity, pressure faces
//   momenta::FixedPressureMomentum<d,o>::VELOCITY   one field per pressure face orientation
//   momenta::FixedVelocityMomentumGeneric::VELOCITY velocity boundaries, their edges and corners
// A save writes them from the lattice's own parameters with the values
// OpenLB's set-up gives them: defineRhoU's values on the core cells, the
// value of the momenta's initialize() on overlap cells that are periodic
// images of cells with that dynamics (RHO = 1 for BounceBack and
// BounceBackVelocity, which defineDynamics puts on the overlap too; the
// boundary setters skip the outermost layer), zero elsewhere.  A load takes the populations of every cell, including the
// overlap, and checks on the core cells each value that the block's dynamics
// read against the lattice's parameters.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_local.hpp"
#include "lbm_storage.hpp"
#include "lbmh_common.hpp"

namespace lbmh {

// Stored column (S, with the storage scale) -> OpenLB value type T.
template <typename S, typename T>
__global__ __launch_bounds__(256) void column_export_kernel(T* __restrict__ dst, const S* __restrict__ src, uint32_t n)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  using V = typename StorageTraits<S>::V;
  dst[k] = T(load_pop<S, V>(src + k));
}

template <typename S, typename T>
__global__ __launch_bounds__(256) void column_import_kernel(S* __restrict__ dst, const T* __restrict__ src, uint32_t n)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  using V = typename StorageTraits<S>::V;
  store_pop<S, V>(dst + k, V(src[k]));
}

// What the checkpoint code needs to know of a block.
struct CheckpointBlock {
  int q = 19;
  int n[3] = {0, 0, 0};
  int periodic[3] = {0, 0, 0};
  int storage = LBMH_F32;
  int kind = LBMH_BGK;
  uint32_t n_padded = 0;
  uint64_t stride = 0;            // elements between columns
  void* d_pop = nullptr;          // column 0 on the device
  int64_t shift[27] = {};         // PS shifts in [0, n_padded)
  int64_t dist[27] = {};          // c_i . projection
  const uint8_t* ids = nullptr;   // host ids [nx][ny][nz]
  const double* walls = nullptr;  // [id][4] rho, u
  const uint32_t* info = nullptr; // packed LocalInfo per id
  int n_ids = 0;                  // entries of walls / info
  double force[3] = {0, 0, 0};
  const double* force_field = nullptr;  // per-cell FORCE [3][ncore], or null (force)
  // per padded cell: the role of the cell an overlap cell holds a copy of,
  // where another block (or this one across a periodic face) fills it; -1
  // where no halo does (then the block's own periodic images apply)
  const int8_t* overlap_role = nullptr;
};

// Role of a dynamics id in OpenLB's field set.
enum CheckpointRole : int { kRoleNone, kRoleBulk, kRoleBounceBack, kRoleMovingWall, kRoleVelocity, kRolePressure };

inline int checkpoint_role(const CheckpointBlock& b, int id)
{
  if (id == 0) return kRoleNone;
  if (id == 1) return kRoleBulk;
  if (id == 2) return kRoleBounceBack;
  if (id >= b.n_ids) return kRoleNone;
  const uint32_t k = b.info[id] & 0xffu;
  switch (k) {
  case kMovingWall: return kRoleMovingWall;
  case kFlatPressure:
  case kFdPressure: return kRolePressure;
  case kDeadWall: return kRoleNone;
  default: return kRoleVelocity;
  }
}

// (direction, orientation) of a pressure face id: the axis and sign of its normal.
inline void pressure_orientation(const CheckpointBlock& b, int id, int& dir, int& orient)
{
  const uint32_t w = b.info[id];
  const int nrm[3] = {int((w >> 8) & 0xfu) - 8, int((w >> 12) & 0xfu) - 8, int((w >> 16) & 0xfu) - 8};
  dir = nrm[0] != 0 ? 0 : (nrm[1] != 0 ? 1 : 2);
  orient = nrm[dir];
}

// A non-population field of the block.
struct CheckpointField {
  std::string key;   // typeid(olb::Array<FIELD>).name() under the Itanium C++ ABI
  int components = 1;
  int which = 0;     // 0 FORCE, 1 descriptors::VELOCITY, 2 RHO, 3 pressure VELOCITY, 4 velocity-boundary VELOCITY
  int dir = 0, orient = 0;
};

inline std::string mangled_int(int v) { return v < 0 ? "n" + std::to_string(-v) : std::to_string(v); }

// The fields OpenLB allocates for this block, in its serialization order.
inline std::vector<CheckpointField> checkpoint_fields(const CheckpointBlock& b)
{
  bool role[6] = {};
  std::vector<std::pair<int, int>> faces;
  const size_t ncore = size_t(b.n[0]) * b.n[1] * b.n[2];
  bool seen[256] = {};
  for (size_t c = 0; c < ncore; ++c) seen[b.ids[c]] = true;
  if (b.overlap_role) {
    // OpenLB defines dynamics on the overlap too, so walls seen only there
    // allocate their fields as well
    for (uint32_t k = 0; k < b.n_padded; ++k) {
      const int r = b.overlap_role[k];
      if (r == kRoleBounceBack || r == kRoleMovingWall) role[r] = true;
    }
  }
  for (int id = 0; id < 256; ++id) {
    if (!seen[id]) continue;
    const int r = checkpoint_role(b, id);
    role[r] = true;
    if (r == kRolePressure) {
      int d, o;
      pressure_orientation(b, id, d, o);
      if (std::find(faces.begin(), faces.end(), std::make_pair(d, o)) == faces.end()) faces.emplace_back(d, o);
    }
  }
  std::vector<CheckpointField> f;
  const std::string pre = "N3olb5ArrayINS_";
  if (b.kind == LBMH_FORCED_BGK || b.kind == LBMH_FORCED_TRT) f.push_back({pre + "11descriptors5FORCEEEE", 3, 0});
  if (role[kRoleMovingWall]) f.push_back({pre + "11descriptors8VELOCITYEEE", 3, 1});
  if (role[kRoleBounceBack] || role[kRoleMovingWall] || role[kRolePressure])
    f.push_back({pre + "7momenta12FixedDensity3RHOEEE", 1, 2});
  for (auto [d, o] : faces) {
    f.push_back({pre + "7momenta21FixedPressureMomentumILi" + mangled_int(d) + "ELi" + mangled_int(o) +
                     "EE8VELOCITYEEE",
                 3, 3, d, o});
  }
  if (role[kRoleVelocity]) f.push_back({pre + "7momenta28FixedVelocityMomentumGeneric8VELOCITYEEE", 3, 4});
  // POPULATION ("...11descriptors10POPULATIONEEE") sorts before all of these
  std::sort(f.begin(), f.end(), [](const CheckpointField& x, const CheckpointField& y) { return x.key < y.key; });
  return f;
}

inline size_t checkpoint_bytes(const CheckpointBlock& b, int real)
{
  const size_t t = real == LBMH_F64 ? 8 : 4;
  size_t cols = 0;
  for (const auto& f : checkpoint_fields(b)) cols += size_t(f.components);
  return size_t(b.q) * (16 + t * b.n_padded) + cols * (8 + t * b.n_padded);
}

// OpenLB's signed shift of a column whose shifts started at zero: rotate()
// (sisd/column.h:193-205) keeps it in (-count, 0] when c_i . projection > 0
// and in [0, count) otherwise.
inline int64_t openlb_shift(int64_t shift, int64_t dist, int64_t count)
{
  return (dist > 0 && shift != 0) ? shift - count : shift;
}

// Core-cell value of component `comp` of field `f` at core cell `cell` (flat
// core index), and whether the block's dynamics read it.
inline double field_value(const CheckpointBlock& b, const CheckpointField& f, int id, int comp, bool& read,
                          size_t cell)
{
  read = false;
  const int r = checkpoint_role(b, id);
  const double* w = b.walls + size_t(id) * 4;
  switch (f.which) {
  case 0:
    if (r != kRoleBulk) return 0.0;
    read = true;
    return b.force_field ? b.force_field[size_t(comp) * b.n[0] * b.n[1] * b.n[2] + cell] : b.force[comp];
  case 1:
    if (r != kRoleMovingWall) return 0.0;
    read = true;
    return w[1 + comp];
  case 2:
    if (r == kRoleBounceBack) return 1.0;  // FixedDensity::initialize; BounceBack never reads it
    if (r != kRoleMovingWall && r != kRolePressure) return 0.0;
    read = true;
    return w[0];
  case 3: {
    if (r != kRolePressure) return 0.0;
    int d, o;
    pressure_orientation(b, id, d, o);
    if (d != f.dir || o != f.orient) return 0.0;
    read = comp != d;  // the normal component is computed from the populations
    return w[1 + comp];
  }
  default:
    if (r != kRoleVelocity) return 0.0;
    read = true;
    return w[1 + comp];
  }
}

// Role of the core cell a padded cell is (or is a periodic image of); -1 outside.
inline int image_id(const CheckpointBlock& b, int px, int py, int pz)
{
  int g[3] = {px - 1, py - 1, pz - 1};
  for (int d = 0; d < 3; ++d) {
    if (g[d] >= 0 && g[d] < b.n[d]) continue;
    if (!b.periodic[d]) return -1;
    g[d] = (g[d] + b.n[d]) % b.n[d];
  }
  return b.ids[(size_t(g[0]) * b.n[1] + g[1]) * b.n[2] + g[2]];
}

template <typename T>
void fill_field_column(const CheckpointBlock& b, const CheckpointField& f, int comp, T* out)
{
  const int P[3] = {b.n[0] + 2, b.n[1] + 2, b.n[2] + 2};
  for (int x = 0; x < P[0]; ++x)
    for (int y = 0; y < P[1]; ++y)
      for (int z = 0; z < P[2]; ++z) {
        const size_t k = (size_t(x) * P[1] + y) * P[2] + z;
        const bool core = x >= 1 && x <= b.n[0] && y >= 1 && y <= b.n[1] && z >= 1 && z <= b.n[2];
        double v = 0.0;
        if (core) {
          bool read;
          const size_t cell = (size_t(x - 1) * b.n[1] + (y - 1)) * b.n[2] + (z - 1);
          v = field_value(b, f, b.ids[cell], comp, read, cell);
        } else if (f.which == 2) {
          // defineDynamics covers the overlap, boundary::set does not (setBoundary.h:87-88)
          int r;
          if (b.overlap_role && b.overlap_role[k] >= 0) {
            r = b.overlap_role[k];
          } else {
            const int id = image_id(b, x, y, z);
            r = id < 0 ? kRoleNone : checkpoint_role(b, id);
          }
          if (r == kRoleBounceBack || r == kRoleMovingWall) v = 1.0;
        }
        out[k] = T(v);
      }
}

template <typename T>
inline void put(uint8_t*& p, T v)
{
  std::memcpy(p, &v, sizeof(T));
  p += sizeof(T);
}

template <typename T>
inline T take(const uint8_t*& p)
{
  T v;
  std::memcpy(&v, p, sizeof(T));
  p += sizeof(T);
  return v;
}

template <typename S, typename T>
int checkpoint_save_typed(const CheckpointBlock& b, uint8_t* out)
{
  const uint32_t n = b.n_padded;
  uint8_t* p = out;
  DeviceBuffer staging;
  if (!std::is_same_v<S, T>) LBMH_HIP(hipMalloc(&staging.p, size_t(n) * sizeof(T)));
  LBMH_HIP(hipDeviceSynchronize());
  for (int i = 0; i < b.q; ++i) {
    put<int64_t>(p, openlb_shift(b.shift[i], b.dist[i], n));
    put<uint64_t>(p, n);
    const S* col = static_cast<const S*>(b.d_pop) + uint64_t(i) * b.stride;
    if constexpr (std::is_same_v<S, T>) {
      LBMH_HIP(hipMemcpy(p, col, size_t(n) * sizeof(T), hipMemcpyDeviceToHost));
    } else {
      hipLaunchKernelGGL((column_export_kernel<S, T>), dim3(grid_for(n, 256)), dim3(256), 0, 0,
                         static_cast<T*>(staging.p), col, n);
      LBMH_HIP(hipGetLastError());
      LBMH_HIP(hipMemcpy(p, staging.p, size_t(n) * sizeof(T), hipMemcpyDeviceToHost));
    }
    p += size_t(n) * sizeof(T);
  }
  std::vector<T> column(n);
  for (const auto& f : checkpoint_fields(b)) {
    for (int d = 0; d < f.components; ++d) {
      put<uint64_t>(p, n);
      fill_field_column<T>(b, f, d, column.data());
      std::memcpy(p, column.data(), size_t(n) * sizeof(T));
      p += size_t(n) * sizeof(T);
    }
  }
  return LBMH_OK;
}

static const char* const kFieldNames[5] = {"FORCE", "VELOCITY", "FixedDensity::RHO", "FixedPressureMomentum::VELOCITY",
                                           "FixedVelocityMomentumGeneric::VELOCITY"};

// Validates the whole stream before it changes anything; on success the
// populations are on the device and new_shift holds the shifts.
template <typename S, typename T>
int checkpoint_load_typed(const CheckpointBlock& b, const uint8_t* in, int64_t* new_shift)
{
  const uint32_t n = b.n_padded;
  const uint8_t* p = in;
  std::vector<const T*> data(b.q);
  for (int i = 0; i < b.q; ++i) {
    const int64_t s = take<int64_t>(p);
    const uint64_t count = take<uint64_t>(p);
    if (count != n) return fail(LBMH_E_INVALID, "checkpoint population column " + std::to_string(i) + " holds " +
                                                    std::to_string(count) + " cells, the block " + std::to_string(n));
    if (s <= -int64_t(n) || s >= int64_t(n))
      return fail(LBMH_E_INVALID, "checkpoint shift of population " + std::to_string(i) + " out of range");
    new_shift[i] = s < 0 ? s + n : s;
    data[i] = reinterpret_cast<const T*>(p);
    p += size_t(n) * sizeof(T);
  }
  std::vector<T> column(n);
  for (const auto& f : checkpoint_fields(b)) {
    for (int d = 0; d < f.components; ++d) {
      const uint64_t count = take<uint64_t>(p);
      if (count != n) return fail(LBMH_E_INVALID, std::string("checkpoint field ") + kFieldNames[f.which] +
                                                      " does not match the block's size");
      std::memcpy(column.data(), p, size_t(n) * sizeof(T));
      p += size_t(n) * sizeof(T);
      const int P[3] = {b.n[0] + 2, b.n[1] + 2, b.n[2] + 2};
      for (int x = 0; x < b.n[0]; ++x)
        for (int y = 0; y < b.n[1]; ++y)
          for (int z = 0; z < b.n[2]; ++z) {
            bool read;
            const size_t cell = (size_t(x) * b.n[1] + y) * b.n[2] + z;
            const double v = field_value(b, f, b.ids[cell], d, read, cell);
            const size_t k = (size_t(x + 1) * P[1] + (y + 1)) * P[2] + (z + 1);
            if (read && !(column[k] == T(v))) {
              return fail(LBMH_E_INVALID, std::string("checkpoint field ") + kFieldNames[f.which] + " at cell (" +
                                              std::to_string(x) + ", " + std::to_string(y) + ", " +
                                              std::to_string(z) + ") differs from the lattice's set-up");
            }
          }
    }
  }
  DeviceBuffer staging;
  if (!std::is_same_v<S, T>) LBMH_HIP(hipMalloc(&staging.p, size_t(n) * sizeof(T)));
  LBMH_HIP(hipDeviceSynchronize());
  for (int i = 0; i < b.q; ++i) {
    S* col = static_cast<S*>(b.d_pop) + uint64_t(i) * b.stride;
    if constexpr (std::is_same_v<S, T>) {
      LBMH_HIP(hipMemcpy(col, data[i], size_t(n) * sizeof(T), hipMemcpyHostToDevice));
    } else {
      LBMH_HIP(hipMemcpy(staging.p, data[i], size_t(n) * sizeof(T), hipMemcpyHostToDevice));
      hipLaunchKernelGGL((column_import_kernel<S, T>), dim3(grid_for(n, 256)), dim3(256), 0, 0, col,
                         static_cast<const T*>(staging.p), n);
      LBMH_HIP(hipGetLastError());
    }
  }
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

template <typename F>
int with_storage(int storage, F&& f)
{
  switch (storage) {
  case LBMH_F64: return f(double{});
  case LBMH_F32: return f(float{});
  case LBMH_F16: return f(_Float16{});
  default: return fail(LBMH_E_INVALID, "unknown storage");
  }
}

inline int checkpoint_save(const CheckpointBlock& b, int real, void* out, size_t bytes)
{
  if (real != LBMH_F64 && real != LBMH_F32) return fail(LBMH_E_INVALID, "checkpoint value type must be F64 or F32");
  if (!out) return fail(LBMH_E_INVALID, "null buffer");
  if (bytes != checkpoint_bytes(b, real)) return fail(LBMH_E_INVALID, "checkpoint buffer size differs from lbmh_checkpoint_size");
  uint8_t* p = static_cast<uint8_t*>(out);
  return with_storage(b.storage, [&](auto s) -> int {
    using S = decltype(s);
    return real == LBMH_F64 ? checkpoint_save_typed<S, double>(b, p) : checkpoint_save_typed<S, float>(b, p);
  });
}

inline int checkpoint_load(const CheckpointBlock& b, int real, const void* in, size_t bytes, int64_t* new_shift)
{
  if (real != LBMH_F64 && real != LBMH_F32) return fail(LBMH_E_INVALID, "checkpoint value type must be F64 or F32");
  if (!in) return fail(LBMH_E_INVALID, "null buffer");
  if (bytes != checkpoint_bytes(b, real))
    return fail(LBMH_E_INVALID, "checkpoint holds " + std::to_string(bytes) + " bytes, a block of this set-up " +
                                    std::to_string(checkpoint_bytes(b, real)));
  const uint8_t* p = static_cast<const uint8_t*>(in);
  return with_storage(b.storage, [&](auto s) -> int {
    using S = decltype(s);
    return real == LBMH_F64 ? checkpoint_load_typed<S, double>(b, p, new_shift)
                            : checkpoint_load_typed<S, float>(b, p, new_shift);
  });
}

}  // namespace lbmh
