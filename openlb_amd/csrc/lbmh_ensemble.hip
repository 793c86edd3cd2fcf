// C ABI of lattice ensembles (declarations in include/lbmh.h).
//
// Members live in two device arenas (populations, dynamics ids) carved at
// creation, so admitting or retiring a member never calls hipMalloc/hipFree
// (hipFree would also wait for the whole device).  The member table and the
// chunk -> member map are device arrays; the host keeps the arena bookkeeping
// and rebuilds the map when membership changes.  Shifts and step counters are
// advanced on the device, so a step loop can be captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "boundary_classify.hpp"
#include "checkpoint.hpp"
#include "lbm_fields.hpp"
#include "lbm_ensemble.hpp"
#include "lbmh.h"
#include "lbmh_common.hpp"

using namespace lbmh;

namespace {

constexpr uint64_t kAlign = 256;

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, capacity) with coalescing of free ranges.
class Arena {
 public:
  explicit Arena(uint64_t capacity = 0) : capacity_(capacity)
  {
    if (capacity) free_[0] = capacity;
  }
  bool alloc(uint64_t bytes, uint64_t* off)
  {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const uint64_t rest = it->second - bytes, start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes)
  {
    bytes = align_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t used() const { return used_; }
  uint64_t capacity() const { return capacity_; }

 private:
  uint64_t capacity_ = 0, used_ = 0;
  std::map<uint64_t, uint64_t> free_;
};

struct Member {
  bool active = false;
  int n[3] = {0, 0, 0};
  int periodic[3] = {0, 0, 0};
  uint64_t pop_off = 0, pop_bytes = 0;   // bytes in the population arena
  uint64_t ids_off = 0, ids_bytes = 0;
  uint32_t ncore = 0;
  lbmh_bulk_params params{};
  // local boundaries: host ids, per-id descriptors, cell lists (core indices)
  std::vector<uint8_t> h_ids;
  uint32_t info[kEnsembleWalls] = {};
  bool id_defined[kEnsembleWalls] = {};
  std::vector<uint32_t> local, edges, corners, planes;
  bool pressure_planes = false;  // InterpolatedPressure faces: its planes run the two-phase pass
};

}  // namespace

struct lbmh_ensemble {
  int device = 0, q = 19, storage = LBMH_F32, kind = LBMH_BGK;
  int max_members = 0;
  void* d_pop = nullptr;
  uint8_t* d_ids = nullptr;
  double* d_walls = nullptr;
  void* d_table = nullptr;
  uint32_t* d_chunk = nullptr;
  uint32_t chunk_capacity = 0, total_chunks = 0;
  size_t entry_bytes = 0;
  Arena pop_arena, id_arena;
  std::vector<Member> members;
  std::vector<unsigned char> h_table;   // host mirror of the member table
  lbmh_launch_config launch{0, 4};
  bool launch_chosen = false;     // set by lbmh_ensemble_set_launch_config: keep it
  uint64_t version = 1;
  hipStream_t capture_stream = nullptr;
  hipGraphExec_t graph = nullptr;
  uint64_t graph_version = 0;
  int64_t graph_steps = 0;
  // local boundaries: per-member id descriptors and (member, cell) lists
  uint32_t* d_info = nullptr;
  uint2* d_local = nullptr;
  uint2* d_edges = nullptr;
  uint2* d_corners = nullptr;
  uint2* d_planes = nullptr;    // plane cells of members without pressure planes: one pass
  uint2* d_pplanes = nullptr;   // plane cells of members with pressure planes: two-phase pass
  size_t boundary_capacity = 0;
  // the two-phase plane pass's populations, boundary_capacity x Q values,
  // allocated with the first member that has InterpolatedPressure faces
  void* d_plane_scratch = nullptr;
  uint32_t n_local = 0, n_edges = 0, n_corners = 0, n_planes = 0, n_pplanes = 0;
  bool has_local = false;            // some member ever had local boundaries
};

namespace {

template <int Q>
EnsembleEntry<Q>& entry(lbmh_ensemble* E, int m)
{
  return *reinterpret_cast<EnsembleEntry<Q>*>(E->h_table.data() + size_t(m) * E->entry_bytes);
}

template <typename F>
int with_q(const lbmh_ensemble* E, F&& f)
{
  return E->q == 19 ? f(std::integral_constant<int, 19>()) : f(std::integral_constant<int, 27>());
}

template <typename F>
int with_types(const lbmh_ensemble* E, F&& f)
{
  if (E->q == 19) {
    if (E->storage == LBMH_F64) return f(std::integral_constant<int, 19>(), double());
    if (E->storage == LBMH_F32) return f(std::integral_constant<int, 19>(), float());
    return f(std::integral_constant<int, 19>(), _Float16());
  }
  if (E->storage == LBMH_F64) return f(std::integral_constant<int, 27>(), double());
  if (E->storage == LBMH_F32) return f(std::integral_constant<int, 27>(), float());
  return f(std::integral_constant<int, 27>(), _Float16());
}

int pull_table(lbmh_ensemble* E)
{
  LBMH_HIP(hipDeviceSynchronize());
  LBMH_HIP(hipMemcpy(E->h_table.data(), E->d_table, E->h_table.size(), hipMemcpyDeviceToHost));
  return LBMH_OK;
}

// Assign chunk ranges to active members in slot order and upload table + map.
int push_table(lbmh_ensemble* E)
{
  std::vector<uint32_t> map;
  map.reserve(E->total_chunks);
  const int rc = with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    for (int m = 0; m < E->max_members; ++m) {
      EnsembleEntry<Q>& e = entry<Q>(E, m);
      if (!E->members[m].active) {
        e.chunk_begin = 0;
        e.steps_left = 0;
        continue;
      }
      e.chunk_begin = uint32_t(map.size());
      const uint32_t chunks = grid_for(e.ncore, kBlock);
      map.insert(map.end(), chunks, uint32_t(m));
    }
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  if (map.size() > E->chunk_capacity) return fail(LBMH_E_MEMORY, "chunk map capacity exceeded");
  E->total_chunks = uint32_t(map.size());
  LBMH_HIP(hipMemcpy(E->d_table, E->h_table.data(), E->h_table.size(), hipMemcpyHostToDevice));
  if (!map.empty()) LBMH_HIP(hipMemcpy(E->d_chunk, map.data(), map.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  E->version += 1;
  return LBMH_OK;
}

int check_member(const lbmh_ensemble* E, int m)
{
  if (!E) return fail(LBMH_E_INVALID, "null ensemble");
  if (m < 0 || m >= E->max_members || !E->members[m].active) return fail(LBMH_E_INVALID, "no such member");
  return LBMH_OK;
}

// Concatenate the local-boundary lists of the active members into the device lists.
int rebuild_lists(lbmh_ensemble* E)
{
  std::vector<uint2> all[5];
  for (int m = 0; m < E->max_members; ++m) {
    const Member& mem = E->members[m];
    if (!mem.active) continue;
    const std::vector<uint32_t>* src[4] = {&mem.local, &mem.edges, &mem.corners, &mem.planes};
    for (int k = 0; k < 4; ++k) {
      const int to = (k == 3 && mem.pressure_planes) ? 4 : k;
      for (uint32_t c : *src[k]) all[to].push_back(make_uint2(uint32_t(m), c));
    }
  }
  for (int k = 0; k < 5; ++k)
    if (all[k].size() > E->boundary_capacity) return fail(LBMH_E_MEMORY, "boundary list capacity exceeded");
  uint2* dst[5] = {E->d_local, E->d_edges, E->d_corners, E->d_planes, E->d_pplanes};
  for (int k = 0; k < 5; ++k)
    if (!all[k].empty()) LBMH_HIP(hipMemcpy(dst[k], all[k].data(), all[k].size() * sizeof(uint2), hipMemcpyHostToDevice));
  E->n_local = uint32_t(all[0].size());
  E->n_edges = uint32_t(all[1].size());
  E->n_corners = uint32_t(all[2].size());
  E->n_planes = uint32_t(all[3].size());
  E->n_pplanes = uint32_t(all[4].size());
  E->version += 1;
  return LBMH_OK;
}

template <int Q>
CoreGeom geom_of(const EnsembleEntry<Q>& e)
{
  CoreGeom g;
  g.ncore = e.ncore;
  g.nx = e.nx;
  g.ny = e.ny;
  g.nz = e.nz;
  g.sx = e.sx;
  g.sy = e.sy;
  g.n_padded = e.n_padded;
  g.stride = e.stride;
  g.div_z = e.div_z;
  g.div_y = e.div_y;
  return g;
}

template <int Q>
Shifts<Q> shifts_of(const EnsembleEntry<Q>& e)
{
  Shifts<Q> s;
  for (int i = 0; i < Q; ++i) s.s[i] = e.shift[i];
  return s;
}

// PostStream stage of the local boundaries: edges, then corners, of the
// members that stepped (only < 0) or of member `only`.
template <int Q, typename S>
int launch_post(lbmh_ensemble* E, const EnsembleEntry<Q>* table, S* arena, const EnsembleLimits& limits,
                hipStream_t stream, int only)
{
  const int forced = (E->kind == LBMH_FORCED_BGK || E->kind == LBMH_FORCED_TRT) ? 1 : 0;
  if (E->n_planes) {
    hipLaunchKernelGGL((ensemble_post_kernel<Q, S, kPostPlane>), dim3(grid_for(E->n_planes, 64 / kPostLanes)), dim3(64),
                       0, stream, table, E->d_planes, E->n_planes, arena, E->d_ids, E->d_walls, E->d_info, limits, only,
                       forced);
    LBMH_HIP(hipGetLastError());
  }
  if (E->n_pplanes) {
    // members with pressure faces: every plane cell into the workspace, then stored
    using V = typename StorageTraits<S>::V;
    V* scratch = static_cast<V*>(E->d_plane_scratch);
    hipLaunchKernelGGL((ensemble_post_kernel<Q, S, kPostPlane, true>), dim3(grid_for(E->n_pplanes, 64 / kPostLanes)),
                       dim3(64), 0, stream, table, E->d_pplanes, E->n_pplanes, arena, E->d_ids, E->d_walls, E->d_info,
                       limits, only, forced, scratch);
    LBMH_HIP(hipGetLastError());
    hipLaunchKernelGGL((ensemble_plane_store_kernel<Q, S>), dim3(grid_for(E->n_pplanes, 64)), dim3(64), 0, stream,
                       table, E->d_pplanes, E->n_pplanes, arena, static_cast<const V*>(scratch), limits, only);
    LBMH_HIP(hipGetLastError());
  }
  if (E->n_edges) {
    hipLaunchKernelGGL((ensemble_post_kernel<Q, S, kPostEdge>), dim3(grid_for(E->n_edges, 64 / kPostLanes)), dim3(64), 0,
                       stream, table,
                       E->d_edges, E->n_edges, arena, E->d_ids, E->d_walls, E->d_info, limits, only, forced);
    LBMH_HIP(hipGetLastError());
  }
  if (E->n_corners) {
    hipLaunchKernelGGL((ensemble_post_kernel<Q, S, kPostCorner>), dim3(grid_for(E->n_corners, 64 / kPostLanes)), dim3(64), 0,
                       stream, table, E->d_corners, E->n_corners, arena, E->d_ids, E->d_walls, E->d_info, limits, only,
                       forced);
    LBMH_HIP(hipGetLastError());
  }
  return LBMH_OK;
}

int launch_steps(lbmh_ensemble* E, int64_t n, hipStream_t stream)
{
  // With local-boundary members the step kernel collides their flat faces
  // too, in divergent waves that need more occupancy than the FP64/FP32 cap
  // gives: 24 KB instead (FP32, 96 local-velocity cavities: 458 us per step
  // against 509 us; velocity-inlet ducts 266 against 269 us).  FP16 has no cap.
  const bool relax = E->has_local && !E->launch_chosen && E->storage != LBMH_F16;
  const unsigned lds = relax ? 24576u : unsigned(E->launch.lds_bytes);
  const uint32_t group = uint32_t(E->launch.xcd_group);
  const uint32_t adv_grid = grid_for(uint32_t(E->max_members), 256);
  return with_types(E, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    auto* table = static_cast<EnsembleEntry<Q>*>(E->d_table);
    S* arena = static_cast<S*>(E->d_pop);
    const EnsembleLimits limits{E->pop_arena.capacity() / sizeof(S), E->id_arena.capacity(),
                                uint32_t(E->max_members)};
    const bool local = E->has_local;
    for (int64_t k = 0; k < n; ++k) {
      if (E->total_chunks) {
#define LBMH_ENSEMBLE_STEP(KIND_)                                                                               \
  hipLaunchKernelGGL((ensemble_step_kernel<Q, S, KIND_>), dim3(E->total_chunks), dim3(kBlock), lds, stream, table, \
                     E->d_chunk, arena, E->d_ids, E->d_walls, E->d_info, group, limits)
#define LBMH_ENSEMBLE_STEP_LOCAL(KIND_)                                                                         \
  hipLaunchKernelGGL((ensemble_step_kernel<Q, S, KIND_, kLocalFlat>), dim3(E->total_chunks), dim3(kBlock), lds,    \
                     stream, table, E->d_chunk, arena, E->d_ids, E->d_walls, E->d_info, group, limits)
        if (local) {
          // the boundary cells collide with their BGK mixin whatever the bulk kind
          switch (E->kind) {
          case LBMH_BGK: LBMH_ENSEMBLE_STEP_LOCAL(kBGK); break;
          case LBMH_FORCED_BGK: LBMH_ENSEMBLE_STEP_LOCAL(kForcedBGK); break;
          case LBMH_TRT: LBMH_ENSEMBLE_STEP_LOCAL(kTRT); break;
          case LBMH_FORCED_TRT: LBMH_ENSEMBLE_STEP_LOCAL(kForcedTRT); break;
          default: LBMH_ENSEMBLE_STEP_LOCAL(kSmagorinsky); break;
          }
        } else {
          switch (E->kind) {
          case LBMH_BGK: LBMH_ENSEMBLE_STEP(kBGK); break;
          case LBMH_FORCED_BGK: LBMH_ENSEMBLE_STEP(kForcedBGK); break;
          case LBMH_TRT: LBMH_ENSEMBLE_STEP(kTRT); break;
          case LBMH_FORCED_TRT: LBMH_ENSEMBLE_STEP(kForcedTRT); break;
          default: LBMH_ENSEMBLE_STEP(kSmagorinsky); break;
          }
        }
#undef LBMH_ENSEMBLE_STEP_LOCAL
#undef LBMH_ENSEMBLE_STEP
        LBMH_HIP(hipGetLastError());
      }
      if (local && E->n_local) {
#define LBMH_ENSEMBLE_LOCAL(KIND_)                                                                              \
  hipLaunchKernelGGL((ensemble_local_kernel<Q, S, KIND_>), dim3(grid_for(E->n_local, 64)), dim3(64), 0, stream,      \
                     table, E->d_local, E->n_local, arena, E->d_ids, E->d_walls, E->d_info, limits)
        switch (E->kind) {
        case LBMH_BGK: LBMH_ENSEMBLE_LOCAL(kBGK); break;
        case LBMH_FORCED_BGK: LBMH_ENSEMBLE_LOCAL(kForcedBGK); break;
        case LBMH_TRT: LBMH_ENSEMBLE_LOCAL(kTRT); break;
        case LBMH_FORCED_TRT: LBMH_ENSEMBLE_LOCAL(kForcedTRT); break;
        default: LBMH_ENSEMBLE_LOCAL(kSmagorinsky); break;
        }
#undef LBMH_ENSEMBLE_LOCAL
        LBMH_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL((ensemble_advance_kernel<Q>), dim3(adv_grid), dim3(256), 0, stream, table,
                         uint32_t(E->max_members));
      LBMH_HIP(hipGetLastError());
      const int rc = launch_post(E, table, arena, limits, stream, -1);
      if (rc != LBMH_OK) return rc;
    }
    return LBMH_OK;
  });
}

}  // namespace

extern "C" {

int lbmh_ensemble_create(const lbmh_ensemble_desc* d, lbmh_ensemble** out)
{
  if (!d || !out) return fail(LBMH_E_INVALID, "null argument");
  *out = nullptr;
  if (d->q != 19 && d->q != 27) return fail(LBMH_E_INVALID, "q must be 19 or 27");
  if (d->storage < LBMH_F64 || d->storage > LBMH_F16) return fail(LBMH_E_INVALID, "unknown storage type");
  if (d->bulk_kind < LBMH_BGK || d->bulk_kind > LBMH_SMAGORINSKY) return fail(LBMH_E_INVALID, "unknown bulk kind");
  if (d->max_members < 1 || d->max_members > (1 << 20)) return fail(LBMH_E_INVALID, "max_members out of range");
  if (d->population_bytes < kAlign || d->cell_capacity < 1) return fail(LBMH_E_INVALID, "arena sizes must be positive");
  auto* E = new (std::nothrow) lbmh_ensemble();
  if (!E) return fail(LBMH_E_MEMORY, "host allocation failed");
  E->device = d->device;
  E->q = d->q;
  E->storage = d->storage;
  E->kind = d->bulk_kind;
  E->max_members = d->max_members;
  // the device indexes table[m] with this exact stride
  E->entry_bytes = E->q == 19 ? sizeof(EnsembleEntry<19>) : sizeof(EnsembleEntry<27>);
  E->members.resize(size_t(E->max_members));
  E->h_table.assign(E->entry_bytes * size_t(E->max_members), 0);
  E->pop_arena = Arena(align_up(d->population_bytes));
  E->id_arena = Arena(align_up(d->cell_capacity) + kAlign * size_t(E->max_members));
  E->chunk_capacity = uint32_t(d->cell_capacity / kBlock + size_t(E->max_members) + 1);
  // the single-lattice defaults (an ensemble sweep of 256 cavities agrees within 1%)
  E->launch = d->storage == LBMH_F16   ? lbmh_launch_config{0, 8}
              : d->storage == LBMH_F64 ? lbmh_launch_config{49152, 8}
                                       : lbmh_launch_config{40960, 4};
  DeviceGuard guard(E->device);
  hipError_t e = hipMalloc(&E->d_pop, E->pop_arena.capacity());
  if (e == hipSuccess) e = hipMalloc((void**)&E->d_ids, E->id_arena.capacity());
  if (e == hipSuccess) e = hipMalloc((void**)&E->d_walls, sizeof(double) * 4 * kEnsembleWalls * E->max_members);
  if (e == hipSuccess) e = hipMalloc(&E->d_table, E->h_table.size());
  if (e == hipSuccess) e = hipMalloc((void**)&E->d_chunk, sizeof(uint32_t) * E->chunk_capacity);
  E->boundary_capacity = d->boundary_capacity ? d->boundary_capacity : d->cell_capacity / 4 + 64;
  if (e == hipSuccess) e = hipMalloc((void**)&E->d_info, sizeof(uint32_t) * kEnsembleWalls * E->max_members);
  if (e == hipSuccess) e = hipMemset(E->d_info, 0, sizeof(uint32_t) * kEnsembleWalls * E->max_members);
  for (uint2** list : {&E->d_local, &E->d_edges, &E->d_corners, &E->d_planes, &E->d_pplanes})
    if (e == hipSuccess) e = hipMalloc((void**)list, sizeof(uint2) * E->boundary_capacity);
  if (e == hipSuccess) e = hipMemcpy(E->d_table, E->h_table.data(), E->h_table.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&E->capture_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    const int rc = hip_fail(e, "lbmh_ensemble_create");
    lbmh_ensemble_destroy(E);
    return e == hipErrorOutOfMemory ? fail(LBMH_E_MEMORY, g_error) : rc;
  }
  *out = E;
  return LBMH_OK;
}

int lbmh_ensemble_destroy(lbmh_ensemble* E)
{
  if (!E) return LBMH_OK;
  DeviceGuard guard(E->device);
  (void)hipDeviceSynchronize();
  if (E->graph) (void)hipGraphExecDestroy(E->graph);
  if (E->capture_stream) (void)hipStreamDestroy(E->capture_stream);
  if (E->d_pop) (void)hipFree(E->d_pop);
  if (E->d_ids) (void)hipFree(E->d_ids);
  if (E->d_walls) (void)hipFree(E->d_walls);
  if (E->d_table) (void)hipFree(E->d_table);
  if (E->d_chunk) (void)hipFree(E->d_chunk);
  if (E->d_info) (void)hipFree(E->d_info);
  if (E->d_local) (void)hipFree(E->d_local);
  if (E->d_edges) (void)hipFree(E->d_edges);
  if (E->d_corners) (void)hipFree(E->d_corners);
  if (E->d_planes) (void)hipFree(E->d_planes);
  if (E->d_pplanes) (void)hipFree(E->d_pplanes);
  if (E->d_plane_scratch) (void)hipFree(E->d_plane_scratch);
  delete E;
  return LBMH_OK;
}

int lbmh_ensemble_member_bytes(const lbmh_ensemble* E, const int extent[3], size_t* pop_bytes, size_t* id_bytes)
{
  if (!E || !extent) return fail(LBMH_E_INVALID, "null argument");
  for (int d = 0; d < 3; ++d)
    if (extent[d] < 1) return fail(LBMH_E_INVALID, "extent must be >= 1");
  const uint64_t npad = uint64_t(extent[0] + 2) * (extent[1] + 2) * (extent[2] + 2);
  const uint64_t elem = storage_size(E->storage);
  const uint64_t stride = align_up(npad * elem) / elem;
  if (pop_bytes) *pop_bytes = size_t(align_up(stride * elem * E->q));
  if (id_bytes) *id_bytes = size_t(align_up(uint64_t(extent[0]) * extent[1] * extent[2]));
  return LBMH_OK;
}

int lbmh_ensemble_add(lbmh_ensemble* E, const int extent[3], const int periodic[3], const uint8_t* ids,
                      const lbmh_bulk_params* params, int64_t steps, int* member)
{
  if (!E || !extent || !periodic || !ids || !params || !member) return fail(LBMH_E_INVALID, "null argument");
  *member = -1;
  for (int d = 0; d < 3; ++d)
    if (extent[d] < 1) return fail(LBMH_E_INVALID, "extent must be >= 1");
  const uint64_t sx = uint64_t(extent[0]) + 2, sy = uint64_t(extent[1]) + 2, sz = uint64_t(extent[2]) + 2;
  const uint64_t npad = sx * sy * sz;
  if (npad >= (uint64_t(1) << 31)) return fail(LBMH_E_INVALID, "padded cell count must be below 2^31");
  if (npad * storage_size(E->storage) >= (uint64_t(1) << 31))
    return fail(LBMH_E_INVALID, "a population column must stay below 2 GiB (padded cells x bytes per population)");
  const uint32_t ncore = uint32_t(uint64_t(extent[0]) * extent[1] * extent[2]);
  for (uint32_t c = 0; c < ncore; ++c)
    if (ids[c] >= kEnsembleWalls) return fail(LBMH_E_INVALID, "ensemble dynamics ids must be below 64");
  if (!valid_bulk_params(E->kind, *params)) return LBMH_E_INVALID;
  int slot = -1;
  for (int m = 0; m < E->max_members; ++m)
    if (!E->members[m].active) {
      slot = m;
      break;
    }
  if (slot < 0) return fail(LBMH_E_MEMORY, "no free member slot");
  size_t pop_bytes = 0, id_bytes = 0;
  lbmh_ensemble_member_bytes(E, extent, &pop_bytes, &id_bytes);
  Member mem;
  if (!E->pop_arena.alloc(pop_bytes, &mem.pop_off)) return fail(LBMH_E_MEMORY, "population arena full");
  if (!E->id_arena.alloc(id_bytes, &mem.ids_off)) {
    E->pop_arena.release(mem.pop_off, pop_bytes);
    return fail(LBMH_E_MEMORY, "id arena full");
  }
  mem.pop_bytes = pop_bytes;
  mem.ids_bytes = id_bytes;
  mem.active = true;
  mem.ncore = ncore;
  mem.params = *params;
  for (int d = 0; d < 3; ++d) {
    mem.n[d] = extent[d];
    mem.periodic[d] = periodic[d] ? 1 : 0;
  }
  DeviceGuard guard(E->device);
  auto undo = [&](int rc) {
    E->pop_arena.release(mem.pop_off, mem.pop_bytes);
    E->id_arena.release(mem.ids_off, mem.ids_bytes);
    return rc;
  };
  int rc = pull_table(E);
  if (rc != LBMH_OK) return undo(rc);
  hipError_t e = hipMemset(static_cast<char*>(E->d_pop) + mem.pop_off, 0, pop_bytes);
  if (e == hipSuccess) e = hipMemcpy(E->d_ids + mem.ids_off, ids, ncore, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    std::vector<double> walls(4 * kEnsembleWalls, 0.0);
    for (int w = 0; w < kEnsembleWalls; ++w) walls[4 * w] = 1.0;
    e = hipMemcpy(E->d_walls + size_t(slot) * 4 * kEnsembleWalls, walls.data(), walls.size() * sizeof(double),
                  hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMemcpy(E->d_info + size_t(slot) * kEnsembleWalls, mem.info, sizeof(mem.info),
                                     hipMemcpyHostToDevice);
  mem.h_ids.assign(ids, ids + ncore);
  if (e != hipSuccess) return undo(hip_fail(e, "lbmh_ensemble_add"));
  const uint64_t elem = storage_size(E->storage);
  with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    EnsembleEntry<Q>& en = entry<Q>(E, slot);
    std::memset(&en, 0, sizeof(en));
    en.pop_offset = mem.pop_off / elem;
    en.stride = align_up(npad * elem) / elem;
    en.ids_offset = mem.ids_off;
    en.n_padded = uint32_t(npad);
    en.ncore = ncore;
    en.nx = uint32_t(extent[0]);
    en.ny = uint32_t(extent[1]);
    en.nz = uint32_t(extent[2]);
    en.sx = uint32_t(sy * sz);
    en.sy = uint32_t(sz);
    en.div_z = FastDiv(uint32_t(extent[2]));
    en.div_y = FastDiv(uint32_t(extent[1]));
    en.periodic = uint32_t(mem.periodic[0] | (mem.periodic[1] << 1) | (mem.periodic[2] << 2));
    en.steps_left = steps;
    en.steps_done = 0;
    en.bp = derive_bulk_params<double>(E->q, *params);
    for (int i = 0; i < Q; ++i) {
      const int64_t dist = neighbor_distance<Q>(i, int64_t(sy * sz), int64_t(sz));
      int64_t dm = dist % int64_t(npad);
      if (dm < 0) dm += int64_t(npad);
      en.dmod[i] = uint32_t(dm);
      en.shift[i] = 0;
    }
    return LBMH_OK;
  });
  E->members[slot] = mem;
  rc = push_table(E);
  if (rc != LBMH_OK) {
    E->members[slot].active = false;
    return undo(rc);
  }
  *member = slot;
  return LBMH_OK;
}

int lbmh_ensemble_remove(lbmh_ensemble* E, int m)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  Member& mem = E->members[m];
  E->pop_arena.release(mem.pop_off, mem.pop_bytes);
  E->id_arena.release(mem.ids_off, mem.ids_bytes);
  mem = Member();
  with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    std::memset(&entry<Q>(E, m), 0, sizeof(EnsembleEntry<Q>));
    return LBMH_OK;
  });
  rc = push_table(E);
  if (rc != LBMH_OK) return rc;
  return rebuild_lists(E);
}

int lbmh_ensemble_set_wall(lbmh_ensemble* E, int m, int id, double rho, const double u[3])
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!u) return fail(LBMH_E_INVALID, "null velocity");
  if (id < LBMH_ID_FIRST_WALL || id >= kEnsembleWalls) return fail(LBMH_E_INVALID, "ensemble wall id must be in [3, 63]");
  DeviceGuard guard(E->device);
  const double w[4] = {rho, u[0], u[1], u[2]};
  LBMH_HIP(hipDeviceSynchronize());
  LBMH_HIP(hipMemcpy(E->d_walls + (size_t(m) * kEnsembleWalls + id) * 4, w, sizeof(w), hipMemcpyHostToDevice));
  Member& mem = E->members[m];
  mem.id_defined[id] = true;
  mem.info[id] = pack_local_info(kMovingWall, 0, 0, 0);
  LBMH_HIP(hipMemcpy(E->d_info + size_t(m) * kEnsembleWalls + id, &mem.info[id], sizeof(uint32_t), hipMemcpyHostToDevice));
  return LBMH_OK;
}

int lbmh_ensemble_set_local_boundary(lbmh_ensemble* E, int m, int id, int kind, double rho, const double u[3])
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!u) return fail(LBMH_E_INVALID, "null velocity");
  if (kind < LBMH_LOCAL_VELOCITY || kind > LBMH_INTERP_PRESSURE) return fail(LBMH_E_INVALID, "unknown local boundary kind");
  if (id < 2 || id >= kEnsembleWalls) return fail(LBMH_E_INVALID, "ensemble boundary id must be in [2, 63]");
  if ((kind == LBMH_LOCAL_PRESSURE || kind == LBMH_INTERP_PRESSURE) && !(rho > 0))
    return fail(LBMH_E_INVALID, "pressure boundary density must be positive");
  Member& mem = E->members[m];
  const int n[3] = {mem.n[0], mem.n[1], mem.n[2]};
  bool used[kEnsembleWalls] = {};
  for (uint8_t v : mem.h_ids) used[v] = true;
  for (int w = 0; w < kEnsembleWalls; ++w) used[w] = used[w] || mem.id_defined[w];
  used[0] = used[1] = used[2] = true;
  std::map<std::tuple<int, int, int, int>, int> entry_ids;
  std::vector<uint8_t> next = mem.h_ids;
  std::vector<uint32_t> cells, edges, corners, planes;
  uint32_t info[kEnsembleWalls];
  double walls[kEnsembleWalls][4];
  std::memcpy(info, mem.info, sizeof(info));
  int free_id = kEnsembleWalls - 1;
  for (int x = 0; x < n[0]; ++x)
    for (int y = 0; y < n[1]; ++y)
      for (int z = 0; z < n[2]; ++z) {
        const uint32_t c = uint32_t((int64_t(x) * n[1] + y) * n[2] + z);
        if (mem.h_ids[c] != id) continue;
        int type, nrm[3];
        classify_boundary_cell(n, mem.periodic, mem.h_ids.data(), x, y, z, type, nrm);
        const bool flat = type == kTypeFlat && (nrm[0] || nrm[1] || nrm[2]);
        const uint32_t k = local_kind(kind, type, flat);
        if (k == kInvalidKind) {
          return fail(LBMH_E_INVALID, kind == LBMH_LOCAL_PRESSURE ? "local pressure boundary on a non-flat wall cell"
                                                                 : "interpolated boundary on a wall cell without normal");
        }
        int a = nrm[0], b = nrm[1], cc = nrm[2];
        if (k == kOuterEdge || k == kInnerEdge) {
          const int plane = nrm[0] == 0 ? 0 : (nrm[1] == 0 ? 1 : 2);
          a = plane;
          b = nrm[(plane + 1) % 3];
          cc = nrm[(plane + 2) % 3];
        }
        const auto key = std::make_tuple(int(k), a, b, cc);
        auto it = entry_ids.find(key);
        if (it == entry_ids.end()) {
          while (free_id >= 3 && used[free_id]) --free_id;
          if (free_id < 3) return fail(LBMH_E_INVALID, "no free ensemble dynamics ids left (ids 3..63)");
          used[free_id] = true;
          info[free_id] = pack_local_info(k, a, b, cc);
          walls[free_id][0] = rho;
          for (int d = 0; d < 3; ++d) walls[free_id][d + 1] = u[d];
          it = entry_ids.emplace(key, free_id).first;
        }
        next[c] = uint8_t(it->second);
        if (k == kInnerEdge || k == kInnerCorner || k == kDeadWall) cells.push_back(c);
        if (k == kFdVelocity || k == kFdPressure) planes.push_back(c);
        if (k == kOuterEdge) edges.push_back(c);
        if (k == kOuterCorner) corners.push_back(c);
      }
  if (entry_ids.empty()) return LBMH_OK;
  DeviceGuard guard(E->device);
  LBMH_HIP(hipDeviceSynchronize());
  if (kind == LBMH_INTERP_PRESSURE && !planes.empty() && !E->d_plane_scratch) {
    const size_t elem = E->storage == LBMH_F64 ? sizeof(double) : sizeof(float);
    LBMH_HIP(hipMalloc(&E->d_plane_scratch, elem * size_t(E->q) * E->boundary_capacity));
  }
  if (kind == LBMH_INTERP_PRESSURE && !planes.empty()) mem.pressure_planes = true;
  for (const auto& kv : entry_ids) {
    const int w = kv.second;
    mem.id_defined[w] = true;
    LBMH_HIP(hipMemcpy(E->d_walls + (size_t(m) * kEnsembleWalls + w) * 4, walls[w], sizeof(walls[w]),
                       hipMemcpyHostToDevice));
  }
  std::memcpy(mem.info, info, sizeof(info));
  mem.h_ids.swap(next);
  mem.local.insert(mem.local.end(), cells.begin(), cells.end());
  mem.edges.insert(mem.edges.end(), edges.begin(), edges.end());
  mem.corners.insert(mem.corners.end(), corners.begin(), corners.end());
  mem.planes.insert(mem.planes.end(), planes.begin(), planes.end());
  LBMH_HIP(hipMemcpy(E->d_info + size_t(m) * kEnsembleWalls, mem.info, sizeof(mem.info), hipMemcpyHostToDevice));
  LBMH_HIP(hipMemcpy(E->d_ids + mem.ids_off, mem.h_ids.data(), mem.ncore, hipMemcpyHostToDevice));
  E->has_local = true;
  return rebuild_lists(E);
}

int lbmh_ensemble_initialize(lbmh_ensemble* E, int m)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!E->n_edges && !E->n_corners && !E->n_planes && !E->n_pplanes) return LBMH_OK;
  DeviceGuard guard(E->device);
  rc = with_types(E, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const EnsembleLimits limits{E->pop_arena.capacity() / sizeof(S), E->id_arena.capacity(),
                                uint32_t(E->max_members)};
    return launch_post<Q, S>(E, static_cast<EnsembleEntry<Q>*>(E->d_table), static_cast<S*>(E->d_pop), limits,
                             nullptr, m);
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_ensemble_set_steps(lbmh_ensemble* E, int m, int64_t steps)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  return with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    EnsembleEntry<Q>& en = entry<Q>(E, m);
    en.steps_left = steps;
    char* dst = static_cast<char*>(E->d_table) + size_t(m) * E->entry_bytes + offsetof(EnsembleEntry<Q>, steps_left);
    LBMH_HIP(hipMemcpy(dst, &en.steps_left, sizeof(int64_t), hipMemcpyHostToDevice));
    return LBMH_OK;
  });
}

int lbmh_ensemble_status(lbmh_ensemble* E, int m, int64_t* done, int64_t* left)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  return with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    const EnsembleEntry<Q>& en = entry<Q>(E, m);
    if (done) *done = en.steps_done;
    if (left) *left = en.steps_left;
    return LBMH_OK;
  });
}

int lbmh_ensemble_status_all(lbmh_ensemble* E, int64_t* done, int64_t* left)
{
  if (!E || !done || !left) return fail(LBMH_E_INVALID, "null argument");
  DeviceGuard guard(E->device);
  const int rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  return with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    for (int m = 0; m < E->max_members; ++m) {
      const EnsembleEntry<Q>& en = entry<Q>(E, m);
      const bool on = E->members[m].active;
      done[m] = on ? en.steps_done : 0;
      left[m] = on ? en.steps_left : 0;
    }
    return LBMH_OK;
  });
}

int lbmh_ensemble_ini_equilibrium(lbmh_ensemble* E, int m, const double* rho_u)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!rho_u) return fail(LBMH_E_INVALID, "null moments");
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  const Member& mem = E->members[m];
  DeviceBuffer staging;
  const size_t bytes = size_t(4) * mem.ncore * sizeof(double);
  LBMH_HIP(hipMalloc(&staging.p, bytes));
  LBMH_HIP(hipMemcpy(staging.p, rho_u, bytes, hipMemcpyHostToDevice));
  const int forced = (E->kind == LBMH_FORCED_BGK || E->kind == LBMH_FORCED_TRT) ? 1 : 0;
  rc = with_types(E, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const EnsembleEntry<Q>& en = entry<Q>(E, m);
    S* pop = static_cast<S*>(E->d_pop) + en.pop_offset;
    hipLaunchKernelGGL((ini_equilibrium_kernel<Q, S>), dim3(grid_for(mem.ncore, 256)), dim3(256), 0, 0,
                       (const double*)staging.p, E->d_ids + mem.ids_off, pop, geom_of(en), shifts_of(en),
                       mem.params.force[0] * 0.5, mem.params.force[1] * 0.5, mem.params.force[2] * 0.5, forced,
                       nullptr);
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_ensemble_upload(lbmh_ensemble* E, int m, const void* host, int dtype)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!host || dtype < LBMH_F64 || dtype > LBMH_F16) return fail(LBMH_E_INVALID, "bad host buffer or dtype");
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  const Member& mem = E->members[m];
  DeviceBuffer staging;
  const size_t bytes = size_t(E->q) * mem.ncore * storage_size(dtype);
  LBMH_HIP(hipMalloc(&staging.p, bytes));
  LBMH_HIP(hipMemcpy(staging.p, host, bytes, hipMemcpyHostToDevice));
  rc = with_types(E, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const EnsembleEntry<Q>& en = entry<Q>(E, m);
    S* pop = static_cast<S*>(E->d_pop) + en.pop_offset;
    const dim3 grid(grid_for(mem.ncore, 256));
    if (dtype == LBMH_F64)
      hipLaunchKernelGGL((scatter_populations<Q, double, S>), grid, dim3(256), 0, 0, (const double*)staging.p, pop, geom_of(en), shifts_of(en));
    else if (dtype == LBMH_F32)
      hipLaunchKernelGGL((scatter_populations<Q, float, S>), grid, dim3(256), 0, 0, (const float*)staging.p, pop, geom_of(en), shifts_of(en));
    else
      hipLaunchKernelGGL((scatter_populations<Q, _Float16, S>), grid, dim3(256), 0, 0, (const _Float16*)staging.p, pop, geom_of(en), shifts_of(en));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_ensemble_download(lbmh_ensemble* E, int m, void* host, int dtype)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!host || dtype < LBMH_F64 || dtype > LBMH_F16) return fail(LBMH_E_INVALID, "bad host buffer or dtype");
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  const Member& mem = E->members[m];
  DeviceBuffer staging;
  const size_t bytes = size_t(E->q) * mem.ncore * storage_size(dtype);
  LBMH_HIP(hipMalloc(&staging.p, bytes));
  rc = with_types(E, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const EnsembleEntry<Q>& en = entry<Q>(E, m);
    const S* pop = static_cast<const S*>(E->d_pop) + en.pop_offset;
    const dim3 grid(grid_for(mem.ncore, 256));
    if (dtype == LBMH_F64)
      hipLaunchKernelGGL((gather_populations<Q, double, S>), grid, dim3(256), 0, 0, (double*)staging.p, pop, geom_of(en), shifts_of(en));
    else if (dtype == LBMH_F32)
      hipLaunchKernelGGL((gather_populations<Q, float, S>), grid, dim3(256), 0, 0, (float*)staging.p, pop, geom_of(en), shifts_of(en));
    else
      hipLaunchKernelGGL((gather_populations<Q, _Float16, S>), grid, dim3(256), 0, 0, (_Float16*)staging.p, pop, geom_of(en), shifts_of(en));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipMemcpy(host, staging.p, bytes, hipMemcpyDeviceToHost));
  return LBMH_OK;
}

namespace {

// The member as one block for the checkpoint code; its wall table comes from
// the device into `walls`.  Call after pull_table.
int member_block(lbmh_ensemble* E, int m, CheckpointBlock& b, std::vector<double>& walls)
{
  const Member& mem = E->members[m];
  walls.resize(size_t(4) * kEnsembleWalls);
  LBMH_HIP(hipMemcpy(walls.data(), E->d_walls + size_t(m) * 4 * kEnsembleWalls, walls.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
  b.q = E->q;
  b.storage = E->storage;
  b.kind = E->kind;
  for (int d = 0; d < 3; ++d) {
    b.n[d] = mem.n[d];
    b.periodic[d] = mem.periodic[d];
    b.force[d] = mem.params.force[d];
  }
  b.ids = mem.h_ids.data();
  b.walls = walls.data();
  b.info = mem.info;
  b.n_ids = kEnsembleWalls;
  return with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    const EnsembleEntry<Q>& en = entry<Q>(E, m);
    b.n_padded = en.n_padded;
    b.stride = en.stride;
    b.d_pop = static_cast<char*>(E->d_pop) + en.pop_offset * storage_size(E->storage);
    for (int i = 0; i < Q; ++i) {
      b.shift[i] = en.shift[i];
      b.dist[i] = neighbor_distance<Q>(i, int64_t(en.sx), int64_t(en.sy));
    }
    return LBMH_OK;
  });
}

}  // namespace

int lbmh_ensemble_checkpoint_size(lbmh_ensemble* E, int m, int real, size_t* bytes)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!bytes) return fail(LBMH_E_INVALID, "null argument");
  if (real != LBMH_F64 && real != LBMH_F32) return fail(LBMH_E_INVALID, "checkpoint value type must be F64 or F32");
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  CheckpointBlock b;
  std::vector<double> walls;
  rc = member_block(E, m, b, walls);
  if (rc != LBMH_OK) return rc;
  *bytes = checkpoint_bytes(b, real);
  return LBMH_OK;
}

int lbmh_ensemble_checkpoint_save(lbmh_ensemble* E, int m, int real, void* buffer, size_t bytes)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  CheckpointBlock b;
  std::vector<double> walls;
  rc = member_block(E, m, b, walls);
  if (rc != LBMH_OK) return rc;
  return checkpoint_save(b, real, buffer, bytes);
}

int lbmh_ensemble_checkpoint_load(lbmh_ensemble* E, int m, int real, const void* buffer, size_t bytes)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  CheckpointBlock b;
  std::vector<double> walls;
  rc = member_block(E, m, b, walls);
  if (rc != LBMH_OK) return rc;
  int64_t shift[27];
  rc = checkpoint_load(b, real, buffer, bytes, shift);
  if (rc != LBMH_OK) return rc;
  // only the member's shifts change in the table
  return with_q(E, [&](auto qc) -> int {
    constexpr int Q = decltype(qc)::value;
    EnsembleEntry<Q>& en = entry<Q>(E, m);
    for (int i = 0; i < Q; ++i) en.shift[i] = uint32_t(shift[i]);
    char* dst = static_cast<char*>(E->d_table) + size_t(m) * E->entry_bytes + offsetof(EnsembleEntry<Q>, shift);
    LBMH_HIP(hipMemcpy(dst, en.shift, sizeof(en.shift), hipMemcpyHostToDevice));
    return LBMH_OK;
  });
}

int lbmh_ensemble_field_export(lbmh_ensemble* E, int m, int field, int overlap, const double conversion[3],
                               float* host, size_t floats)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!host) return fail(LBMH_E_INVALID, "null argument");
  if (field < LBMH_FIELD_DENSITY || field > LBMH_FIELD_PHYS_PRESSURE) return fail(LBMH_E_INVALID, "unknown field");
  if (overlap < 0 || overlap > 1) return fail(LBMH_E_INVALID, "overlap must be 0 or 1 (the block's padding)");
  if ((field == LBMH_FIELD_PHYS_VELOCITY || field == LBMH_FIELD_PHYS_PRESSURE) && !conversion)
    return fail(LBMH_E_INVALID, "physical fields need the unit conversion factors");
  const Member& mem = E->members[m];
  const uint64_t points = uint64_t(mem.n[0] + 2 * overlap) * (mem.n[1] + 2 * overlap) * (mem.n[2] + 2 * overlap);
  const size_t want = size_t(points) * field_components(field);
  if (floats != want)
    return fail(LBMH_E_INVALID, "field buffer holds " + std::to_string(floats) + " floats, the export " +
                                    std::to_string(want));
  if (points > 0xffffffffull || !field_export_grid_fits(field_export_blocks(mem.n[0], mem.n[1], mem.n[2], overlap)))
    return fail(LBMH_E_INVALID, "member too large for one export (its 32z x 8x tiles exceed 2^32 work-items)");
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  DeviceBuffer out;
  LBMH_HIP(hipMalloc(&out.p, want * sizeof(float)));
  FieldConversion conv{0.0, 0.0, 0.0};
  if (conversion) conv = FieldConversion{conversion[0], conversion[1], conversion[2]};
  const int forced = (E->kind == LBMH_FORCED_BGK || E->kind == LBMH_FORCED_TRT) ? 1 : 0;
  const double* F = mem.params.force;
  rc = with_types(E, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const EnsembleEntry<Q>& en = entry<Q>(E, m);
    PostArgs<Q> a;
    a.pop = static_cast<S*>(E->d_pop) + en.pop_offset;
    a.stride = en.stride;
    a.n_padded = en.n_padded;
    for (int i = 0; i < Q; ++i) a.shift[i] = en.shift[i];
    a.ids = E->d_ids + en.ids_offset;
    a.walls = E->d_walls + size_t(m) * kEnsembleWalls * 4;
    a.info = E->d_info + size_t(m) * kEnsembleWalls;
    a.cells = nullptr;
    a.remote = nullptr;
    a.forced = 0;
    a.force = nullptr;
    a.count = 0;
    a.nx = int(en.nx);
    a.ny = int(en.ny);
    a.nz = int(en.nz);
    a.sx = en.sx;
    a.sy = en.sy;
    a.periodic = en.periodic;
    a.div_z = en.div_z;
    a.div_y = en.div_y;
    a.omega = en.bp.omega;
    hipLaunchKernelGGL((field_export_kernel<Q, S>), dim3(uint32_t(field_export_blocks(a.nx, a.ny, a.nz, overlap))),
                       dim3(kExportTileZ * kExportTileX), 0, 0, a,
                       overlap, field, conv, forced, F[0] * 0.5, F[1] * 0.5, F[2] * 0.5,
                       nullptr, static_cast<float*>(out.p), uint32_t(points));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipMemcpy(host, out.p, want * sizeof(float), hipMemcpyDeviceToHost));
  return LBMH_OK;
}

int lbmh_ensemble_moments(lbmh_ensemble* E, int m, double* rho_u)
{
  int rc = check_member(E, m);
  if (rc != LBMH_OK) return rc;
  if (!rho_u) return fail(LBMH_E_INVALID, "null output");
  DeviceGuard guard(E->device);
  rc = pull_table(E);
  if (rc != LBMH_OK) return rc;
  const Member& mem = E->members[m];
  DeviceBuffer out;
  const size_t bytes = size_t(4) * mem.ncore * sizeof(double);
  LBMH_HIP(hipMalloc(&out.p, bytes));
  const int forced = (E->kind == LBMH_FORCED_BGK || E->kind == LBMH_FORCED_TRT) ? 1 : 0;
  rc = with_types(E, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const EnsembleEntry<Q>& en = entry<Q>(E, m);
    const S* pop = static_cast<const S*>(E->d_pop) + en.pop_offset;
    hipLaunchKernelGGL((moments_kernel<Q, S>), dim3(grid_for(mem.ncore, 256)), dim3(256), 0, 0, (double*)out.p,
                       E->d_ids + mem.ids_off, pop, geom_of(en), shifts_of(en), mem.params.force[0] * 0.5,
                       mem.params.force[1] * 0.5, mem.params.force[2] * 0.5, forced, nullptr);
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipMemcpy(rho_u, out.p, bytes, hipMemcpyDeviceToHost));
  return LBMH_OK;
}

int lbmh_ensemble_step(lbmh_ensemble* E, int64_t n, void* stream)
{
  if (!E) return fail(LBMH_E_INVALID, "null ensemble");
  if (n < 0) return fail(LBMH_E_INVALID, "negative step count");
  DeviceGuard guard(E->device);
  return launch_steps(E, n, static_cast<hipStream_t>(stream));
}

int lbmh_ensemble_step_graph(lbmh_ensemble* E, int64_t n, void* stream)
{
  if (!E) return fail(LBMH_E_INVALID, "null ensemble");
  if (n < 0) return fail(LBMH_E_INVALID, "negative step count");
  DeviceGuard guard(E->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  constexpr int64_t kGraphSteps = 32;
  if (n >= kGraphSteps && (!E->graph || E->graph_version != E->version)) {
    if (E->graph) {
      LBMH_HIP(hipGraphExecDestroy(E->graph));
      E->graph = nullptr;
    }
    LBMH_HIP(hipStreamBeginCapture(E->capture_stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch_steps(E, kGraphSteps, E->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(E->capture_stream, &g);
    if (rc != LBMH_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    const hipError_t ei = hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
    E->graph_version = E->version;
    E->graph_steps = kGraphSteps;
  }
  int64_t done = 0;
  if (E->graph && E->graph_version == E->version) {
    for (; done + E->graph_steps <= n; done += E->graph_steps) LBMH_HIP(hipGraphLaunch(E->graph, s));
  }
  return launch_steps(E, n - done, s);
}

int lbmh_ensemble_set_launch_config(lbmh_ensemble* E, const lbmh_launch_config* cfg)
{
  if (!E || !cfg) return fail(LBMH_E_INVALID, "null argument");
  if (cfg->lds_bytes < 0 || cfg->lds_bytes > 65536) return fail(LBMH_E_INVALID, "lds_bytes must be in [0, 65536]");
  if (cfg->xcd_group < 0 || cfg->xcd_group > 64) return fail(LBMH_E_INVALID, "xcd_group must be in [0, 64]");
  E->launch = *cfg;
  E->launch_chosen = true;
  E->version += 1;
  return LBMH_OK;
}

int lbmh_ensemble_autotune(lbmh_ensemble* E, int trial_steps, lbmh_launch_config* chosen)
{
  if (!E) return fail(LBMH_E_INVALID, "null ensemble");
  if (trial_steps < 1) return fail(LBMH_E_INVALID, "trial_steps must be >= 1");
  DeviceGuard guard(E->device);
  // The trials run real steps: the members' populations and the device table
  // (shifts, step counters) are copied aside first and put back after each
  // candidate, so every candidate times the same members and the ensemble
  // ends as it began.
  uint64_t end = 0;
  for (const Member& m : E->members)
    if (m.active) end = std::max(end, m.pop_off + m.pop_bytes);
  struct Scratch {
    void* pop = nullptr;
    void* table = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scratch()
    {
      if (pop) (void)hipFree(pop);
      if (table) (void)hipFree(table);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } sc;
  const size_t table_bytes = E->h_table.size();
  LBMH_HIP(hipDeviceSynchronize());
  if (end) {
    const hipError_t e = hipMalloc(&sc.pop, end);
    if (e != hipSuccess) return fail(LBMH_E_MEMORY, "autotune needs device memory for a copy of the members' populations");
    LBMH_HIP(hipMemcpy(sc.pop, E->d_pop, end, hipMemcpyDeviceToDevice));
  }
  LBMH_HIP(hipMalloc(&sc.table, table_bytes));
  LBMH_HIP(hipMemcpy(sc.table, E->d_table, table_bytes, hipMemcpyDeviceToDevice));
  LBMH_HIP(hipEventCreate(&sc.e0));
  LBMH_HIP(hipEventCreate(&sc.e1));
  auto restore = [&]() -> int {
    LBMH_HIP(hipDeviceSynchronize());
    if (end) LBMH_HIP(hipMemcpy(E->d_pop, sc.pop, end, hipMemcpyDeviceToDevice));
    LBMH_HIP(hipMemcpy(E->d_table, sc.table, table_bytes, hipMemcpyDeviceToDevice));
    return LBMH_OK;
  };
  static const int kLds[] = {0, 16384, 24576, 32768, 40960, 49152};
  static const int kGroup[] = {4, 8};
  const lbmh_launch_config before = E->launch;
  const bool before_chosen = E->launch_chosen;
  lbmh_launch_config best = before;
  float best_ms = 0.f;
  bool have = false;
  int rc = LBMH_OK;
  for (int group : kGroup) {
    for (int lds : kLds) {
      E->launch = lbmh_launch_config{lds, group};
      E->launch_chosen = true;  // time exactly this shape
      rc = launch_steps(E, 1, nullptr);
      if (rc == LBMH_OK && hipEventRecord(sc.e0, nullptr) != hipSuccess) rc = fail(LBMH_E_DEVICE, "hipEventRecord");
      if (rc == LBMH_OK) rc = launch_steps(E, trial_steps, nullptr);
      if (rc == LBMH_OK && hipEventRecord(sc.e1, nullptr) != hipSuccess) rc = fail(LBMH_E_DEVICE, "hipEventRecord");
      if (rc == LBMH_OK && hipEventSynchronize(sc.e1) != hipSuccess) rc = fail(LBMH_E_DEVICE, "hipEventSynchronize");
      float ms = 0.f;
      if (rc == LBMH_OK && hipEventElapsedTime(&ms, sc.e0, sc.e1) != hipSuccess) rc = fail(LBMH_E_DEVICE, "hipEventElapsedTime");
      const int rr = restore();
      if (rc == LBMH_OK) rc = rr;
      if (rc != LBMH_OK) {
        E->launch = before;
        E->launch_chosen = before_chosen;
        E->version += 1;
        return rc;
      }
      if (!have || ms < best_ms) {
        best_ms = ms;
        best = E->launch;
        have = true;
      }
    }
  }
  E->launch = best;
  E->launch_chosen = true;
  E->version += 1;  // a captured step loop is rebuilt with the chosen shape
  if (chosen) *chosen = best;
  return LBMH_OK;
}

int lbmh_ensemble_usage(const lbmh_ensemble* E, int* members, size_t* cells, size_t* used, size_t* capacity)
{
  if (!E) return fail(LBMH_E_INVALID, "null ensemble");
  int count = 0;
  size_t ncells = 0;
  for (const Member& m : E->members)
    if (m.active) {
      ++count;
      ncells += m.ncore;
    }
  if (members) *members = count;
  if (cells) *cells = ncells;
  if (used) *used = size_t(E->pop_arena.used() + E->id_arena.used());
  if (capacity) *capacity = size_t(E->pop_arena.capacity() + E->id_arena.capacity());
  return LBMH_OK;
}

}  // extern "C"
