// Kernels of the lattice ensemble: many independent blocks of uneven extent,
// each with its own padded PS columns, dynamics ids, parameters and walls,
// advanced together by one launch per time step.
//
// The launch covers the concatenated 256-cell chunks of every active lattice.
// A chunk finds its lattice through `chunk_lattice` and reads that lattice's
// entry from the table with uniform (scalar) loads, then runs exactly the
// cell body of the single-block kernel.  The shifts and step counters live in
// the table and are advanced on the device by a one-thread-per-lattice kernel,
// so a sequence of steps needs no host involvement and can be captured in a
// HIP graph.
#pragma once

#include "lbm_kernels.hpp"

namespace lbmh {

constexpr int kEnsembleWalls = 64;  // wall-table entries per lattice (ids 0..63)

template <int Q>
struct EnsembleEntry {
  uint64_t pop_offset;      // element offset of column 0 inside the population arena
  uint64_t stride;          // elements between columns
  uint64_t ids_offset;      // byte offset of the lattice's ids inside the id arena
  uint32_t n_padded, ncore, nx, ny, nz, sx, sy;
  FastDiv div_z, div_y;
  uint32_t periodic;
  uint32_t chunk_begin;     // first chunk of this lattice in the launch
  int64_t steps_left;       // 0: idle, < 0: unbounded
  int64_t steps_done;
  uint32_t stepped;         // 1 if the last advance moved this lattice (its post-processors run)
  uint32_t pad_;
  BulkParams<double> bp;
  uint32_t shift[Q];        // PS shifts, in [0, n_padded)
  uint32_t dmod[Q];         // (c_i . projection) mod n_padded
};

// Sizes of the arenas, so that a workgroup can refuse an entry that does not fit.
struct EnsembleLimits {
  uint64_t pop_elements;    // elements in the population arena
  uint64_t id_bytes;        // bytes in the id arena
  uint32_t members;         // table slots
};

// Whole-member arguments of the single-block cell body, from a table entry.
template <int Q, typename S>
__device__ __forceinline__ StepArgs<Q, typename StorageTraits<S>::V> member_args(
    const EnsembleEntry<Q>& e, uint32_t lat, S* arena, const uint8_t* ids, const double* walls, const uint32_t* info)
{
  using V = typename StorageTraits<S>::V;
  StepArgs<Q, V> a;
  a.pop = arena + e.pop_offset;
  a.stride = e.stride;
  a.n_padded = e.n_padded;
#pragma unroll
  for (int i = 0; i < Q; ++i) a.shift[i] = e.shift[i];
  a.ids = ids + e.ids_offset;
  a.ncore = e.ncore;
  a.nx = e.nx;
  a.ny = e.ny;
  a.nz = e.nz;
  a.sx = e.sx;
  a.sy = e.sy;
  a.div_z = e.div_z;
  a.div_y = e.div_y;
  a.periodic = e.periodic;
  a.xcd_group = 0;
  a.walls = walls + uint64_t(lat) * kEnsembleWalls * 4;
  a.info = info ? info + uint64_t(lat) * kEnsembleWalls : nullptr;
  a.force = nullptr;  // members take the uniform force
  a.bp.omega = V(e.bp.omega);
  a.bp.omega2 = V(e.bp.omega2);
  a.bp.guo = V(e.bp.guo);
  for (int d = 0; d < 3; ++d) {
    a.bp.force[d] = V(e.bp.force[d]);
    a.bp.half_force[d] = V(e.bp.half_force[d]);
  }
  a.bp.sm_a = V(e.bp.sm_a);
  a.bp.sm_b = V(e.bp.sm_b);
  a.bp.sm_k = V(e.bp.sm_k);
  a.bp.tau = V(e.bp.tau);
  a.bp.half_tau = V(e.bp.half_tau);
  return a;
}

// uniform guard: every column and the id range of the entry inside the arenas
template <int Q>
__device__ __forceinline__ bool entry_fits(const EnsembleEntry<Q>& e, const EnsembleLimits& limits)
{
  return !(e.stride < e.n_padded || e.pop_offset + uint64_t(Q) * e.stride > limits.pop_elements ||
           e.ids_offset + e.ncore > limits.id_bytes);
}

// LOCAL = kLocalSkip when some member has local boundaries: their cells are
// left to ensemble_local_kernel.
template <int Q, typename S, int KIND, int LOCAL = kLocalNone>
__global__ __launch_bounds__(kBlock) void ensemble_step_kernel(const EnsembleEntry<Q>* __restrict__ table,
                                                               const uint32_t* __restrict__ chunk_lattice,
                                                               S* __restrict__ arena, const uint8_t* __restrict__ ids,
                                                               const double* __restrict__ walls,
                                                               const uint32_t* __restrict__ info, uint32_t xcd_group,
                                                               EnsembleLimits limits)
{
  uint32_t b = blockIdx.x;
  if (xcd_group) b = remap_xcd(b, gridDim.x, xcd_group);
  const uint32_t lat = chunk_lattice[b];
  if (lat >= limits.members) return;
  const EnsembleEntry<Q>& e = table[lat];
  if (e.steps_left == 0) return;
  if (!entry_fits(e, limits)) return;
  const auto a = member_args<Q, S>(e, lat, arena, ids, walls, LOCAL == kLocalNone ? nullptr : info);
  collide_stream_cell<Q, S, KIND, true, LOCAL>(a, (b - e.chunk_begin) * kBlock + threadIdx.x);
}

// The local-boundary cells of every stepping member: list entries are
// (member, core cell) pairs.
template <int Q, typename S, int KIND>
__global__ __launch_bounds__(64) void ensemble_local_kernel(const EnsembleEntry<Q>* __restrict__ table,
                                                            const uint2* __restrict__ cells, uint32_t count,
                                                            S* __restrict__ arena, const uint8_t* __restrict__ ids,
                                                            const double* __restrict__ walls,
                                                            const uint32_t* __restrict__ info, EnsembleLimits limits)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint2 lc = cells[k];
  if (lc.x >= limits.members) return;
  const EnsembleEntry<Q>& e = table[lc.x];
  if (e.steps_left == 0 || !entry_fits(e, limits) || lc.y >= e.ncore) return;
  const auto a = member_args<Q, S>(e, lc.x, arena, ids, walls, info);
  collide_stream_cell<Q, S, KIND, true, kLocalCollide>(a, lc.y);
}

// Plane, edge or corner post-processors (KIND) of the members that stepped
// last (or of member `only` when only >= 0: initialize()).  Eight lanes per
// (member, cell) entry, as outer_post_kernel.  The SCRATCH plane pass
// computes entry k's populations into scratch[k * Q ...] and
// ensemble_plane_store_kernel writes them afterwards: pressure faces read the
// velocity of their face neighbours from populations the pass rewrites
// (plane_scratch_kernel / plane_store_kernel of a lattice).  The two are
// separate instantiations, as the lattice's kernels are, so that each
// compiles the reconstruction exactly as the lattice does.
template <int Q, typename S, int KIND, bool SCRATCH = false>
__global__ __launch_bounds__(64) void ensemble_post_kernel(const EnsembleEntry<Q>* __restrict__ table,
                                                           const uint2* __restrict__ cells, uint32_t count,
                                                           S* __restrict__ arena, const uint8_t* __restrict__ ids,
                                                           const double* __restrict__ walls,
                                                           const uint32_t* __restrict__ info, EnsembleLimits limits,
                                                           int only, int forced,
                                                           typename StorageTraits<S>::V* scratch = nullptr)
{
  using V = typename StorageTraits<S>::V;
  __shared__ V nb[64 / kPostLanes][kPostSlots][4];
  const int slot = int(threadIdx.x % kPostLanes), j = int(threadIdx.x / kPostLanes);
  const uint32_t k = blockIdx.x * (64 / kPostLanes) + uint32_t(j);
  bool valid = k < count;
  uint2 lc = make_uint2(0, 0);
  PostArgs<Q> a;
  if (valid) {
    lc = cells[k];
    valid = lc.x < limits.members;
  }
  if (valid) {
    const EnsembleEntry<Q>& e = table[lc.x];
    valid = (only >= 0 ? int(lc.x) == only : e.stepped != 0) && entry_fits(e, limits) && lc.y < e.ncore;
    if (valid) {
      a.pop = arena + e.pop_offset;
      a.stride = e.stride;
      a.n_padded = e.n_padded;
#pragma unroll
      for (int i = 0; i < Q; ++i) a.shift[i] = e.shift[i];
      a.ids = ids + e.ids_offset;
      a.walls = walls + uint64_t(lc.x) * kEnsembleWalls * 4;
      a.info = info + uint64_t(lc.x) * kEnsembleWalls;
      a.cells = nullptr;
      a.recs = nullptr;
      a.remote = nullptr;
      // a forced bulk neighbour reports u + F/2 (the member's uniform FORCE)
      a.forced = uint32_t(forced);
      a.force = nullptr;
      a.half_force[0] = e.bp.half_force[0];
      a.half_force[1] = e.bp.half_force[1];
      a.half_force[2] = e.bp.half_force[2];
      a.count = 0;
      a.nx = int(e.nx);
      a.ny = int(e.ny);
      a.nz = int(e.nz);
      a.sx = e.sx;
      a.sy = e.sy;
      a.periodic = e.periodic;
      a.div_z = e.div_z;
      a.div_y = e.div_y;
      a.omega = e.bp.omega;
    }
  }
  if (valid && slot < kPostSlots) post_gather<Q, S, KIND>(a, lc.y, slot, nb[j][slot]);
  __syncthreads();
  if (valid && slot == 0) {
    if constexpr (SCRATCH)
      post_finish<Q, S, KIND>(a, lc.y, int(a.ids[lc.y]), nb[j], scratch + uint64_t(k) * Q);
    else
      post_finish<Q, S, KIND>(a, lc.y, int(a.ids[lc.y]), nb[j]);
  }
}

// Second half of a two-phase plane pass: entry k's populations from scratch.
template <int Q, typename S>
__global__ __launch_bounds__(64) void ensemble_plane_store_kernel(const EnsembleEntry<Q>* __restrict__ table,
                                                                  const uint2* __restrict__ cells, uint32_t count,
                                                                  S* __restrict__ arena,
                                                                  const typename StorageTraits<S>::V* scratch,
                                                                  EnsembleLimits limits, int only)
{
  using V = typename StorageTraits<S>::V;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint2 lc = cells[k];
  if (lc.x >= limits.members) return;
  const EnsembleEntry<Q>& e = table[lc.x];
  if (!((only >= 0 ? int(lc.x) == only : e.stepped != 0) && entry_fits(e, limits) && lc.y < e.ncore)) return;
  int x, y, z;
  core_coords(e.div_z, e.div_y, int(e.ny), int(e.nz), lc.y, x, y, z);
  const uint32_t n = uint32_t(x + 1) * e.sx + uint32_t(y + 1) * e.sy + uint32_t(z + 1);
  S* pop = arena + e.pop_offset;
#pragma unroll
  for (int i = 0; i < Q; ++i)
    store_pop<S, V>(column_at(pop + uint64_t(i) * e.stride, wrap_index(n, e.shift[i], e.n_padded)),
                    scratch[uint64_t(k) * Q + i]);
}

// Streaming of every lattice that stepped: rotate its columns, count the step.
// One thread owns one lattice, so the read of steps_left and its update cannot race.
template <int Q>
__global__ void ensemble_advance_kernel(EnsembleEntry<Q>* table, uint32_t nslots)
{
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nslots) return;
  EnsembleEntry<Q>& e = table[l];
  e.stepped = e.steps_left != 0 ? 1u : 0u;
  if (e.steps_left == 0) return;
  for (int i = 0; i < Q; ++i) {
    const uint32_t s = e.shift[i], d = e.dmod[i];
    e.shift[i] = s >= d ? s - d : s + (e.n_padded - d);
  }
  if (e.steps_left > 0) e.steps_left -= 1;
  e.steps_done += 1;
}

}  // namespace lbmh
