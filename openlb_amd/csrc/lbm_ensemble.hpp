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

constexpr int kEnsembleWalls = 16;  // wall-table entries per lattice (ids 0..15)

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

template <int Q, typename S, int KIND>
__global__ __launch_bounds__(kBlock) void ensemble_step_kernel(const EnsembleEntry<Q>* __restrict__ table,
                                                               const uint32_t* __restrict__ chunk_lattice,
                                                               S* __restrict__ arena, const uint8_t* __restrict__ ids,
                                                               const double* __restrict__ walls, uint32_t xcd_group,
                                                               EnsembleLimits limits)
{
  using V = typename StorageTraits<S>::V;
  uint32_t b = blockIdx.x;
  if (xcd_group) b = remap_xcd(b, gridDim.x, xcd_group);
  const uint32_t lat = chunk_lattice[b];
  if (lat >= limits.members) return;
  const EnsembleEntry<Q>& e = table[lat];
  if (e.steps_left == 0) return;
  // uniform guard: every column and the id range of the entry inside the arenas
  if (e.stride < e.n_padded || e.pop_offset + uint64_t(Q) * e.stride > limits.pop_elements ||
      e.ids_offset + e.ncore > limits.id_bytes)
    return;
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
  a.info = nullptr;
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
  collide_stream_cell<Q, S, KIND, true>(a, (b - e.chunk_begin) * kBlock + threadIdx.x);
}

// Streaming of every lattice that stepped: rotate its columns, count the step.
// One thread owns one lattice, so the read of steps_left and its update cannot race.
template <int Q>
__global__ void ensemble_advance_kernel(EnsembleEntry<Q>* table, uint32_t nslots)
{
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nslots) return;
  EnsembleEntry<Q>& e = table[l];
  if (e.steps_left == 0) return;
  for (int i = 0; i < Q; ++i) {
    const uint32_t s = e.shift[i], d = e.dmod[i];
    e.shift[i] = s >= d ? s - d : s + (e.n_padded - d);
  }
  if (e.steps_left > 0) e.steps_left -= 1;
  e.steps_done += 1;
}

}  // namespace lbmh
