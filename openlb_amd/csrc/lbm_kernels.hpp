// HIP kernels of the block-lattice time step for gfx950 (wave64).
//
// Memory layout in HBM: one column per discrete velocity i, each holding all
// n_padded cells of the padded block (padding 1, x-major, z contiguous,
// src/core/blockStructure.h:65-78), columns `stride` elements apart.  A column
// is addressed through its periodic shift: logical cell n of column i lives at
// element (n + shift_i) mod n_padded (src/core/platform/cpu/sisd/column.h:
// 160-204).  Streaming is shift_i -= c_i . (sx, sy, 1) on the host, so a time
// step is a single pass that reads and writes every population once.
//
// One lane owns one core cell.  Lanes of a wave take consecutive z, so each of
// the Q loads and stores of a wave is one contiguous 64-element run of a
// column (256 B at FP32).  The periodic overlap refresh that OpenLB performs
// as a separate copy after collision (superLattice.hh:632-637) is folded into
// the same pass: a boundary cell also writes each population that will be
// pulled across a periodic face into the halo slot it is pulled from.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "fastdiv.hpp"
#include "lbm_collision.hpp"
#include "lbm_local.hpp"
#include "lbm_storage.hpp"

namespace lbmh {

constexpr int kBlock = 256;
constexpr int kMaxWalls = 256;
constexpr int kIdBulk = 1;

template <int Q, typename V>
struct StepArgs {
  void* pop;
  uint64_t stride;          // elements between population columns
  uint32_t n_padded;        // cells per column (modulus of the shifts)
  uint32_t shift[Q];        // PS shift of each column, in [0, n_padded)
  const uint8_t* ids;       // per core cell dynamics id (flat core index)
  uint32_t ncore;
  uint32_t nx, ny, nz;
  uint32_t sx, sy;          // padded strides of x and y
  FastDiv div_z, div_y;
  uint32_t periodic;        // bit d set: axis d periodic
  uint32_t xcd_group;       // runs of consecutive chunks per XCD (0: identity)
  const double* walls;      // [kMaxWalls][4] rho, ux, uy, uz
  const uint32_t* info;     // [kMaxWalls] packed LocalInfo (local-boundary lattices only)
  const V* force;           // per-cell FORCE [3][ncore] of the forced kinds, or null (bp.force)
  BulkParams<V> bp;
};

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), so neighbouring cell chunks land in different L2s and the cache line
// a PS-shifted column shares between two chunks is fetched twice (measured:
// ~12% extra HBM reads at 256 cells per chunk).  Grouped remap: XCD x takes
// runs of `group` consecutive chunks, while the eight XCDs together stay
// inside a window of 8*group chunks, so the number of concurrent DRAM streams
// does not grow the way it does when each XCD walks its own eighth of the
// lattice.  Blocks past the last whole window keep the identity mapping, which
// keeps the map a bijection for any grid size.
__device__ __forceinline__ uint32_t remap_xcd(uint32_t b, uint32_t nblocks, uint32_t group)
{
  constexpr uint32_t kXcd = 8;
  const uint32_t window = kXcd * group;
  if (b >= (nblocks / window) * window) return b;
  const uint32_t x = b % kXcd, k = b / kXcd;
  return (k / group) * window + x * group + (k % group);
}

// Write populations that cross a periodic face into the halo cell they are
// pulled from next step.  For velocity i of core cell r the target core cell
// is m = wrap(r + c_i); if m != r + c_i the value must sit at m - c_i.  Per
// axis d that slot differs from r by +n_d (r_d = 0, c_id = -1) or -n_d
// (r_d = n_d - 1, c_id = +1) on a periodic axis and by 0 otherwise, so the
// padded index of the slot is n + dn_i with dn_i a sum of per-lane terms
// picked by the compile-time c_i; a velocity that leaves through a
// non-periodic face is not pushed (the overlap there is never pulled).
template <int Q, typename S, typename V>
__device__ __forceinline__ void push_periodic_halo(const StepArgs<Q, V>& a, S* pop, const V (&f)[Q],
                                                int x, int y, int z)
{
  using L = Lattice<Q>;
  const int r[3] = {x, y, z};
  const int n[3] = {int(a.nx), int(a.ny), int(a.nz)};
  const int stride[3] = {int(a.sx), int(a.sy), 1};
  int lo[3], hi[3];         // padded-index offset of the slot when leaving through the low / high face
  bool lo_out[3], hi_out[3]; // leaving through a non-periodic low / high face
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool per = (a.periodic >> d) & 1u;
    const bool at_lo = r[d] == 0, at_hi = r[d] == n[d] - 1;
    lo[d] = per && at_lo ? n[d] * stride[d] : 0;
    hi[d] = per && at_hi ? -n[d] * stride[d] : 0;
    lo_out[d] = !per && at_lo;
    hi_out[d] = !per && at_hi;
  }
  const uint32_t base = uint32_t(x + 1) * a.sx + uint32_t(y + 1) * a.sy + uint32_t(z + 1);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    int dn = 0;
    bool out = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (L::c[i][d] < 0) {
        dn += lo[d];
        out = out || lo_out[d];
      } else if (L::c[i][d] > 0) {
        dn += hi[d];
        out = out || hi_out[d];
      }
    }
    if (dn != 0 && !out) {
      const uint32_t k = wrap_index(base + uint32_t(dn), a.shift[i], a.n_padded);
      store_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, k), f[i]);
    }
  }
}

// How a pass treats the ids >= 3 that carry local-boundary dynamics: the
// whole-lattice pass leaves them to a second pass over the list of those
// cells, so the bulk kernel keeps its register budget (and occupancy).
enum LocalMode : int { kLocalNone = 0, kLocalSkip = 1, kLocalCollide = 2, kLocalFlat = 3 };

// Time step of core cell c of one block: collide in place, refresh the
// periodic overlap.  The caller advances the shifts afterwards.
template <int Q, typename S, int KIND, bool IDS, int LOCAL = kLocalNone>
__device__ __forceinline__ void collide_stream_cell(const StepArgs<Q, typename StorageTraits<S>::V>& a, uint32_t c)
{
  using V = typename StorageTraits<S>::V;
  if (c >= a.ncore) return;
  const uint32_t t = a.div_z.div(c);
  const uint32_t z = c - t * a.nz;
  const uint32_t x = a.div_y.div(t);
  const uint32_t y = t - x * a.ny;
  const bool edge = a.periodic &&
                    ((((a.periodic & 1u) != 0) && (x == 0 || x + 1 == a.nx)) ||
                     (((a.periodic & 2u) != 0) && (y == 0 || y + 1 == a.ny)) ||
                     (((a.periodic & 4u) != 0) && (z == 0 || z + 1 == a.nz)));
  int id = kIdBulk;
  if constexpr (IDS) {
    id = a.ids[c];
    // NoDynamics cells are not collided, but on a periodic face their
    // populations still cross into the overlap like every other core cell's
    if (id == 0 && !edge) return;
    if constexpr (LOCAL == kLocalSkip) {
      if (id >= 3 && (a.info[id] & 0xffu) != kMovingWall) return;
    }
    if constexpr (LOCAL == kLocalFlat) {
      // flat faces, and the outer edges and corners (BGK at their fixed
      // velocity), stay in this pass; inner edges and corners go to the list pass
      if (id >= 3) {
        const uint32_t k = a.info[id] & 0xffu;
        if (k == kInnerEdge || k == kInnerCorner || k == kDeadWall) return;
      }
    }
  }
  const uint32_t n = (x + 1) * a.sx + (y + 1) * a.sy + (z + 1);
  S* pop = static_cast<S*>(a.pop);

  V f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = load_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, wrap_index(n, a.shift[i], a.n_padded)));
  }
  // The store addresses are recomputed after the collision from an opaque
  // copy of n instead of held through it (Q fewer live registers: 30-55 VGPRs
  // less per kernel, at 4 VALU per population).
  uint32_t ns = n;
  asm volatile("" : "+v"(ns));
  if (!IDS || id != 0) {
    if (!IDS || id == kIdBulk) {
      if constexpr (KIND == kForcedBGK || KIND == kForcedTRT) {
        // ForcedBGK / ForcedTRT read the cell's own FORCE (momenta/elements.h:1480-1488)
        BulkParams<V> p = a.bp;
        if (a.force) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            p.force[d] = a.force[uint64_t(d) * a.ncore + c];
            p.half_force[d] = p.force[d] * V(0.5);
          }
        }
        collide_bulk<Q, V, KIND>(f, p);
      } else {
        collide_bulk<Q, V, KIND>(f, a.bp);
      }
    } else if (id == 2) {
      revert<Q, V>(f);
    } else {
      const double* w = a.walls + 4 * id;
      bool wall = true;
      if constexpr (LOCAL == kLocalCollide) {
        const LocalInfo li = unpack_local_info(a.info[id]);
        if (li.kind != kMovingWall) {
          wall = false;
          if (li.kind != kDeadWall) {
            V rho, usq;
            collide_local<Q, V>(f, li, w, a.bp.omega, rho, usq);
          }
        }
      }
      if constexpr (LOCAL == kLocalFlat) {
        const LocalInfo li = unpack_local_info(a.info[id]);
        if (li.kind == kFdVelocity || li.kind == kFdPressure || li.kind == kOuterEdge || li.kind == kOuterCorner) {
          // BGK at the momenta of the cell's type: CombinedRLB's plain case
          wall = false;
          collide_dirichlet<Q, V>(f, li, w, a.bp.omega);
        } else if (li.kind != kMovingWall) {
          wall = false;
          collide_flat<Q, V>(f, li, w, a.bp.omega);
        }
      }
      if (wall) {
        const V uw[3] = {V(w[1]), V(w[2]), V(w[3])};
        moving_wall<Q, V>(f, V(w[0]), uw);
      }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      store_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, wrap_index(ns, a.shift[i], a.n_padded)), f[i]);
    }
  }
  if (edge) {
    // opaque coordinates: what the push derives from them cannot be hoisted
    // above the collision, where it would hold registers through it
    uint32_t hx = x, hy = y, hz = z;
    asm volatile("" : "+v"(hx), "+v"(hy), "+v"(hz));
    push_periodic_halo<Q, S, V>(a, pop, f, int(hx), int(hy), int(hz));
  }
}

// One time step of a block, one lane per core cell.
template <int Q, typename S, int KIND, bool IDS, int LOCAL = kLocalNone>
__global__ __launch_bounds__(kBlock) void collide_stream_kernel(StepArgs<Q, typename StorageTraits<S>::V> a)
{
  uint32_t b = blockIdx.x;
  if (a.xcd_group) b = remap_xcd(b, gridDim.x, a.xcd_group);
  collide_stream_cell<Q, S, KIND, IDS, LOCAL>(a, b * kBlock + threadIdx.x);
}

// Measurement kernel, not part of a time step: the step kernel's access
// pattern with its arithmetic taken out.  Each lane reads the Q populations of
// its cell at their shifted slots and writes the same values back in place,
// with the step kernel's cell mapping, addressing and launch shape.  The values
// pass through an empty asm so the stores cannot be dropped as no-ops.  Its
// rate is the ceiling of the PS in-place read-modify-write on this device,
// against which the step kernel is priced beside the vendor peak.
template <int Q, typename S>
__global__ __launch_bounds__(kBlock) void pattern_ceiling_kernel(StepArgs<Q, float> a)
{
  uint32_t b = blockIdx.x;
  if (a.xcd_group) b = remap_xcd(b, gridDim.x, a.xcd_group);
  const uint32_t c = b * kBlock + threadIdx.x;
  if (c >= a.ncore) return;
  const uint32_t t = a.div_z.div(c);
  const uint32_t z = c - t * a.nz;
  const uint32_t x = a.div_y.div(t);
  const uint32_t y = t - x * a.ny;
  const uint32_t n = (x + 1) * a.sx + (y + 1) * a.sy + (z + 1);
  S* pop = static_cast<S*>(a.pop);
  S f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = *column_at(pop + uint64_t(i) * a.stride, wrap_index(n, a.shift[i], a.n_padded));
  uint32_t ns = n;
  asm volatile("" : "+v"(ns));
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    asm volatile("" : "+v"(f[i]));
    *column_at(pop + uint64_t(i) * a.stride, wrap_index(ns, a.shift[i], a.n_padded)) = f[i];
  }
}

// The same time step for the listed local-boundary cells (the cells the
// whole-lattice pass skipped).  The two passes touch disjoint populations.
template <int Q, typename S, int KIND>
__global__ __launch_bounds__(64) void local_boundary_kernel(StepArgs<Q, typename StorageTraits<S>::V> a,
                                                            const uint32_t* cells, uint32_t count)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  collide_stream_cell<Q, S, KIND, true, kLocalCollide>(a, cells[k]);
}

// ---- data movement and evaluation kernels (not on the time-step path) ----

struct CoreGeom {
  uint32_t ncore, nx, ny, nz, sx, sy, n_padded;
  uint64_t stride;
  FastDiv div_z, div_y;
};

__device__ __forceinline__ uint32_t padded_of_core(const CoreGeom& g, uint32_t c)
{
  const uint32_t t = g.div_z.div(c);
  const uint32_t z = c - t * g.nz;
  const uint32_t x = g.div_y.div(t);
  const uint32_t y = t - x * g.ny;
  return (x + 1) * g.sx + (y + 1) * g.sy + (z + 1);
}

template <int Q>
struct Shifts { uint32_t s[Q]; };

// host layout [q][ncore] (type H) -> PS storage (type S)
template <int Q, typename H, typename S>
__global__ void scatter_populations(const H* src, S* pop, CoreGeom g, Shifts<Q> sh)
{
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncore) return;
  const uint32_t n = padded_of_core(g, c);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    using V = typename StorageTraits<S>::V;
    store_pop<S, V>(pop + uint64_t(i) * g.stride + wrap_index(n, sh.s[i], g.n_padded),
                    V(src[uint64_t(i) * g.ncore + c]));
  }
}

template <int Q, typename H, typename S>
__global__ void gather_populations(H* dst, const S* pop, CoreGeom g, Shifts<Q> sh)
{
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncore) return;
  const uint32_t n = padded_of_core(g, c);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    using V = typename StorageTraits<S>::V;
    dst[uint64_t(i) * g.ncore + c] = H(load_pop<S, V>(pop + uint64_t(i) * g.stride + wrap_index(n, sh.s[i], g.n_padded)));
  }
}

// Overlap refresh between two blocks of one device (HomogeneousCopyTask::copy,
// src/core/blockLattice.hh:752-787): all Q stored values of source core cell
// p.y go to destination overlap cell p.x, each through its own column shift.
// blockIdx.y is the population, so a wave copies consecutive plan entries of
// one column.
template <int Q, typename S>
__global__ __launch_bounds__(256) void halo_copy_kernel(S* dst, uint64_t dst_stride, uint32_t dst_n, Shifts<Q> dsh,
                                                        const S* src, uint64_t src_stride, uint32_t src_n,
                                                        Shifts<Q> ssh, const uint2* pairs, uint32_t count)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = blockIdx.y;
  if (k >= count || i >= uint32_t(Q)) return;
  const uint2 p = pairs[k];
  dst[uint64_t(i) * dst_stride + wrap_index(p.x, dsh.s[i], dst_n)] =
      src[uint64_t(i) * src_stride + wrap_index(p.y, ssh.s[i], src_n)];
}

// One halo of a batched copy: both blocks' columns and shifts, its pair
// list, and where its pairs start in the batch.
template <int Q, typename S>
struct HaloCopyDesc {
  S* dst;
  const S* src;
  const uint2* pairs;
  uint64_t dst_stride, src_stride;
  uint32_t dst_n, src_n;
  uint32_t first, count;
  uint32_t dst_shift[Q];
  uint32_t src_shift[Q];
};

// The copies of halo_copy_kernel for many halos in one launch: lane k of
// the batch finds its halo by bisection over the halos' first pairs.
template <int Q, typename S>
__global__ __launch_bounds__(256) void halo_copy_many_kernel(const HaloCopyDesc<Q, S>* __restrict__ descs,
                                                             uint32_t halos, uint32_t total)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = blockIdx.y;
  if (k >= total || i >= uint32_t(Q)) return;
  uint32_t lo = 0, hi = halos;  // the last halo whose first pair is <= k
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (descs[mid].first <= k) lo = mid;
    else hi = mid;
  }
  const HaloCopyDesc<Q, S>& d = descs[lo];
  const uint32_t j = k - d.first;
  if (j >= d.count) return;
  const uint2 p = d.pairs[j];
  d.dst[uint64_t(i) * d.dst_stride + wrap_index(p.x, d.dst_shift[i], d.dst_n)] =
      d.src[uint64_t(i) * d.src_stride + wrap_index(p.y, d.src_shift[i], d.src_n)];
}

// Equilibrium from per-cell (rho, u), iniEquilibrium semantics.
template <int Q, typename S>
__global__ void ini_equilibrium_kernel(const double* rho_u, const uint8_t* ids, S* pop, CoreGeom g,
                                       Shifts<Q> sh, double hfx, double hfy, double hfz, int forced,
                                       const typename StorageTraits<S>::V* force)
{
  using V = typename StorageTraits<S>::V;
  using L = Lattice<Q>;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncore) return;
  const int id = ids ? ids[c] : 1;
  if (id == 0) return;
  const V rho = V(rho_u[c]);
  V u[3] = {V(rho_u[uint64_t(g.ncore) + c]), V(rho_u[2 * uint64_t(g.ncore) + c]),
            V(rho_u[3 * uint64_t(g.ncore) + c])};
  if (forced && id == 1) {
    u[0] -= half_force<V>(force, g.ncore, c, 0, hfx);
    u[1] -= half_force<V>(force, g.ncore, c, 1, hfy);
    u[2] -= half_force<V>(force, g.ncore, c, 2, hfz);
  }
  const V drho = rho - V(1);
  const V usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const V A = drho - V(1.5) * rho * usq;
  V f[Q];
  f[0] = weight<V, Q>(0) * A;
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    V even, odd;
    equilibrium_pair<Q, V>(i, A, rho, u, even, odd);
    f[i] = even + odd;
    f[i + L::half] = even - odd;
  }
  const uint32_t n = padded_of_core(g, c);
#pragma unroll
  for (int i = 0; i < Q; ++i) store_pop<S, V>(pop + uint64_t(i) * g.stride + wrap_index(n, sh.s[i], g.n_padded), f[i]);
}

// Per-cell rho, u in FP64 ([4][ncore]); u includes F/2 on forced bulk cells.
template <int Q, typename S>
__global__ void moments_kernel(double* out, const uint8_t* ids, const S* pop, CoreGeom g, Shifts<Q> sh,
                               double hfx, double hfy, double hfz, int forced,
                               const typename StorageTraits<S>::V* force)
{
  using V = typename StorageTraits<S>::V;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncore) return;
  const uint32_t n = padded_of_core(g, c);
  V f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = load_pop<S, V>(pop + uint64_t(i) * g.stride + wrap_index(n, sh.s[i], g.n_padded));
  V drho, j[3];
  moments<Q, V>(f, drho, j);
  const V rho = V(1) + drho;
  double u[3] = {double(j[0] / rho), double(j[1] / rho), double(j[2] / rho)};
  const int id = ids ? ids[c] : 1;
  if (forced && id == 1) {
    u[0] += double(half_force<V>(force, g.ncore, c, 0, hfx));
    u[1] += double(half_force<V>(force, g.ncore, c, 1, hfy));
    u[2] += double(half_force<V>(force, g.ncore, c, 2, hfz));
  }
  out[c] = double(rho);
  out[uint64_t(g.ncore) + c] = u[0];
  out[2 * uint64_t(g.ncore) + c] = u[1];
  out[3 * uint64_t(g.ncore) + c] = u[2];
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// First level of the FP64 conservation reduction over bulk cells: one
// partial {rho, jx, jy, jz} per workgroup, fixed order (deterministic).
template <int Q, typename S>
__global__ __launch_bounds__(kBlock) void reduce_moments_kernel(double* partial, const uint8_t* ids, const S* pop,
                                                                CoreGeom g, Shifts<Q> sh)
{
  using V = typename StorageTraits<S>::V;
  double acc[4] = {0, 0, 0, 0};
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < g.ncore; c += gridDim.x * kBlock) {
    if (ids && ids[c] != 1) continue;
    const uint32_t n = padded_of_core(g, c);
    V f[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = load_pop<S, V>(pop + uint64_t(i) * g.stride + wrap_index(n, sh.s[i], g.n_padded));
    V drho, j[3];
    moments<Q, V>(f, drho, j);
    acc[0] += 1.0 + double(drho);
    acc[1] += double(j[0]);
    acc[2] += double(j[1]);
    acc[3] += double(j[2]);
  }
  __shared__ double lds[4][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) lds[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[threadIdx.x][w];
    partial[blockIdx.x * 4 + threadIdx.x] = s;
  }
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Lattice statistics of OpenLB (LatticeStatistics::Aggregatable, src/core/
// latticeStatistics.h:49-71): over the cells whose collision reports a
// statistic -- bulk cells {rho, |u|^2} (u with the F/2 shift when forced) and
// moving walls {rho_w, |u_w|^2} (collision.h:343); BounceBack reports none.
// Partial per workgroup: {count, sum rho, sum |u|^2, max |u|^2}.
template <int Q, typename S>
__global__ __launch_bounds__(kBlock) void statistics_kernel(double* partial, const uint8_t* ids, const S* pop,
                                                            CoreGeom g, Shifts<Q> sh, const double* walls,
                                                            const uint32_t* info, double omega,
                                                            double hfx, double hfy, double hfz, int forced,
                                                            const typename StorageTraits<S>::V* force)
{
  using V = typename StorageTraits<S>::V;
  double acc[4] = {0, 0, 0, 0};
  for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < g.ncore; c += gridDim.x * kBlock) {
    const int id = ids ? ids[c] : 1;
    if (id == 0 || id == 2) continue;
    double rho, usq;
    const LocalInfo li = info ? unpack_local_info(info[id < 3 ? 0 : id]) : LocalInfo{kMovingWall, 0, 0, 0};
    if (id >= 3 && li.kind == kDeadWall) continue;
    if (id >= 3 && li.kind != kMovingWall) {
      // the statistic the local boundary's collision reports
      const uint32_t n = padded_of_core(g, c);
      V f[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) f[i] = load_pop<S, V>(pop + uint64_t(i) * g.stride + wrap_index(n, sh.s[i], g.n_padded));
      V r, u2;
      collide_local<Q, V>(f, li, walls + 4 * id, V(omega), r, u2);
      rho = double(r);
      usq = double(u2);
    } else if (id >= 3) {
      const double* w = walls + 4 * id;
      rho = w[0];
      usq = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
    } else {
      const uint32_t n = padded_of_core(g, c);
      V f[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) f[i] = load_pop<S, V>(pop + uint64_t(i) * g.stride + wrap_index(n, sh.s[i], g.n_padded));
      V drho, j[3];
      moments<Q, V>(f, drho, j);
      const V r = V(1) + drho;
      V u[3] = {j[0] / r, j[1] / r, j[2] / r};
      if (forced) {
        u[0] += half_force<V>(force, g.ncore, c, 0, hfx);
        u[1] += half_force<V>(force, g.ncore, c, 1, hfy);
        u[2] += half_force<V>(force, g.ncore, c, 2, hfz);
      }
      rho = double(r);
      usq = double(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    }
    acc[0] += 1.0;
    acc[1] += rho;
    acc[2] += usq;
    acc[3] = fmax(acc[3], usq);
  }
  __shared__ double lds[4][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) lds[k][wave] = s;
  }
  {
    const double m = wave_max(acc[3]);
    if (lane == 0) lds[3][wave] = m;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w)
      s = threadIdx.x == 3 ? fmax(s, lds[3][w]) : s + lds[threadIdx.x][w];
    partial[blockIdx.x * 4 + threadIdx.x] = s;
  }
}

static __global__ __launch_bounds__(kBlock) void statistics_partials_kernel(double* out, const double* partial,
                                                                            uint32_t nparts)
{
  __shared__ double lds[4][kBlock / 64];
  double acc[4] = {0, 0, 0, 0};
  for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) {
    for (int k = 0; k < 3; ++k) acc[k] += partial[p * 4 + k];
    acc[3] = fmax(acc[3], partial[p * 4 + 3]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) lds[k][wave] = s;
  }
  {
    const double m = wave_max(acc[3]);
    if (lane == 0) lds[3][wave] = m;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w)
      s = threadIdx.x == 3 ? fmax(s, lds[3][w]) : s + lds[threadIdx.x][w];
    out[threadIdx.x] = s;
  }
}

// Second level: one workgroup sums the partials in a fixed order.
static __global__ __launch_bounds__(kBlock) void reduce_partials_kernel(double* out, const double* partial,
                                                                        uint32_t nparts)
{
  __shared__ double lds[4][kBlock / 64];
  double acc[4] = {0, 0, 0, 0};
  for (uint32_t p = threadIdx.x; p < nparts; p += kBlock)
    for (int k = 0; k < 4; ++k) acc[k] += partial[p * 4 + k];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) lds[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[threadIdx.x][w];
    out[threadIdx.x] = s;
  }
}

}  // namespace lbmh
