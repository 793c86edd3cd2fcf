// C ABI of the MI355X block-lattice backend (declarations in include/lbmh.h).
//
// Host side of one OpenLB block on one GPU: allocation of the population
// columns, the per-cell dynamics ids and the wall table, the PS shifts (kept
// on the host and passed to every launch), launches of the fused time step,
// and the data-movement / evaluation kernels.  Errors are returned as status
// codes with a per-thread message; nothing throws across the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "boundary_classify.hpp"
#include "checkpoint.hpp"
#include "lbm_fields.hpp"
#include "lbm_kernels.hpp"
#include "lbmh.h"
#include "lbmh_common.hpp"

using namespace lbmh;

thread_local std::string lbmh::g_error;

// Workspace of lbmh_reduce_moments_noalloc: the per-workgroup partial sums of
// the FP64 moment reduction, kScratchPartials workgroups x 4 doubles (32 KiB
// per device, held until the process ends).  Host code never uses the name
// `scratch` as an address: each device's copy is reached through the pointer
// hipGetSymbolAddress returns for it (scratch_address).
static constexpr uint32_t kScratchPartials = 1024;
static __device__ __align__(256) double scratch[4 * kScratchPartials];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device `device`, looked up once
// per device with that device current.
int scratch_address(int device, double** out)
{
  if (device < 0 || device >= kMaxDevices) return fail(LBMH_E_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(LBMH_E_DEVICE, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return LBMH_OK;
}

}  // namespace

struct lbmh_lattice {
  int device = 0;
  int q = 19;
  int n[3] = {0, 0, 0};
  int periodic[3] = {0, 0, 0};
  int storage = LBMH_F32;
  int kind = LBMH_BGK;
  uint32_t size[3] = {0, 0, 0};
  uint32_t n_padded = 0;
  uint64_t stride = 0;
  uint32_t ncore = 0;
  int64_t steps = 0;
  int64_t shift[27] = {};
  int64_t dist[27] = {};
  void* d_pop = nullptr;
  uint8_t* d_ids = nullptr;
  double* d_walls = nullptr;
  double* d_partials = nullptr;
  uint32_t n_partials = 0;
  bool uniform_bulk = false;   // every core cell has id 1: skip the id read
  lbmh_launch_config launch{49152, 4};
  bool launch_chosen = false;     // set by lbmh_set_launch_config / lbmh_autotune: keep it
  lbmh_bulk_params params{};
  double walls[kMaxWalls][4] = {};
  size_t footprint = 0;
  // local (regularized) boundaries
  std::vector<uint8_t> h_ids;        // host copy of the dynamics ids
  std::vector<lbmh_halo*> halos_in;  // halos that fill this block's overlap
  std::vector<lbmh_halo*> halos_out; // halos that read this block
  bool id_defined[kMaxWalls] = {};   // ids given a wall or boundary entry
  uint8_t material_of[kMaxWalls] = {};  // the material (id of lbmh_set_dynamics_ids) each dynamics id stands for
  uint32_t info[kMaxWalls] = {};     // packed LocalInfo per id
  uint32_t* d_info = nullptr;
  bool has_local = false;
  // per-cell FORCE of the forced kinds: [3][ncore] on the device in the
  // arithmetic type (double for FP64 storage, float otherwise) and in FP64 on
  // the host (checkpoints); empty: the uniform params.force
  void* d_force = nullptr;
  std::vector<double> h_force;
  std::vector<uint32_t> h_edges, h_corners, h_local, h_planes;
  uint32_t* d_edges = nullptr;
  uint32_t* d_corners = nullptr;
  uint32_t* d_local = nullptr;     // local-boundary cells of the list pass (edges, corners, inner types)
  uint32_t* d_planes = nullptr;    // interpolated-boundary faces (plane finite-difference processor)
  bool pressure_planes = false;    // some plane is a pressure face: compute into d_scratch first
  void* d_scratch = nullptr;
  // the post-processed lists with their neighbours resolved (PostCell records)
  PostCell* d_edge_recs = nullptr;
  PostCell* d_corner_recs = nullptr;
  PostCell* d_plane_recs = nullptr;
  // block of a split lattice: materials of the first overlap layer (padded
  // map, core entries unused), empty for a lone block
  std::vector<uint8_t> h_ring;
  // post-processor neighbours beyond the core on non-periodic axes of such a
  // block, in block coordinates, and the post-processor stage that reads each:
  // they lie in other blocks and are read from d_remote (rho, u: 4 values
  // each, in the arithmetic type), which the halos into this block fill
  std::vector<std::array<int, 3>> remote_pos;
  std::vector<uint8_t> remote_stage;
  bool remote_resolved = true;
  void* d_remote = nullptr;
  int64_t init_step = -1;  // step count at which lbmh_initialize asked for the post-processors
  // descriptor tables of the batched halo copies and remote fetches this block leads
  void* d_halo_batch = nullptr;
  size_t halo_batch_bytes = 0;
  void* d_remote_batch = nullptr;
  size_t remote_batch_bytes = 0;
};

// Library default launch shape, from sweeps of (LDS reservation, XCD group)
// on the porous channel, the turbulent channel and a Taylor-Green box.  FP32:
// a 40 KB LDS reservation caps the resident workgroups at 4 per CU (fewer
// concurrent DRAM streams), XCD groups of 4.  FP64: 48 KB (3 per CU), groups
// of 8, 1-2% ahead of the FP32 shape.  FP16: no cap, groups of 8.  With
// local boundaries the whole-lattice pass collides the flat faces too, and
// its divergent waves need the occupancy: no cap (201^3 cavity, FP32: 298 us
// against 324 us).
static lbmh_launch_config default_launch(int storage, bool local)
{
  if (storage == LBMH_F16) return lbmh_launch_config{0, 8};
  if (local) return lbmh_launch_config{0, 4};
  return storage == LBMH_F64 ? lbmh_launch_config{49152, 8} : lbmh_launch_config{40960, 4};
}

struct lbmh_halo {
  lbmh_lattice* dst = nullptr;  // null once the destination block is destroyed
  const lbmh_lattice* src = nullptr;  // null once the source block is destroyed
  int device = 0;
  uint2* d_pairs = nullptr;   // (destination overlap cell, source core cell), padded ids
  std::vector<uint2> pairs;   // the same on the host: which cell each overlap cell copies
  uint32_t count = 0;
  // post-processor neighbours of dst that src owns, by post-processor stage
  std::vector<RemotePair> remote[kPostStages];
  RemotePair* d_remote[kPostStages] = {};
};

namespace {

CoreGeom core_geom(const lbmh_lattice* L)
{
  CoreGeom g;
  g.ncore = L->ncore;
  g.nx = L->n[0];
  g.ny = L->n[1];
  g.nz = L->n[2];
  g.sx = L->size[1] * L->size[2];
  g.sy = L->size[2];
  g.n_padded = L->n_padded;
  g.stride = L->stride;
  g.div_z = FastDiv(uint32_t(L->n[2]));
  g.div_y = FastDiv(uint32_t(L->n[1]));
  return g;
}

template <int Q>
Shifts<Q> shifts_of(const lbmh_lattice* L)
{
  Shifts<Q> s;
  for (int i = 0; i < Q; ++i) s.s[i] = uint32_t(L->shift[i]);
  return s;
}

// The lattice's per-cell FORCE in the arithmetic type of storage S, or null.
template <typename S>
const typename StorageTraits<S>::V* force_field(const lbmh_lattice* L)
{
  return static_cast<const typename StorageTraits<S>::V*>(L->d_force);
}

template <typename V>
BulkParams<V> bulk_params(const lbmh_lattice* L)
{
  return derive_bulk_params<V>(L->q, L->params);
}

// Whole-block arguments of the per-cell momenta evaluation (post-processors,
// field export); `cells` and `count` are left for the caller.
template <int Q>
PostArgs<Q> post_args(const lbmh_lattice* L)
{
  PostArgs<Q> a;
  a.pop = L->d_pop;
  a.stride = L->stride;
  a.n_padded = L->n_padded;
  for (int i = 0; i < Q; ++i) a.shift[i] = uint32_t(L->shift[i]);
  a.ids = L->d_ids;
  a.walls = L->d_walls;
  a.info = L->d_info;
  a.cells = nullptr;
  a.recs = nullptr;
  a.remote = L->d_remote;
  a.forced = 0;
  a.force = nullptr;
  a.half_force[0] = a.half_force[1] = a.half_force[2] = 0.0;
  a.count = 0;
  a.nx = L->n[0];
  a.ny = L->n[1];
  a.nz = L->n[2];
  a.sx = L->size[1] * L->size[2];
  a.sy = L->size[2];
  a.periodic = uint32_t(L->periodic[0] | (L->periodic[1] << 1) | (L->periodic[2] << 2));
  a.div_z = FastDiv(uint32_t(L->n[2]));
  a.div_y = FastDiv(uint32_t(L->n[1]));
  a.omega = L->params.omega;
  return a;
}

// The post-processors read a forced bulk neighbour's velocity with F/2 added,
// as its momenta report it
template <int Q>
void post_forcing(const lbmh_lattice* L, PostArgs<Q>& a)
{
  if (L->kind != LBMH_FORCED_BGK && L->kind != LBMH_FORCED_TRT) return;
  a.forced = 1;
  a.force = L->d_force;
  for (int d = 0; d < 3; ++d) a.half_force[d] = L->params.force[d] * 0.5;
}

// The cells of a post-processor list with their neighbours resolved, as the
// device's cell_rho_u would find them: offsets of the processor's slots,
// periodic axes wrapped, kPostOutside beyond a non-periodic face.  A block of
// a split lattice (one with overlap materials) has neighbours there instead:
// each such slot becomes an entry of L->remote_pos, read from the remote
// buffer.
std::vector<PostCell> post_records(lbmh_lattice* L, const std::vector<uint32_t>& cells, int kind)
{
  const int n[3] = {L->n[0], L->n[1], L->n[2]};
  const uint32_t sx = L->size[1] * L->size[2], sy = L->size[2];
  std::vector<PostCell> out(cells.size());
  for (size_t k = 0; k < cells.size(); ++k) {
    const uint32_t c = cells[k];
    const int r[3] = {int(c / (uint32_t(n[1]) * uint32_t(n[2]))), int((c / uint32_t(n[2])) % uint32_t(n[1])),
                      int(c % uint32_t(n[2]))};
    PostCell& p = out[k];
    p.c = c;
    p.id = L->h_ids[c];
    const LocalInfo li = unpack_local_info(L->info[p.id]);
    for (int slot = 0; slot < kPostSlots; ++slot) {
      int o[3];
      post_offset(li, kind, slot, o);
      int q[3];
      bool outside = false;
      for (int d = 0; d < 3; ++d) {
        q[d] = r[d] + o[d];
        if (q[d] < 0 || q[d] >= n[d]) {
          if (!L->periodic[d]) outside = true;
          else q[d] = (q[d] % n[d] + n[d]) % n[d];
        }
      }
      if (outside && !L->h_ring.empty()) {
        p.nb[slot] = kPostRemote | uint32_t(L->remote_pos.size());
        p.nb_id[slot] = 0;
        L->remote_pos.push_back({q[0], q[1], q[2]});
        L->remote_stage.push_back(uint8_t(post_stage(kind)));
      } else if (outside) {
        p.nb[slot] = kPostOutside;
        p.nb_id[slot] = 0;
      } else {
        const uint32_t nc = (uint32_t(q[0]) * uint32_t(n[1]) + uint32_t(q[1])) * uint32_t(n[2]) + uint32_t(q[2]);
        p.nb[slot] = uint32_t(q[0] + 1) * sx + uint32_t(q[1] + 1) * sy + uint32_t(q[2] + 1);
        p.nb_id[slot] = L->h_ids[nc];
      }
    }
  }
  return out;
}

size_t arith_size(int storage) { return storage == LBMH_F64 ? sizeof(double) : sizeof(float); }

void drop_remote(lbmh_lattice* L)
{
  for (lbmh_halo* H : L->halos_in) {
    for (int st = 0; st < kPostStages; ++st) {
      H->remote[st].clear();
      if (H->d_remote[st]) (void)hipFree(H->d_remote[st]);
      H->d_remote[st] = nullptr;
    }
  }
  if (L->d_remote) (void)hipFree(L->d_remote);
  L->d_remote = nullptr;
  L->remote_resolved = L->remote_pos.empty();
}

// Finds the block that owns each remote neighbour of L: the halo source of
// the first overlap layer cell on the way to it, stepped on by the rest of
// the offset inside that block.  Neighbours with no halo source lie outside
// the domain and keep the moments of a cell without dynamics (rho 1, u 0).
int resolve_remote(lbmh_lattice* L)
{
  if (L->remote_resolved) return LBMH_OK;
  DeviceGuard guard(L->device);
  drop_remote(L);
  const size_t count = L->remote_pos.size();
  std::unordered_map<uint32_t, std::pair<lbmh_halo*, uint32_t>> src_of;
  for (lbmh_halo* H : L->halos_in) {
    if (!H->src) continue;
    for (const uint2& pr : H->pairs) src_of[pr.x] = {H, pr.y};
  }
  const uint32_t sx = L->size[1] * L->size[2], sy = L->size[2];
  for (size_t k = 0; k < count; ++k) {
    const auto& q = L->remote_pos[k];
    int r[3], delta[3];
    for (int d = 0; d < 3; ++d) {
      r[d] = std::min(std::max(q[d], -1), L->n[d]);
      delta[d] = q[d] - r[d];
    }
    const uint32_t pr = uint32_t(r[0] + 1) * sx + uint32_t(r[1] + 1) * sy + uint32_t(r[2] + 1);
    auto it = src_of.find(pr);
    if (it == src_of.end()) continue;
    lbmh_halo* H = it->second.first;
    const lbmh_lattice* S = H->src;
    const uint32_t s = it->second.second;
    const uint32_t ssx = S->size[1] * S->size[2], ssy = S->size[2];
    int t[3] = {int(s / ssx) - 1 + delta[0], int((s % ssx) / ssy) - 1 + delta[1], int(s % ssy) - 1 + delta[2]};
    for (int d = 0; d < 3; ++d) {
      if (t[d] < 0 || t[d] >= S->n[d])
        return fail(LBMH_E_INVALID, "a post-processor neighbour lies beyond the block next to this one");
    }
    const uint32_t tp = uint32_t(t[0] + 1) * ssx + uint32_t(t[1] + 1) * ssy + uint32_t(t[2] + 1);
    const uint32_t tc = (uint32_t(t[0]) * uint32_t(S->n[1]) + uint32_t(t[1])) * uint32_t(S->n[2]) + uint32_t(t[2]);
    H->remote[L->remote_stage[k]].push_back(RemotePair{tp, tc, uint32_t(k)});
  }
  if (count) {
    // every entry starts at the moments of a cell without dynamics
    const size_t es = arith_size(L->storage);
    std::vector<uint8_t> init(count * 4 * es);
    for (size_t k = 0; k < count; ++k) {
      if (es == sizeof(double)) {
        const double v[4] = {1.0, 0.0, 0.0, 0.0};
        std::memcpy(init.data() + k * 4 * es, v, sizeof(v));
      } else {
        const float v[4] = {1.0f, 0.0f, 0.0f, 0.0f};
        std::memcpy(init.data() + k * 4 * es, v, sizeof(v));
      }
    }
    LBMH_HIP(hipMalloc(&L->d_remote, init.size()));
    LBMH_HIP(hipMemcpy(L->d_remote, init.data(), init.size(), hipMemcpyHostToDevice));
  }
  for (lbmh_halo* H : L->halos_in) {
    for (int st = 0; st < kPostStages; ++st) {
      const auto& v = H->remote[st];
      if (v.empty()) continue;
      LBMH_HIP(hipMalloc((void**)&H->d_remote[st], sizeof(RemotePair) * v.size()));
      LBMH_HIP(hipMemcpy(H->d_remote[st], v.data(), sizeof(RemotePair) * v.size(), hipMemcpyHostToDevice));
    }
  }
  L->remote_resolved = true;
  return LBMH_OK;
}

// Moments of H's remote neighbours read by post-processor stage `stage`, as
// H->src holds them now, into H->dst's remote buffer.  A source without
// post-processors may not have streamed this step yet (step `at`): it is read
// through the shifts it will have, its post-collision populations are final.
template <int Q>
int remote_source_args(const lbmh_halo* H, int64_t at, PostArgs<Q>& a)
{
  const lbmh_lattice* src = H->src;
  a = post_args<Q>(src);
  post_forcing<Q>(src, a);
  if (src->steps + 1 == at) {
    const int64_t N = src->n_padded;
    for (int i = 0; i < Q; ++i) {
      int64_t v = (src->shift[i] - src->dist[i]) % N;
      if (v < 0) v += N;
      a.shift[i] = uint32_t(v);
    }
  } else if (src->steps != at) {
    return fail(LBMH_E_STATE, "blocks of a split lattice are out of step");
  }
  return LBMH_OK;
}

// The fetches of one post-processor stage for every block of a group in one
// launch; the descriptor table is kept by the block that runs the group.
template <int Q, typename S>
int launch_remote_stage(lbmh_lattice* runner, const std::vector<lbmh_lattice*>& group, int stage,
                        hipStream_t stream)
{
  std::vector<RemoteBatchDesc<Q>> descs;
  uint32_t total = 0;
  for (lbmh_lattice* X : group) {
    for (lbmh_halo* H : X->halos_in) {
      if (H->remote[stage].empty()) continue;
      if (!H->src) return fail(LBMH_E_INVALID, "a block of this halo was destroyed");
      RemoteBatchDesc<Q> d{};
      const int rc = remote_source_args<Q>(H, X->steps, d.a);
      if (rc != LBMH_OK) return rc;
      d.pairs = H->d_remote[stage];
      d.out = X->d_remote;
      d.first = total;
      d.count = uint32_t(H->remote[stage].size());
      total += d.count;
      descs.push_back(d);
    }
  }
  if (descs.empty()) return LBMH_OK;
  const size_t bytes = sizeof(RemoteBatchDesc<Q>) * descs.size();
  if (runner->remote_batch_bytes < bytes) {
    if (runner->d_remote_batch) LBMH_HIP(hipFree(runner->d_remote_batch));
    runner->d_remote_batch = nullptr;
    runner->remote_batch_bytes = 0;
    LBMH_HIP(hipMalloc(&runner->d_remote_batch, bytes));
    runner->remote_batch_bytes = bytes;
  }
  LBMH_HIP(hipMemcpyAsync(runner->d_remote_batch, descs.data(), bytes, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL((remote_moments_many_kernel<Q, S>), dim3(grid_for(total, 64)), dim3(64), 0, stream,
                     static_cast<const RemoteBatchDesc<Q>*>(runner->d_remote_batch), uint32_t(descs.size()), total);
  LBMH_HIP(hipGetLastError());
  return LBMH_OK;
}

bool has_post(const lbmh_lattice* L)
{
  return L->has_local && (!L->h_edges.empty() || !L->h_corners.empty() || !L->h_planes.empty());
}

// The blocks whose post-processors must run together with L's: those joined
// to L through halos, directly or through one another, that have
// post-processors of their own.
std::vector<lbmh_lattice*> post_group(lbmh_lattice* L)
{
  std::vector<lbmh_lattice*> group{L};
  for (size_t k = 0; k < group.size(); ++k) {
    lbmh_lattice* X = group[k];
    auto visit = [&](lbmh_lattice* Y) {
      if (Y && has_post(Y) && std::find(group.begin(), group.end(), Y) == group.end()) group.push_back(Y);
    };
    for (lbmh_halo* H : X->halos_in) visit(const_cast<lbmh_lattice*>(H->src));
    for (lbmh_halo* H : X->halos_out) visit(H->dst);
  }
  return group;
}

// One PostStream stage of the local boundaries of one block: the plane
// finite-difference processors (stage 0), outer edges (1), outer corners (2),
// on the populations after the shift.
template <int Q, typename S>
int launch_post_stage(lbmh_lattice* L, int stage, hipStream_t stream)
{
  PostArgs<Q> a = post_args<Q>(L);
  post_forcing<Q>(L, a);
  if (stage == kStagePlanes && !L->h_planes.empty()) {
    using V = typename StorageTraits<S>::V;
    a.cells = L->d_planes;
    a.recs = L->d_plane_recs;
    a.count = uint32_t(L->h_planes.size());
    if (L->pressure_planes) {
      hipLaunchKernelGGL((plane_scratch_kernel<Q, S>), dim3(grid_for(a.count, 64 / kPostLanes)), dim3(64), 0, stream,
                         a, static_cast<V*>(L->d_scratch));
      LBMH_HIP(hipGetLastError());
      hipLaunchKernelGGL((plane_store_kernel<Q, S>), dim3(grid_for(a.count, 64)), dim3(64), 0, stream, a,
                         static_cast<const V*>(L->d_scratch));
    } else {
      hipLaunchKernelGGL((outer_post_kernel<Q, S, kPostPlane>), dim3(grid_for(a.count, 64 / kPostLanes)), dim3(64), 0,
                         stream, a);
    }
    LBMH_HIP(hipGetLastError());
  }
  if (stage == kStageEdges && !L->h_edges.empty()) {
    a.cells = L->d_edges;
    a.recs = L->d_edge_recs;
    a.count = uint32_t(L->h_edges.size());
    hipLaunchKernelGGL((outer_post_kernel<Q, S, kPostEdge>), dim3(grid_for(a.count, 64 / kPostLanes)), dim3(64), 0,
                       stream, a);
    LBMH_HIP(hipGetLastError());
  }
  if (stage == kStageCorners && !L->h_corners.empty()) {
    a.cells = L->d_corners;
    a.recs = L->d_corner_recs;
    a.count = uint32_t(L->h_corners.size());
    hipLaunchKernelGGL((outer_post_kernel<Q, S, kPostCorner>), dim3(grid_for(a.count, 64 / kPostLanes)), dim3(64), 0,
                       stream, a);
    LBMH_HIP(hipGetLastError());
  }
  return LBMH_OK;
}

// The PostStream post-processors of L and of every block they are coupled to
// (post_group), stage by stage: before each stage every block fetches the
// moments its processors of that stage read in other blocks, then every
// block runs the stage.  A neighbour in another block is thus read after the
// earlier stages and before its own stage, exactly as a neighbour inside the
// block -- the reference's blocks get the same by post-processing their
// overlap copies too (boundary::set covers the overlap cells within the
// padding, src/boundary/setBoundary.h:83-101, after the PostStream exchange,
// src/core/superLattice.hh:639-640).  The group runs once every member has
// reached the call (lbmh_stream of its step, or lbmh_initialize); a member
// that calls earlier only records it.
template <int Q, typename S>
int run_post(lbmh_lattice* L, hipStream_t stream, bool init)
{
  if (!has_post(L)) return LBMH_OK;
  if (init) L->init_step = L->steps;
  const std::vector<lbmh_lattice*> group = post_group(L);
  for (lbmh_lattice* X : group) {
    if (X->steps > L->steps) return fail(LBMH_E_STATE, "blocks of a split lattice are out of step");
    if (X->steps < L->steps) return LBMH_OK;
    if (init && X->init_step != L->steps) return LBMH_OK;
  }
  for (lbmh_lattice* X : group) {
    if (X->remote_pos.empty()) continue;
    const int rc = resolve_remote(X);
    if (rc != LBMH_OK) return rc;
  }
  for (int stage = 0; stage < kPostStages; ++stage) {
    const int rc = launch_remote_stage<Q, S>(L, group, stage, stream);
    if (rc != LBMH_OK) return rc;
    for (lbmh_lattice* X : group) {
      const int rc = launch_post_stage<Q, S>(X, stage, stream);
      if (rc != LBMH_OK) return rc;
    }
  }
  if (init) {
    for (lbmh_lattice* X : group) X->init_step = -1;
  }
  return LBMH_OK;
}

// Phases of a time step: the collision launches, and the stream (shift of
// every column, post-processors).  A block that exchanges overlap cells with
// other blocks runs them apart, with the copies in between.
enum StepPhase : int { kPhaseCollide = 1, kPhaseStream = 2, kPhaseBoth = 3 };

// Arguments of a whole-block pass with the block's current shifts and launch shape.
template <int Q, typename V>
StepArgs<Q, V> step_args(const lbmh_lattice* L)
{
  StepArgs<Q, V> a;
  a.pop = L->d_pop;
  a.stride = L->stride;
  a.n_padded = L->n_padded;
  a.ids = L->d_ids;
  a.ncore = L->ncore;
  a.nx = L->n[0];
  a.ny = L->n[1];
  a.nz = L->n[2];
  a.sx = L->size[1] * L->size[2];
  a.sy = L->size[2];
  a.div_z = FastDiv(uint32_t(L->n[2]));
  a.div_y = FastDiv(uint32_t(L->n[1]));
  a.periodic = uint32_t(L->periodic[0] | (L->periodic[1] << 1) | (L->periodic[2] << 2));
  a.xcd_group = uint32_t(L->launch.xcd_group);
  a.walls = L->d_walls;
  a.info = L->d_info;
  a.force = static_cast<const V*>(L->d_force);
  a.bp = bulk_params<V>(L);
  for (int i = 0; i < Q; ++i) a.shift[i] = uint32_t(L->shift[i]);
  return a;
}

// stream: shift_i -= c_i . projection (blockLattice.hh:317-339)
void advance_shifts(lbmh_lattice* L)
{
  const int64_t N = L->n_padded;
  for (int i = 0; i < L->q; ++i) {
    int64_t s = (L->shift[i] - L->dist[i]) % N;
    if (s < 0) s += N;
    L->shift[i] = s;
  }
}

template <int Q, typename S, int KIND>
int launch_kind(lbmh_lattice* L, hipStream_t stream, int phase)
{
  using V = typename StorageTraits<S>::V;
  const StepArgs<Q, V> a = step_args<Q, V>(L);
  const uint32_t grid = (L->ncore + kBlock - 1) / kBlock;
  // the dynamic LDS is never touched: it only limits resident workgroups per CU
  const unsigned lds = unsigned(L->launch.lds_bytes);
  if (phase & kPhaseCollide) {
  if (L->has_local) {
    // the boundary cells collide with their BGK mixin (OMEGA) whatever the bulk kind
    hipLaunchKernelGGL((collide_stream_kernel<Q, S, KIND, true, kLocalFlat>), dim3(grid), dim3(kBlock), lds, stream, a);
    LBMH_HIP(hipGetLastError());
    const uint32_t count = uint32_t(L->h_local.size());
    if (count) {
      hipLaunchKernelGGL((local_boundary_kernel<Q, S, KIND>), dim3(grid_for(count, 64)), dim3(64), 0, stream, a,
                         static_cast<const uint32_t*>(L->d_local), count);
    }
  } else if (L->uniform_bulk) {
    hipLaunchKernelGGL((collide_stream_kernel<Q, S, KIND, false>), dim3(grid), dim3(kBlock), lds, stream, a);
  } else {
    hipLaunchKernelGGL((collide_stream_kernel<Q, S, KIND, true>), dim3(grid), dim3(kBlock), lds, stream, a);
  }
  LBMH_HIP(hipGetLastError());
  }
  if (!(phase & kPhaseStream)) return LBMH_OK;
  advance_shifts(L);
  L->steps += 1;
  if (L->has_local) return run_post<Q, S>(L, stream, false);
  return LBMH_OK;
}

template <int Q, typename S>
int launch_step(lbmh_lattice* L, hipStream_t stream, int phase)
{
  switch (L->kind) {
  case LBMH_BGK: return launch_kind<Q, S, kBGK>(L, stream, phase);
  case LBMH_FORCED_BGK: return launch_kind<Q, S, kForcedBGK>(L, stream, phase);
  case LBMH_TRT: return launch_kind<Q, S, kTRT>(L, stream, phase);
  case LBMH_FORCED_TRT: return launch_kind<Q, S, kForcedTRT>(L, stream, phase);
  case LBMH_SMAGORINSKY: return launch_kind<Q, S, kSmagorinsky>(L, stream, phase);
  default: return fail(LBMH_E_INVALID, "unknown bulk kind");
  }
}

template <int Q>
int step_q(lbmh_lattice* L, hipStream_t stream, int phase = kPhaseBoth)
{
  switch (L->storage) {
  case LBMH_F64: return launch_step<Q, double>(L, stream, phase);
  case LBMH_F32: return launch_step<Q, float>(L, stream, phase);
  default: return launch_step<Q, _Float16>(L, stream, phase);
  }
}

// Dispatch of a functor over (Q, storage type).
template <typename F>
int with_types(const lbmh_lattice* L, F&& f)
{
  if (L->q == 19) {
    if (L->storage == LBMH_F64) return f(std::integral_constant<int, 19>(), double());
    if (L->storage == LBMH_F32) return f(std::integral_constant<int, 19>(), float());
    return f(std::integral_constant<int, 19>(), _Float16());
  }
  if (L->storage == LBMH_F64) return f(std::integral_constant<int, 27>(), double());
  if (L->storage == LBMH_F32) return f(std::integral_constant<int, 27>(), float());
  return f(std::integral_constant<int, 27>(), _Float16());
}


}  // namespace

extern "C" {

const char* lbmh_last_error(void) { return g_error.c_str(); }

int lbmh_abi_version(void) { return LBMH_ABI_VERSION; }

int lbmh_device_count(int* count)
{
  if (!count) return fail(LBMH_E_INVALID, "count is null");
  LBMH_HIP(hipGetDeviceCount(count));
  return LBMH_OK;
}

int lbmh_device_arch(int device, char* name, size_t len)
{
  if (!name || len == 0) return fail(LBMH_E_INVALID, "name buffer is null");
  hipDeviceProp_t prop;
  LBMH_HIP(hipGetDeviceProperties(&prop, device));
  std::snprintf(name, len, "%s", prop.gcnArchName);
  return LBMH_OK;
}

int lbmh_lattice_create(const lbmh_lattice_desc* desc, lbmh_lattice** out)
{
  if (!desc || !out) return fail(LBMH_E_INVALID, "null argument");
  *out = nullptr;
  if (desc->q != 19 && desc->q != 27) return fail(LBMH_E_INVALID, "q must be 19 or 27");
  if (desc->storage < LBMH_F64 || desc->storage > LBMH_F16) return fail(LBMH_E_INVALID, "unknown storage type");
  if (desc->bulk_kind < LBMH_BGK || desc->bulk_kind > LBMH_SMAGORINSKY) return fail(LBMH_E_INVALID, "unknown bulk kind");
  if (desc->q == 27 && desc->bulk_kind == LBMH_FORCED_TRT) {
    // keep the kernel set to the dynamics OpenLB defines for both descriptors
  }
  for (int d = 0; d < 3; ++d) {
    if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1) return fail(LBMH_E_INVALID, "extent must be >= 1");
  }
  const uint64_t sx = uint64_t(desc->nx) + 2, sy = uint64_t(desc->ny) + 2, sz = uint64_t(desc->nz) + 2;
  const uint64_t npad = sx * sy * sz;
  // 32-bit cell indices, with room for index + shift < 2^32 (CellID is uint32 in
  // OpenLB, src/core/blockStructure.h:36,80-82)
  if (npad >= (uint64_t(1) << 31)) return fail(LBMH_E_INVALID, "padded cell count must be below 2^31");
  // kernels address a population column with 32-bit byte offsets and wrap
  // them with an unsigned min, which needs a column below 2 GiB
  if (npad * storage_size(desc->storage) >= (uint64_t(1) << 31))
    return fail(LBMH_E_INVALID, "a population column must stay below 2 GiB (padded cells x bytes per population)");
  auto* L = new (std::nothrow) lbmh_lattice();
  if (!L) return fail(LBMH_E_MEMORY, "host allocation failed");
  L->device = desc->device;
  L->q = desc->q;
  L->n[0] = desc->nx;
  L->n[1] = desc->ny;
  L->n[2] = desc->nz;
  for (int d = 0; d < 3; ++d) L->periodic[d] = desc->periodic[d] ? 1 : 0;
  L->storage = desc->storage;
  L->kind = desc->bulk_kind;
  L->size[0] = uint32_t(sx);
  L->size[1] = uint32_t(sy);
  L->size[2] = uint32_t(sz);
  L->n_padded = uint32_t(npad);
  L->ncore = uint32_t(uint64_t(desc->nx) * desc->ny * desc->nz);
  // columns start on 256-byte boundaries
  const uint64_t align = 256 / storage_size(L->storage);
  L->stride = (npad + align - 1) / align * align;
  for (int i = 0; i < L->q; ++i) {
    L->shift[i] = 0;
    L->dist[i] = L->q == 19 ? neighbor_distance<19>(i, int64_t(sy * sz), int64_t(sz))
                            : neighbor_distance<27>(i, int64_t(sy * sz), int64_t(sz));
  }
  L->params.omega = 1.0;
  // measured best shapes on MI355X for the 256^3 D3Q19 step: FP16 columns
  // want full occupancy, FP32/FP64 three to four workgroups per CU
  L->launch = default_launch(L->storage, false);
  for (int w = 0; w < kMaxWalls; ++w) {
    L->walls[w][0] = 1.0;
    L->material_of[w] = uint8_t(w);
  }

  DeviceGuard guard(L->device);
  const size_t pop_bytes = size_t(L->stride) * L->q * storage_size(L->storage);
  L->n_partials = 1024;
  hipError_t e = hipMalloc(&L->d_pop, pop_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&L->d_ids, L->ncore);
  if (e == hipSuccess) e = hipMalloc((void**)&L->d_walls, sizeof(double) * 4 * kMaxWalls);
  if (e == hipSuccess) e = hipMalloc((void**)&L->d_partials, sizeof(double) * 4 * (L->n_partials + 1));
  if (e == hipSuccess) e = hipMalloc((void**)&L->d_info, sizeof(uint32_t) * kMaxWalls);
  if (e == hipSuccess) e = hipMemcpy(L->d_info, L->info, sizeof(L->info), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(L->d_pop, 0, pop_bytes);
  if (e == hipSuccess) e = hipMemset(L->d_ids, 0, L->ncore);
  if (e == hipSuccess) e = hipMemcpy(L->d_walls, L->walls, sizeof(L->walls), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    const int rc = hip_fail(e, "lbmh_lattice_create");
    lbmh_lattice_destroy(L);
    return rc == LBMH_E_DEVICE && e == hipErrorOutOfMemory ? fail(LBMH_E_MEMORY, g_error) : rc;
  }
  L->footprint = pop_bytes + L->ncore + sizeof(double) * 4 * (kMaxWalls + L->n_partials + 1);
  *out = L;
  return LBMH_OK;
}

int lbmh_lattice_destroy(lbmh_lattice* L)
{
  if (!L) return LBMH_OK;
  DeviceGuard guard(L->device);
  for (lbmh_halo* H : L->halos_in) H->dst = nullptr;
  for (lbmh_halo* H : L->halos_out) {
    H->src = nullptr;
    if (H->dst) H->dst->remote_resolved = H->dst->remote_pos.empty();
  }
  if (L->d_remote) (void)hipFree(L->d_remote);
  if (L->d_halo_batch) (void)hipFree(L->d_halo_batch);
  if (L->d_remote_batch) (void)hipFree(L->d_remote_batch);
  if (L->d_pop) (void)hipFree(L->d_pop);
  if (L->d_force) (void)hipFree(L->d_force);
  if (L->d_ids) (void)hipFree(L->d_ids);
  if (L->d_walls) (void)hipFree(L->d_walls);
  if (L->d_partials) (void)hipFree(L->d_partials);
  if (L->d_info) (void)hipFree(L->d_info);
  if (L->d_edges) (void)hipFree(L->d_edges);
  if (L->d_corners) (void)hipFree(L->d_corners);
  if (L->d_local) (void)hipFree(L->d_local);
  if (L->d_planes) (void)hipFree(L->d_planes);
  if (L->d_scratch) (void)hipFree(L->d_scratch);
  if (L->d_edge_recs) (void)hipFree(L->d_edge_recs);
  if (L->d_corner_recs) (void)hipFree(L->d_corner_recs);
  if (L->d_plane_recs) (void)hipFree(L->d_plane_recs);
  delete L;
  return LBMH_OK;
}

int lbmh_set_dynamics_ids(lbmh_lattice* L, const uint8_t* ids)
{
  if (!L || !ids) return fail(LBMH_E_INVALID, "null argument");
  DeviceGuard guard(L->device);
  bool uniform = true;
  for (uint32_t c = 0; c < L->ncore; ++c) {
    if (ids[c] != 1) {
      uniform = false;
      break;
    }
  }
  LBMH_HIP(hipMemcpy(L->d_ids, ids, L->ncore, hipMemcpyHostToDevice));
  L->uniform_bulk = uniform;
  // a new id map drops the local boundaries defined on the previous one
  L->h_ids.assign(ids, ids + L->ncore);
  L->has_local = false;
  if (!L->launch_chosen) L->launch = default_launch(L->storage, false);
  L->h_edges.clear();
  L->h_corners.clear();
  L->h_local.clear();
  L->h_planes.clear();
  L->pressure_planes = false;
  L->remote_pos.clear();
  L->remote_stage.clear();
  drop_remote(L);
  for (int w = 0; w < kMaxWalls; ++w) {
    L->info[w] = pack_local_info(kMovingWall, 0, 0, 0);
    L->id_defined[w] = false;
    L->material_of[w] = uint8_t(w);
  }
  LBMH_HIP(hipMemcpy(L->d_info, L->info, sizeof(L->info), hipMemcpyHostToDevice));
  return LBMH_OK;
}

int lbmh_set_bulk_params(lbmh_lattice* L, const lbmh_bulk_params* p)
{
  if (!L || !p) return fail(LBMH_E_INVALID, "null argument");
  if (!valid_bulk_params(L->kind, *p)) return LBMH_E_INVALID;
  L->params = *p;
  return LBMH_OK;
}

int lbmh_set_force_field(lbmh_lattice* L, const double* force)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  if (L->kind != LBMH_FORCED_BGK && L->kind != LBMH_FORCED_TRT)
    return fail(LBMH_E_INVALID, "a FORCE field needs a forced bulk kind (LBMH_FORCED_BGK / LBMH_FORCED_TRT)");
  DeviceGuard guard(L->device);
  LBMH_HIP(hipDeviceSynchronize());
  const size_t n = size_t(3) * L->ncore;
  const size_t es = L->storage == LBMH_F64 ? sizeof(double) : sizeof(float);
  if (!force) {
    if (L->d_force) {
      LBMH_HIP(hipFree(L->d_force));
      L->footprint -= n * es;
    }
    L->d_force = nullptr;
    L->h_force.clear();
    return LBMH_OK;
  }
  for (size_t k = 0; k < n; ++k)
    if (!std::isfinite(force[k])) return fail(LBMH_E_INVALID, "FORCE field holds a non-finite value");
  if (!L->d_force) {
    LBMH_HIP(hipMalloc(&L->d_force, n * es));
    L->footprint += n * es;
  }
  if (L->storage == LBMH_F64) {
    LBMH_HIP(hipMemcpy(L->d_force, force, n * es, hipMemcpyHostToDevice));
  } else {
    std::vector<float> v(force, force + n);
    LBMH_HIP(hipMemcpy(L->d_force, v.data(), n * es, hipMemcpyHostToDevice));
  }
  L->h_force.assign(force, force + n);
  return LBMH_OK;
}

int lbmh_set_wall(lbmh_lattice* L, int id, double rho, const double u[3])
{
  if (!L || !u) return fail(LBMH_E_INVALID, "null argument");
  if (id < LBMH_ID_FIRST_WALL || id >= kMaxWalls) return fail(LBMH_E_INVALID, "wall id must be in [3, 255]");
  DeviceGuard guard(L->device);
  L->walls[id][0] = rho;
  for (int d = 0; d < 3; ++d) L->walls[id][d + 1] = u[d];
  L->info[id] = pack_local_info(kMovingWall, 0, 0, 0);
  L->id_defined[id] = true;
  LBMH_HIP(hipMemcpy(L->d_walls + 4 * id, L->walls[id], sizeof(double) * 4, hipMemcpyHostToDevice));
  LBMH_HIP(hipMemcpy(L->d_info + id, &L->info[id], sizeof(uint32_t), hipMemcpyHostToDevice));
  return LBMH_OK;
}

int lbmh_classify_boundary(const int extent[3], const int periodic[3], const uint8_t* ids, int8_t* out)
{
  if (!extent || !periodic || !ids || !out) return fail(LBMH_E_INVALID, "null argument");
  if (extent[0] < 1 || extent[1] < 1 || extent[2] < 1) return fail(LBMH_E_INVALID, "extent must be >= 1");
  const int n[3] = {extent[0], extent[1], extent[2]};
  const int per[3] = {periodic[0] ? 1 : 0, periodic[1] ? 1 : 0, periodic[2] ? 1 : 0};
  const int64_t nc = int64_t(n[0]) * n[1] * n[2];
  for (int x = 0; x < n[0]; ++x)
    for (int y = 0; y < n[1]; ++y)
      for (int z = 0; z < n[2]; ++z) {
        const int64_t c = (int64_t(x) * n[1] + y) * n[2] + z;
        out[c] = -1;
        out[nc + c] = out[2 * nc + c] = out[3 * nc + c] = 0;
        if (ids[c] == 0 || ids[c] == 1) continue;
        int type, nrm[3];
        classify_boundary_cell(n, per, ids, x, y, z, type, nrm);
        out[c] = int8_t(type);
        for (int d = 0; d < 3; ++d) out[(d + 1) * nc + c] = int8_t(nrm[d]);
      }
  return LBMH_OK;
}

int lbmh_classify_boundary_overlap(const int extent[3], const int periodic[3], const uint8_t* ids,
                                   const uint8_t* ring, int8_t* out)
{
  if (!extent || !periodic || !ids || !ring || !out) return fail(LBMH_E_INVALID, "null argument");
  if (extent[0] < 1 || extent[1] < 1 || extent[2] < 1) return fail(LBMH_E_INVALID, "extent must be >= 1");
  const int n[3] = {extent[0], extent[1], extent[2]};
  const int per[3] = {periodic[0] ? 1 : 0, periodic[1] ? 1 : 0, periodic[2] ? 1 : 0};
  const int64_t nc = int64_t(n[0]) * n[1] * n[2];
  for (int x = 0; x < n[0]; ++x)
    for (int y = 0; y < n[1]; ++y)
      for (int z = 0; z < n[2]; ++z) {
        const int64_t c = (int64_t(x) * n[1] + y) * n[2] + z;
        out[c] = -1;
        out[nc + c] = out[2 * nc + c] = out[3 * nc + c] = 0;
        if (ids[c] == 0 || ids[c] == 1) continue;
        int type, nrm[3];
        classify_boundary_cell(n, per, ids, x, y, z, type, nrm, ring);
        out[c] = int8_t(type);
        for (int d = 0; d < 3; ++d) out[(d + 1) * nc + c] = int8_t(nrm[d]);
      }
  return LBMH_OK;
}

int lbmh_set_overlap_ids(lbmh_lattice* L, const uint8_t* ring)
{
  if (!L || !ring) return fail(LBMH_E_INVALID, "null argument");
  if (L->has_local) return fail(LBMH_E_STATE, "set the overlap ids before the local boundaries");
  L->h_ring.assign(ring, ring + L->n_padded);
  return LBMH_OK;
}

int lbmh_set_local_boundary(lbmh_lattice* L, int id, int kind, double rho, const double u[3])
{
  if (!L || !u) return fail(LBMH_E_INVALID, "null argument");
  if (kind < LBMH_LOCAL_VELOCITY || kind > LBMH_INTERP_PRESSURE) return fail(LBMH_E_INVALID, "unknown local boundary kind");
  if (id < 2 || id >= kMaxWalls) return fail(LBMH_E_INVALID, "boundary id must be in [2, 255]");
  if (L->h_ids.size() != L->ncore) return fail(LBMH_E_INVALID, "set the dynamics ids first");
  if ((kind == LBMH_LOCAL_PRESSURE || kind == LBMH_INTERP_PRESSURE) && !(rho > 0))
    return fail(LBMH_E_INVALID, "pressure boundary density must be positive");
  DeviceGuard guard(L->device);
  const int n[3] = {L->n[0], L->n[1], L->n[2]};
  const int per[3] = {L->periodic[0], L->periodic[1], L->periodic[2]};
  // ids free for the boundary's per-type entries: absent from the map and not yet defined
  bool used[kMaxWalls] = {};
  for (uint8_t v : L->h_ids) used[v] = true;
  for (int w = 0; w < kMaxWalls; ++w) used[w] = used[w] || L->id_defined[w];
  used[0] = used[1] = used[2] = true;
  std::map<std::tuple<int, int, int, int>, int> entry;  // (kind, a, b, c) -> id
  std::vector<uint8_t> next = L->h_ids;
  std::vector<uint32_t> edges, corners, cells, planes;
  int free_id = kMaxWalls - 1;
  for (int x = 0; x < n[0]; ++x)
    for (int y = 0; y < n[1]; ++y)
      for (int z = 0; z < n[2]; ++z) {
        const uint32_t c = uint32_t((int64_t(x) * n[1] + y) * n[2] + z);
        if (L->h_ids[c] != id) continue;
        int type, nrm[3];
        classify_boundary_cell(n, per, L->h_ids.data(), x, y, z, type, nrm,
                               L->h_ring.empty() ? nullptr : L->h_ring.data());
        const bool flat = type == kTypeFlat && (nrm[0] || nrm[1] || nrm[2]);
        const uint32_t k = local_kind(kind, type, flat);
        if (k == kInvalidKind) {
          return fail(LBMH_E_INVALID, kind == LBMH_LOCAL_PRESSURE ? "local pressure boundary on a non-flat wall cell"
                                                                 : "interpolated boundary on a wall cell without normal");
        }
        int a = nrm[0], b = nrm[1], cc = nrm[2];
        if (k == kOuterEdge || k == kInnerEdge) {
          // (plane, n[(plane+1)%3], n[(plane+2)%3]), plane the axis the normal
          // does not point along (setBoundary3D.h:583-625)
          const int plane = nrm[0] == 0 ? 0 : (nrm[1] == 0 ? 1 : 2);
          a = plane;
          b = nrm[(plane + 1) % 3];
          cc = nrm[(plane + 2) % 3];
        }
        const auto key = std::make_tuple(int(k), a, b, cc);
        auto it = entry.find(key);
        if (it == entry.end()) {
          while (free_id >= 3 && used[free_id]) --free_id;
          if (free_id < 3) return fail(LBMH_E_INVALID, "no free dynamics ids left for the boundary's cell types");
          used[free_id] = true;
          L->info[free_id] = pack_local_info(k, a, b, cc);
          L->material_of[free_id] = uint8_t(id);
          L->walls[free_id][0] = rho;
          for (int d = 0; d < 3; ++d) L->walls[free_id][d + 1] = u[d];
          L->id_defined[free_id] = true;
          it = entry.emplace(key, free_id).first;
        }
        next[c] = uint8_t(it->second);
        // faces, outer edges and corners are collided by the whole-lattice pass,
        // inner edges and corners (and cells without dynamics) by the list pass
        if (k == kInnerEdge || k == kInnerCorner || k == kDeadWall) cells.push_back(c);
        if (k == kFdVelocity || k == kFdPressure) planes.push_back(c);
        if (k == kFdPressure) L->pressure_planes = true;
        if (k == kOuterEdge) edges.push_back(c);
        if (k == kOuterCorner) corners.push_back(c);
      }
  if (entry.empty()) return LBMH_OK;
  L->h_ids.swap(next);
  L->h_edges.insert(L->h_edges.end(), edges.begin(), edges.end());
  L->h_corners.insert(L->h_corners.end(), corners.begin(), corners.end());
  L->h_local.insert(L->h_local.end(), cells.begin(), cells.end());
  L->h_planes.insert(L->h_planes.end(), planes.begin(), planes.end());
  LBMH_HIP(hipMemcpy(L->d_ids, L->h_ids.data(), L->ncore, hipMemcpyHostToDevice));
  LBMH_HIP(hipMemcpy(L->d_walls, L->walls, sizeof(L->walls), hipMemcpyHostToDevice));
  LBMH_HIP(hipMemcpy(L->d_info, L->info, sizeof(L->info), hipMemcpyHostToDevice));
  auto upload_list = [&](const std::vector<uint32_t>& h, uint32_t*& d) -> int {
    if (d) {
      LBMH_HIP(hipFree(d));
      d = nullptr;
    }
    if (h.empty()) return LBMH_OK;
    LBMH_HIP(hipMalloc((void**)&d, sizeof(uint32_t) * h.size()));
    LBMH_HIP(hipMemcpy(d, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice));
    return LBMH_OK;
  };
  int rc = upload_list(L->h_edges, L->d_edges);
  if (rc == LBMH_OK) rc = upload_list(L->h_corners, L->d_corners);
  if (rc == LBMH_OK) rc = upload_list(L->h_local, L->d_local);
  if (rc == LBMH_OK) rc = upload_list(L->h_planes, L->d_planes);
  // every list is resolved again: this boundary may have relabelled
  // neighbours of the cells of an earlier one
  auto upload_records = [&](const std::vector<uint32_t>& h, int kind, PostCell*& d) -> int {
    if (d) {
      LBMH_HIP(hipFree(d));
      d = nullptr;
    }
    if (h.empty()) return LBMH_OK;
    const std::vector<PostCell> recs = post_records(L, h, kind);
    LBMH_HIP(hipMalloc((void**)&d, sizeof(PostCell) * recs.size()));
    LBMH_HIP(hipMemcpy(d, recs.data(), sizeof(PostCell) * recs.size(), hipMemcpyHostToDevice));
    return LBMH_OK;
  };
  L->remote_pos.clear();
  L->remote_stage.clear();
  drop_remote(L);
  if (rc == LBMH_OK) rc = upload_records(L->h_edges, kPostEdge, L->d_edge_recs);
  if (rc == LBMH_OK) rc = upload_records(L->h_corners, kPostCorner, L->d_corner_recs);
  if (rc == LBMH_OK) rc = upload_records(L->h_planes, kPostPlane, L->d_plane_recs);
  L->remote_resolved = L->remote_pos.empty();
  if (rc == LBMH_OK && L->pressure_planes) {
    if (L->d_scratch) {
      LBMH_HIP(hipFree(L->d_scratch));
      L->d_scratch = nullptr;
    }
    const size_t elem = L->storage == LBMH_F64 ? sizeof(double) : sizeof(float);
    LBMH_HIP(hipMalloc(&L->d_scratch, elem * L->q * L->h_planes.size()));
  }
  if (rc != LBMH_OK) return rc;
  L->has_local = true;
  L->uniform_bulk = false;
  if (!L->launch_chosen) L->launch = default_launch(L->storage, true);
  return LBMH_OK;
}

int lbmh_initialize(lbmh_lattice* L)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  if (!L->has_local) return LBMH_OK;
  DeviceGuard guard(L->device);
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    return run_post<Q, S>(L, nullptr, true);
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_populations_upload(lbmh_lattice* L, const void* host, int host_dtype)
{
  if (!L || !host) return fail(LBMH_E_INVALID, "null argument");
  if (host_dtype < LBMH_F64 || host_dtype > LBMH_F16) return fail(LBMH_E_INVALID, "unknown host dtype");
  DeviceGuard guard(L->device);
  const size_t bytes = size_t(L->q) * L->ncore * storage_size(host_dtype);
  DeviceBuffer staging;
  LBMH_HIP(hipMalloc(&staging.p, bytes));
  LBMH_HIP(hipMemcpy(staging.p, host, bytes, hipMemcpyHostToDevice));
  const CoreGeom g = core_geom(L);
  const uint32_t grid = grid_for(L->ncore, 256);
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const Shifts<Q> sh = shifts_of<Q>(L);
    S* pop = static_cast<S*>(L->d_pop);
    if (host_dtype == LBMH_F64)
      hipLaunchKernelGGL((scatter_populations<Q, double, S>), dim3(grid), dim3(256), 0, 0, (const double*)staging.p, pop, g, sh);
    else if (host_dtype == LBMH_F32)
      hipLaunchKernelGGL((scatter_populations<Q, float, S>), dim3(grid), dim3(256), 0, 0, (const float*)staging.p, pop, g, sh);
    else
      hipLaunchKernelGGL((scatter_populations<Q, _Float16, S>), dim3(grid), dim3(256), 0, 0, (const _Float16*)staging.p, pop, g, sh);
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_populations_download(lbmh_lattice* L, void* host, int host_dtype)
{
  if (!L || !host) return fail(LBMH_E_INVALID, "null argument");
  if (host_dtype < LBMH_F64 || host_dtype > LBMH_F16) return fail(LBMH_E_INVALID, "unknown host dtype");
  DeviceGuard guard(L->device);
  const size_t bytes = size_t(L->q) * L->ncore * storage_size(host_dtype);
  DeviceBuffer staging;
  LBMH_HIP(hipMalloc(&staging.p, bytes));
  const CoreGeom g = core_geom(L);
  const uint32_t grid = grid_for(L->ncore, 256);
  LBMH_HIP(hipDeviceSynchronize());
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    const Shifts<Q> sh = shifts_of<Q>(L);
    const S* pop = static_cast<const S*>(L->d_pop);
    if (host_dtype == LBMH_F64)
      hipLaunchKernelGGL((gather_populations<Q, double, S>), dim3(grid), dim3(256), 0, 0, (double*)staging.p, pop, g, sh);
    else if (host_dtype == LBMH_F32)
      hipLaunchKernelGGL((gather_populations<Q, float, S>), dim3(grid), dim3(256), 0, 0, (float*)staging.p, pop, g, sh);
    else
      hipLaunchKernelGGL((gather_populations<Q, _Float16, S>), dim3(grid), dim3(256), 0, 0, (_Float16*)staging.p, pop, g, sh);
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipMemcpy(host, staging.p, bytes, hipMemcpyDeviceToHost));
  return LBMH_OK;
}

// One population column of a host block in OpenLB's storage order, with the
// host block's own overlap width: element k of the host column holds padded
// cell (k - shift) mod count.  Only the cells of the lattice's padded region
// (core plus one overlap layer) cross; the device column keeps its own shift.
namespace {

struct ColumnMap {
  int64_t count = 0;        // host column length
  std::vector<uint32_t> host_of;  // for each device padded cell (its padded id): host padded id
};

int column_map(const lbmh_lattice* L, int padding, ColumnMap& m)
{
  if (padding < 1) return fail(LBMH_E_INVALID, "host block padding must be at least 1");
  const int64_t P[3] = {L->n[0] + 2 * int64_t(padding), L->n[1] + 2 * int64_t(padding), L->n[2] + 2 * int64_t(padding)};
  m.count = P[0] * P[1] * P[2];
  if (m.count > 0xffffffffll) return fail(LBMH_E_INVALID, "host block too large");
  m.host_of.resize(L->n_padded);
  const int64_t off = padding - 1;
  for (uint32_t x = 0; x < L->size[0]; ++x)
    for (uint32_t y = 0; y < L->size[1]; ++y)
      for (uint32_t z = 0; z < L->size[2]; ++z) {
        const uint64_t d = (uint64_t(x) * L->size[1] + y) * L->size[2] + z;
        m.host_of[d] = uint32_t(((x + off) * P[1] + (y + off)) * P[2] + (z + off));
      }
  return LBMH_OK;
}

inline int64_t wrap64(int64_t k, int64_t n)
{
  k %= n;
  return k < 0 ? k + n : k;
}

}  // namespace

int lbmh_column_upload(lbmh_lattice* L, int i, const void* host, int64_t shift, int padding, int host_dtype)
{
  if (!L || !host) return fail(LBMH_E_INVALID, "null argument");
  if (i < 0 || i >= L->q) return fail(LBMH_E_INVALID, "population index out of range");
  if (host_dtype != LBMH_F64 && host_dtype != LBMH_F32) return fail(LBMH_E_INVALID, "column values must be F64 or F32");
  ColumnMap m;
  int rc = column_map(L, padding, m);
  if (rc != LBMH_OK) return rc;
  if (shift <= -m.count || shift >= m.count) return fail(LBMH_E_INVALID, "column shift out of (-count, count)");
  const int64_t n = L->n_padded, s_dev = L->shift[i];
  DeviceGuard guard(L->device);
  LBMH_HIP(hipDeviceSynchronize());
  return with_storage(L->storage, [&](auto s) -> int {
    using S = decltype(s);
    S* col = static_cast<S*>(L->d_pop) + uint64_t(i) * L->stride;
    auto put = [&](auto t) -> int {
      using T = decltype(t);
      const T* src = static_cast<const T*>(host);
      std::vector<T> dev(static_cast<size_t>(n), T(0));
      for (int64_t d = 0; d < n; ++d) dev[size_t(wrap64(d + s_dev, n))] = src[size_t(wrap64(m.host_of[size_t(d)] + shift, m.count))];
      if constexpr (std::is_same_v<S, T>) {
        LBMH_HIP(hipMemcpy(col, dev.data(), size_t(n) * sizeof(T), hipMemcpyHostToDevice));
      } else {
        DeviceBuffer staging;
        LBMH_HIP(hipMalloc(&staging.p, size_t(n) * sizeof(T)));
        LBMH_HIP(hipMemcpy(staging.p, dev.data(), size_t(n) * sizeof(T), hipMemcpyHostToDevice));
        hipLaunchKernelGGL((column_import_kernel<S, T>), dim3(grid_for(uint32_t(n), 256)), dim3(256), 0, 0, col,
                           static_cast<const T*>(staging.p), uint32_t(n));
        LBMH_HIP(hipGetLastError());
        LBMH_HIP(hipDeviceSynchronize());
      }
      return LBMH_OK;
    };
    return host_dtype == LBMH_F64 ? put(double{}) : put(float{});
  });
}

int lbmh_column_download(lbmh_lattice* L, int i, void* host, int64_t shift, int padding, int host_dtype)
{
  if (!L || !host) return fail(LBMH_E_INVALID, "null argument");
  if (i < 0 || i >= L->q) return fail(LBMH_E_INVALID, "population index out of range");
  if (host_dtype != LBMH_F64 && host_dtype != LBMH_F32) return fail(LBMH_E_INVALID, "column values must be F64 or F32");
  ColumnMap m;
  int rc = column_map(L, padding, m);
  if (rc != LBMH_OK) return rc;
  if (shift <= -m.count || shift >= m.count) return fail(LBMH_E_INVALID, "column shift out of (-count, count)");
  const int64_t n = L->n_padded, s_dev = L->shift[i];
  DeviceGuard guard(L->device);
  LBMH_HIP(hipDeviceSynchronize());
  return with_storage(L->storage, [&](auto s) -> int {
    using S = decltype(s);
    const S* col = static_cast<const S*>(L->d_pop) + uint64_t(i) * L->stride;
    auto get = [&](auto t) -> int {
      using T = decltype(t);
      std::vector<T> dev(static_cast<size_t>(n));
      if constexpr (std::is_same_v<S, T>) {
        LBMH_HIP(hipMemcpy(dev.data(), col, size_t(n) * sizeof(T), hipMemcpyDeviceToHost));
      } else {
        DeviceBuffer staging;
        LBMH_HIP(hipMalloc(&staging.p, size_t(n) * sizeof(T)));
        hipLaunchKernelGGL((column_export_kernel<S, T>), dim3(grid_for(uint32_t(n), 256)), dim3(256), 0, 0,
                           static_cast<T*>(staging.p), col, uint32_t(n));
        LBMH_HIP(hipGetLastError());
        LBMH_HIP(hipMemcpy(dev.data(), staging.p, size_t(n) * sizeof(T), hipMemcpyDeviceToHost));
      }
      T* dst = static_cast<T*>(host);
      for (int64_t d = 0; d < n; ++d) dst[size_t(wrap64(m.host_of[size_t(d)] + shift, m.count))] = dev[size_t(wrap64(d + s_dev, n))];
      return LBMH_OK;
    };
    return host_dtype == LBMH_F64 ? get(double{}) : get(float{});
  });
}

int lbmh_device_synchronize(int device)
{
  int count = 0;
  LBMH_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(LBMH_E_INVALID, "no such device");
  DeviceGuard guard(device);
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_ini_equilibrium(lbmh_lattice* L, const double* rho_u)
{
  if (!L || !rho_u) return fail(LBMH_E_INVALID, "null argument");
  DeviceGuard guard(L->device);
  const size_t bytes = size_t(4) * L->ncore * sizeof(double);
  DeviceBuffer staging;
  LBMH_HIP(hipMalloc(&staging.p, bytes));
  LBMH_HIP(hipMemcpy(staging.p, rho_u, bytes, hipMemcpyHostToDevice));
  const CoreGeom g = core_geom(L);
  const uint32_t grid = grid_for(L->ncore, 256);
  const int forced = (L->kind == LBMH_FORCED_BGK || L->kind == LBMH_FORCED_TRT) ? 1 : 0;
  const double* hf = L->params.force;
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    hipLaunchKernelGGL((ini_equilibrium_kernel<Q, S>), dim3(grid), dim3(256), 0, 0, (const double*)staging.p,
                       L->uniform_bulk ? nullptr : L->d_ids, static_cast<S*>(L->d_pop), g, shifts_of<Q>(L),
                       hf[0] * 0.5, hf[1] * 0.5, hf[2] * 0.5, forced, force_field<S>(L));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_step(lbmh_lattice* L, int64_t n_steps, void* stream)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  if (n_steps < 0) return fail(LBMH_E_INVALID, "negative step count");
  DeviceGuard guard(L->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int64_t k = 0; k < n_steps; ++k) {
    const int rc = L->q == 19 ? step_q<19>(L, s) : step_q<27>(L, s);
    if (rc != LBMH_OK) return rc;
  }
  return LBMH_OK;
}

int lbmh_pattern_ceiling(lbmh_lattice* L, int64_t n_steps, void* stream)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  if (n_steps < 0) return fail(LBMH_E_INVALID, "negative step count");
  DeviceGuard guard(L->device);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t grid = (L->ncore + kBlock - 1) / kBlock;
  for (int64_t k = 0; k < n_steps; ++k) {
    const int rc = with_types(L, [&](auto qc, auto s) -> int {
      constexpr int Q = decltype(qc)::value;
      using S = decltype(s);
      hipLaunchKernelGGL((pattern_ceiling_kernel<Q, S>), dim3(grid), dim3(kBlock), unsigned(L->launch.lds_bytes), st,
                         step_args<Q, float>(L));
      LBMH_HIP(hipGetLastError());
      return LBMH_OK;
    });
    if (rc != LBMH_OK) return rc;
    advance_shifts(L);
    L->steps += 1;
  }
  return LBMH_OK;
}

int lbmh_collide(lbmh_lattice* L, void* stream)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  DeviceGuard guard(L->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return L->q == 19 ? step_q<19>(L, s, kPhaseCollide) : step_q<27>(L, s, kPhaseCollide);
}

int lbmh_stream(lbmh_lattice* L, void* stream)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  DeviceGuard guard(L->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return L->q == 19 ? step_q<19>(L, s, kPhaseStream) : step_q<27>(L, s, kPhaseStream);
}

int lbmh_halo_create(lbmh_lattice* dst, const lbmh_lattice* src, const uint32_t* dst_cells,
                     const uint32_t* src_cells, size_t n, lbmh_halo** out)
{
  if (!dst || !src || !out || (n && (!dst_cells || !src_cells))) return fail(LBMH_E_INVALID, "null argument");
  *out = nullptr;
  if (dst->q != src->q || dst->storage != src->storage) return fail(LBMH_E_INVALID, "blocks differ in q or storage");
  if (dst->device != src->device) return fail(LBMH_E_INVALID, "blocks live on different devices");
  if (n > 0xffffffffu) return fail(LBMH_E_INVALID, "too many overlap cells");
  std::vector<uint2> pairs(n);
  for (size_t k = 0; k < n; ++k) {
    if (dst_cells[k] >= dst->n_padded || src_cells[k] >= src->n_padded)
      return fail(LBMH_E_INVALID, "cell id outside its padded block");
    pairs[k] = make_uint2(dst_cells[k], src_cells[k]);
  }
  auto* H = new (std::nothrow) lbmh_halo();
  if (!H) return fail(LBMH_E_MEMORY, "host allocation failed");
  H->dst = dst;
  H->src = src;
  H->device = dst->device;
  H->count = uint32_t(n);
  H->pairs = pairs;
  dst->halos_in.push_back(H);
  dst->remote_resolved = dst->remote_pos.empty();
  const_cast<lbmh_lattice*>(src)->halos_out.push_back(H);
  DeviceGuard guard(dst->device);
  if (n) {
    hipError_t e = hipMalloc((void**)&H->d_pairs, sizeof(uint2) * n);
    if (e == hipSuccess) e = hipMemcpy(H->d_pairs, pairs.data(), sizeof(uint2) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      const int rc = hip_fail(e, "lbmh_halo_create");
      lbmh_halo_destroy(H);
      return rc;
    }
  }
  *out = H;
  return LBMH_OK;
}

int lbmh_halo_destroy(lbmh_halo* H)
{
  if (!H) return LBMH_OK;
  DeviceGuard guard(H->device);
  if (H->dst) {
    H->dst->halos_in.erase(std::remove(H->dst->halos_in.begin(), H->dst->halos_in.end(), H),
                           H->dst->halos_in.end());
    H->dst->remote_resolved = H->dst->remote_pos.empty();
  }
  for (int st = 0; st < kPostStages; ++st) {
    if (H->d_remote[st]) (void)hipFree(H->d_remote[st]);
  }
  if (H->src) {
    auto& out = const_cast<lbmh_lattice*>(H->src)->halos_out;
    out.erase(std::remove(out.begin(), out.end(), H), out.end());
  }
  if (H->d_pairs) (void)hipFree(H->d_pairs);
  delete H;
  return LBMH_OK;
}

int lbmh_halo_copy(lbmh_halo* H, void* stream)
{
  if (!H) return fail(LBMH_E_INVALID, "null halo");
  if (!H->dst || !H->src) return fail(LBMH_E_INVALID, "a block of this halo was destroyed");
  if (!H->count) return LBMH_OK;
  const lbmh_lattice* D = H->dst;
  const lbmh_lattice* S_ = H->src;
  DeviceGuard guard(D->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return with_types(D, [&](auto qc, auto st) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(st);
    hipLaunchKernelGGL((halo_copy_kernel<Q, S>), dim3(grid_for(H->count, 256), Q), dim3(256), 0, s,
                       static_cast<S*>(D->d_pop), D->stride, D->n_padded, shifts_of<Q>(D),
                       static_cast<const S*>(S_->d_pop), S_->stride, S_->n_padded, shifts_of<Q>(S_), H->d_pairs,
                       H->count);
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
}

int lbmh_halo_copy_many(lbmh_halo* const* halos, size_t n, void* stream)
{
  if (n && !halos) return fail(LBMH_E_INVALID, "null argument");
  if (n > 0xffffu) return fail(LBMH_E_INVALID, "too many halos in one batch");
  uint64_t total = 0;
  for (size_t h = 0; h < n; ++h) {
    if (!halos[h]) return fail(LBMH_E_INVALID, "null halo");
    if (!halos[h]->dst || !halos[h]->src) return fail(LBMH_E_INVALID, "a block of this halo was destroyed");
    const lbmh_lattice* D = halos[h]->dst;
    const lbmh_lattice* D0 = halos[0]->dst;
    if (D->device != D0->device || D->q != D0->q || D->storage != D0->storage)
      return fail(LBMH_E_INVALID, "halos of one batch must share the device, q and storage");
    total += halos[h]->count;
  }
  if (total == 0) return LBMH_OK;
  if (total > 0xffffffffu) return fail(LBMH_E_INVALID, "too many overlap cells in one batch");
  lbmh_lattice* owner = halos[0]->dst;
  DeviceGuard guard(owner->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return with_types(owner, [&](auto qc, auto st) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(st);
    using Desc = HaloCopyDesc<Q, S>;
    std::vector<Desc> descs;
    descs.reserve(n);
    uint32_t first = 0;
    for (size_t h = 0; h < n; ++h) {
      const lbmh_halo* H = halos[h];
      if (!H->count) continue;
      Desc d{};
      d.dst = static_cast<S*>(H->dst->d_pop);
      d.src = static_cast<const S*>(H->src->d_pop);
      d.pairs = H->d_pairs;
      d.dst_stride = H->dst->stride;
      d.src_stride = H->src->stride;
      d.dst_n = H->dst->n_padded;
      d.src_n = H->src->n_padded;
      d.first = first;
      d.count = H->count;
      for (int i = 0; i < Q; ++i) {
        d.dst_shift[i] = uint32_t(H->dst->shift[i]);
        d.src_shift[i] = uint32_t(H->src->shift[i]);
      }
      descs.push_back(d);
      first += H->count;
    }
    // the table lives with the first block; the previous batch's kernel on
    // this stream has read it before the copy below runs
    const size_t bytes = sizeof(Desc) * descs.size();
    if (owner->halo_batch_bytes < bytes) {
      if (owner->d_halo_batch) LBMH_HIP(hipFree(owner->d_halo_batch));
      owner->d_halo_batch = nullptr;
      owner->halo_batch_bytes = 0;
      LBMH_HIP(hipMalloc(&owner->d_halo_batch, bytes));
      owner->halo_batch_bytes = bytes;
    }
    LBMH_HIP(hipMemcpyAsync(owner->d_halo_batch, descs.data(), bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((halo_copy_many_kernel<Q, S>), dim3(grid_for(first, 256), Q), dim3(256), 0, s,
                       static_cast<const Desc*>(owner->d_halo_batch), uint32_t(descs.size()), first);
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
}

int lbmh_get_step_count(const lbmh_lattice* L, int64_t* steps)
{
  if (!L || !steps) return fail(LBMH_E_INVALID, "null argument");
  *steps = L->steps;
  return LBMH_OK;
}

int lbmh_get_shifts(const lbmh_lattice* L, int64_t* shifts)
{
  if (!L || !shifts) return fail(LBMH_E_INVALID, "null argument");
  for (int i = 0; i < L->q; ++i) shifts[i] = L->shift[i];
  return LBMH_OK;
}

namespace {

// Role (CheckpointRole) of dynamics id `id` of lattice S
int role_in(const lbmh_lattice* S, int id)
{
  CheckpointBlock c;
  c.info = S->info;
  c.n_ids = kMaxWalls;
  return checkpoint_role(c, id);
}

CheckpointBlock checkpoint_block(const lbmh_lattice* L, std::vector<uint8_t>& no_ids, std::vector<int8_t>& overlap)
{
  CheckpointBlock b;
  b.q = L->q;
  for (int d = 0; d < 3; ++d) {
    b.n[d] = L->n[d];
    b.periodic[d] = L->periodic[d];
    b.force[d] = L->params.force[d];
  }
  b.storage = L->storage;
  b.kind = L->kind;
  b.n_padded = L->n_padded;
  b.stride = L->stride;
  b.d_pop = L->d_pop;
  for (int i = 0; i < L->q; ++i) {
    b.shift[i] = L->shift[i];
    b.dist[i] = L->dist[i];
  }
  if (L->h_ids.size() == L->ncore) {
    b.ids = L->h_ids.data();
  } else {
    no_ids.assign(L->ncore, 0);  // ids never set: every cell has id 0
    b.ids = no_ids.data();
  }
  b.walls = &L->walls[0][0];
  b.info = L->info;
  b.force_field = L->h_force.empty() ? nullptr : L->h_force.data();
  b.n_ids = kMaxWalls;
  if (!L->halos_in.empty()) {
    overlap.assign(L->n_padded, int8_t(-1));
    for (const lbmh_halo* H : L->halos_in) {
      const lbmh_lattice* S = H->src;
      if (!S) continue;
      for (const uint2& p : H->pairs) {
        const uint32_t sy = uint32_t(S->n[1] + 2), sz = uint32_t(S->n[2] + 2);
        const int x = int(p.y / (sy * sz)) - 1, y = int(p.y / sz % sy) - 1, z = int(p.y % sz) - 1;
        const bool core = x >= 0 && x < S->n[0] && y >= 0 && y < S->n[1] && z >= 0 && z < S->n[2];
        const size_t c = (size_t(x) * S->n[1] + y) * S->n[2] + z;
        overlap[p.x] = int8_t(core && S->h_ids.size() == S->ncore ? role_in(S, S->h_ids[c]) : kRoleNone);
      }
    }
    b.overlap_role = overlap.data();
  }
  return b;
}

}  // namespace

int lbmh_checkpoint_size(const lbmh_lattice* L, int real, size_t* bytes)
{
  if (!L || !bytes) return fail(LBMH_E_INVALID, "null argument");
  if (real != LBMH_F64 && real != LBMH_F32) return fail(LBMH_E_INVALID, "checkpoint value type must be F64 or F32");
  std::vector<uint8_t> no_ids;
  std::vector<int8_t> overlap;
  *bytes = checkpoint_bytes(checkpoint_block(L, no_ids, overlap), real);
  return LBMH_OK;
}

int lbmh_checkpoint_save(lbmh_lattice* L, int real, void* buffer, size_t bytes)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  DeviceGuard guard(L->device);
  std::vector<uint8_t> no_ids;
  std::vector<int8_t> overlap;
  return checkpoint_save(checkpoint_block(L, no_ids, overlap), real, buffer, bytes);
}

int lbmh_checkpoint_load(lbmh_lattice* L, int real, const void* buffer, size_t bytes)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  DeviceGuard guard(L->device);
  std::vector<uint8_t> no_ids;
  std::vector<int8_t> overlap;
  int64_t shift[27];
  const int rc = checkpoint_load(checkpoint_block(L, no_ids, overlap), real, buffer, bytes, shift);
  if (rc != LBMH_OK) return rc;
  for (int i = 0; i < L->q; ++i) L->shift[i] = shift[i];
  return LBMH_OK;
}

int lbmh_field_export(lbmh_lattice* L, int field, int overlap, const double conversion[3], float* host, size_t floats)
{
  if (!L || !host) return fail(LBMH_E_INVALID, "null argument");
  if (field < LBMH_FIELD_DENSITY || field > LBMH_FIELD_PHYS_PRESSURE) return fail(LBMH_E_INVALID, "unknown field");
  if (overlap < 0 || overlap > 1) return fail(LBMH_E_INVALID, "overlap must be 0 or 1 (the block's padding)");
  if ((field == LBMH_FIELD_PHYS_VELOCITY || field == LBMH_FIELD_PHYS_PRESSURE) && !conversion)
    return fail(LBMH_E_INVALID, "physical fields need the unit conversion factors");
  const uint64_t points = uint64_t(L->n[0] + 2 * overlap) * (L->n[1] + 2 * overlap) * (L->n[2] + 2 * overlap);
  const size_t want = size_t(points) * field_components(field);
  if (floats != want)
    return fail(LBMH_E_INVALID, "field buffer holds " + std::to_string(floats) + " floats, the export " +
                                    std::to_string(want));
  if (points > 0xffffffffull || !field_export_grid_fits(field_export_blocks(L->n[0], L->n[1], L->n[2], overlap)))
    return fail(LBMH_E_INVALID, "block too large for one export (its 32z x 8x tiles exceed 2^32 work-items)");
  DeviceGuard guard(L->device);
  DeviceBuffer out;
  LBMH_HIP(hipMalloc(&out.p, want * sizeof(float)));
  FieldConversion conv{0.0, 0.0, 0.0};
  if (conversion) conv = FieldConversion{conversion[0], conversion[1], conversion[2]};
  const int forced = (L->kind == LBMH_FORCED_BGK || L->kind == LBMH_FORCED_TRT) ? 1 : 0;
  const double* F = L->params.force;
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    hipLaunchKernelGGL((field_export_kernel<Q, S>),
                       dim3(uint32_t(field_export_blocks(L->n[0], L->n[1], L->n[2], overlap))),
                       dim3(kExportTileZ * kExportTileX), 0, 0, post_args<Q>(L), overlap, field, conv, forced, F[0] * 0.5, F[1] * 0.5, F[2] * 0.5,
                       force_field<S>(L), static_cast<float*>(out.p), uint32_t(points));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipMemcpy(host, out.p, want * sizeof(float), hipMemcpyDeviceToHost));
  return LBMH_OK;
}

int lbmh_reduce_field(lbmh_lattice* L, int field, int op, const uint8_t* materials, const double conversion[3],
                      double out[3], uint64_t* cells)
{
  if (!L || !out) return fail(LBMH_E_INVALID, "null argument");
  if (field < LBMH_FIELD_DENSITY || field > LBMH_FIELD_FORCE) return fail(LBMH_E_INVALID, "unknown field");
  if (op < LBMH_REDUCE_SUM || op > LBMH_REDUCE_MAX) return fail(LBMH_E_INVALID, "unknown reduction");
  if ((field == LBMH_FIELD_PHYS_VELOCITY || field == LBMH_FIELD_PHYS_PRESSURE) && !conversion)
    return fail(LBMH_E_INVALID, "physical fields need the unit conversion factors");
  IdMask mask{};
  for (int id = 0; id < 256; ++id) {
    const int m = L->material_of[id];
    if (materials ? materials[m] != 0 : m != 0) mask.bits[id >> 5] |= 1u << (id & 31);
  }
  FieldConversion conv{0.0, 0.0, 0.0};
  if (conversion) conv = FieldConversion{conversion[0], conversion[1], conversion[2]};
  const int forced = (L->kind == LBMH_FORCED_BGK || L->kind == LBMH_FORCED_TRT) ? 1 : 0;
  const double* F = L->params.force;
  DeviceGuard guard(L->device);
  const uint32_t blocks = std::min<uint32_t>(L->n_partials, grid_for(L->ncore, 256));
  LBMH_HIP(hipDeviceSynchronize());
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    hipLaunchKernelGGL((field_reduce_kernel<Q, S>), dim3(blocks), dim3(256), 0, 0, post_args<Q>(L), field, op, conv,
                       forced, F[0] * 0.5, F[1] * 0.5, F[2] * 0.5, force_field<S>(L), mask, L->d_partials);
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  double* result = L->d_partials + 4 * L->n_partials;
  hipLaunchKernelGGL(field_reduce_partials_kernel, dim3(1), dim3(256), 0, 0, result, L->d_partials, blocks, op);
  LBMH_HIP(hipGetLastError());
  double r[4];
  LBMH_HIP(hipMemcpy(r, result, sizeof(r), hipMemcpyDeviceToHost));
  const int comps = (field == LBMH_FIELD_VELOCITY || field == LBMH_FIELD_PHYS_VELOCITY || field == LBMH_FIELD_FORCE) ? 3 : 1;
  for (int d = 0; d < comps; ++d) out[d] = r[d];
  if (cells) *cells = uint64_t(r[3]);
  return LBMH_OK;
}

int lbmh_moments(lbmh_lattice* L, double* rho_u)
{
  if (!L || !rho_u) return fail(LBMH_E_INVALID, "null argument");
  DeviceGuard guard(L->device);
  const size_t bytes = size_t(4) * L->ncore * sizeof(double);
  DeviceBuffer out;
  LBMH_HIP(hipMalloc(&out.p, bytes));
  const CoreGeom g = core_geom(L);
  const uint32_t grid = grid_for(L->ncore, 256);
  const int forced = (L->kind == LBMH_FORCED_BGK || L->kind == LBMH_FORCED_TRT) ? 1 : 0;
  const double* hf = L->params.force;
  LBMH_HIP(hipDeviceSynchronize());
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    hipLaunchKernelGGL((moments_kernel<Q, S>), dim3(grid), dim3(256), 0, 0, (double*)out.p,
                       L->uniform_bulk ? nullptr : L->d_ids, static_cast<const S*>(L->d_pop), g, shifts_of<Q>(L),
                       hf[0] * 0.5, hf[1] * 0.5, hf[2] * 0.5, forced, force_field<S>(L));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  LBMH_HIP(hipMemcpy(rho_u, out.p, bytes, hipMemcpyDeviceToHost));
  return LBMH_OK;
}

int lbmh_reduce_moments(lbmh_lattice* L, double out[4])
{
  if (!L || !out) return fail(LBMH_E_INVALID, "null argument");
  DeviceGuard guard(L->device);
  const CoreGeom g = core_geom(L);
  const uint32_t blocks = std::min<uint32_t>(L->n_partials, grid_for(L->ncore, kBlock));
  LBMH_HIP(hipDeviceSynchronize());
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    hipLaunchKernelGGL((reduce_moments_kernel<Q, S>), dim3(blocks), dim3(kBlock), 0, 0, L->d_partials,
                       L->uniform_bulk ? nullptr : L->d_ids, static_cast<const S*>(L->d_pop), g, shifts_of<Q>(L));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  double* result = L->d_partials + 4 * L->n_partials;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kBlock), 0, 0, result, L->d_partials, blocks);
  LBMH_HIP(hipGetLastError());
  LBMH_HIP(hipMemcpy(out, result, sizeof(double) * 4, hipMemcpyDeviceToHost));
  return LBMH_OK;
}

int lbmh_reduce_moments_noalloc(lbmh_lattice* L, double* out_device, void* stream)
{
  if (!L || !out_device) return fail(LBMH_E_INVALID, "null argument");
  DeviceGuard guard(L->device);
  // the same workgroup count as lbmh_reduce_moments, so the sums agree bit for bit
  const uint32_t blocks = std::min<uint32_t>(L->n_partials, grid_for(L->ncore, kBlock));
  if (blocks > kScratchPartials)
    return fail(LBMH_E_INVALID, "reduction needs " + std::to_string(blocks) + " partial sums, the workspace holds " +
                                    std::to_string(kScratchPartials));
  double* partials = nullptr;
  const int rs = scratch_address(L->device, &partials);
  if (rs != LBMH_OK) return rs;
  const CoreGeom g = core_geom(L);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // every one of the `blocks` partials is written before it is read: nothing to clear
  const int rc = with_types(L, [&](auto qc, auto st) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(st);
    hipLaunchKernelGGL((reduce_moments_kernel<Q, S>), dim3(blocks), dim3(kBlock), 0, s, partials,
                       L->uniform_bulk ? nullptr : L->d_ids, static_cast<const S*>(L->d_pop), g, shifts_of<Q>(L));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kBlock), 0, s, out_device, partials, blocks);
  LBMH_HIP(hipGetLastError());
  return LBMH_OK;
}

int lbmh_statistics(lbmh_lattice* L, double out[4])
{
  if (!L || !out) return fail(LBMH_E_INVALID, "null argument");
  DeviceGuard guard(L->device);
  const CoreGeom g = core_geom(L);
  const uint32_t blocks = std::min<uint32_t>(L->n_partials, grid_for(L->ncore, kBlock));
  const int forced = (L->kind == LBMH_FORCED_BGK || L->kind == LBMH_FORCED_TRT) ? 1 : 0;
  const double* hf = L->params.force;
  LBMH_HIP(hipDeviceSynchronize());
  const int rc = with_types(L, [&](auto qc, auto s) -> int {
    constexpr int Q = decltype(qc)::value;
    using S = decltype(s);
    hipLaunchKernelGGL((statistics_kernel<Q, S>), dim3(blocks), dim3(kBlock), 0, 0, L->d_partials,
                       L->uniform_bulk ? nullptr : L->d_ids, static_cast<const S*>(L->d_pop), g, shifts_of<Q>(L),
                       static_cast<const double*>(L->d_walls), L->has_local ? L->d_info : nullptr,
                       L->params.omega, hf[0] * 0.5, hf[1] * 0.5, hf[2] * 0.5, forced, force_field<S>(L));
    LBMH_HIP(hipGetLastError());
    return LBMH_OK;
  });
  if (rc != LBMH_OK) return rc;
  double* result = L->d_partials + 4 * L->n_partials;
  hipLaunchKernelGGL(statistics_partials_kernel, dim3(1), dim3(kBlock), 0, 0, result, L->d_partials, blocks);
  LBMH_HIP(hipGetLastError());
  double acc[4];
  LBMH_HIP(hipMemcpy(acc, result, sizeof(double) * 4, hipMemcpyDeviceToHost));
  out[0] = acc[0];
  out[1] = acc[0] > 0 ? acc[1] / acc[0] : 0.0;
  out[2] = acc[0] > 0 ? 0.5 * acc[2] / acc[0] : 0.0;
  out[3] = std::sqrt(acc[3]);
  return LBMH_OK;
}

int lbmh_synchronize(lbmh_lattice* L)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  DeviceGuard guard(L->device);
  LBMH_HIP(hipDeviceSynchronize());
  return LBMH_OK;
}

int lbmh_set_launch_config(lbmh_lattice* L, const lbmh_launch_config* cfg)
{
  if (!L || !cfg) return fail(LBMH_E_INVALID, "null argument");
  if (cfg->lds_bytes < 0 || cfg->lds_bytes > 65536) return fail(LBMH_E_INVALID, "lds_bytes must be in [0, 65536]");
  if (cfg->xcd_group < 0 || cfg->xcd_group > 64) return fail(LBMH_E_INVALID, "xcd_group must be in [0, 64]");
  L->launch = *cfg;
  L->launch_chosen = true;
  return LBMH_OK;
}

int lbmh_get_launch_config(const lbmh_lattice* L, lbmh_launch_config* cfg)
{
  if (!L || !cfg) return fail(LBMH_E_INVALID, "null argument");
  *cfg = L->launch;
  return LBMH_OK;
}

int lbmh_autotune(lbmh_lattice* L, int trial_steps, lbmh_launch_config* chosen)
{
  if (!L) return fail(LBMH_E_INVALID, "null lattice");
  if (trial_steps < 1) return fail(LBMH_E_INVALID, "trial_steps must be >= 1");
  DeviceGuard guard(L->device);
  lbmh_lattice_desc desc{};
  desc.q = L->q;
  desc.nx = L->n[0];
  desc.ny = L->n[1];
  desc.nz = L->n[2];
  for (int d = 0; d < 3; ++d) desc.periodic[d] = L->periodic[d];
  desc.storage = L->storage;
  desc.bulk_kind = L->kind;
  desc.device = L->device;
  lbmh_lattice* T = nullptr;
  int rc = lbmh_lattice_create(&desc, &T);
  if (rc != LBMH_OK) return rc;
  struct Cleanup {
    lbmh_lattice* t;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Cleanup()
    {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      lbmh_lattice_destroy(t);
    }
  } cleanup{T};
  LBMH_HIP(hipMemcpy(T->d_ids, L->d_ids, L->ncore, hipMemcpyDeviceToDevice));
  LBMH_HIP(hipMemcpy(T->d_walls, L->d_walls, sizeof(double) * 4 * kMaxWalls, hipMemcpyDeviceToDevice));
  LBMH_HIP(hipMemcpy(T->d_info, L->d_info, sizeof(uint32_t) * kMaxWalls, hipMemcpyDeviceToDevice));
  T->uniform_bulk = L->uniform_bulk;
  T->has_local = L->has_local;  // times the whole-lattice pass; the clone has no boundary lists
  T->params = L->params;
  if (!L->h_force.empty()) {  // the forced step reads the FORCE field: time that too
    rc = lbmh_set_force_field(T, L->h_force.data());
    if (rc != LBMH_OK) return rc;
  }
  LBMH_HIP(hipEventCreate(&cleanup.e0));
  LBMH_HIP(hipEventCreate(&cleanup.e1));
  static const int kLds[] = {0, 24576, 32768, 40960, 49152, 65536};
  static const int kGroup[] = {0, 2, 4, 8, 16};
  float best_ms = 0.f;
  lbmh_launch_config best = L->launch;
  bool have = false;
  for (int group : kGroup) {
    for (int lds : kLds) {
      T->launch = lbmh_launch_config{lds, group};
      rc = lbmh_step(T, 2, nullptr);
      if (rc != LBMH_OK) return rc;
      LBMH_HIP(hipEventRecord(cleanup.e0, nullptr));
      rc = lbmh_step(T, trial_steps, nullptr);
      if (rc != LBMH_OK) return rc;
      LBMH_HIP(hipEventRecord(cleanup.e1, nullptr));
      LBMH_HIP(hipEventSynchronize(cleanup.e1));
      float ms = 0.f;
      LBMH_HIP(hipEventElapsedTime(&ms, cleanup.e0, cleanup.e1));
      if (!have || ms < best_ms) {
        best_ms = ms;
        best = T->launch;
        have = true;
      }
    }
  }
  L->launch = best;
  L->launch_chosen = true;
  if (chosen) *chosen = best;
  return LBMH_OK;
}

int lbmh_lattice_footprint(const lbmh_lattice* L, size_t* bytes)
{
  if (!L || !bytes) return fail(LBMH_E_INVALID, "null argument");
  *bytes = L->footprint;
  return LBMH_OK;
}

}  // extern "C"
