// Local (regularized) velocity and pressure boundaries, and the finite-
// difference post-processors OpenLB attaches to their outer edges and corners.
//
// OpenLB assigns these per wall cell from its discrete normal
// (src/boundary/localVelocity3D.h:36-116, localPressure3D.h:36-66):
//   flat face        CombinedRLBdynamics<BGK, RegularizedVelocityBoundaryTuple<dir,orient>>
//                    (pressure: RegularizedPressureBoundaryTuple<dir,orient>)
//   outer edge       BGK with FixedVelocityBoundaryTuple + OuterVelocityEdgeProcessor3D
//   outer corner     BGK with FixedVelocityBoundaryTuple + OuterVelocityCornerProcessor3D
//   inner edge       CombinedRLBdynamics<BGK, InnerEdgeVelocityTuple3D<plane,n1,n2>>
//   inner corner     CombinedRLBdynamics<BGK, InnerCornerVelocityTuple3D<nx,ny,nz>>
// and boundary::InterpolatedVelocity / InterpolatedPressure (interpolatedVelocity3D.h:
// 34-90, interpolatedPressure3D.h:37-90) give flat faces plain BGK with Dirichlet
// momenta plus a PlaneFdBoundaryProcessor3D (boundaryPostProcessors3D.hh:38-92).
// Here each such cell carries a dynamics id >= 3 whose entry holds the kind,
// the normal (or plane and normals for edges) and the defined rho / u.  The
// collisions run inside the fused step kernel; the edge and corner
// post-processors run afterwards, as OpenLB's PostStream stage does, over
// lists of those cells (edges first, priority 0, then corners, priority 1).
//
// These cells are a thin surface layer, so the code favours following
// OpenLB's operation order over the pair-wise deviation form of the bulk path:
// the equilibrium is rho t_i (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) - t_i
// (lbm.h:50-64) and every sum runs over populations in index order.
#pragma once

#include "fastdiv.hpp"
#include "lbm_collision.hpp"
#include "lbm_storage.hpp"

namespace lbmh {

enum LocalKind : uint32_t {
  kMovingWall = 0,     // BounceBackVelocity (per-id rho, u)
  kFlatVelocity = 1,
  kFlatPressure = 2,
  kOuterEdge = 3,
  kOuterCorner = 4,
  kInnerEdge = 5,
  kInnerCorner = 6,
  kDeadWall = 7,       // wall cell without a matching boundary type: no dynamics
  kFdVelocity = 8,     // InterpolatedVelocity face: BGK with BasicDirichletVelocityBoundaryTuple
  kFdPressure = 9      // InterpolatedPressure face: BGK with BasicDirichletPressureBoundaryTuple
};

constexpr uint32_t kInvalidKind = 0xffu;

// Dynamics kind a wall cell gets from its boundary type (0 flat, 1 external
// corner, 2 internal corner, 3 external edge, 4 internal edge) under boundary
// condition `bc` (0 LocalVelocity, 1 LocalPressure, 2 InterpolatedVelocity,
// 3 InterpolatedPressure); `flat` means a flat type with a non-zero normal.
// kInvalidKind where OpenLB throws (localPressure3D.h:46-58 for non-flat
// cells; a flat cell without normal under the interpolated boundaries,
// setBoundary3D.h:199-223).
inline uint32_t local_kind(int bc, int type, bool flat)
{
  static const uint32_t kOthers[5] = {0, kOuterCorner, kInnerCorner, kOuterEdge, kInnerEdge};
  switch (bc) {
  case 1: return flat ? kFlatPressure : kInvalidKind;
  case 3:
    if (type == 0) return flat ? kFdPressure : kInvalidKind;
    return kDeadWall;  // interpolatedPressure3D.h: no dynamics off flat faces
  case 2:
    if (type == 0) return flat ? kFdVelocity : kInvalidKind;
    return kOthers[type];
  default:
    if (type == 0) return flat ? kFlatVelocity : kDeadWall;
    return kOthers[type];
  }
}

// Per-id packed description: kind in bits 0-7, then three signed 4-bit
// fields (normal components, or plane / n1 / n2 for edges) stored + 8.
struct LocalInfo {
  uint32_t kind;
  int a, b, c;
};

__host__ __device__ inline uint32_t pack_local_info(uint32_t kind, int a, int b, int c)
{
  return kind | (uint32_t(a + 8) << 8) | (uint32_t(b + 8) << 12) | (uint32_t(c + 8) << 16);
}

__host__ __device__ inline LocalInfo unpack_local_info(uint32_t w)
{
  LocalInfo r;
  r.kind = w & 0xffu;
  r.a = int((w >> 8) & 0xfu) - 8;
  r.b = int((w >> 12) & 0xfu) - 8;
  r.c = int((w >> 16) & 0xfu) - 8;
  return r;
}

// Component d of velocity i for a run-time axis d (i is a compile-time
// constant after unrolling, so this is a select among three constants).
template <int Q>
__device__ __forceinline__ int cdir(int i, int d)
{
  using L = Lattice<Q>;
  return d == 0 ? L::c[i][0] : (d == 1 ? L::c[i][1] : L::c[i][2]);
}

// equilibrium<D>::secondOrder on shifted populations      lbm.h:50-64
template <int Q, typename V>
__device__ __forceinline__ V feq_shifted(int i, V rho, const V (&u)[3], V usq)
{
  using L = Lattice<Q>;
  V cu = V(0);
#pragma unroll
  for (int d = 0; d < 3; ++d) cu += V(L::c[i][d]) * u[d];
  const V t = weight<V, Q>(i);
  return rho * t * (V(1) + V(3) * cu + V(3) * V(3) * V(0.5) * cu * cu - V(3) * V(0.5) * usq) - t;
}

template <typename V>
__device__ __forceinline__ V norm_sqr(const V (&u)[3])
{
  return u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
}

// velocityBMRho<dir,orient>: density from the populations with c[dir] == 0
// and c[dir] == orient                                  momenta/elements.h:54-70
template <int Q, typename V>
__device__ __forceinline__ V bm_rho(const V (&f)[Q], int dir, int orient, const V (&u)[3])
{
  using L = Lattice<Q>;
  V on = V(0), nrm = V(0);
#pragma unroll
  for (int i = 0; i < Q; ++i)
    if (cdir<Q>(i, dir) == 0) on += f[i];
#pragma unroll
  for (int i = 0; i < Q; ++i)
    if (cdir<Q>(i, dir) == orient) nrm += f[i];
  return (V(2) * nrm + on + V(1)) / (V(1) + V(orient) * u[dir]);
}

// BoundaryHelpers<dir,orient>::computeStress           boundary/helper.h:31-80
template <int Q, typename V>
__device__ __forceinline__ void boundary_stress(const V (&f)[Q], int dir, int orient, V rho, const V (&u)[3],
                                                V (&pi)[6])
{
  using L = Lattice<Q>;
  const V usq = norm_sqr(u);
  V fneq[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int ci = cdir<Q>(i, dir);
    fneq[i] = (ci == 0 && i == 0) ? V(0) : f[i] - feq_shifted<Q, V>(i, rho, u, usq);
  }
  int k = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = a; b < 3; ++b) {
      V s = V(0);
#pragma unroll
      for (int i = 0; i < Q; ++i)
        if (cdir<Q>(i, dir) == 0) s += V(L::c[i][a] * L::c[i][b]) * fneq[i];
#pragma unroll
      for (int i = 0; i < Q; ++i)
        if (cdir<Q>(i, dir) == orient) s += V(2) * V(L::c[i][a] * L::c[i][b]) * fneq[i];
      pi[k++] = s;
    }
  }
}

// lbm<D>::computeStress                                          lbm.h:311-329
template <int Q, typename V>
__device__ __forceinline__ void bulk_stress(const V (&f)[Q], V rho, const V (&u)[3], V (&pi)[6])
{
  using L = Lattice<Q>;
  int k = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = a; b < 3; ++b) {
      V s = V(0);
#pragma unroll
      for (int i = 0; i < Q; ++i) s += V(L::c[i][a] * L::c[i][b]) * f[i];
      s -= rho * u[a] * u[b];
      if (a == b) s -= V(1) / V(3) * (rho - V(1));
      pi[k++] = s;
    }
  }
}

// equilibrium<D>::fromPiToFneq                                  lbm.h:225-247
template <int Q, typename V>
__device__ __forceinline__ V pi_to_fneq(int i, const V (&pi)[6])
{
  using L = Lattice<Q>;
  V fneq = V(0);
  int k = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = a; b < 3; ++b) {
      V add = V(L::c[i][a] * L::c[i][b]);
      if (a == b) add -= V(1) / V(3);
      else add *= V(2);
      add *= pi[k++];
      fneq += add;
    }
  }
  fneq *= weight<V, Q>(i) * V(3) * V(3) / V(2);
  return fneq;
}

__host__ __device__ inline int flat_dir(const LocalInfo& li)
{
  return li.a ? 0 : (li.b ? 1 : 2);
}

__host__ __device__ inline int info_component(const LocalInfo& li, int d)
{
  return d == 0 ? li.a : (d == 1 ? li.b : li.c);
}

// rho and u of a local-boundary cell through its boundary momenta
//   flat velocity  VelocityBoundaryDensity + FixedVelocityMomentumGeneric   elements.h:325-342, 528-571
//   flat pressure  FixedDensity + FixedPressureMomentum                      elements.h:237-262, 575-640
//   outer edges and corners: BulkDensity + fixed velocity                   aliases.h:188-193
//   inner edges / corners: InnerEdgeDensity3D / InnerCornerDensity3D        elements.h:400-445
template <int Q, typename V>
__device__ __forceinline__ void local_rho_u(const V (&f)[Q], const LocalInfo& li, const double* w, V& rho, V (&u)[3])
{
  using L = Lattice<Q>;
  u[0] = V(w[1]);
  u[1] = V(w[2]);
  u[2] = V(w[3]);
  switch (li.kind) {
  case kFlatVelocity:
  case kFdVelocity: {
    const int dir = flat_dir(li);
    rho = bm_rho<Q, V>(f, dir, info_component(li, dir), u);
    break;
  }
  case kFlatPressure:
  case kFdPressure: {
    const int dir = flat_dir(li), orient = info_component(li, dir);
    const V r = V(w[0]);
    V on = V(0), nrm = V(0);
#pragma unroll
    for (int i = 0; i < Q; ++i)
      if (cdir<Q>(i, dir) == 0) on += f[i];
#pragma unroll
    for (int i = 0; i < Q; ++i)
      if (cdir<Q>(i, dir) == orient) nrm += f[i];
    u[dir] = V(orient) * ((V(2) * nrm + on + V(1)) / r - V(1));
    rho = r;
    break;
  }
  case kOuterEdge:
  case kOuterCorner: {
    V s = V(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) s += f[i];
    rho = s + V(1);
    break;
  }
  case kInnerEdge: {
    const V r1 = bm_rho<Q, V>(f, (li.a + 1) % 3, li.b, u);
    const V r2 = bm_rho<Q, V>(f, (li.a + 2) % 3, li.c, u);
    rho = (r1 + r2) / V(2);
    break;
  }
  case kInnerCorner: {
    const V rx = bm_rho<Q, V>(f, 0, li.a, u);
    const V ry = bm_rho<Q, V>(f, 1, li.b, u);
    const V rz = bm_rho<Q, V>(f, 2, li.c, u);
    rho = (rx + ry + rz) / V(3);
    break;
  }
  default:
    rho = V(1);
    u[0] = u[1] = u[2] = V(0);
    break;
  }
}

// InnerEdgeStress3D / InnerCornerStress3D: bounce back the non-equilibrium
// part of the unknown populations, then the bulk stress    elements.h:1711-1765
template <int Q, typename V>
__device__ __forceinline__ void inner_stress(const V (&f)[Q], const LocalInfo& li, V rho, const V (&u)[3],
                                             V (&pi)[6])
{
  using L = Lattice<Q>;
  const V usq = norm_sqr(u);
  V g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) g[i] = f[i];
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    bool unknown;
    if (li.kind == kInnerEdge) {
      unknown = cdir<Q>(i, (li.a + 1) % 3) == -li.b && cdir<Q>(i, (li.a + 2) % 3) == -li.c;
    } else {
      unknown = L::c[i][0] == -li.a && L::c[i][1] == -li.b && L::c[i][2] == -li.c;
    }
    if (unknown) {
      const int o = opposite<Q>(i);
      g[i] = f[o] - feq_shifted<Q, V>(o, rho, u, usq) + feq_shifted<Q, V>(i, rho, u, usq);
    }
  }
  bulk_stress<Q, V>(g, rho, u, pi);
}

// Collision of a local-boundary cell: CombinedRLBdynamics::collide
// (dynamics.h:420-432) -- regularize from the boundary momenta, then BGK with
// the momenta recomputed -- or plain BGK at (rho = sum f + 1, u fixed) on outer
// edges and corners.  Returns the statistic the collision reports.
template <int Q, typename V>
__device__ __forceinline__ void collide_local(V (&f)[Q], const LocalInfo& li, const double* w, V omega, V& rho_out,
                                              V& usq_out)
{
  V rho, u[3];
  const bool plain = li.kind == kOuterEdge || li.kind == kOuterCorner || li.kind == kFdVelocity ||
                     li.kind == kFdPressure;
  if (!plain) {
    V pi[6];
    local_rho_u<Q, V>(f, li, w, rho, u);
    if (li.kind == kFlatVelocity || li.kind == kFlatPressure) {
      const int dir = flat_dir(li);
      boundary_stress<Q, V>(f, dir, info_component(li, dir), rho, u, pi);
    } else {
      inner_stress<Q, V>(f, li, rho, u, pi);
    }
    const V usq = norm_sqr(u);
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = feq_shifted<Q, V>(i, rho, u, usq) + pi_to_fneq<Q, V>(i, pi);
  }
  local_rho_u<Q, V>(f, li, w, rho, u);
  const V usq = norm_sqr(u);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const V e = feq_shifted<Q, V>(i, rho, u, usq);
    f[i] *= V(1) - omega;
    f[i] += omega * e;
  }
  rho_out = rho;
  usq_out = usq;
}

// Regularized flat velocity / pressure boundary with a small register
// footprint, for the whole-lattice kernel: the same operations in the same
// order as collide_local, with the non-equilibrium parts folded into the six
// stress accumulators instead of kept per population.  The face axis DIR is a
// compile-time constant, so the on-wall set is known at compile time and only
// the normal set depends on the run-time orientation.
template <int Q, typename V, int DIR>
__device__ __forceinline__ void collide_flat_axis(V (&f)[Q], int orient, bool pressure, const double* w, V omega)
{
  using L = Lattice<Q>;
  V u[3], rho;
  auto momenta = [&]() {
    u[0] = V(w[1]);
    u[1] = V(w[2]);
    u[2] = V(w[3]);
    V on = V(0), nrm = V(0);
#pragma unroll
    for (int i = 0; i < Q; ++i)
      if (L::c[i][DIR] == 0) on += f[i];
#pragma unroll
    for (int i = 0; i < Q; ++i)
      if (L::c[i][DIR] == orient) nrm += f[i];
    if (pressure) {
      rho = V(w[0]);
      u[DIR] = V(orient) * ((V(2) * nrm + on + V(1)) / rho - V(1));
    } else {
      rho = (V(2) * nrm + on + V(1)) / (V(1) + V(orient) * u[DIR]);
    }
  };
  momenta();
  V usq = norm_sqr(u);
  V pi[6] = {V(0), V(0), V(0), V(0), V(0), V(0)};
  // BoundaryHelpers::computeStress: on-wall populations, then the normal ones
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (L::c[i][DIR] != 0) continue;
    const V fneq = i == 0 ? V(0) : f[i] - feq_shifted<Q, V>(i, rho, u, usq);
    int k = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) pi[k++] += V(L::c[i][a] * L::c[i][b]) * fneq;
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (L::c[i][DIR] == 0 || L::c[i][DIR] != orient) continue;
    const V fneq = f[i] - feq_shifted<Q, V>(i, rho, u, usq);
    int k = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) pi[k++] += V(2) * V(L::c[i][a] * L::c[i][b]) * fneq;
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = feq_shifted<Q, V>(i, rho, u, usq) + pi_to_fneq<Q, V>(i, pi);
  momenta();
  usq = norm_sqr(u);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const V e = feq_shifted<Q, V>(i, rho, u, usq);
    f[i] *= V(1) - omega;
    f[i] += omega * e;
  }
}

// BGK with the Dirichlet momenta of an interpolated boundary face
template <int Q, typename V>
__device__ __forceinline__ void collide_dirichlet(V (&f)[Q], const LocalInfo& li, const double* w, V omega)
{
  V rho, u[3];
  local_rho_u<Q, V>(f, li, w, rho, u);
  const V usq = norm_sqr(u);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const V e = feq_shifted<Q, V>(i, rho, u, usq);
    f[i] *= V(1) - omega;
    f[i] += omega * e;
  }
}

template <int Q, typename V>
__device__ __forceinline__ void collide_flat(V (&f)[Q], const LocalInfo& li, const double* w, V omega)
{
  const int dir = flat_dir(li), orient = info_component(li, dir);
  const bool pressure = li.kind == kFlatPressure;
  if (dir == 0) collide_flat_axis<Q, V, 0>(f, orient, pressure, w, omega);
  else if (dir == 1) collide_flat_axis<Q, V, 1>(f, orient, pressure, w, omega);
  else collide_flat_axis<Q, V, 2>(f, orient, pressure, w, omega);
}

// ---- post-processors (PostStream stage) ----------------------------------

// Half the FORCE of core cell c, component d: the per-cell field [3][ncore]
// where the lattice has one, else the uniform value (F/2 given in FP64).
template <typename V>
__device__ __forceinline__ V half_force(const V* force, uint32_t ncore, uint32_t c, int d, double uniform_half)
{
  return force ? force[uint64_t(d) * ncore + c] * V(0.5) : V(uniform_half);
}

template <int Q>
struct PostArgs {
  void* pop;
  uint64_t stride;
  uint32_t n_padded;
  uint32_t shift[Q];
  const uint8_t* ids;
  const double* walls;      // [256][4]
  const uint32_t* info;     // [256] packed LocalInfo
  const uint32_t* cells;    // core indices of the cells to process
  const struct PostCell* recs;  // the same cells with their neighbours resolved (lattices), or null
  const void* remote;       // rho, u of the neighbours in other blocks (4 values each, arithmetic type), or null
  // bulk cells of a forced kind report u + F/2 (forcing::Guo's momenta,
  // momenta/elements.h:1473-1488): the per-cell FORCE [3][ncore] in the
  // arithmetic type, or null and the uniform F/2
  uint32_t forced;
  const void* force;
  double half_force[3];
  uint32_t count;
  int nx, ny, nz;
  uint32_t sx, sy;
  uint32_t periodic;
  FastDiv div_z, div_y;
  double omega;
};

template <int Q, typename S, typename V>
__device__ __forceinline__ void load_core_cell(const PostArgs<Q>& a, int x, int y, int z, V (&f)[Q])
{
  const S* pop = static_cast<const S*>(a.pop);
  const uint32_t n = uint32_t(x + 1) * a.sx + uint32_t(y + 1) * a.sy + uint32_t(z + 1);
#pragma unroll
  for (int i = 0; i < Q; ++i)
    f[i] = load_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, wrap_index(n, a.shift[i], a.n_padded)));
}

// computeRho / computeU of the cell at padded index n with dynamics id `id`
// through its dynamics' momenta; n == kPostOutside is the overlap beyond a
// non-periodic face (no dynamics: OneDensity, ZeroMomentum).
constexpr uint32_t kPostOutside = 0xffffffffu;
// A record slot kPostRemote | k names entry k of the block's remote buffer:
// a neighbour that lies in another block of a split lattice.
constexpr uint32_t kPostRemote = 0x80000000u;

template <int Q, typename S, typename V>
__device__ void rho_u_at(const PostArgs<Q>& a, uint32_t n, int id, V& rho, V (&u)[3])
{
  rho = V(1);
  u[0] = u[1] = u[2] = V(0);
  if (n == kPostOutside || id == 0 || id == 2) return;  // NoDynamics: OneDensity; BounceBack: FixedDensity 1, ZeroMomentum
  const S* pop = static_cast<const S*>(a.pop);
  if (id == 1) {
    // lbm<D>::computeRhoU                                      lbm.h:266-307
    using L = Lattice<Q>;
    V f[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)
      f[i] = load_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, wrap_index(n, a.shift[i], a.n_padded)));
    V s = V(0), j[3] = {V(0), V(0), V(0)};
#pragma unroll
    for (int i = 0; i < Q; ++i) s += f[i];
#pragma unroll
    for (int i = 0; i < Q; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) j[d] += V(L::c[i][d]) * f[i];
    rho = s + V(1);
    u[0] = j[0] / rho;
    u[1] = j[1] / rho;
    u[2] = j[2] / rho;
    if (a.forced) {
      // the core index of padded cell n, for a per-cell FORCE
      const uint32_t x = n / a.sx - 1, y = (n % a.sx) / a.sy - 1, z = n % a.sy - 1;
      const uint32_t ncore = uint32_t(a.nx) * uint32_t(a.ny) * uint32_t(a.nz);
      const uint32_t c = (x * uint32_t(a.ny) + y) * uint32_t(a.nz) + z;
      const V* force = static_cast<const V*>(a.force);
#pragma unroll
      for (int d = 0; d < 3; ++d) u[d] += half_force<V>(force, ncore, c, d, a.half_force[d]);
    }
    return;
  }
  const double* w = a.walls + 4 * id;
  const LocalInfo li = unpack_local_info(a.info[id]);
  if (li.kind == kMovingWall) {
    rho = V(w[0]);
    u[0] = V(w[1]);
    u[1] = V(w[2]);
    u[2] = V(w[3]);
    return;
  }
  if (li.kind == kDeadWall) return;
  V f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i)
    f[i] = load_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, wrap_index(n, a.shift[i], a.n_padded)));
  local_rho_u<Q, V>(f, li, w, rho, u);
}

// computeRho / computeU of the cell at core coordinates (x, y, z) through its
// dynamics' momenta.  Periodic axes wrap; beyond a non-periodic face lies the
// overlap, whose cells have no dynamics (OneDensity, ZeroMomentum).
template <int Q, typename S, typename V>
__device__ void cell_rho_u(const PostArgs<Q>& a, int x, int y, int z, V& rho, V (&u)[3])
{
  int r[3] = {x, y, z};
  const int n[3] = {a.nx, a.ny, a.nz};
  rho = V(1);
  u[0] = u[1] = u[2] = V(0);
  for (int d = 0; d < 3; ++d) {
    if (r[d] < 0 || r[d] >= n[d]) {
      if (!((a.periodic >> d) & 1u)) return;
      r[d] = (r[d] % n[d] + n[d]) % n[d];
    }
  }
  const uint32_t c = (uint32_t(r[0]) * uint32_t(a.ny) + uint32_t(r[1])) * uint32_t(a.nz) + uint32_t(r[2]);
  const uint32_t np = uint32_t(r[0] + 1) * a.sx + uint32_t(r[1] + 1) * a.sy + uint32_t(r[2] + 1);
  rho_u_at<Q, S, V>(a, np, a.ids[c], rho, u);
}

// fd::boundaryGradient                                    finiteDifference.h:40-44
template <typename V>
__device__ __forceinline__ V boundary_gradient(V u0, V u1, V u2)
{
  return (-V(3) * u0 + V(4) * u1 - V(1) * u2) / V(2);
}

// Regularized reconstruction from rho, the cell's fixed u and the velocity
// gradient dA[a][b] = d_a u_b          boundaryPostProcessors3D.hh:190-211, 270-290
// The populations go to the lattice, or to `scratch` (Q values) when given.
template <int Q, typename S, typename V>
__device__ void store_regularized(const PostArgs<Q>& a, int x, int y, int z, const V (&u)[3], V rho,
                                  const V (&dA)[3][3], V* scratch = nullptr)
{
  const V s = -rho / V(3) / V(a.omega);
  V pi[6];
  pi[0] = V(2) * dA[0][0] * s;
  pi[3] = V(2) * dA[1][1] * s;
  pi[5] = V(2) * dA[2][2] * s;
  pi[1] = (dA[0][1] + dA[1][0]) * s;
  pi[2] = (dA[0][2] + dA[2][0]) * s;
  pi[4] = (dA[1][2] + dA[2][1]) * s;
  const V usq = norm_sqr(u);
  S* pop = static_cast<S*>(a.pop);
  const uint32_t n = uint32_t(x + 1) * a.sx + uint32_t(y + 1) * a.sy + uint32_t(z + 1);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const V fi = feq_shifted<Q, V>(i, rho, u, usq) + pi_to_fneq<Q, V>(i, pi);
    if (scratch) scratch[i] = fi;
    else store_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, wrap_index(n, a.shift[i], a.n_padded)), fi);
  }
}

// ... at the cell's fixed velocity w[1..3]
template <int Q, typename S, typename V>
__device__ void store_reconstruction(const PostArgs<Q>& a, int x, int y, int z, const double* w, V rho,
                                     const V (&dA)[3][3])
{
  const V u[3] = {V(w[1]), V(w[2]), V(w[3])};
  store_regularized<Q, S, V>(a, x, y, z, u, rho, dA);
}

__device__ __forceinline__ void core_coords(const FastDiv& div_z, const FastDiv& div_y, int ny, int nz, uint32_t c,
                                            int& x, int& y, int& z)
{
  const uint32_t t = div_z.div(c);
  z = int(c - t * uint32_t(nz));
  const uint32_t xx = div_y.div(t);
  y = int(t - xx * uint32_t(ny));
  x = int(xx);
}

// The edge and corner post-processors each read six neighbours (rho and u
// through the neighbour's own momenta).  A cell is handled by eight lanes:
// six evaluate one neighbour each, so the dependent loads of the neighbours
// overlap, and one lane then rebuilds the cell.
//   edge (plane p, n1, n2)  OuterVelocityEdgeProcessor3D::apply, boundaryPostProcessors3D.hh:166-238
//     slots: -n1 e_d1, -2 n1 e_d1, -n2 e_d2, -2 n2 e_d2, +e_p, -e_p  (d1 = p+1, d2 = p+2 mod 3)
//   corner (nx, ny, nz)     OuterVelocityCornerProcessor3D::apply, boundaryPostProcessors3D.hh:242-292
//     slots: -n_d e_d and -2 n_d e_d for d = x, y, z
constexpr int kPostLanes = 8;
constexpr int kPostSlots = 6;

// A post-processed cell of a lattice with its neighbours resolved on the host
// when the boundary is set: its dynamics id, and the padded index and id of
// each neighbour slot (periodic axes wrapped, kPostOutside beyond a
// non-periodic face).  A list kernel then starts from one record instead of
// the chain core index -> id -> table entry -> neighbour id -> ...
struct PostCell {
  uint32_t c;                   // core index
  uint32_t id;
  uint32_t nb[kPostSlots];
  uint32_t nb_id[kPostSlots];
};
//   plane (dir, orient)     PlaneFdBoundaryProcessor3D::apply, boundaryPostProcessors3D.hh:38-92
//     slots, axis d = 0..2: -orient e_d and -2 orient e_d along the normal, +e_d and -e_d along the face
enum PostKind : int { kPostEdge = 0, kPostCorner = 1, kPostPlane = 2 };
// The order in which a lattice runs the kinds after the shift: the plane
// processors, then edges, then corners (PostStream priorities).
enum PostStage : int { kStagePlanes = 0, kStageEdges = 1, kStageCorners = 2, kPostStages = 3 };
__host__ __device__ constexpr int post_stage(int kind)
{
  return kind == kPostPlane ? kStagePlanes : (kind == kPostEdge ? kStageEdges : kStageCorners);
}

__host__ __device__ inline void post_offset(const LocalInfo& li, int kind, int slot, int (&o)[3])
{
  o[0] = o[1] = o[2] = 0;
  if (kind == kPostPlane) {
    const int d = slot >> 1, dir = flat_dir(li);
    if (d == dir) {
      const int orient = info_component(li, dir);
      o[d] = (slot & 1) ? -2 * orient : -orient;
    } else {
      o[d] = (slot & 1) ? -1 : 1;
    }
    return;
  }
  if (kind == kPostCorner) {
    const int d = slot >> 1, n = info_component(li, d);
    o[d] = (slot & 1) ? -2 * n : -n;
    return;
  }
  const int p = li.a, d1 = (p + 1) % 3, d2 = (p + 2) % 3;
  switch (slot) {
  case 0: o[d1] = -li.b; break;
  case 1: o[d1] = -2 * li.b; break;
  case 2: o[d2] = -li.c; break;
  case 3: o[d2] = -2 * li.c; break;
  case 4: o[p] = 1; break;
  default: o[p] = -1; break;
  }
}

// rho and u of neighbour `slot` of post-processed cell c
template <int Q, typename S, int KIND>
__device__ void post_gather(const PostArgs<Q>& a, uint32_t c, int slot, typename StorageTraits<S>::V (&out)[4])
{
  using V = typename StorageTraits<S>::V;
  int x, y, z;
  core_coords(a.div_z, a.div_y, a.ny, a.nz, c, x, y, z);
  const LocalInfo li = unpack_local_info(a.info[a.ids[c]]);
  int o[3];
  post_offset(li, KIND, slot, o);
  V rho, u[3];
  cell_rho_u<Q, S, V>(a, x + o[0], y + o[1], z + o[2], rho, u);
  out[0] = rho;
  out[1] = u[0];
  out[2] = u[1];
  out[3] = u[2];
}

// rho and u of neighbour `slot` of a resolved post-processed cell
template <int Q, typename S>
__device__ void post_gather_rec(const PostArgs<Q>& a, const PostCell* r, int slot,
                                typename StorageTraits<S>::V (&out)[4])
{
  using V = typename StorageTraits<S>::V;
  const uint32_t n = r->nb[slot];
  if (n != kPostOutside && (n & kPostRemote)) {
    const V* m = static_cast<const V*>(a.remote) + uint64_t(n & ~kPostRemote) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = m[k];
    return;
  }
  V rho, u[3];
  rho_u_at<Q, S, V>(a, n, int(r->nb_id[slot]), rho, u);
  out[0] = rho;
  out[1] = u[0];
  out[2] = u[1];
  out[3] = u[2];
}

// Density and velocity gradient from the six neighbour evaluations, then the
// regularized reconstruction of cell c
template <int Q, typename S, int KIND>
__device__ void post_finish(const PostArgs<Q>& a, uint32_t c, int id, const typename StorageTraits<S>::V (*nb)[4],
                            typename StorageTraits<S>::V* scratch = nullptr)
{
  using V = typename StorageTraits<S>::V;
  int x, y, z;
  core_coords(a.div_z, a.div_y, a.ny, a.nz, c, x, y, z);
  const double* w = a.walls + 4 * id;
  const LocalInfo li = unpack_local_info(a.info[id]);
  V u0[3] = {V(w[1]), V(w[2]), V(w[3])};
  V rho, dA[3][3];
  if constexpr (KIND == kPostPlane) {
    // rho and u from the cell's own Dirichlet momenta
    V f[Q];
    load_core_cell<Q, S, V>(a, x, y, z, f);
    local_rho_u<Q, V>(f, li, w, rho, u0);
    const int dir = flat_dir(li), orient = info_component(li, dir);
    for (int d = 0; d < 3; ++d)
      for (int b = 0; b < 3; ++b)
        dA[d][b] = d == dir ? V(-orient) * boundary_gradient(u0[b], nb[2 * d][1 + b], nb[2 * d + 1][1 + b])
                            : (nb[2 * d][1 + b] - nb[2 * d + 1][1 + b]) / V(2);
    store_regularized<Q, S, V>(a, x, y, z, u0, rho, dA, scratch);
    return;
  } else if constexpr (KIND == kPostCorner) {
    rho = V(4) / V(9) * (nb[4][0] + nb[2][0] + nb[0][0]) - V(1) / V(9) * (nb[5][0] + nb[3][0] + nb[1][0]);
    for (int d = 0; d < 3; ++d) {
      const int n = info_component(li, d);
      for (int b = 0; b < 3; ++b) dA[d][b] = V(-n) * boundary_gradient(u0[b], nb[2 * d][1 + b], nb[2 * d + 1][1 + b]);
    }
  } else {
    const int p = li.a, n1 = li.b, n2 = li.c, d1 = (p + 1) % 3, d2 = (p + 2) % 3;
    // rho10 = slot 0, rho20 = slot 1, rho01 = slot 2, rho02 = slot 3
    rho = V(2) / V(3) * (nb[2][0] + nb[0][0]) - V(1) / V(6) * (nb[3][0] + nb[1][0]);
    for (int b = 0; b < 3; ++b) {
      dA[p][b] = (nb[4][1 + b] - nb[5][1 + b]) / V(2);                      // centralGradient
      dA[d1][b] = V(-n1) * boundary_gradient(u0[b], nb[0][1 + b], nb[1][1 + b]);
      dA[d2][b] = V(-n2) * boundary_gradient(u0[b], nb[2][1 + b], nb[3][1 + b]);
    }
  }
  store_reconstruction<Q, S, V>(a, x, y, z, w, rho, dA);
}

// One launch per post-processor kind over the cell list: 8 cells per workgroup.
// Pressure faces read the normal velocity of their face neighbours from those
// neighbours' populations, which the same pass rewrites: their plane pass
// computes every cell into `scratch` first, and plane_store_kernel writes the
// populations afterwards.
template <int Q, typename S>
__global__ __launch_bounds__(64) void plane_scratch_kernel(PostArgs<Q> a, typename StorageTraits<S>::V* scratch)
{
  using V = typename StorageTraits<S>::V;
  __shared__ V nb[64 / kPostLanes][kPostSlots][4];
  const int slot = int(threadIdx.x % kPostLanes), j = int(threadIdx.x / kPostLanes);
  const uint32_t k = blockIdx.x * (64 / kPostLanes) + uint32_t(j);
  const bool valid = k < a.count;
  if (valid && slot < kPostSlots) post_gather_rec<Q, S>(a, a.recs + k, slot, nb[j][slot]);
  __syncthreads();
  if (valid && slot == 0)
    post_finish<Q, S, kPostPlane>(a, a.recs[k].c, int(a.recs[k].id), nb[j], scratch + uint64_t(k) * Q);
}

template <int Q, typename S>
__global__ __launch_bounds__(64) void plane_store_kernel(PostArgs<Q> a, const typename StorageTraits<S>::V* scratch)
{
  using V = typename StorageTraits<S>::V;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.count) return;
  int x, y, z;
  core_coords(a.div_z, a.div_y, a.ny, a.nz, a.cells[k], x, y, z);
  S* pop = static_cast<S*>(a.pop);
  const uint32_t n = uint32_t(x + 1) * a.sx + uint32_t(y + 1) * a.sy + uint32_t(z + 1);
#pragma unroll
  for (int i = 0; i < Q; ++i)
    store_pop<S, V>(column_at(pop + uint64_t(i) * a.stride, wrap_index(n, a.shift[i], a.n_padded)),
                    scratch[uint64_t(k) * Q + i]);
}

template <int Q, typename S, int KIND>
__global__ __launch_bounds__(64) void outer_post_kernel(PostArgs<Q> a)
{
  using V = typename StorageTraits<S>::V;
  __shared__ V nb[64 / kPostLanes][kPostSlots][4];
  const int slot = int(threadIdx.x % kPostLanes), j = int(threadIdx.x / kPostLanes);
  const uint32_t k = blockIdx.x * (64 / kPostLanes) + uint32_t(j);
  const bool valid = k < a.count;
  if (valid && slot < kPostSlots) post_gather_rec<Q, S>(a, a.recs + k, slot, nb[j][slot]);
  __syncthreads();
  if (valid && slot == 0) post_finish<Q, S, KIND>(a, a.recs[k].c, int(a.recs[k].id), nb[j]);
}

// Neighbours of post-processed cells that lie in another block: entry `slot`
// of the consuming block's remote buffer takes rho and u of padded cell
// `src_padded` (core index `src_core`) of the block that owns it.
struct RemotePair {
  uint32_t src_padded;
  uint32_t src_core;
  uint32_t slot;
};

// The PostStream exchange of the post-processors' neighbourhood
// (superLattice.hh:639-640, the overlap cells boundary::set requests through
// addPoints2CommBC): rho and u of the listed cells of each source block, as
// its post-processors would read them, into the consumers' remote buffers --
// many halos in one launch, entry k of the batch finding its halo by
// bisection over the halos' first entries.
template <int Q>
struct RemoteBatchDesc {
  PostArgs<Q> a;             // the source block, at the shifts to read it with
  const RemotePair* pairs;
  void* out;                 // the consumer's remote buffer
  uint32_t first, count;
};

template <int Q, typename S>
__global__ __launch_bounds__(64) void remote_moments_many_kernel(const RemoteBatchDesc<Q>* __restrict__ descs,
                                                                 uint32_t halos, uint32_t total)
{
  using V = typename StorageTraits<S>::V;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  uint32_t lo = 0, hi = halos;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (descs[mid].first <= k) lo = mid;
    else hi = mid;
  }
  const RemoteBatchDesc<Q>& d = descs[lo];
  const uint32_t j = k - d.first;
  if (j >= d.count) return;
  const RemotePair p = d.pairs[j];
  V rho, u[3];
  rho_u_at<Q, S, V>(d.a, p.src_padded, int(d.a.ids[p.src_core]), rho, u);
  V* o = static_cast<V*>(d.out) + uint64_t(p.slot) * 4;
  o[0] = rho;
  o[1] = u[0];
  o[2] = u[1];
  o[3] = u[2];
}

}  // namespace lbmh
