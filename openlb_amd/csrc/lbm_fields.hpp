// Moment fields of a block, evaluated on the device in the point order of a
// VTK ImageData piece, for the output writers (OpenLB's SuperVTMwriter3D
// buffer, src/io/superVtmWriter3D.hh:437-458: z outermost, x fastest, over
// [-overlap, extent + overlap) per axis, every component of a point together,
// converted to float).
//
// Each point reports what OpenLB's functor reads from the cell there:
//   density        cell.computeRho()                 functors/lattice/latticeDensity3D.hh
//   velocity       cell.computeU(u)                  latticeVelocity3D.hh
//   physVelocity   converter.getPhysVelocity(u)      latticePhysVelocity3D.hh:84-96
//   physPressure   converter.getPhysPressure((rho - 1) / invCs2)   latticePhysPressure3D.hh:71-78
// through the cell's own momenta (cell_rho_u: bulk populations, fixed wall
// values, the regularized boundaries' density and velocity); bulk cells of a
// forced lattice report u + F/2 (forcing::Guo's momenta).
//
// Overlap points are what OpenLB's overlap cells hold after the writer's
// communicate() (superVtmWriter3D.hh:71-75): on a periodic axis the image's
// populations and FORCE, the descriptor's fields, which is the image's value
// for a bulk cell; every other cell there reports rho 1, u 0 -- the wall
// dynamics' own fields are not communicated and keep initialize()'s values
// (FixedDensity 1, velocity 0), the boundary setters leave the outermost
// layer without dynamics (setBoundary.h:87-88), and beyond a non-periodic
// face there are no dynamics.
#pragma once

#include "lbm_kernels.hpp"
#include "lbm_local.hpp"

namespace lbmh {

enum ExportField : int { kFieldDensity = 0, kFieldVelocity = 1, kFieldPhysVelocity = 2, kFieldPhysPressure = 3 };

inline int field_components(int field) { return (field == kFieldVelocity || field == kFieldPhysVelocity) ? 3 : 1; }

struct FieldConversion {
  double velocity;         // UnitConverter::getConversionFactorVelocity
  double pressure;         // UnitConverter::getConversionFactorPressure
  double pressure_offset;  // charPhysPressure
};

// Lanes run along z, the axis the population columns are contiguous in: a
// workgroup of 256 lanes covers 32 z x 8 x points of one y row, so every
// population column is read in whole 128-byte lines, and the eight waves'
// float stores into the same lines of the x-fastest output merge in the L2.
// (One lane per point in output order puts consecutive lanes one x plane
// apart, i.e. one cache line per lane and population.)
constexpr uint32_t kExportTileZ = 32, kExportTileX = 8;

// Workgroups of the export grid.  The tiles pad x to a multiple of 8 and z to
// a multiple of 32, so the grid can hold several times the points of a thin
// block; the caller refuses a grid beyond HIP's 2^32 - 1 work-items per launch.
inline uint64_t field_export_blocks(int nx, int ny, int nz, int overlap)
{
  const uint64_t ex = uint64_t(nx + 2 * overlap), ey = uint64_t(ny + 2 * overlap), ez = uint64_t(nz + 2 * overlap);
  return ((ex + kExportTileX - 1) / kExportTileX) * ey * ((ez + kExportTileZ - 1) / kExportTileZ);
}

inline bool field_export_grid_fits(uint64_t blocks)
{
  return blocks * (kExportTileZ * kExportTileX) <= 0xffffffffull;
}

// launched with field_export_blocks(...) workgroups of kExportTileZ * kExportTileX lanes
template <int Q, typename S>
__global__ __launch_bounds__(256) void field_export_kernel(PostArgs<Q> a, int overlap, int field, FieldConversion conv,
                                                           int forced, double hfx, double hfy, double hfz,
                                                           const typename StorageTraits<S>::V* force,
                                                           float* __restrict__ out, uint32_t points)
{
  static_assert(kExportTileZ * kExportTileX == 256, "one workgroup is one tile");
  using V = typename StorageTraits<S>::V;
  const uint32_t ex = uint32_t(a.nx + 2 * overlap), ey = uint32_t(a.ny + 2 * overlap);
  const uint32_t ez = uint32_t(a.nz + 2 * overlap);
  const uint32_t tiles_x = (ex + kExportTileX - 1) / kExportTileX;
  const uint32_t rest = blockIdx.x / tiles_x;
  const uint32_t px = (blockIdx.x - rest * tiles_x) * kExportTileX + threadIdx.x / kExportTileZ;
  const uint32_t py = rest % ey;
  const uint32_t pz = (rest / ey) * kExportTileZ + threadIdx.x % kExportTileZ;
  if (px >= ex || py >= ey || pz >= ez) return;
  const uint32_t p = (pz * ey + py) * ex + px;
  if (p >= points) return;
  const int x = int(px) - overlap, y = int(py) - overlap, z = int(pz) - overlap;
  int r[3] = {x, y, z};
  const int n[3] = {a.nx, a.ny, a.nz};
  bool inside = true, image = false;
  for (int d = 0; d < 3; ++d) {
    if (r[d] >= 0 && r[d] < n[d]) continue;
    image = true;
    if (!((a.periodic >> d) & 1u)) inside = false;
    r[d] = (r[d] + n[d]) % n[d];
  }
  const uint32_t core = (uint32_t(r[0]) * uint32_t(a.ny) + uint32_t(r[1])) * uint32_t(a.nz) + uint32_t(r[2]);
  const int id = inside ? a.ids[core] : 0;
  V rho = V(1), u[3] = {V(0), V(0), V(0)};
  if (inside && (!image || id == 1)) cell_rho_u<Q, S, V>(a, r[0], r[1], r[2], rho, u);
  if (forced && inside && id == 1) {
    const uint32_t ncore = uint32_t(a.nx) * uint32_t(a.ny) * uint32_t(a.nz);
    u[0] += half_force<V>(force, ncore, core, 0, hfx);
    u[1] += half_force<V>(force, ncore, core, 1, hfy);
    u[2] += half_force<V>(force, ncore, core, 2, hfz);
  }
  switch (field) {
  case kFieldDensity:
    out[p] = float(rho);
    break;
  case kFieldVelocity:
    out[3 * uint64_t(p)] = float(u[0]);
    out[3 * uint64_t(p) + 1] = float(u[1]);
    out[3 * uint64_t(p) + 2] = float(u[2]);
    break;
  case kFieldPhysVelocity:
    for (int d = 0; d < 3; ++d) out[3 * uint64_t(p) + d] = float(V(conv.velocity) * u[d]);
    break;
  default: {
    // (rho - 1.0) / invCs2 is evaluated in double and stored in the value
    // type; the conversion a*b+c is one fused multiply-add, as the compiled
    // OpenLB functor evaluates it
    const V lattice_pressure = V((double(rho) - 1.0) / 3.0);
    if constexpr (sizeof(V) == 4)
      out[p] = __builtin_fmaf(float(conv.pressure), lattice_pressure, float(conv.pressure_offset));
    else
      out[p] = float(__builtin_fma(conv.pressure, lattice_pressure, conv.pressure_offset));
  }
  }
}

// ---- reductions of a cell field ---------------------------------------------

enum ReduceOp : int { kReduceSum = 0, kReduceMin = 1, kReduceMax = 2 };
constexpr int kFieldForce = 4;

// The dynamics ids a reduction visits: bit id of bits[id / 32].
struct IdMask {
  uint32_t bits[8];
};

__device__ __forceinline__ double reduce_combine(int op, double a, double b)
{
  return op == kReduceSum ? a + b : (op == kReduceMin ? fmin(a, b) : fmax(a, b));
}

// SumO / MinO / MaxO's reset values (blockLatticeFieldReductionO.h:54-100)
__host__ __device__ __forceinline__ double reduce_identity(int op)
{
  return op == kReduceSum ? 0.0 : (op == kReduceMin ? 1.7976931348623157e308 : -1.7976931348623157e308);
}

__device__ __forceinline__ double wave_reduce(int op, double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = reduce_combine(op, v, __shfl_down(v, off, 64));
  return v;
}

// First level of a field reduction: core cells whose id is in `mask`, each
// cell's value as lbmh_field_export evaluates it at that core point (in the
// arithmetic type, not rounded to float) or its FORCE; one partial per
// workgroup and component plus the count of visited cells, in a fixed order.
template <int Q, typename S>
__global__ __launch_bounds__(256) void field_reduce_kernel(PostArgs<Q> a, int field, int op, FieldConversion conv,
                                                           int forced, double hfx, double hfy, double hfz,
                                                           const typename StorageTraits<S>::V* force, IdMask mask,
                                                           double* partial)
{
  using V = typename StorageTraits<S>::V;
  const uint32_t ncore = uint32_t(a.nx) * uint32_t(a.ny) * uint32_t(a.nz);
  const double init = reduce_identity(op);
  double acc[3] = {init, init, init};
  double count = 0.0;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncore; c += gridDim.x * blockDim.x) {
    const int id = a.ids[c];
    if (!((mask.bits[id >> 5] >> (id & 31)) & 1u)) continue;
    count += 1.0;
    double v[3] = {0.0, 0.0, 0.0};
    if (field == kFieldForce) {
      if (id == kIdBulk) {
        v[0] = double(half_force<V>(force, ncore, c, 0, hfx) * V(2));
        v[1] = double(half_force<V>(force, ncore, c, 1, hfy) * V(2));
        v[2] = double(half_force<V>(force, ncore, c, 2, hfz) * V(2));
      }
    } else {
      int x, y, z;
      core_coords(a.div_z, a.div_y, a.ny, a.nz, c, x, y, z);
      V rho = V(1), u[3] = {V(0), V(0), V(0)};
      if (id != 0) cell_rho_u<Q, S, V>(a, x, y, z, rho, u);
      if (forced && id == kIdBulk) {
        u[0] += half_force<V>(force, ncore, c, 0, hfx);
        u[1] += half_force<V>(force, ncore, c, 1, hfy);
        u[2] += half_force<V>(force, ncore, c, 2, hfz);
      }
      if (field == kFieldDensity) {
        v[0] = double(rho);
      } else if (field == kFieldVelocity) {
        for (int d = 0; d < 3; ++d) v[d] = double(u[d]);
      } else if (field == kFieldPhysVelocity) {
        for (int d = 0; d < 3; ++d) v[d] = double(V(conv.velocity) * u[d]);
      } else {
        const V lattice_pressure = V((double(rho) - 1.0) / 3.0);
        if constexpr (sizeof(V) == 4)
          v[0] = double(__builtin_fmaf(float(conv.pressure), lattice_pressure, float(conv.pressure_offset)));
        else
          v[0] = __builtin_fma(conv.pressure, lattice_pressure, conv.pressure_offset);
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] = reduce_combine(op, acc[d], v[d]);
  }
  __shared__ double lds[4][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double r = wave_reduce(op, acc[d]);
    if (lane == 0) lds[d][wave] = r;
  }
  const double n = wave_reduce(kReduceSum, count);
  if (lane == 0) lds[3][wave] = n;
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = int(threadIdx.x), kop = k == 3 ? kReduceSum : op;
    double r = lds[k][0];
    for (int w = 1; w < 4; ++w) r = reduce_combine(kop, r, lds[k][w]);
    partial[blockIdx.x * 4 + k] = r;
  }
}

// Second level: the partials in a fixed order (deterministic).
static __global__ __launch_bounds__(256) void field_reduce_partials_kernel(double* out, const double* partial, uint32_t nparts,
                                                                    int op)
{
  __shared__ double lds[4][4];
  double acc[4] = {reduce_identity(op), reduce_identity(op), reduce_identity(op), 0.0};
  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x)
    for (int k = 0; k < 4; ++k) acc[k] = reduce_combine(k == 3 ? kReduceSum : op, acc[k], partial[p * 4 + k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double r = wave_reduce(k == 3 ? kReduceSum : op, acc[k]);
    if (lane == 0) lds[k][wave] = r;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = int(threadIdx.x), kop = k == 3 ? kReduceSum : op;
    double r = lds[k][0];
    for (int w = 1; w < 4; ++w) r = reduce_combine(kop, r, lds[k][w]);
    out[k] = r;
  }
}

}  // namespace lbmh
