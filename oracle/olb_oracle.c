/* Oracle: CPU restatement of OpenLB's block-lattice time step.
 * TEST INFRASTRUCTURE ONLY -- see olb_oracle.h for scope and citations. */
#include "olb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct olo_lattice {
  int q, n[3], periodic[3], storage, pad;
  int64_t size[3], proj[3], ncells, ncore;
  int64_t shift[27], dist[27];
  int c[27][3], opp[27], tnum[27], tden[27];
  void* pop;
  uint8_t* ids;
  int64_t nhalo;
  int64_t *halo_dst, *halo_src;
  int kind;
  double omega, magic, smago, force[3];
  double* force_field;  /* per padded cell, 3 values (NULL: the uniform force) */
  double wall_rho[256], wall_u[256][3];
  /* local boundaries, per padded cell: kind (OLK_*; 0 = none), discrete
   * normal or (plane, n1, n2), and the defined rho / u */
  uint8_t* lkind;
  int8_t* lnorm;
  double* lval;
  int64_t n_edge, n_corner, n_plane;
  int64_t *edge_cells, *corner_cells, *plane_cells;
};

/* per-cell local boundary kinds */
enum {
  OLK_NONE = 0,
  OLK_FLAT_VELOCITY = 1,  /* CombinedRLB + RegularizedVelocityBoundaryTuple   momenta/aliases.h:127-132 */
  OLK_FLAT_PRESSURE = 2,  /* CombinedRLB + RegularizedPressureBoundaryTuple   momenta/aliases.h:135-141 */
  OLK_OUTER_EDGE = 3,     /* BGK + FixedVelocityBoundaryTuple, OuterVelocityEdgeProcessor3D */
  OLK_OUTER_CORNER = 4,   /* BGK + FixedVelocityBoundaryTuple, OuterVelocityCornerProcessor3D */
  OLK_INNER_EDGE = 5,     /* CombinedRLB + InnerEdgeVelocityTuple3D          momenta/aliases.h:204-209 */
  OLK_INNER_CORNER = 6,   /* CombinedRLB + InnerCornerVelocityTuple3D        momenta/aliases.h:212-217 */
  OLK_DEAD = 7,           /* boundary-material cell without a matching type: no dynamics */
  OLK_FD_VELOCITY = 8,    /* BGK + BasicDirichletVelocityBoundaryTuple, PlaneFdBoundaryProcessor3D */
  OLK_FD_PRESSURE = 9     /* BGK + BasicDirichletPressureBoundaryTuple, PlaneFdBoundaryProcessor3D */
};

static char g_err[256];
static void set_err(const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); }
const char* olo_last_error(void) { return g_err; }

/* ---- descriptors: src/descriptor/definition/common.h:120-166 (D3Q19), 313-366 (D3Q27) */
static const int C19[19][3] = {
  {0, 0, 0},
  {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
  {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
  {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
  {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
static const int C27[27][3] = {
  {0, 0, 0},
  {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
  {-1, 0, 1}, {0, -1, -1}, {0, -1, 1}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
  {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
  {1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};

static void fill_descriptor(olo_lattice* L)
{
  const int q = L->q;
  for (int i = 0; i < q; ++i) {
    int nz = 0;
    for (int d = 0; d < 3; ++d) {
      L->c[i][d] = (q == 19) ? C19[i][d] : C27[i][d];
      nz += L->c[i][d] != 0;
    }
    /* opposite: i <-> i + q/2 for i >= 1 */
    L->opp[i] = (i == 0) ? 0 : (i <= q / 2 ? i + q / 2 : i - q / 2);
    if (q == 19) {
      L->tnum[i] = 1;
      L->tden[i] = (nz == 0) ? 3 : (nz == 1 ? 18 : 36);
    } else {
      static const int num[4] = {8, 2, 1, 1}, den[4] = {27, 27, 54, 216};
      L->tnum[i] = num[nz];
      L->tden[i] = den[nz];
    }
  }
}

/* ---- IEEE binary16 conversion (round to nearest even), own implementation */
static float half_to_float(uint16_t h)
{
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t exp = (h >> 10) & 0x1fu;
  const uint32_t man = h & 0x3ffu;
  float out;
  if (exp == 0) {
    out = ldexpf((float)man, -24);
  } else if (exp == 31) {
    out = man ? NAN : INFINITY;
  } else {
    out = ldexpf((float)(man | 0x400u), (int)exp - 25);
  }
  uint32_t bits;
  memcpy(&bits, &out, 4);
  bits |= sign;
  memcpy(&out, &bits, 4);
  return out;
}

static uint16_t float_to_half(float f)
{
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u);
  if (ax >= 0x477ff000u) return sign | 0x7c00u; /* rounds to >= 65520 -> inf */
  if (ax < 0x38800000u) {
    /* subnormal half: value = m * 2^-24, round |f| * 2^24 to nearest even */
    float a;
    memcpy(&a, &ax, 4);
    const float scaled = a * 16777216.0f; /* exact: power of two */
    float r = nearbyintf(scaled);          /* default rounding mode: ties to even */
    return sign | (uint16_t)r;
  }
  /* normal: keep 10 mantissa bits, round to nearest even on the dropped 13 */
  uint32_t e = (ax >> 23) - 127 + 15;
  uint32_t m = ax & 0x7fffffu;
  uint32_t h = (e << 10) | (m >> 13);
  const uint32_t rest = m & 0x1fffu;
  if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h += 1; /* may carry into exponent */
  return sign | (uint16_t)h;
}

/* FORCE field of the cell at padded index n: its own value where a per-cell
 * field is set (ForcedBGKdynamics reads the cell's FORCE, momenta/elements.h:
 * 1480-1488, forcing.h:37-74), else the lattice's uniform force */
static const double* cell_force(const olo_lattice* L, int64_t n)
{
  return L->force_field ? L->force_field + 3 * n : L->force;
}

/* ---- instantiations */
#define SFX f64
#define STORE_T double
#define V double
#define LOAD(p) (p)
#define STOREV(v) (v)
#include "olb_oracle_impl.h"
#undef SFX
#undef STORE_T
#undef V
#undef LOAD
#undef STOREV

#define SFX f32
#define STORE_T float
#define V float
#define LOAD(p) (p)
#define STOREV(v) (v)
#include "olb_oracle_impl.h"
#undef SFX
#undef STORE_T
#undef V
#undef LOAD
#undef STOREV

#define SFX f16
#define STORE_T uint16_t
#define V float
/* FP16 keeps fb_i * 2^12 (exact power-of-two scaling, same as the device) */
#define LOAD(p) (half_to_float(p) * (1.0f / 4096.0f))
#define STOREV(v) float_to_half((v) * 4096.0f)
#include "olb_oracle_impl.h"
#undef SFX
#undef STORE_T
#undef V
#undef LOAD
#undef STOREV

/* ---- boundary classification (computeBoundaryTypeAndNormal,
 * src/geometry/discreteNormals.hh:174-1190).  OpenLB tests neighbour patterns
 * in a fixed order: flat faces, external corners, internal corners, external
 * edges, internal edges.  The first match of the first four groups is kept;
 * an internal-edge match replaces whatever was found before.  Each rule
 * lists the neighbours it tests: 'W' wall (neither fluid nor outside),
 * 'F' fluid, 'N' not fluid.  The resulting normal points out of the fluid. */
typedef struct { signed char dx, dy, dz; char pred; } olo_cond;
typedef struct { signed char type, n[3], overwrite, ncond; olo_cond c[8]; } olo_rule;
static const olo_rule RULES[46] = {
  {0, {-1, 0, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {1, 0, 0, 'F'}}},
  {0, {1, 0, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {-1, 0, 0, 'F'}}},
  {0, {0, -1, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {1, 0, 0, 'W'}, {-1, 0, 0, 'W'}, {0, 1, 0, 'F'}}},
  {0, {0, 1, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {1, 0, 0, 'W'}, {-1, 0, 0, 'W'}, {0, -1, 0, 'F'}}},
  {0, {0, 0, -1}, 0, 5, {{1, 0, 0, 'W'}, {-1, 0, 0, 'W'}, {0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, 1, 'F'}}},
  {0, {0, 0, 1}, 0, 5, {{1, 0, 0, 'W'}, {-1, 0, 0, 'W'}, {0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, -1, 'F'}}},
  {1, {-1, -1, -1}, 0, 7, {{1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {1, 0, 1, 'W'}, {0, 1, 0, 'W'}, {1, 1, 0, 'W'}, {0, 1, 1, 'W'}, {1, 1, 1, 'F'}}},
  {1, {-1, 1, -1}, 0, 7, {{1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {1, 0, 1, 'W'}, {0, -1, 0, 'W'}, {1, -1, 0, 'W'}, {0, -1, 1, 'W'}, {1, -1, 1, 'F'}}},
  {1, {-1, 1, 1}, 0, 7, {{1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {1, 0, -1, 'W'}, {0, -1, 0, 'W'}, {1, -1, 0, 'W'}, {0, -1, -1, 'W'}, {1, -1, -1, 'F'}}},
  {1, {-1, -1, 1}, 0, 7, {{1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {1, 0, -1, 'W'}, {0, 1, 0, 'W'}, {1, 1, 0, 'W'}, {0, 1, -1, 'W'}, {1, 1, -1, 'F'}}},
  {1, {1, 1, 1}, 0, 7, {{-1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {-1, 0, -1, 'W'}, {0, -1, 0, 'W'}, {0, -1, -1, 'W'}, {-1, -1, 0, 'W'}, {-1, -1, -1, 'F'}}},
  {1, {1, -1, 1}, 0, 7, {{-1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {-1, 0, -1, 'W'}, {0, 1, 0, 'W'}, {0, 1, -1, 'W'}, {-1, 1, 0, 'W'}, {-1, 1, -1, 'F'}}},
  {1, {1, -1, -1}, 0, 7, {{-1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {-1, 0, 1, 'W'}, {0, 1, 0, 'W'}, {0, 1, 1, 'W'}, {-1, 1, 0, 'W'}, {-1, 1, 1, 'F'}}},
  {1, {1, 1, -1}, 0, 7, {{-1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {-1, 0, 1, 'W'}, {0, -1, 0, 'W'}, {0, -1, 1, 'W'}, {-1, -1, 0, 'W'}, {-1, -1, 1, 'F'}}},
  {2, {1, 1, 1}, 0, 6, {{-1, 0, 0, 'F'}, {0, 0, -1, 'F'}, {0, 0, 1, 'W'}, {1, 0, 0, 'W'}, {0, -1, 0, 'F'}, {0, 1, 0, 'W'}}},
  {2, {1, -1, 1}, 0, 6, {{-1, 0, 0, 'F'}, {0, 0, -1, 'F'}, {0, 0, 1, 'W'}, {1, 0, 0, 'W'}, {0, 1, 0, 'F'}, {0, -1, 0, 'W'}}},
  {2, {1, -1, -1}, 0, 6, {{-1, 0, 0, 'F'}, {0, 0, 1, 'F'}, {0, 0, -1, 'W'}, {1, 0, 0, 'W'}, {0, 1, 0, 'F'}, {0, -1, 0, 'W'}}},
  {2, {1, 1, -1}, 0, 6, {{-1, 0, 0, 'F'}, {0, 0, 1, 'F'}, {0, 0, -1, 'W'}, {1, 0, 0, 'W'}, {0, -1, 0, 'F'}, {0, 1, 0, 'W'}}},
  {2, {-1, 1, 1}, 0, 6, {{1, 0, 0, 'F'}, {0, 0, -1, 'F'}, {-1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {0, -1, 0, 'F'}, {0, 1, 0, 'W'}}},
  {2, {-1, -1, 1}, 0, 6, {{1, 0, 0, 'F'}, {0, 0, -1, 'F'}, {-1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {0, 1, 0, 'F'}, {0, -1, 0, 'W'}}},
  {2, {-1, 1, -1}, 0, 6, {{1, 0, 0, 'F'}, {0, 0, 1, 'F'}, {-1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {0, -1, 0, 'F'}, {0, 1, 0, 'W'}}},
  {2, {-1, -1, -1}, 0, 6, {{1, 0, 0, 'F'}, {0, 0, 1, 'F'}, {-1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {0, 1, 0, 'F'}, {0, -1, 0, 'W'}}},
  {3, {0, 1, -1}, 0, 7, {{-1, 0, 0, 'W'}, {1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {1, 0, 1, 'N'}, {-1, 0, 1, 'N'}, {0, -1, 1, 'F'}, {0, -1, 0, 'W'}}},
  {3, {0, -1, -1}, 0, 7, {{-1, 0, 0, 'W'}, {1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {1, 0, 1, 'N'}, {-1, 0, 1, 'N'}, {0, 1, 1, 'F'}, {0, 1, 0, 'W'}}},
  {3, {0, -1, 1}, 0, 7, {{-1, 0, 0, 'W'}, {1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {1, 0, -1, 'N'}, {-1, 0, -1, 'N'}, {0, 1, -1, 'F'}, {0, 1, 0, 'N'}}},
  {3, {0, 1, 1}, 0, 7, {{-1, 0, 0, 'W'}, {1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {1, 0, -1, 'N'}, {-1, 0, -1, 'N'}, {0, -1, -1, 'F'}, {0, -1, 0, 'N'}}},
  {3, {-1, 0, -1}, 0, 7, {{0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, 1, 'W'}, {1, 0, 1, 'F'}, {1, 0, 0, 'W'}, {-1, 0, 0, 'N'}, {0, 0, -1, 'N'}}},
  {3, {1, 0, -1}, 0, 7, {{0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, 1, 'W'}, {-1, 0, 1, 'F'}, {-1, 0, 0, 'W'}, {1, 0, 0, 'N'}, {0, 0, -1, 'N'}}},
  {3, {-1, 0, 1}, 0, 7, {{0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, -1, 'W'}, {1, 0, -1, 'F'}, {1, 0, 0, 'W'}, {-1, 0, 0, 'N'}, {0, 0, 1, 'N'}}},
  {3, {1, 0, 1}, 0, 7, {{0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, -1, 'W'}, {-1, 0, -1, 'F'}, {-1, 0, 0, 'W'}, {1, 0, 0, 'N'}, {0, 0, 1, 'N'}}},
  {3, {-1, -1, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {0, 1, 0, 'W'}, {1, 0, 0, 'N'}, {1, 1, 0, 'F'}}},
  {3, {1, -1, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {0, 1, 0, 'W'}, {-1, 0, 0, 'N'}, {-1, 1, 0, 'F'}}},
  {3, {1, 1, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {0, -1, 0, 'W'}, {-1, 0, 0, 'N'}, {-1, -1, 0, 'F'}}},
  {3, {-1, 1, 0}, 0, 5, {{0, 0, 1, 'W'}, {0, 0, -1, 'W'}, {0, -1, 0, 'W'}, {1, 0, 0, 'N'}, {1, -1, 0, 'F'}}},
  {4, {0, -1, -1}, 1, 8, {{-1, 0, 0, 'W'}, {1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {0, 0, 1, 'F'}, {0, 1, 0, 'F'}, {0, -1, 0, 'W'}, {-1, -1, 0, 'W'}, {1, -1, 0, 'W'}}},
  {4, {0, 1, -1}, 1, 8, {{-1, 0, 0, 'W'}, {1, 0, 0, 'W'}, {0, 0, -1, 'W'}, {0, 0, 1, 'F'}, {0, -1, 0, 'F'}, {0, 1, 0, 'W'}, {-1, 1, 0, 'W'}, {1, 1, 0, 'W'}}},
  {4, {0, -1, 1}, 1, 8, {{1, 0, 0, 'W'}, {-1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {0, 0, -1, 'F'}, {0, 1, 0, 'F'}, {-1, -1, 0, 'W'}, {0, -1, 0, 'W'}, {1, -1, 0, 'W'}}},
  {4, {0, 1, 1}, 1, 8, {{1, 0, 0, 'W'}, {-1, 0, 0, 'W'}, {0, 0, 1, 'W'}, {0, 0, -1, 'F'}, {0, -1, 0, 'F'}, {-1, 1, 0, 'W'}, {0, 1, 0, 'W'}, {1, 1, 0, 'W'}}},
  {4, {1, 0, 1}, 1, 8, {{-1, 0, 0, 'F'}, {1, 0, 0, 'W'}, {0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, -1, 'F'}, {0, 0, 1, 'W'}, {0, 1, 1, 'W'}, {0, -1, 1, 'W'}}},
  {4, {1, 0, -1}, 1, 8, {{-1, 0, 0, 'F'}, {1, 0, 0, 'W'}, {0, 1, 0, 'W'}, {0, -1, 0, 'W'}, {0, 0, 1, 'F'}, {0, 0, -1, 'W'}, {0, 1, -1, 'W'}, {0, -1, -1, 'W'}}},
  {4, {-1, 0, 1}, 1, 8, {{1, 0, 0, 'F'}, {-1, 0, 0, 'W'}, {0, -1, 0, 'W'}, {0, 1, 0, 'W'}, {0, 0, -1, 'F'}, {0, 0, 1, 'W'}, {0, 1, 1, 'W'}, {0, -1, 1, 'W'}}},
  {4, {-1, 0, -1}, 1, 8, {{1, 0, 0, 'F'}, {-1, 0, 0, 'W'}, {0, -1, 0, 'W'}, {0, 1, 0, 'W'}, {0, 0, 1, 'F'}, {0, 0, -1, 'W'}, {0, 1, -1, 'W'}, {0, -1, -1, 'W'}}},
  {4, {1, 1, 0}, 1, 6, {{0, -1, 0, 'F'}, {0, 0, -1, 'W'}, {0, 0, 1, 'W'}, {0, 1, 0, 'W'}, {-1, 0, 0, 'F'}, {-1, -1, 0, 'F'}}},
  {4, {-1, 1, 0}, 1, 6, {{0, -1, 0, 'F'}, {0, 0, -1, 'W'}, {0, 0, 1, 'W'}, {0, 1, 0, 'W'}, {1, 0, 0, 'F'}, {1, -1, 0, 'F'}}},
  {4, {1, -1, 0}, 1, 6, {{0, 1, 0, 'F'}, {0, 0, -1, 'W'}, {0, 0, 1, 'W'}, {0, -1, 0, 'W'}, {-1, 0, 0, 'F'}, {-1, 1, 0, 'F'}}},
  {4, {-1, -1, 0}, 1, 6, {{0, 1, 0, 'F'}, {0, 0, -1, 'W'}, {0, 0, 1, 'W'}, {0, -1, 0, 'W'}, {1, 0, 0, 'F'}, {1, 1, 0, 'F'}}},
};

/* material of cell (x,y,z): periodic axes wrap, elsewhere beyond the block is outside */
static int material_at(const int n[3], const int per[3], const uint8_t* m, int x, int y, int z)
{
  int r[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    if (r[d] < 0 || r[d] >= n[d]) {
      if (!per[d]) return 0;
      r[d] = (r[d] % n[d] + n[d]) % n[d];
    }
  }
  return m[((int64_t)r[0] * n[1] + r[1]) * n[2] + r[2]];
}

static void classify_cell(const int n[3], const int per[3], const uint8_t* m, int x, int y, int z,
                          int* type, int nrm[3])
{
  *type = 0;
  nrm[0] = nrm[1] = nrm[2] = 0;
  int found = 0;
  for (int r = 0; r < 46; ++r) {
    const olo_rule* R = &RULES[r];
    if (found && !R->overwrite) continue;
    int ok = 1;
    for (int k = 0; k < R->ncond && ok; ++k) {
      const int mat = material_at(n, per, m, x + R->c[k].dx, y + R->c[k].dy, z + R->c[k].dz);
      const int fluid = mat == 1, outside = mat == 0;
      switch (R->c[k].pred) {
      case 'W': ok = !fluid && !outside; break;
      case 'F': ok = fluid; break;
      default: ok = !fluid; break;
      }
    }
    if (!ok) continue;
    *type = R->type;
    for (int d = 0; d < 3; ++d) nrm[d] = R->n[d];
    found = 1;
  }
}

int olo_classify(int nx, int ny, int nz, int px, int py, int pz, const uint8_t* materials, int8_t* out)
{
  const int n[3] = {nx, ny, nz}, per[3] = {px, py, pz};
  const int64_t nc = (int64_t)nx * ny * nz;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t c = ((int64_t)x * ny + y) * nz + z;
        out[c] = -1;
        out[nc + c] = out[2 * nc + c] = out[3 * nc + c] = 0;
        if (materials[c] == 0 || materials[c] == 1) continue;
        int type, nrm[3];
        classify_cell(n, per, materials, x, y, z, &type, nrm);
        out[c] = (int8_t)type;
        for (int d = 0; d < 3; ++d) out[(d + 1) * nc + c] = (int8_t)nrm[d];
      }
  return 0;
}

static int append_cell(int64_t** list, int64_t* count, int64_t cell)
{
  int64_t* p = (int64_t*)realloc(*list, sizeof(int64_t) * (size_t)(*count + 1));
  if (!p) return -1;
  p[*count] = cell;
  *list = p;
  *count += 1;
  return 0;
}

/* boundary::set<LocalVelocity/LocalPressure> (setBoundary.h:83-101) with the
 * dynamics localVelocity3D.h:47-98 / localPressure3D.h:46-58 choose per type;
 * OuterVelocityEdge/CornerProcessor3D are registered for external edges and
 * corners (localVelocity3D.h:101-112).  The edge processor's template
 * arguments are (plane, n[(plane+1)%3], n[(plane+2)%3]) with plane the axis
 * the normal does not point along (setBoundary3D.h:583-625). */
int olo_define_local(olo_lattice* L, const uint8_t* materials, int material, int kind,
                     double rho, double ux, double uy, double uz)
{
  if (kind < OLO_LOCAL_VELOCITY || kind > OLO_INTERP_PRESSURE) { set_err("olo_define_local: kind"); return -1; }
  const int p = L->pad;
  if (!L->lkind) {
    L->lkind = (uint8_t*)calloc((size_t)L->ncells, 1);
    L->lnorm = (int8_t*)calloc((size_t)L->ncells * 3, 1);
    L->lval = (double*)calloc((size_t)L->ncells * 4, sizeof(double));
    if (!L->lkind || !L->lnorm || !L->lval) { set_err("olo_define_local: out of memory"); return -1; }
  }
  for (int x = 0; x < L->n[0]; ++x)
    for (int y = 0; y < L->n[1]; ++y)
      for (int z = 0; z < L->n[2]; ++z) {
        const int64_t c = ((int64_t)x * L->n[1] + y) * L->n[2] + z;
        if (materials[c] != material) continue;
        const int64_t cell = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p;
        int type, nrm[3], k;
        classify_cell(L->n, L->periodic, materials, x, y, z, &type, nrm);
        const int flat = type == 0 && (nrm[0] || nrm[1] || nrm[2]);
        if (kind == OLO_LOCAL_PRESSURE) {
          if (!flat) { set_err("olo_define_local: LocalPressure needs flat boundaries"); return -1; }
          k = OLK_FLAT_PRESSURE;
        } else if (kind == OLO_INTERP_PRESSURE) {
          /* interpolatedPressure3D.h:46-90: flat faces only, other types keep no dynamics */
          if (type == 0 && !flat) { set_err("olo_define_local: InterpolatedPressure on a cell without normal"); return -1; }
          k = flat ? OLK_FD_PRESSURE : OLK_DEAD;
        } else if (type == 0) {
          if (kind == OLO_INTERP_VELOCITY && !flat) {
            set_err("olo_define_local: InterpolatedVelocity on a cell without normal");
            return -1;
          }
          k = flat ? (kind == OLO_INTERP_VELOCITY ? OLK_FD_VELOCITY : OLK_FLAT_VELOCITY) : OLK_DEAD;
        } else {
          static const int map[5] = {0, OLK_OUTER_CORNER, OLK_INNER_CORNER, OLK_OUTER_EDGE, OLK_INNER_EDGE};
          k = map[type];
        }
        int8_t* ln = &L->lnorm[3 * cell];
        if (k == OLK_OUTER_EDGE || k == OLK_INNER_EDGE) {
          const int plane = nrm[0] == 0 ? 0 : (nrm[1] == 0 ? 1 : 2);
          ln[0] = (int8_t)plane;
          ln[1] = (int8_t)nrm[(plane + 1) % 3];
          ln[2] = (int8_t)nrm[(plane + 2) % 3];
        } else {
          for (int d = 0; d < 3; ++d) ln[d] = (int8_t)nrm[d];
        }
        L->lkind[cell] = (uint8_t)k;
        L->lval[4 * cell + 0] = rho;
        L->lval[4 * cell + 1] = ux;
        L->lval[4 * cell + 2] = uy;
        L->lval[4 * cell + 3] = uz;
        if (k == OLK_OUTER_EDGE && append_cell(&L->edge_cells, &L->n_edge, cell) != 0) return -1;
        if (k == OLK_OUTER_CORNER && append_cell(&L->corner_cells, &L->n_corner, cell) != 0) return -1;
        if ((k == OLK_FD_VELOCITY || k == OLK_FD_PRESSURE) && append_cell(&L->plane_cells, &L->n_plane, cell) != 0)
          return -1;
      }
  return 0;
}

static size_t storage_bytes(int storage) { return storage == OLO_F64 ? 8 : (storage == OLO_F32 ? 4 : 2); }

/* overlap-1 halo list: a cell outside the core is refreshed from the core cell
 * it wraps onto, if every out-of-range axis is periodic */
static int build_halo(olo_lattice* L)
{
  const int p = L->pad;
  int64_t cap = 0;
  for (int pass = 0; pass < 2; ++pass) {
    int64_t k = 0;
    for (int x = -1; x <= L->n[0]; ++x)
      for (int y = -1; y <= L->n[1]; ++y)
        for (int z = -1; z <= L->n[2]; ++z) {
          const int r[3] = {x, y, z};
          int outside = 0, ok = 1, w[3];
          for (int d = 0; d < 3; ++d) {
            w[d] = r[d];
            if (r[d] < 0 || r[d] >= L->n[d]) {
              outside = 1;
              if (!L->periodic[d]) ok = 0;
              w[d] = (r[d] + L->n[d]) % L->n[d];
            }
          }
          if (!outside || !ok) continue;
          if (pass == 1) {
            L->halo_dst[k] = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p;
            L->halo_src[k] = (int64_t)(w[0] + p) * L->proj[0] + (int64_t)(w[1] + p) * L->proj[1] + w[2] + p;
          }
          ++k;
        }
    if (pass == 0) {
      cap = k;
      L->halo_dst = (int64_t*)malloc(sizeof(int64_t) * (size_t)(cap ? cap : 1));
      L->halo_src = (int64_t*)malloc(sizeof(int64_t) * (size_t)(cap ? cap : 1));
      if (!L->halo_dst || !L->halo_src) return -1;
    }
  }
  L->nhalo = cap;
  return 0;
}

olo_lattice* olo_create(int q, int nx, int ny, int nz, int px, int py, int pz, int storage)
{
  if ((q != 19 && q != 27) || nx < 1 || ny < 1 || nz < 1 || storage < 0 || storage > 2) {
    set_err("olo_create: invalid arguments");
    return NULL;
  }
  olo_lattice* L = (olo_lattice*)calloc(1, sizeof(olo_lattice));
  if (!L) return NULL;
  L->q = q;
  L->n[0] = nx; L->n[1] = ny; L->n[2] = nz;
  L->periodic[0] = px; L->periodic[1] = py; L->periodic[2] = pz;
  L->storage = storage;
  L->pad = 1;
  for (int d = 0; d < 3; ++d) L->size[d] = L->n[d] + 2 * L->pad;
  L->proj[0] = L->size[1] * L->size[2];
  L->proj[1] = L->size[2];
  L->proj[2] = 1;
  L->ncells = L->size[0] * L->size[1] * L->size[2];
  L->ncore = (int64_t)nx * ny * nz;
  fill_descriptor(L);
  for (int i = 0; i < q; ++i) {
    L->shift[i] = 0;
    L->dist[i] = L->c[i][0] * L->proj[0] + L->c[i][1] * L->proj[1] + L->c[i][2];
  }
  /* new T[count]{} : zero-initialised storage (sisd/column.h:141-147) */
  L->pop = calloc((size_t)(q * L->ncells), storage_bytes(storage));
  L->ids = (uint8_t*)calloc((size_t)L->ncells, 1);
  if (!L->pop || !L->ids || build_halo(L) != 0) {
    set_err("olo_create: out of memory");
    olo_destroy(L);
    return NULL;
  }
  L->kind = OLO_BGK;
  L->omega = 1.0;
  for (int i = 0; i < 256; ++i) L->wall_rho[i] = 1.0;
  return L;
}

void olo_destroy(olo_lattice* L)
{
  if (!L) return;
  free(L->pop);
  free(L->force_field);
  free(L->ids);
  free(L->halo_dst);
  free(L->halo_src);
  free(L->lkind);
  free(L->lnorm);
  free(L->lval);
  free(L->edge_cells);
  free(L->corner_cells);
  free(L->plane_cells);
  free(L);
}

int olo_set_ids(olo_lattice* L, const uint8_t* ids)
{
  const int p = L->pad;
  for (int x = 0; x < L->n[0]; ++x)
    for (int y = 0; y < L->n[1]; ++y)
      for (int z = 0; z < L->n[2]; ++z)
        L->ids[(int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p] =
          ids[((int64_t)x * L->n[1] + y) * L->n[2] + z];
  return 0;
}

int olo_set_bulk(olo_lattice* L, int kind, double omega, double magic, double smago,
                 double fx, double fy, double fz)
{
  if (kind < OLO_BGK || kind > OLO_SMAGORINSKY) { set_err("olo_set_bulk: kind"); return -1; }
  L->kind = kind;
  L->omega = omega;
  L->magic = magic;
  L->smago = smago;
  L->force[0] = fx; L->force[1] = fy; L->force[2] = fz;
  return 0;
}

int olo_set_force_field(olo_lattice* L, const double* force)
{
  free(L->force_field);
  L->force_field = NULL;
  if (!force) return 0;
  L->force_field = (double*)calloc((size_t)L->ncells * 3, sizeof(double));
  if (!L->force_field) { set_err("olo_set_force_field: out of memory"); return -1; }
  const int p = L->pad;
  const int64_t nc = L->ncore;
  for (int x = 0; x < L->n[0]; ++x)
    for (int y = 0; y < L->n[1]; ++y)
      for (int z = 0; z < L->n[2]; ++z) {
        const int64_t c = ((int64_t)x * L->n[1] + y) * L->n[2] + z;
        const int64_t n = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p;
        for (int d = 0; d < 3; ++d) L->force_field[3 * n + d] = force[(int64_t)d * nc + c];
      }
  return 0;
}

int olo_set_wall(olo_lattice* L, int id, double rho, double ux, double uy, double uz)
{
  if (id < 3 || id > 255) { set_err("olo_set_wall: id must be in [3,255]"); return -1; }
  L->wall_rho[id] = rho;
  L->wall_u[id][0] = ux; L->wall_u[id][1] = uy; L->wall_u[id][2] = uz;
  return 0;
}

int olo_upload(olo_lattice* L, const double* f)
{
  switch (L->storage) {
  case OLO_F64: upload_f64(L, f); break;
  case OLO_F32: upload_f32(L, f); break;
  default: upload_f16(L, f); break;
  }
  return 0;
}

int olo_download(olo_lattice* L, double* f)
{
  switch (L->storage) {
  case OLO_F64: download_f64(L, f); break;
  case OLO_F32: download_f32(L, f); break;
  default: download_f16(L, f); break;
  }
  return 0;
}

int olo_step(olo_lattice* L, int64_t steps, int threads)
{
  if (threads < 1) threads = 1;
  for (int64_t s = 0; s < steps; ++s) {
    switch (L->storage) {
    case OLO_F64: step_f64(L, threads); break;
    case OLO_F32: step_f32(L, threads); break;
    default: step_f16(L, threads); break;
    }
  }
  return 0;
}

int olo_initialize(olo_lattice* L)
{
  switch (L->storage) {
  case OLO_F64: post_process_f64(L); break;
  case OLO_F32: post_process_f32(L); break;
  default: post_process_f16(L); break;
  }
  return 0;
}

int olo_statistics(const olo_lattice* L, double out[4])
{
  switch (L->storage) {
  case OLO_F64: statistics_f64(L, out); break;
  case OLO_F32: statistics_f32(L, out); break;
  default: statistics_f16(L, out); break;
  }
  return 0;
}

int olo_get_shifts(const olo_lattice* L, int64_t* shifts)
{
  for (int i = 0; i < L->q; ++i) shifts[i] = L->shift[i];
  return 0;
}

int olo_field_export(const olo_lattice* L, int field, int overlap, const double* conv, float* out)
{
  if (field < 0 || field > 3 || overlap < 0 || overlap > L->pad) { set_err("olo_field_export: field or overlap"); return -1; }
  switch (L->storage) {
  case OLO_F64: field_export_f64(L, field, overlap, conv, out); break;
  case OLO_F32: field_export_f32(L, field, overlap, conv, out); break;
  default: field_export_f16(L, field, overlap, conv, out); break;
  }
  return 0;
}

int olo_moments(const olo_lattice* L, double* rho_u)
{
  switch (L->storage) {
  case OLO_F64: moments_f64(L, rho_u); break;
  case OLO_F32: moments_f32(L, rho_u); break;
  default: moments_f16(L, rho_u); break;
  }
  return 0;
}

int olo_equilibrium(int q, int kind, int storage, int64_t n, const double* rho_u,
                    double fx, double fy, double fz, double* f)
{
  return olo_equilibrium_field(q, kind, storage, n, rho_u, fx, fy, fz, NULL, f);
}

int olo_equilibrium_field(int q, int kind, int storage, int64_t n, const double* rho_u,
                          double fx, double fy, double fz, const double* force, double* f)
{
  olo_lattice L;
  memset(&L, 0, sizeof L);
  if (q != 19 && q != 27) { set_err("olo_equilibrium: q"); return -1; }
  L.q = q;
  L.kind = kind;
  L.force[0] = fx; L.force[1] = fy; L.force[2] = fz;
  fill_descriptor(&L);
  switch (storage) {
  case OLO_F64: equilibrium_f64(&L, n, rho_u, force, f); break;
  case OLO_F32: equilibrium_f32(&L, n, rho_u, force, f); break;
  default: equilibrium_f16(&L, n, rho_u, force, f); break;
  }
  return 0;
}

int olo_round_to_storage(int storage, int64_t n, double* values)
{
  for (int64_t k = 0; k < n; ++k) {
    if (storage == OLO_F32) values[k] = (double)(float)values[k];
    else if (storage == OLO_F16)
      values[k] = (double)(half_to_float(float_to_half((float)values[k] * 4096.0f)) * (1.0f / 4096.0f));
  }
  return 0;
}
