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
  double wall_rho[256], wall_u[256][3];
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
  free(L->ids);
  free(L->halo_dst);
  free(L->halo_src);
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

int olo_get_shifts(const olo_lattice* L, int64_t* shifts)
{
  for (int i = 0; i < L->q; ++i) shifts[i] = L->shift[i];
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
  olo_lattice L;
  memset(&L, 0, sizeof L);
  if (q != 19 && q != 27) { set_err("olo_equilibrium: q"); return -1; }
  L.q = q;
  L.kind = kind;
  L.force[0] = fx; L.force[1] = fy; L.force[2] = fz;
  fill_descriptor(&L);
  switch (storage) {
  case OLO_F64: equilibrium_f64(&L, n, rho_u, f); break;
  case OLO_F32: equilibrium_f32(&L, n, rho_u, f); break;
  default: equilibrium_f16(&L, n, rho_u, f); break;
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
