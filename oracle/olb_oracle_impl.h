/* Storage-type instantiation of the oracle time step.  Included by
 * olb_oracle.c once per storage type with these macros defined:
 *   SFX        function-name suffix (f64, f32, f16)
 *   STORE_T    stored population type
 *   V          arithmetic type (the reference's value_t)
 *   LOAD(p)    STORE_T -> V
 *   STOREV(v)  V -> STORE_T
 * TEST INFRASTRUCTURE ONLY (see olb_oracle.h).
 */
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

static inline V FN(tw)(const olo_lattice* L, int i)
{
  /* Fraction::as<V>() = V(num) / V(den)   utilities/fraction.h:57-61 */
  return (V)L->tnum[i] / (V)L->tden[i];
}

/* Bulk momenta: rho = sum f + 1, j = sum c f, u = j / rho        lbm.h:266-307
 * Forced momenta add F/2 to u                           momenta/elements.h:1480-1488 */
static void FN(rho_u)(const olo_lattice* L, const V* f, int forced, V* rho, V* u)
{
  V r = 0;
  for (int i = 0; i < L->q; ++i) r += f[i];
  r += (V)1;
  V j[3] = {0, 0, 0};
  for (int i = 0; i < L->q; ++i)
    for (int d = 0; d < 3; ++d) j[d] += f[i] * (V)L->c[i][d];
  for (int d = 0; d < 3; ++d) u[d] = j[d] / r;
  if (forced)
    for (int d = 0; d < 3; ++d) u[d] += (V)L->force[d] * (V)0.5;
  *rho = r;
}

/* equilibrium<D>::secondOrder                                      lbm.h:50-64 */
static V FN(feq)(const olo_lattice* L, int i, V rho, const V* u, V usq)
{
  V cu = 0;
  for (int d = 0; d < 3; ++d) cu += (V)L->c[i][d] * u[d];
  const V inv = (V)3;
  const V t = FN(tw)(L, i);
  return rho * t * ((V)1 + inv * cu + inv * inv * (V)0.5 * cu * cu - inv * (V)0.5 * usq) - t;
}

static V FN(usqr)(const V* u)
{
  V s = 0;
  for (int d = 0; d < 3; ++d) s += u[d] * u[d];
  return s;
}

/* collision::BGK                                                collision.h:57-80 */
static void FN(bgk)(const olo_lattice* L, V* f, V rho, const V* u, V omega)
{
  const V usq = FN(usqr)(u);
  for (int i = 0; i < L->q; ++i) {
    const V e = FN(feq)(L, i, rho, u, usq);
    f[i] *= (V)1 - omega;
    f[i] += omega * e;
  }
}

/* collision::TRT                                               collision.h:209-247 */
static void FN(trt)(const olo_lattice* L, V* f, V rho, const V* u, V omega, V magic)
{
  const V usq = FN(usqr)(u);
  V e[27], fp[27], fm[27], ep[27], em[27];
  for (int i = 0; i < L->q; ++i) e[i] = FN(feq)(L, i, rho, u, usq);
  for (int i = 0; i < L->q; ++i) {
    const int o = L->opp[i];
    fp[i] = (V)0.5 * (f[i] + f[o]);
    fm[i] = (V)0.5 * (f[i] - f[o]);
  }
  for (int i = 0; i < L->q; ++i) {
    const int o = L->opp[i];
    ep[i] = (V)0.5 * (e[i] + e[o]);
    em[i] = (V)0.5 * (e[i] - e[o]);
  }
  const V omega2 = (V)1 / (magic / ((V)1 / omega - (V)0.5) + (V)0.5);
  for (int i = 0; i < L->q; ++i) f[i] -= omega * (fp[i] - ep[i]) + omega2 * (fm[i] - em[i]);
}

/* lbm<D>::addExternalForce (Guo source)                           lbm.h:545-566 */
static void FN(guo)(const olo_lattice* L, V* f, V rho, const V* u, V omega)
{
  const V inv = (V)3;
  for (int i = 0; i < L->q; ++i) {
    V cu = 0;
    for (int d = 0; d < 3; ++d) cu += (V)L->c[i][d] * u[d];
    cu *= inv * inv;
    V term = 0;
    for (int d = 0; d < 3; ++d)
      term += (((V)L->c[i][d] - u[d]) * inv + cu * (V)L->c[i][d]) * (V)L->force[d];
    term *= FN(tw)(L, i);
    term *= (V)1 - omega * (V)0.5;
    term *= rho;
    f[i] += term;
  }
}

/* SmagorinskyEffectiveOmega::computeEffectiveOmega          collisionLES.h:58-80
 * with computePiNeqNormSqr                          momenta/interface.h:164-173
 * and lbm<D>::computeStress                                     lbm.h:310-328 */
static V FN(smago_omega)(const olo_lattice* L, const V* f, V rho, const V* u, V omega, V cs)
{
  V pi[6];
  int k = 0;
  for (int a = 0; a < 3; ++a) {
    for (int b = a; b < 3; ++b) {
      V s = 0;
      for (int i = 0; i < L->q; ++i) s += (V)(L->c[i][a] * L->c[i][b]) * f[i];
      s -= rho * u[a] * u[b];
      if (a == b) s -= (V)1 / (V)3 * (rho - (V)1);
      pi[k++] = s;
    }
  }
  if (L->q == 19) {
    /* D3Q19 runs OpenLB's generated kernel (src/cse/dynamics/d7a4bdcd...cse.h:110):
     * omega_eff = 1 / (K1 sqrt(C^2 |Pi|/(sqrt2 rho) + K2 tau^2) + tau/2) with the
     * generator's constants K1 = 3.00000046417339, K2 = 0.0277777691819762
     * (3 and 1/36 up to 1.5e-7 and 3.1e-7 relative). */
    const V nh = (V)sqrt((double)((V)0.5 * (pi[0] * pi[0]) + (V)0.5 * (pi[3] * pi[3]) +
                                  (V)0.5 * (pi[5] * pi[5]) + pi[1] * pi[1] + pi[2] * pi[2] +
                                  pi[4] * pi[4]));
    const V inner = (V)1 / rho * (cs * cs) * nh + (V)0.0277777691819762 / (omega * omega);
    return (V)1 / ((V)3.00000046417339 * (V)sqrt((double)inner) + (V)0.5 / omega);
  }
  /* the reference's literal 2. promotes these sums to double */
  V nsq = (V)((double)pi[0] * pi[0] + 2. * pi[1] * pi[1] + (double)pi[2] * pi[2]);
  nsq = (V)(nsq + ((double)pi[2] * pi[2] + (double)pi[3] * pi[3] + 2. * pi[4] * pi[4] + (double)pi[5] * pi[5]));
  const V norm = (V)sqrt((double)nsq);
  const V pre = (V)((double)(cs * cs * (V)3 * (V)3 * (V)2) * sqrt(2.0));
  const V tauMol = (V)1 / omega;
  const V tauTurb = (V)0.5 * ((V)sqrt((double)(tauMol * tauMol + pre / rho * norm)) - tauMol);
  const V tauEff = tauMol + tauTurb;
  return (V)1 / tauEff;
}

/* collision::Revert                                            collision.h:293-312 */
static void FN(revert)(const olo_lattice* L, V* f)
{
  for (int i = 1; i <= L->q / 2; ++i) {
    const V tmp = f[i];
    f[i] = f[L->opp[i]];
    f[L->opp[i]] = tmp;
  }
}

/* NguyenLaddCorrection<Revert> with FixedDensity + FixedVelocityMomentum
 *                                                              collision.h:323-348 */
static void FN(moving_wall)(const olo_lattice* L, V* f, int id)
{
  FN(revert)(L, f);
  const V rho = (V)L->wall_rho[id];
  const V u[3] = {(V)L->wall_u[id][0], (V)L->wall_u[id][1], (V)L->wall_u[id][2]};
  for (int i = 1; i < L->q; ++i)
    for (int d = 0; d < 3; ++d)
      f[i] += (V)2 * rho * u[d] * (V)L->c[i][d] * FN(tw)(L, i) * (V)3;
}

static void FN(collide_cell)(const olo_lattice* L, V* f, int id)
{
  if (id == 2) { FN(revert)(L, f); return; }
  if (id >= 3) { FN(moving_wall)(L, f, id); return; }
  const V omega = (V)L->omega;
  V rho, u[3];
  switch (L->kind) {
  case OLO_BGK:
    FN(rho_u)(L, f, 0, &rho, u);
    FN(bgk)(L, f, rho, u, omega);
    break;
  case OLO_FORCED_BGK:
    FN(rho_u)(L, f, 1, &rho, u);
    FN(bgk)(L, f, rho, u, omega);
    FN(guo)(L, f, rho, u, omega);
    break;
  case OLO_TRT:
    FN(rho_u)(L, f, 0, &rho, u);
    FN(trt)(L, f, rho, u, omega, (V)L->magic);
    break;
  case OLO_FORCED_TRT:
    FN(rho_u)(L, f, 1, &rho, u);
    FN(trt)(L, f, rho, u, omega, (V)L->magic);
    FN(guo)(L, f, rho, u, omega);
    break;
  case OLO_SMAGORINSKY: {
    FN(rho_u)(L, f, 0, &rho, u);
    const V w = FN(smago_omega)(L, f, rho, u, omega, (V)L->smago);
    FN(bgk)(L, f, rho, u, w);
    break;
  }
  default:
    break;
  }
}

/* One collideAndStream(): collide core cells in place, refresh the overlap-1
 * halo from the periodic images, rotate every column. */
static void FN(step)(olo_lattice* L, int threads)
{
  STORE_T* pop = (STORE_T*)L->pop;
  const int64_t N = L->ncells;
  const int q = L->q;
  const int p = L->pad;
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int x = 0; x < L->n[0]; ++x) {
    V f[27];
    int64_t idx[27];
    for (int y = 0; y < L->n[1]; ++y) {
      const int64_t row = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1];
      for (int z = 0; z < L->n[2]; ++z) {
        const int64_t n = row + z + p;
        const int id = L->ids[n];
        if (id == 0) continue;
        for (int i = 0; i < q; ++i) {
          int64_t k = n + L->shift[i];
          if (k >= N) k -= N;
          idx[i] = (int64_t)i * N + k;
          f[i] = LOAD(pop[idx[i]]);
        }
        FN(collide_cell)(L, f, id);
        for (int i = 0; i < q; ++i) pop[idx[i]] = STOREV(f[i]);
      }
    }
  }
  /* halo copy: all q populations of each overlap cell from its periodic image */
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t h = 0; h < L->nhalo; ++h) {
    const int64_t dst = L->halo_dst[h], src = L->halo_src[h];
    for (int i = 0; i < q; ++i) {
      int64_t kd = dst + L->shift[i], ks = src + L->shift[i];
      if (kd >= N) kd -= N;
      if (ks >= N) ks -= N;
      pop[(int64_t)i * N + kd] = pop[(int64_t)i * N + ks];
    }
  }
  for (int i = 0; i < q; ++i) {
    int64_t s = L->shift[i] - L->dist[i];
    s %= N;
    if (s < 0) s += N;
    L->shift[i] = s;
  }
}

static void FN(upload)(olo_lattice* L, const double* f)
{
  STORE_T* pop = (STORE_T*)L->pop;
  const int64_t N = L->ncells, nc = L->ncore;
  const int p = L->pad;
  for (int i = 0; i < L->q; ++i)
    for (int x = 0; x < L->n[0]; ++x)
      for (int y = 0; y < L->n[1]; ++y)
        for (int z = 0; z < L->n[2]; ++z) {
          const int64_t c = ((int64_t)x * L->n[1] + y) * L->n[2] + z;
          const int64_t n = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p;
          int64_t k = n + L->shift[i];
          if (k >= N) k -= N;
          pop[(int64_t)i * N + k] = STOREV((V)f[(int64_t)i * nc + c]);
        }
}

static void FN(download)(const olo_lattice* L, double* f)
{
  const STORE_T* pop = (const STORE_T*)L->pop;
  const int64_t N = L->ncells, nc = L->ncore;
  const int p = L->pad;
  for (int i = 0; i < L->q; ++i)
    for (int x = 0; x < L->n[0]; ++x)
      for (int y = 0; y < L->n[1]; ++y)
        for (int z = 0; z < L->n[2]; ++z) {
          const int64_t c = ((int64_t)x * L->n[1] + y) * L->n[2] + z;
          const int64_t n = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p;
          int64_t k = n + L->shift[i];
          if (k >= N) k -= N;
          f[(int64_t)i * nc + c] = (double)LOAD(pop[(int64_t)i * N + k]);
        }
}

static void FN(moments)(const olo_lattice* L, double* out)
{
  const STORE_T* pop = (const STORE_T*)L->pop;
  const int64_t N = L->ncells, nc = L->ncore;
  const int p = L->pad;
  const int forced = (L->kind == OLO_FORCED_BGK || L->kind == OLO_FORCED_TRT);
  for (int x = 0; x < L->n[0]; ++x)
    for (int y = 0; y < L->n[1]; ++y)
      for (int z = 0; z < L->n[2]; ++z) {
        const int64_t c = ((int64_t)x * L->n[1] + y) * L->n[2] + z;
        const int64_t n = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p;
        V f[27], rho, u[3];
        for (int i = 0; i < L->q; ++i) {
          int64_t k = n + L->shift[i];
          if (k >= N) k -= N;
          f[i] = LOAD(pop[(int64_t)i * N + k]);
        }
        FN(rho_u)(L, f, forced && L->ids[n] == 1, &rho, u);
        out[c] = (double)rho;
        for (int d = 0; d < 3; ++d) out[(int64_t)(d + 1) * nc + c] = (double)u[d];
      }
}

static void FN(equilibrium)(olo_lattice* L, int64_t n, const double* rho_u, double* f)
{
  const int forced = (L->kind == OLO_FORCED_BGK || L->kind == OLO_FORCED_TRT);
  for (int64_t c = 0; c < n; ++c) {
    const V rho = (V)rho_u[c];
    V u[3];
    for (int d = 0; d < 3; ++d) {
      u[d] = (V)rho_u[(int64_t)(d + 1) * n + c];
      /* ForcedMomentum::inverseShift, dynamics/interface.h:111-121 */
      if (forced) u[d] -= (V)L->force[d] * (V)0.5;
    }
    const V usq = FN(usqr)(u);
    for (int i = 0; i < L->q; ++i) {
      const STORE_T s = STOREV(FN(feq)(L, i, rho, u, usq));
      f[(int64_t)i * n + c] = (double)LOAD(s);
    }
  }
}

#undef FN
#undef CAT
#undef CAT2
