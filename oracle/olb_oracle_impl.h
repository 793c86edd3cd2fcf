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
 * Forced momenta add F/2 to u, F the cell's FORCE field   momenta/elements.h:1480-1488
 * (F == NULL: plain momenta) */
static void FN(rho_u)(const olo_lattice* L, const V* f, const double* F, V* rho, V* u)
{
  V r = 0;
  for (int i = 0; i < L->q; ++i) r += f[i];
  r += (V)1;
  V j[3] = {0, 0, 0};
  for (int i = 0; i < L->q; ++i)
    for (int d = 0; d < 3; ++d) j[d] += f[i] * (V)L->c[i][d];
  for (int d = 0; d < 3; ++d) u[d] = j[d] / r;
  if (F)
    for (int d = 0; d < 3; ++d) u[d] += (V)F[d] * (V)0.5;
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
static void FN(guo)(const olo_lattice* L, V* f, V rho, const V* u, V omega, const double* F)
{
  const V inv = (V)3;
  for (int i = 0; i < L->q; ++i) {
    V cu = 0;
    for (int d = 0; d < 3; ++d) cu += (V)L->c[i][d] * u[d];
    cu *= inv * inv;
    V term = 0;
    for (int d = 0; d < 3; ++d)
      term += (((V)L->c[i][d] - u[d]) * inv + cu * (V)L->c[i][d]) * (V)F[d];
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

/* ---- local (regularized) boundaries ------------------------------------ */

/* velocityBMRho<dir,orient>                          momenta/elements.h:54-70 */
static V FN(bm_rho)(const olo_lattice* L, const V* f, int dir, int orient, const V* u)
{
  V on = 0, nrm = 0;
  for (int i = 0; i < L->q; ++i)
    if (L->c[i][dir] == 0) on += f[i];
  for (int i = 0; i < L->q; ++i)
    if (L->c[i][dir] == orient) nrm += f[i];
  return ((V)2 * nrm + on + (V)1) / ((V)1 + (V)orient * u[dir]);
}

/* BoundaryHelpers<dir,orient>::computeStress                boundary/helper.h:31-80 */
static void FN(bc_stress)(const olo_lattice* L, const V* f, int dir, int orient, V rho, const V* u, V* pi)
{
  const V usq = FN(usqr)(u);
  V fneq[27];
  for (int i = 0; i < L->q; ++i) {
    fneq[i] = 0;
    if (L->c[i][dir] == 0) fneq[i] = (i == 0) ? (V)0 : f[i] - FN(feq)(L, i, rho, u, usq);
  }
  for (int i = 0; i < L->q; ++i)
    if (L->c[i][dir] == orient) fneq[i] = f[i] - FN(feq)(L, i, rho, u, usq);
  int k = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      V s = 0;
      for (int i = 0; i < L->q; ++i)
        if (L->c[i][dir] == 0) s += (V)(L->c[i][a] * L->c[i][b]) * fneq[i];
      for (int i = 0; i < L->q; ++i)
        if (L->c[i][dir] == orient) s += (V)2 * (V)(L->c[i][a] * L->c[i][b]) * fneq[i];
      pi[k++] = s;
    }
}

/* lbm<D>::computeStress on shifted populations                   lbm.h:311-329 */
static void FN(bulk_stress)(const olo_lattice* L, const V* f, V rho, const V* u, V* pi)
{
  int k = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      V s = 0;
      for (int i = 0; i < L->q; ++i) s += (V)(L->c[i][a] * L->c[i][b]) * f[i];
      s -= rho * u[a] * u[b];
      if (a == b) s -= (V)1 / (V)3 * (rho - (V)1);
      pi[k++] = s;
    }
}

/* equilibrium<D>::fromPiToFneq                                   lbm.h:225-247 */
static V FN(pi_to_fneq)(const olo_lattice* L, int i, const V* pi)
{
  V fneq = 0;
  int k = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      V add = (V)(L->c[i][a] * L->c[i][b]);
      if (a == b) add -= (V)1 / (V)3;
      else add *= (V)2;
      add *= pi[k++];
      fneq += add;
    }
  fneq *= FN(tw)(L, i) * (V)3 * (V)3 / (V)2;
  return fneq;
}

/* InnerEdgeDensity3D / InnerCornerDensity3D              momenta/elements.h:400-445 */
static V FN(inner_rho)(const olo_lattice* L, const V* f, int k, const int8_t* ln, const V* u)
{
  if (k == OLK_INNER_EDGE) {
    const int p = ln[0];
    const V r1 = FN(bm_rho)(L, f, (p + 1) % 3, ln[1], u);
    const V r2 = FN(bm_rho)(L, f, (p + 2) % 3, ln[2], u);
    return (r1 + r2) / (V)2;
  }
  const V rx = FN(bm_rho)(L, f, 0, ln[0], u);
  const V ry = FN(bm_rho)(L, f, 1, ln[1], u);
  const V rz = FN(bm_rho)(L, f, 2, ln[2], u);
  return (rx + ry + rz) / (V)3;
}

/* InnerEdgeStress3D / InnerCornerStress3D: the unknown populations are
 * replaced by bounce-back of their non-equilibrium part, then bulk stress
 *                                                   momenta/elements.h:1711-1765 */
static void FN(inner_stress)(const olo_lattice* L, const V* f, int k, const int8_t* ln, V rho, const V* u, V* pi)
{
  const V usq = FN(usqr)(u);
  V g[27];
  for (int i = 0; i < L->q; ++i) g[i] = f[i];
  for (int i = 0; i < L->q; ++i) {
    int unknown;
    if (k == OLK_INNER_EDGE) {
      const int p = ln[0];
      unknown = L->c[i][(p + 1) % 3] == -ln[1] && L->c[i][(p + 2) % 3] == -ln[2];
    } else {
      unknown = L->c[i][0] == -ln[0] && L->c[i][1] == -ln[1] && L->c[i][2] == -ln[2];
    }
    if (!unknown) continue;
    const int o = L->opp[i];
    g[i] = g[o] - FN(feq)(L, o, rho, u, usq) + FN(feq)(L, i, rho, u, usq);
  }
  FN(bulk_stress)(L, g, rho, u, pi);
}

/* rho and u of a local-boundary cell, from its boundary momenta */
static void FN(local_rho_u)(const olo_lattice* L, const V* f, int64_t n, V* rho, V* u)
{
  const int k = L->lkind[n];
  const int8_t* ln = &L->lnorm[3 * n];
  const double* lv = &L->lval[4 * n];
  for (int d = 0; d < 3; ++d) u[d] = (V)lv[1 + d];
  switch (k) {
  case OLK_FLAT_VELOCITY:
  case OLK_FD_VELOCITY: {
    const int dir = ln[0] ? 0 : (ln[1] ? 1 : 2);
    *rho = FN(bm_rho)(L, f, dir, ln[dir], u);
    break;
  }
  case OLK_FLAT_PRESSURE:
  case OLK_FD_PRESSURE: {
    /* FixedDensity + FixedPressureMomentum::computeU      elements.h:589-613 */
    const int dir = ln[0] ? 0 : (ln[1] ? 1 : 2), orient = ln[dir];
    const V r = (V)lv[0];
    V on = 0, nrm = 0;
    for (int i = 0; i < L->q; ++i)
      if (L->c[i][dir] == 0) on += f[i];
    for (int i = 0; i < L->q; ++i)
      if (L->c[i][dir] == orient) nrm += f[i];
    u[dir] = (V)orient * (((V)2 * nrm + on + (V)1) / r - (V)1);
    *rho = r;
    break;
  }
  case OLK_OUTER_EDGE:
  case OLK_OUTER_CORNER: {
    V r = 0;
    for (int i = 0; i < L->q; ++i) r += f[i];
    *rho = r + (V)1;
    break;
  }
  case OLK_INNER_EDGE:
  case OLK_INNER_CORNER:
    *rho = FN(inner_rho)(L, f, k, ln, u);
    break;
  default:
    *rho = (V)1;
    u[0] = u[1] = u[2] = 0;
    break;
  }
}

/* CombinedRLBdynamics::collide (dynamics.h:420-432), or plain BGK with
 * FixedVelocityBoundaryTuple on outer edges and corners; returns the
 * statistic the collision reports (rho, |u|^2), rho < 0 for none */
static V FN(collide_local)(const olo_lattice* L, V* f, int64_t n, V* usq_out)
{
  const int k = L->lkind[n];
  const int8_t* ln = &L->lnorm[3 * n];
  const V omega = (V)L->omega;
  V rho, u[3];
  if (k == OLK_DEAD || k == OLK_NONE) return (V)-1;
  if (k != OLK_OUTER_EDGE && k != OLK_OUTER_CORNER && k != OLK_FD_VELOCITY && k != OLK_FD_PRESSURE) {
    V pi[6];
    FN(local_rho_u)(L, f, n, &rho, u);
    if (k == OLK_FLAT_VELOCITY || k == OLK_FLAT_PRESSURE) {
      const int dir = ln[0] ? 0 : (ln[1] ? 1 : 2);
      FN(bc_stress)(L, f, dir, ln[dir], rho, u, pi);
    } else {
      FN(inner_stress)(L, f, k, ln, rho, u, pi);
    }
    const V usq = FN(usqr)(u);
    for (int i = 0; i < L->q; ++i) f[i] = FN(feq)(L, i, rho, u, usq) + FN(pi_to_fneq)(L, i, pi);
  }
  FN(local_rho_u)(L, f, n, &rho, u);
  FN(bgk)(L, f, rho, u, omega);
  *usq_out = FN(usqr)(u);
  return rho;
}

static void FN(collide_cell)(const olo_lattice* L, V* f, int id, int64_t n)
{
  const double* F = cell_force(L, n);
  if (id == 2) { FN(revert)(L, f); return; }
  if (id >= 3) { FN(moving_wall)(L, f, id); return; }
  const V omega = (V)L->omega;
  V rho, u[3];
  switch (L->kind) {
  case OLO_BGK:
    FN(rho_u)(L, f, NULL, &rho, u);
    FN(bgk)(L, f, rho, u, omega);
    break;
  case OLO_FORCED_BGK:
    FN(rho_u)(L, f, F, &rho, u);
    FN(bgk)(L, f, rho, u, omega);
    FN(guo)(L, f, rho, u, omega, F);
    break;
  case OLO_TRT:
    FN(rho_u)(L, f, NULL, &rho, u);
    FN(trt)(L, f, rho, u, omega, (V)L->magic);
    break;
  case OLO_FORCED_TRT:
    FN(rho_u)(L, f, F, &rho, u);
    FN(trt)(L, f, rho, u, omega, (V)L->magic);
    FN(guo)(L, f, rho, u, omega, F);
    break;
  case OLO_SMAGORINSKY: {
    FN(rho_u)(L, f, NULL, &rho, u);
    const V w = FN(smago_omega)(L, f, rho, u, omega, (V)L->smago);
    FN(bgk)(L, f, rho, u, w);
    break;
  }
  default:
    break;
  }
}

static void FN(load_cell)(const olo_lattice* L, int64_t n, V* f)
{
  const STORE_T* pop = (const STORE_T*)L->pop;
  const int64_t N = L->ncells;
  for (int i = 0; i < L->q; ++i) {
    int64_t k = n + L->shift[i];
    if (k >= N) k -= N;
    f[i] = LOAD(pop[(int64_t)i * N + k]);
  }
}

static void FN(store_cell)(olo_lattice* L, int64_t n, const V* f)
{
  STORE_T* pop = (STORE_T*)L->pop;
  const int64_t N = L->ncells;
  for (int i = 0; i < L->q; ++i) {
    int64_t k = n + L->shift[i];
    if (k >= N) k -= N;
    pop[(int64_t)i * N + k] = STOREV(f[i]);
  }
}

/* padded index of cell n moved by (dx,dy,dz); periodic axes wrap inside the core */
static int64_t FN(neighbor)(const olo_lattice* L, int64_t n, int dx, int dy, int dz)
{
  const int p = L->pad;
  int r[3] = {(int)(n / L->proj[0]) - p, (int)((n % L->proj[0]) / L->proj[1]) - p, (int)(n % L->proj[1]) - p};
  const int o[3] = {dx, dy, dz};
  for (int d = 0; d < 3; ++d) {
    r[d] += o[d];
    if (L->periodic[d] && (r[d] < 0 || r[d] >= L->n[d])) r[d] = (r[d] % L->n[d] + L->n[d]) % L->n[d];
  }
  return (int64_t)(r[0] + p) * L->proj[0] + (int64_t)(r[1] + p) * L->proj[1] + r[2] + p;
}

/* computeRho / computeU of any cell through its dynamics' momenta */
static void FN(cell_rho_u)(const olo_lattice* L, int64_t n, V* rho, V* u)
{
  const int id = L->ids[n];
  if (L->lkind && L->lkind[n]) {
    V f[27];
    FN(load_cell)(L, n, f);
    FN(local_rho_u)(L, f, n, rho, u);
    return;
  }
  if (id == 1) {
    V f[27];
    FN(load_cell)(L, n, f);
    FN(rho_u)(L, f, (L->kind == OLO_FORCED_BGK || L->kind == OLO_FORCED_TRT) ? cell_force(L, n) : NULL, rho, u);
    return;
  }
  u[0] = u[1] = u[2] = 0;
  *rho = (V)1;  /* NoDynamics: OneDensity; BounceBack: FixedDensity defined as 1 */
  if (id >= 3) {
    *rho = (V)L->wall_rho[id];
    for (int d = 0; d < 3; ++d) u[d] = (V)L->wall_u[id][d];
  }
}

/* The float buffer SuperVTMwriter3D::dataArray fills (io/superVtmWriter3D.hh:
 * 437-458): points z-outer, x-fastest over [-overlap, n + overlap), each the
 * functor's value on that cell -- density / velocity / physVelocity /
 * physPressure (functors/lattice/latticeDensity3D.hh, latticeVelocity3D.hh,
 * latticePhysVelocity3D.hh:84-96, latticePhysPressure3D.hh:71-78).  A point
 * on a periodic axis holds its image's populations and FORCE after the
 * writer's communicate() (superVtmWriter3D.hh:71-75), so an image of a bulk
 * cell reports the bulk cell's moments; any other overlap cell reports rho 1,
 * u 0: wall dynamics keep their initialize() fields there (RHO 1, velocity
 * 0; those fields are not communicated), boundary::set leaves the outermost
 * layer without dynamics (setBoundary.h:87-88), and beyond a non-periodic
 * face there are none. */
static void FN(field_export)(const olo_lattice* L, int field, int overlap, const double* conv, float* out)
{
  const int p = L->pad;
  int64_t k = 0;
  for (int z = -overlap; z < L->n[2] + overlap; ++z)
    for (int y = -overlap; y < L->n[1] + overlap; ++y)
      for (int x = -overlap; x < L->n[0] + overlap; ++x, ++k) {
        int r[3] = {x, y, z};
        int inside = 1, image = 0;
        for (int d = 0; d < 3; ++d) {
          if (r[d] >= 0 && r[d] < L->n[d]) continue;
          image = 1;
          if (!L->periodic[d]) inside = 0;
          else r[d] = (r[d] % L->n[d] + L->n[d]) % L->n[d];
        }
        V rho = (V)1, u[3] = {0, 0, 0};
        if (inside) {
          const int64_t n = (int64_t)(r[0] + p) * L->proj[0] + (int64_t)(r[1] + p) * L->proj[1] + r[2] + p;
          if (!image || L->ids[n] == 1) FN(cell_rho_u)(L, n, &rho, u);
        }
        const int dim = (field == 1 || field == 2) ? 3 : 1;
        float* o = out + k * dim;
        if (field == 0) {
          o[0] = (float)rho;
        } else if (field == 1) {
          for (int d = 0; d < 3; ++d) o[d] = (float)u[d];
        } else if (field == 2) {
          for (int d = 0; d < 3; ++d) o[d] = (float)((V)conv[0] * u[d]);
        } else {
          /* getPhysPressure's a*b+c is contracted to one fused multiply-add
           * in the compiled reference (its FP32 output matches the fused
           * form on every point, the separately rounded one not) */
          const V lp = (V)(((double)rho - 1.0) / 3.0);
          const V cp = (V)conv[1], p0 = (V)conv[2];
          o[0] = (float)(sizeof(V) == 4 ? (V)fmaf((float)cp, (float)lp, (float)p0) : (V)fma(cp, lp, p0));
        }
      }
}

/* fd::boundaryGradient / centralGradient                  finiteDifference.h:33-44 */
static V FN(bgrad)(V u0, V u1, V u2) { return (-(V)3 * u0 + (V)4 * u1 - (V)1 * u2) / (V)2; }

/* regularized reconstruction from rho, u and the velocity gradient
 * dA_uB[a][b] = d_a u_b (boundaryPostProcessors3D.hh:190-211, 270-290) */
static void FN(fd_reconstruct)(olo_lattice* L, int64_t n, V rho, V (*dA)[3])
{
  const V omega = (V)L->omega;
  const V s = -rho / (V)3 / omega;
  V pi[6];
  pi[0] = (V)2 * dA[0][0] * s;
  pi[3] = (V)2 * dA[1][1] * s;
  pi[5] = (V)2 * dA[2][2] * s;
  pi[1] = (dA[0][1] + dA[1][0]) * s;
  pi[2] = (dA[0][2] + dA[2][0]) * s;
  pi[4] = (dA[1][2] + dA[2][1]) * s;
  V u[3];
  for (int d = 0; d < 3; ++d) u[d] = (V)L->lval[4 * n + 1 + d];
  const V usq = FN(usqr)(u);
  V f[27];
  for (int i = 0; i < L->q; ++i) f[i] = FN(feq)(L, i, rho, u, usq) + FN(pi_to_fneq)(L, i, pi);
  FN(store_cell)(L, n, f);
}

static void FN(neighbor_u)(const olo_lattice* L, int64_t n, const int* off, V* u)
{
  V r;
  FN(cell_rho_u)(L, FN(neighbor)(L, n, off[0], off[1], off[2]), &r, u);
}

static V FN(neighbor_rho)(const olo_lattice* L, int64_t n, const int* off)
{
  V r, u[3];
  FN(cell_rho_u)(L, FN(neighbor)(L, n, off[0], off[1], off[2]), &r, u);
  return r;
}

/* OuterVelocityEdgeProcessor3D<plane,n1,n2>::apply    boundaryPostProcessors3D.hh:166-238 */
static void FN(outer_edge)(olo_lattice* L, int64_t n)
{
  const int8_t* ln = &L->lnorm[3 * n];
  const int p = ln[0], n1 = ln[1], n2 = ln[2];
  const int d1 = (p + 1) % 3, d2 = (p + 2) % 3;
  int o[3];
  o[0] = o[1] = o[2] = 0; o[d1] = -n1;     const V rho10 = FN(neighbor_rho)(L, n, o);
  o[0] = o[1] = o[2] = 0; o[d2] = -n2;     const V rho01 = FN(neighbor_rho)(L, n, o);
  o[0] = o[1] = o[2] = 0; o[d1] = -2 * n1; const V rho20 = FN(neighbor_rho)(L, n, o);
  o[0] = o[1] = o[2] = 0; o[d2] = -2 * n2; const V rho02 = FN(neighbor_rho)(L, n, o);
  const V rho = (V)2 / (V)3 * (rho01 + rho10) - (V)1 / (V)6 * (rho02 + rho20);
  V dA[3][3], u0[3], u1[3], u2[3];
  for (int d = 0; d < 3; ++d) u0[d] = (V)L->lval[4 * n + 1 + d];
  /* along the edge: central difference */
  o[0] = o[1] = o[2] = 0; o[p] = 1;  FN(neighbor_u)(L, n, o, u1);
  o[0] = o[1] = o[2] = 0; o[p] = -1; FN(neighbor_u)(L, n, o, u2);
  for (int b = 0; b < 3; ++b) dA[p][b] = (u1[b] - u2[b]) / (V)2;
  /* into the fluid: one-sided second-order differences */
  const int dirs[2] = {d1, d2}, nrm[2] = {n1, n2};
  for (int k = 0; k < 2; ++k) {
    o[0] = o[1] = o[2] = 0; o[dirs[k]] = -nrm[k];     FN(neighbor_u)(L, n, o, u1);
    o[0] = o[1] = o[2] = 0; o[dirs[k]] = -2 * nrm[k]; FN(neighbor_u)(L, n, o, u2);
    for (int b = 0; b < 3; ++b) dA[dirs[k]][b] = (V)(-nrm[k]) * FN(bgrad)(u0[b], u1[b], u2[b]);
  }
  FN(fd_reconstruct)(L, n, rho, dA);
}

/* OuterVelocityCornerProcessor3D<nx,ny,nz>::apply     boundaryPostProcessors3D.hh:242-292 */
static void FN(outer_corner)(olo_lattice* L, int64_t n)
{
  const int8_t* ln = &L->lnorm[3 * n];
  V r1[3], r2[3];
  for (int d = 0; d < 3; ++d) {
    int o[3] = {0, 0, 0};
    o[d] = -ln[d];
    r1[d] = FN(neighbor_rho)(L, n, o);
    o[d] = -2 * ln[d];
    r2[d] = FN(neighbor_rho)(L, n, o);
  }
  const V rho = (V)4 / (V)9 * (r1[2] + r1[1] + r1[0]) - (V)1 / (V)9 * (r2[2] + r2[1] + r2[0]);
  V dA[3][3], u0[3], u1[3], u2[3];
  for (int d = 0; d < 3; ++d) u0[d] = (V)L->lval[4 * n + 1 + d];
  for (int a = 0; a < 3; ++a) {
    int o[3] = {0, 0, 0};
    o[a] = -ln[a];     FN(neighbor_u)(L, n, o, u1);
    o[a] = -2 * ln[a]; FN(neighbor_u)(L, n, o, u2);
    for (int b = 0; b < 3; ++b) dA[a][b] = (V)(-ln[a]) * FN(bgrad)(u0[b], u1[b], u2[b]);
  }
  FN(fd_reconstruct)(L, n, rho, dA);
}

/* PlaneFdBoundaryProcessor3D<dir,orient>::apply            boundaryPostProcessors3D.hh:38-92
 * rho and u from the cell's Dirichlet momenta, the velocity gradient by a
 * one-sided second-order difference along the normal and central differences
 * along the face (finiteDifference3D.h:73-215), then the regularized
 * reconstruction.  The result goes to `out` (q values) instead of the lattice
 * so that the caller decides when neighbours see it. */
static void FN(plane_fd)(const olo_lattice* L, int64_t n, V* out)
{
  const int8_t* ln = &L->lnorm[3 * n];
  const int dir = ln[0] ? 0 : (ln[1] ? 1 : 2), orient = ln[dir];
  V f[27], rho, u[3];
  FN(load_cell)(L, n, f);
  FN(local_rho_u)(L, f, n, &rho, u);
  V dA[3][3], u1[3], u2[3];
  for (int d = 0; d < 3; ++d) {
    int o[3] = {0, 0, 0};
    if (d == dir) {
      o[d] = -orient;     FN(neighbor_u)(L, n, o, u1);
      o[d] = -2 * orient; FN(neighbor_u)(L, n, o, u2);
      for (int b = 0; b < 3; ++b) dA[d][b] = (V)(-orient) * FN(bgrad)(u[b], u1[b], u2[b]);
    } else {
      o[d] = 1;  FN(neighbor_u)(L, n, o, u1);
      o[d] = -1; FN(neighbor_u)(L, n, o, u2);
      for (int b = 0; b < 3; ++b) dA[d][b] = (u1[b] - u2[b]) / (V)2;
    }
  }
  const V omega = (V)L->omega;
  const V s = -rho / (V)3 / omega;
  V pi[6];
  pi[0] = (V)2 * dA[0][0] * s;
  pi[3] = (V)2 * dA[1][1] * s;
  pi[5] = (V)2 * dA[2][2] * s;
  pi[1] = (dA[0][1] + dA[1][0]) * s;
  pi[2] = (dA[0][2] + dA[2][0]) * s;
  pi[4] = (dA[1][2] + dA[2][1]) * s;
  const V usq = FN(usqr)(u);
  for (int i = 0; i < L->q; ++i) out[i] = FN(feq)(L, i, rho, u, usq) + FN(pi_to_fneq)(L, i, pi);
}

/* The plane processors run one cell after the other in cell order, as the CPU
 * platform applies them.  Their order against each other and against the
 * edge processors (all priority 0) changes the result only by rounding: the
 * regularized reconstruction keeps the partial population sums the Dirichlet
 * momenta of the neighbours read. */
static void FN(post_process)(olo_lattice* L)
{
  V out[27];
  for (int64_t k = 0; k < L->n_plane; ++k) {
    FN(plane_fd)(L, L->plane_cells[k], out);
    FN(store_cell)(L, L->plane_cells[k], out);
  }
  for (int64_t k = 0; k < L->n_edge; ++k) FN(outer_edge)(L, L->edge_cells[k]);
  for (int64_t k = 0; k < L->n_corner; ++k) FN(outer_corner)(L, L->corner_cells[k]);
}

/* statistics the next collision would report (Aggregatable, latticeStatistics.h:49-71):
 * out = {cells, sum rho, sum |u|^2, max |u|^2}, summed in FP64 */
static void FN(statistics)(const olo_lattice* L, double* out)
{
  out[0] = out[1] = out[2] = out[3] = 0;
  const int p = L->pad;
  const int forced = (L->kind == OLO_FORCED_BGK || L->kind == OLO_FORCED_TRT);
  for (int x = 0; x < L->n[0]; ++x)
    for (int y = 0; y < L->n[1]; ++y)
      for (int z = 0; z < L->n[2]; ++z) {
        const int64_t n = (int64_t)(x + p) * L->proj[0] + (int64_t)(y + p) * L->proj[1] + z + p;
        const int id = L->ids[n];
        V rho, usq, u[3], f[27];
        if (id == 0 || (id == 2 && !(L->lkind && L->lkind[n]))) continue;
        FN(load_cell)(L, n, f);
        if (L->lkind && L->lkind[n]) {
          rho = FN(collide_local)(L, f, n, &usq);
          if (rho < 0) continue;
        } else if (id == 1) {
          FN(rho_u)(L, f, forced ? cell_force(L, n) : NULL, &rho, u);
          usq = FN(usqr)(u);
        } else {
          rho = (V)L->wall_rho[id];
          for (int d = 0; d < 3; ++d) u[d] = (V)L->wall_u[id][d];
          usq = FN(usqr)(u);
        }
        out[0] += 1;
        out[1] += (double)rho;
        out[2] += (double)usq;
        if ((double)usq > out[3]) out[3] = (double)usq;
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
        if (L->lkind && L->lkind[n]) {
          V usq;
          (void)FN(collide_local)(L, f, n, &usq);
        } else {
          FN(collide_cell)(L, f, id, n);
        }
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
  FN(post_process)(L);
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
        FN(rho_u)(L, f, (forced && L->ids[n] == 1) ? cell_force(L, n) : NULL, &rho, u);
        out[c] = (double)rho;
        for (int d = 0; d < 3; ++d) out[(int64_t)(d + 1) * nc + c] = (double)u[d];
      }
}

/* force: NULL (the lattice's uniform FORCE) or [3][n] per cell */
static void FN(equilibrium)(olo_lattice* L, int64_t n, const double* rho_u, const double* force, double* f)
{
  const int forced = (L->kind == OLO_FORCED_BGK || L->kind == OLO_FORCED_TRT);
  for (int64_t c = 0; c < n; ++c) {
    const V rho = (V)rho_u[c];
    V u[3];
    for (int d = 0; d < 3; ++d) {
      u[d] = (V)rho_u[(int64_t)(d + 1) * n + c];
      /* ForcedMomentum::inverseShift, dynamics/interface.h:111-121 */
      const double Fd = force ? force[(int64_t)d * n + c] : L->force[d];
      if (forced) u[d] -= (V)Fd * (V)0.5;
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
