/* Oracle: CPU restatement of OpenLB's block-lattice time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * backend: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  The product path (openlb_amd/) never links or calls it.
 *
 * What it restates (reference files, OpenLB 1.8):
 *   block layout, padding, cell ids     src/core/blockStructure.h:65-78,150-182
 *   periodic-shift (PS) columns         src/core/platform/cpu/sisd/column.h:128-221
 *   stream = rotate every column        src/core/blockLattice.hh:317-339
 *   collide loop over core cells        src/core/platform/cpu/sisd/operator.h:143-228
 *   time-step order                     src/core/superLattice.hh:601-663
 *   overlap-1 halo after collision      src/core/superLattice.hh:138-143,
 *                                       src/communication/blockCommunicationNeighborhood.hh:66-107,
 *                                       src/geometry/cuboidDecomposition.hh:155-177
 *   moments, 2nd-order equilibrium      src/dynamics/lbm.h:50-64,266-307,545-566
 *   BGK / TRT / Revert / Nguyen-Ladd    src/dynamics/collision.h:57-80,209-247,293-348
 *   Smagorinsky effective omega         src/dynamics/collisionLES.h:58-89
 *   Guo forcing (u += F/2)              src/dynamics/forcing.h:37-74,
 *                                       src/dynamics/momenta/elements.h:1473-1512
 *   descriptors D3Q19 / D3Q27           src/descriptor/definition/common.h:120-166,313-366
 *
 * Pinning: every function here is checked against golden vectors written by
 * the compiled reference (tests/golden/), and against closed-form invariants
 * (tests/test_oracle_invariants.py).
 */
#ifndef OLB_ORACLE_H
#define OLB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* population storage types */
enum { OLO_F64 = 0, OLO_F32 = 1, OLO_F16 = 2 };

/* bulk collision kinds (one per lattice, applied to cells with id 1) */
enum {
  OLO_BGK = 0,          /* BGKdynamics            dynamics.h:89-98   */
  OLO_FORCED_BGK = 1,   /* ForcedBGKdynamics      dynamics.h:213-220 */
  OLO_TRT = 2,          /* TRTdynamics            dynamics.h:445-451 */
  OLO_FORCED_TRT = 3,   /* ForcedTRTdynamics      dynamics.h:457-464 */
  OLO_SMAGORINSKY = 4   /* SmagorinskyBGKdynamics smagorinskyBGKdynamics.h:38-44 */
};

/* per-cell dynamics ids: 0 NoDynamics, 1 bulk, 2 BounceBack,
 * 3..255 BounceBackVelocity with a per-id wall (rho, u) */

typedef struct olo_lattice olo_lattice;

olo_lattice* olo_create(int q, int nx, int ny, int nz, int px, int py, int pz, int storage);
void olo_destroy(olo_lattice* L);
int olo_set_ids(olo_lattice* L, const uint8_t* ids);
int olo_set_bulk(olo_lattice* L, int kind, double omega, double magic, double smago,
                 double fx, double fy, double fz);
int olo_set_wall(olo_lattice* L, int id, double rho, double ux, double uy, double uz);
/* populations of the core cells, shifted values, layout [q][nx][ny][nz] */
int olo_upload(olo_lattice* L, const double* f);
int olo_download(olo_lattice* L, double* f);
int olo_step(olo_lattice* L, int64_t steps, int threads);
int olo_get_shifts(const olo_lattice* L, int64_t* shifts);
/* rho, ux, uy, uz of every core cell as the bulk momenta define them, [4][nx][ny][nz] */
int olo_moments(const olo_lattice* L, double* rho_u);
/* shifted equilibrium populations f_i - t_i for given rho/u, iniEquilibrium
 * semantics (forced kinds inverse-shift u by F/2), layout [q][n] */
int olo_equilibrium(int q, int kind, int storage, int64_t n, const double* rho_u,
                    double fx, double fy, double fz, double* f);
/* round a double array through the storage type (f64 / f32 / f16) */
int olo_round_to_storage(int storage, int64_t n, double* values);
const char* olo_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
