/* lbmh.h -- C ABI of the MI355X (gfx950) block-lattice backend.
 *
 * One lattice handle is one OpenLB block: a padded core of nx*ny*nz cells
 * (padding 1), one periodic-shift (PS) population column per discrete
 * velocity in HBM, shifted storage f_i - t_i in FP64, FP32 or FP16, and one
 * dynamics-id byte per core cell.  A time step is one fused kernel launch:
 * collision of every core cell as an in-place read-modify-write, the
 * periodic overlap refresh folded into the same pass, and streaming as a
 * shift of every column.
 *
 * Every function returns 0 on success or a negative LBMH_E* code; the text
 * of the last error of the calling thread is lbmh_last_error().  No C++
 * exception crosses this boundary.  Handles are not thread-safe: one host
 * thread drives a lattice at a time.
 *
 * Each entry point names the OpenLB 1.8 interface it replaces (paths are
 * relative to the OpenLB source tree).
 */
#ifndef LBMH_H
#define LBMH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBMH_ABI_VERSION 1

/* status codes */
#define LBMH_OK 0
#define LBMH_E_INVALID (-1)   /* bad argument                               */
#define LBMH_E_DEVICE (-2)    /* HIP runtime error                          */
#define LBMH_E_MEMORY (-3)    /* allocation failed                          */
#define LBMH_E_STATE (-4)     /* call not valid in the handle's state       */

/* population storage (shifted deviations f_i - t_i) */
#define LBMH_F64 0
#define LBMH_F32 1
#define LBMH_F16 2            /* FP16 storage, FP32 collision arithmetic    */

/* bulk collision kinds, applied to cells whose dynamics id is 1 */
#define LBMH_BGK 0            /* BGKdynamics             src/dynamics/dynamics.h:89-98   */
#define LBMH_FORCED_BGK 1     /* ForcedBGKdynamics (Guo)  src/dynamics/dynamics.h:213-220 */
#define LBMH_TRT 2            /* TRTdynamics              src/dynamics/dynamics.h:445-451 */
#define LBMH_FORCED_TRT 3     /* ForcedTRTdynamics        src/dynamics/dynamics.h:457-464 */
#define LBMH_SMAGORINSKY 4    /* SmagorinskyBGKdynamics   src/dynamics/smagorinskyBGKdynamics.h:38-44 */

/* per-cell dynamics ids (one byte per core cell):
 *   0       NoDynamics               src/dynamics/dynamics.h:50-62   (not collided)
 *   1       the lattice's bulk kind
 *   2       BounceBack               src/dynamics/dynamics.h:472-483
 *   3..255  BounceBackVelocity wall  src/dynamics/dynamics.h:550-555, each id
 *           with its own fixed density and wall velocity (lbmh_set_wall) */
#define LBMH_ID_NONE 0
#define LBMH_ID_BULK 1
#define LBMH_ID_BOUNCE_BACK 2
#define LBMH_ID_FIRST_WALL 3

typedef struct lbmh_lattice lbmh_lattice;

typedef struct lbmh_lattice_desc {
  int q;            /* 19 (D3Q19) or 27 (D3Q27)                               */
  int nx, ny, nz;   /* core extent; z is contiguous in memory                 */
  int periodic[3];  /* periodic overlap refresh per axis (0/1)                */
  int storage;      /* LBMH_F64 / LBMH_F32 / LBMH_F16                         */
  int bulk_kind;    /* LBMH_BGK ... LBMH_SMAGORINSKY                          */
  int device;       /* HIP device ordinal                                     */
} lbmh_lattice_desc;

typedef struct lbmh_bulk_params {
  double omega;     /* descriptors::OMEGA = 1/tau                             */
  double magic;     /* collision::TRT::MAGIC (TRT kinds)                      */
  double smago;     /* collision::LES::SMAGORINSKY (Smagorinsky kind)         */
  double force[3];  /* descriptors::FORCE, uniform over the bulk cells         */
} lbmh_bulk_params;

/* Create a block lattice.  Replaces the GPU_CUDA construction path of
 * ConcreteBlockLattice (src/core/blockLattice.hh:393-416), which allocates one
 * CyclicColumn per population (src/core/platform/gpu/cuda/column.hh:217-267).
 * Populations start at zero (f_i = t_i), every cell has id 0.  Limits: fewer
 * than 2^31 padded cells, and one population column ((nx+2)(ny+2)(nz+2) x
 * bytes per population) below 2 GiB, i.e. up to 2^28 padded cells in FP64,
 * 2^29 in FP32 and 2^30 in FP16 (the kernels use 32-bit byte offsets inside a
 * column).  Ensemble members obey the same limits. */
int lbmh_lattice_create(const lbmh_lattice_desc* desc, lbmh_lattice** out);
int lbmh_lattice_destroy(lbmh_lattice* lattice);

/* Per-cell dynamics ids, host array [nx][ny][nz].  Replaces the dynamics masks
 * and DYNAMICS pointer field of ConcreteBlockCollisionO<GPU_CUDA>
 * (src/core/platform/gpu/cuda/operator.h:46-124, mask.h:31-85). */
int lbmh_set_dynamics_ids(lbmh_lattice* lattice, const uint8_t* ids);

/* Bulk parameters.  Replaces SuperLattice::setParameter<OMEGA/MAGIC/SMAGORINSKY>
 * (src/core/superLattice.h:318-336) and the per-cell FORCE field of
 * ForcedBGKdynamics for a uniform force (src/dynamics/forcing.h:37-74). */
int lbmh_set_bulk_params(lbmh_lattice* lattice, const lbmh_bulk_params* params);

/* Per-cell force of the forced bulk kinds (LBMH_FORCED_BGK, LBMH_FORCED_TRT):
 * host array [3][nx][ny][nz] of lattice-unit forces, replacing the uniform
 * lbmh_bulk_params.force on every bulk cell; NULL returns to the uniform force.
 * Replaces the per-cell descriptors::FORCE field that ForcedBGKdynamics and
 * ForcedTRTdynamics read in their Guo forcing (src/dynamics/forcing.h:37-74,
 * src/dynamics/momenta/elements.h:1480-1488) and that moments, statistics,
 * equilibrium initialisation and field export read for the F/2 velocity shift.
 * Non-finite values, and a lattice whose bulk kind is not forced, are
 * LBMH_E_INVALID. */
int lbmh_set_force_field(lbmh_lattice* lattice, const double* force);

/* Moving wall id (3..255): fixed density rho and wall velocity u.  Replaces the
 * RHO / VELOCITY fields of BounceBackVelocity cells
 * (src/dynamics/momenta/elements.h:237-262, collision.h:323-348). */
int lbmh_set_wall(lbmh_lattice* lattice, int id, double rho, const double u[3]);

/* Populations of the core cells, shifted values f_i - t_i, host layout
 * [q][nx][ny][nz] in host_dtype (LBMH_F64/F32/F16).  Replaces the
 * processing-context copies of CyclicColumn (src/core/platform/gpu/cuda/
 * column.hh:285-296) and ConcreteBlockLattice::setProcessingContext. */
int lbmh_populations_upload(lbmh_lattice* lattice, const void* host, int host_dtype);
int lbmh_populations_download(lbmh_lattice* lattice, void* host, int host_dtype);

/* Population column i of a host block in OpenLB's own storage, as a host
 * CyclicColumn holds it: (nx+2p)(ny+2p)(nz+2p) values for overlap width
 * `padding` p >= 1, element k the value of padded cell (k - shift) mod count
 * (cell id (x+p)(ny+2p)(nz+2p) + (y+p)(nz+2p) + (z+p), src/core/
 * blockStructure.h:65-78), shift in (-count, count), values in OpenLB's value
 * type (LBMH_F64 or LBMH_F32, shifted f_i - t_i).  Upload sets the lattice's
 * core cells and its one overlap layer from the host column; download writes
 * those cells of the host column and leaves its other overlap layers and its
 * shift as they are.  These are the host/device copies of a CyclicColumn's
 * processing-context switch (src/core/platform/gpu/cuda/column.hh:157-187,
 * 285-296), for blocks whose overlap is wider than the lattice's.  Both
 * synchronise the device. */
int lbmh_column_upload(lbmh_lattice* lattice, int i, const void* host, int64_t shift, int padding, int host_dtype);
int lbmh_column_download(lbmh_lattice* lattice, int i, void* host, int64_t shift, int padding, int host_dtype);

/* Equilibrium initialisation on the device from per-cell moments, host layout
 * [4][nx][ny][nz] = rho, ux, uy, uz.  Replaces SuperLattice::iniEquilibrium
 * (src/core/blockLattice.hh:202-220; dynamics/interface.h:111-121): forced
 * bulk cells inverse-shift u by F/2; cells with id 0 are left untouched. */
int lbmh_ini_equilibrium(lbmh_lattice* lattice, const double* rho_u);

/* Advance n_steps time steps on the given HIP stream (NULL = null stream),
 * asynchronously.  Replaces SuperLattice::collideAndStream
 * (src/core/superLattice.hh:601-663) for one block: collide, PostCollide
 * overlap-1 refresh (src/communication/superCommunicator.hh:189-218), stream
 * (src/core/blockLattice.hh:317-339). */
int lbmh_step(lbmh_lattice* lattice, int64_t n_steps, void* stream);

/* The two halves of lbmh_step for blocks that exchange overlap cells with
 * other blocks: lbmh_collide launches the collision (and the periodic
 * self-overlap refresh) without streaming; lbmh_stream shifts every column
 * and runs the post-processors (for blocks whose post-processors are joined,
 * once the last of them has streamed: lbmh_set_overlap_ids).
 * SuperLattice::collideAndStream collides all
 * blocks, communicates, then streams all blocks (src/core/superLattice.hh:
 * 601-663); one lbmh_step equals lbmh_collide followed by lbmh_stream. */
int lbmh_collide(lbmh_lattice* lattice, void* stream);
int lbmh_stream(lbmh_lattice* lattice, void* stream);

/* Overlap refresh between two blocks on the same device.  After both blocks
 * collided and before `dst` streams, lbmh_halo_copy copies all q populations
 * of core cell src_cells[k] of `src` into overlap cell dst_cells[k] of `dst`
 * (padded cell ids, the lists BlockCommunicationNeighborhood builds,
 * src/communication/blockCommunicationNeighborhood.hh:89-107; the copy of
 * ConcreteBlockCommunicator, src/core/blockLattice.hh:752-787, 918-924).
 * Both blocks must share q and storage type; src may equal dst. */
typedef struct lbmh_halo lbmh_halo;
int lbmh_halo_create(lbmh_lattice* dst, const lbmh_lattice* src, const uint32_t* dst_cells,
                     const uint32_t* src_cells, size_t n, lbmh_halo** out);
int lbmh_halo_destroy(lbmh_halo* halo);
int lbmh_halo_copy(lbmh_halo* halo, void* stream);
/* lbmh_halo_copy of n halos in one launch (the copies of OpenLB's block
 * communicators for every block of a SuperLattice at once,
 * src/communication/superCommunicator.hh:189-218).  The halos must share
 * device, q and storage; the batch's descriptor table is kept by the first
 * halo's destination block, so batches led by one block must run on one
 * stream. */
int lbmh_halo_copy_many(lbmh_halo* const* halos, size_t n, void* stream);

/* Number of completed time steps and the PS shift of every column, in the
 * convention of CyclicColumn::_shift, normalised to [0, n_padded_cells)
 * (src/core/platform/cpu/sisd/column.h:193-204). */
int lbmh_get_step_count(const lbmh_lattice* lattice, int64_t* steps);
int lbmh_get_shifts(const lbmh_lattice* lattice, int64_t* shifts);

/* OpenLB checkpoint of the block, as raw bytes in OpenLB's value type `real`
 * (LBMH_F64 for a double lattice, LBMH_F32 for float).  The bytes are those
 * Serializable::save(std::uint8_t*) writes for a ConcreteBlockLattice
 * (src/core/serializer.hh:106-110, io/serializerIO.hh:96-106): POPULATION as
 * q CyclicColumns {int64 shift, uint64 count, T[count]} over the padded block
 * (src/core/platform/cpu/sisd/column.hh:66-89, gpu/cuda/column.hh:298-321),
 * then the other fields OpenLB allocates for the block's dynamics (FORCE,
 * VELOCITY, FixedDensity::RHO, FixedPressureMomentum / FixedVelocityMomentumGeneric
 * VELOCITY) as Columns {uint64 count, T[count]}, in OpenLB's field order
 * (src/core/data.hh:52-62).  Replaces getSerializableSize / save / load of
 * ConcreteBlockLattice (src/core/blockLattice.hh:597-626); one block's bytes
 * are also a one-block SuperLattice's (src/core/superLattice.hh:806-842).
 * lbmh_checkpoint_load takes the populations of every padded cell and the
 * shifts, after checking the sizes and, on the core cells, every field value
 * the block's dynamics read against the lattice's own set-up; on any
 * mismatch it changes nothing and returns LBMH_E_INVALID.  Both synchronise
 * the device. */
int lbmh_checkpoint_size(const lbmh_lattice* lattice, int real, size_t* bytes);
int lbmh_checkpoint_save(lbmh_lattice* lattice, int real, void* buffer, size_t bytes);
int lbmh_checkpoint_load(lbmh_lattice* lattice, int real, const void* buffer, size_t bytes);

/* Per-cell moments in FP64, host layout [4][nx][ny][nz]: rho and u as the
 * cell's momenta define them (bulk: lbm<D>::computeRhoU, src/dynamics/lbm.h:
 * 300-307, forced: u += F/2).  Replaces SuperLatticePhysVelocity3D-style host
 * evaluation after setProcessingContext(Evaluation). */
int lbmh_moments(lbmh_lattice* lattice, double* rho_u);

/* Moment fields for output, evaluated on the device and written to `host` as
 * float in the point order of a VTK ImageData piece: x fastest, then y, then
 * z, over [-overlap, n + overlap) per axis (overlap 0 or 1), all components of
 * a point together -- the buffer OpenLB's SuperVTMwriter3D::dataArray fills
 * before encoding it (src/io/superVtmWriter3D.hh:437-458).  Each point is what
 * OpenLB's functor evaluates on that cell through the cell's own momenta:
 *   LBMH_FIELD_DENSITY        SuperLatticeDensity3D        (1 component)
 *   LBMH_FIELD_VELOCITY       SuperLatticeVelocity3D       (3)
 *   LBMH_FIELD_PHYS_VELOCITY  SuperLatticePhysVelocity3D   (3), u * conversion[0]
 *   LBMH_FIELD_PHYS_PRESSURE  SuperLatticePhysPressure3D   (1),
 *                             (rho - 1)/3 * conversion[1] + conversion[2]
 * (src/functors/lattice/latticeDensity3D.hh, latticeVelocity3D.hh,
 * latticePhysVelocity3D.hh:84-96, latticePhysPressure3D.hh:71-78), with
 * conversion = {UnitConverter velocity factor, pressure factor,
 * charPhysPressure}, NULL for the lattice-unit fields.  Overlap points on a
 * periodic axis are their periodic image (the writer communicates first);
 * beyond a non-periodic face they have no dynamics.  `floats` must be the
 * number of points times the components.  Only these floats cross PCIe, not
 * the populations. */
#define LBMH_FIELD_DENSITY 0
#define LBMH_FIELD_VELOCITY 1
#define LBMH_FIELD_PHYS_VELOCITY 2
#define LBMH_FIELD_PHYS_PRESSURE 3
int lbmh_field_export(lbmh_lattice* lattice, int field, int overlap, const double conversion[3], float* host,
                      size_t floats);

/* Reduction of a cell field over the core cells whose dynamics id is
 * selected: the sum, minimum or maximum of each component, in FP64, and the
 * number of cells visited.  Fields: the four of lbmh_field_export, each cell
 * evaluated through its own momenta as there (in the lattice's arithmetic
 * type, not rounded to float), and LBMH_FIELD_FORCE, the FORCE the bulk
 * cells carry (the per-cell field of lbmh_set_force_field or the uniform
 * force; 0 on other cells).  `materials` holds 256 flags, nonzero for the
 * materials to visit -- the ids given to lbmh_set_dynamics_ids, whose cells
 * lbmh_set_local_boundary relabels but which still count as their material,
 * like OpenLB's SuperGeometry::getMaterialIndicator -- or is NULL for every
 * material but 0 (NoDynamics).  Empty selections give
 * the reduction's identity: 0, +DBL_MAX (min) or -DBL_MAX (max).  `out`
 * holds 1 (density, pressure) or 3 values; `cells` may be NULL.  Replaces
 * SuperLatticeFieldReductionO<T,DESCRIPTOR,FIELD,SumO|MinO|MaxO,CONDITION>::
 * compute() (src/core/superLatticeFieldReductionO.h:41-160, the operators of
 * src/core/blockLatticeFieldReductionO.h:35-100, the GPU kernels of
 * src/core/platform/gpu/cuda/fieldReduction.hh) and the SuperSum3D /
 * SuperMin3D / SuperMax3D / SuperAverage3D functors over a material
 * indicator (src/functors/lattice/integral/superIntegralF3D.h:37-62,
 * superMax3D.h, superMin3D.h, superAverage3D.h; average = sum / cells). */
#define LBMH_FIELD_FORCE 4
#define LBMH_REDUCE_SUM 0
#define LBMH_REDUCE_MIN 1
#define LBMH_REDUCE_MAX 2
int lbmh_reduce_field(lbmh_lattice* lattice, int field, int op, const uint8_t* materials, const double conversion[3],
                      double out[3], uint64_t* cells);

/* FP64 sums over the cells with dynamics id 1: out = {sum rho, sum j_x,
 * sum j_y, sum j_z} (j = rho u without forcing shift).  Replaces
 * SuperLatticeFieldReductionO / reductionFunctionDevice
 * (src/core/platform/gpu/cuda/fieldReduction.hh:152-170). */
int lbmh_reduce_moments(lbmh_lattice* lattice, double out[4]);

/* lbmh_reduce_moments for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in OpenLB.  The same four FP64 sums, bit for bit,
 * written to out_device (4 doubles of device memory the caller owns),
 * asynchronously on `stream` (NULL = null stream), ordered after the work
 * already queued there; nothing is allocated, copied to the host or
 * synchronised.  The per-workgroup partial sums live in a fixed workspace of
 * the library on each device (1024 workgroups x 4 doubles, 32 KiB, held until
 * the process ends); a reduction that would need more partials is refused
 * with LBMH_E_INVALID.  Every call on one device shares that workspace:
 * calls that overlap on one device (different streams, different host
 * threads) race and must be serialised by the caller. */
int lbmh_reduce_moments_noalloc(lbmh_lattice* lattice, double* out_device, void* stream);

/* Local (regularized) boundaries, OpenLB's boundary::LocalVelocity and
 * boundary::LocalPressure with BGKdynamics mixed in, their default mixin
 * (src/boundary/localVelocity3D.h:36-116, localPressure3D.h:36-66, set by
 * boundary::set, src/boundary/setBoundary.h:83-101).  Every core cell that
 * currently carries dynamics id `id` (the boundary's material) is classified
 * from the id map -- id 1 fluid, id 0 outside, any other id wall -- as
 * computeBoundaryTypeAndNormal does (src/geometry/discreteNormals.hh:
 * 174-1190), and gets the dynamics OpenLB assigns to its type:
 *   flat face      regularized velocity (pressure) boundary, CombinedRLBdynamics
 *   outer edge     BGK at fixed velocity + OuterVelocityEdgeProcessor3D
 *   outer corner   BGK at fixed velocity + OuterVelocityCornerProcessor3D
 *   inner edge     CombinedRLBdynamics + InnerEdgeVelocityTuple3D
 *   inner corner   CombinedRLBdynamics + InnerCornerVelocityTuple3D
 * A velocity boundary cell that matches no type gets no dynamics; a pressure
 * boundary accepts flat faces only (LBMH_E_INVALID otherwise, where OpenLB
 * throws).  The cells are relabelled to free ids (counting down from 255), one
 * per (type, normal); rho/u are what defineRhoU gives the material: the
 * velocity of a velocity boundary, the density (and tangential velocity) of a
 * pressure boundary.  The post-processors run after every step's shift, edges
 * before corners (PostStream priorities).  The bulk may run any kind: the
 * boundary cells collide with the BGK mixin at OMEGA (as next to the
 * Smagorinsky bulk of examples/turbulence/aorta3d/aorta3d.cpp:108-127), and
 * the post-processors read a forced bulk neighbour's velocity with F/2 added
 * (momenta/elements.h:1473-1488).  Other mixins are not built.
 * lbmh_set_dynamics_ids drops all local boundaries. */
#define LBMH_LOCAL_VELOCITY 0
#define LBMH_LOCAL_PRESSURE 1
/* boundary::InterpolatedVelocity / InterpolatedPressure
 * (src/boundary/interpolatedVelocity3D.h:34-90, interpolatedPressure3D.h:37-90):
 * flat faces get BGK with Dirichlet momenta and the finite-difference
 * PlaneFdBoundaryProcessor3D (boundaryPostProcessors3D.hh:38-92), which runs
 * before the edge processors; edges, corners and inner types are those of
 * LocalVelocity.  A velocity face cell without a normal is an error (OpenLB
 * throws); InterpolatedPressure leaves non-flat cells without dynamics. */
#define LBMH_INTERP_VELOCITY 2
#define LBMH_INTERP_PRESSURE 3
int lbmh_set_local_boundary(lbmh_lattice* lattice, int id, int kind, double rho, const double u[3]);

/* Boundary type and discrete normal of every core cell of an id map that is
 * neither fluid (1) nor outside (0): out int8 [4][nx][ny][nz] = type (0 flat,
 * 1 external corner, 2 internal corner, 3 external edge, 4 internal edge; -1
 * on fluid and outside cells) and the normal (pointing out of the fluid).
 * Host only.  Replaces computeBoundaryTypeAndNormal
 * (src/geometry/discreteNormals.hh:174-1190). */
int lbmh_classify_boundary(const int extent[3], const int periodic[3], const uint8_t* ids, int8_t* out);

/* The same classification for one block of a split lattice: beyond each
 * non-periodic face lies the block's first overlap layer, whose materials
 * `ring` gives as a padded map uint8 [nx+2][ny+2][nz+2] (core entries
 * unused): the ids of the neighbouring blocks' cells there, 0 where the
 * domain ends.  computeBoundaryTypeAndNormal sees the same on a block
 * geometry whose overlap holds the neighbours' materials
 * (src/boundary/setBoundary.h:83-101).  Host only. */
int lbmh_classify_boundary_overlap(const int extent[3], const int periodic[3], const uint8_t* ids,
                                   const uint8_t* ring, int8_t* out);

/* Materials of the first overlap layer of a block of a split lattice, as
 * lbmh_classify_boundary_overlap takes them.  lbmh_set_local_boundary then
 * classifies the block's boundary cells with them, and reads every
 * post-processor neighbour beyond a non-periodic face from the block that
 * owns it: the halos into this block name the owner of the first-layer cell
 * on the way to the neighbour, and the rest of the offset is taken inside
 * that owner.  A neighbour whose first-layer cell has no halo source is
 * outside the domain (no dynamics: rho 1, u 0).
 * Blocks whose post-processors are joined through halos run them together,
 * stage by stage (planes, edges, corners): before each stage every block
 * fetches the neighbours that stage reads in other blocks.  A neighbour in
 * another block is so read after the earlier stages and before its own, as
 * a neighbour inside the block is; the reference gets the same by running
 * the post-processors on each block's overlap copies too, after the
 * PostStream exchange (src/boundary/setBoundary.h:83-101,
 * src/core/superLattice.hh:639-640).  The group runs when the last of its
 * blocks reaches lbmh_stream for the step (the others' lbmh_stream only
 * shift), or lbmh_initialize; every block must have collided, and the halo
 * copies run, before any of them streams.  Call before
 * lbmh_set_local_boundary (LBMH_E_STATE after). */
int lbmh_set_overlap_ids(lbmh_lattice* lattice, const uint8_t* ring);

/* SuperLattice::initialize() for the local boundaries: run the edge and
 * corner post-processors once on the current populations
 * (src/core/blockLattice.hh:56-67).  No-op without local boundaries.  A block
 * of a split lattice whose post-processors are joined to other blocks' (see
 * lbmh_set_overlap_ids) runs them when the last of those blocks calls. */
int lbmh_initialize(lbmh_lattice* lattice);

/* Lattice statistics in FP64 over the cells whose collision reports one:
 * bulk cells (rho and |u|^2, u with the F/2 shift when forced) and moving
 * walls (the wall's rho and |u_w|^2); NoDynamics and BounceBack cells count
 * none.  out = {cells counted, average rho, average energy 0.5*<|u|^2>,
 * max |u|}.  Evaluated on the current populations, i.e. the values the next
 * collideAndStream would accumulate.  Replaces LatticeStatistics'
 * Aggregatable + reset() (src/core/latticeStatistics.h:49-71,
 * src/core/latticeStatistics.hh:75-90), which OpenLB fills during collision
 * when statistics are enabled. */
int lbmh_statistics(lbmh_lattice* lattice, double out[4]);

/* Wait for all work queued on the lattice's device.  Replaces
 * gpu::cuda::device::synchronize (src/core/platform/gpu/cuda/device.hh:64-69). */
int lbmh_synchronize(lbmh_lattice* lattice);
/* The same for a device ordinal, without a lattice: the synchronisation
 * SuperLattice performs before each post-processing stage on a GPU platform
 * (src/core/superLattice.hh:666-671). */
int lbmh_device_synchronize(int device);

/* Launch shape of the time-step kernel.  No OpenLB counterpart: the CUDA
 * platform launches every kernel with 32-thread blocks (src/core/platform/gpu/
 * cuda/operator.hh:422-429).  lds_bytes reserves dynamic LDS per 256-thread
 * workgroup purely to cap the resident workgroups per CU (160 KB LDS per CU):
 * fewer waves in flight keep fewer DRAM rows open for the 2*Q concurrent
 * population streams.  xcd_group G > 0 hands each XCD runs of G consecutive
 * 256-cell chunks inside a window of 8*G chunks, so a cache line two chunks
 * share is fetched into one L2 instead of two.  Defaults: lds_bytes 40960
 * (4 workgroups per CU) and xcd_group 4 for FP32 storage, lds_bytes 49152
 * and xcd_group 8 for FP64, lds_bytes 0 and xcd_group 8 for FP16 storage;
 * FP64/FP32 lattices with local velocity /
 * pressure boundaries switch to lds_bytes 0 (their whole-lattice pass also
 * collides the flat faces, whose divergent waves need the occupancy).  A
 * shape set here or chosen by lbmh_autotune is kept. */
typedef struct lbmh_launch_config {
  int lds_bytes;    /* 0 .. 65536                                             */
  int xcd_group;    /* 0 (dispatch order) .. 64                               */
} lbmh_launch_config;

int lbmh_set_launch_config(lbmh_lattice* lattice, const lbmh_launch_config* config);
int lbmh_get_launch_config(const lbmh_lattice* lattice, lbmh_launch_config* config);

/* Time candidate launch shapes on a scratch lattice of the same extent,
 * dynamics ids and parameters (the lattice's own populations are untouched),
 * keep the fastest on the lattice and return it in *chosen (may be NULL).
 * trial_steps >= 1 timed steps per candidate.  Needs device memory for one
 * more copy of the lattice. */
int lbmh_autotune(lbmh_lattice* lattice, int trial_steps, lbmh_launch_config* chosen);

/* Measurement only, no OpenLB counterpart: n_steps passes of the time step's
 * memory access pattern with its arithmetic taken out.  Each pass reads every
 * core cell's Q populations at their shifted slots and writes the same values
 * back in place, with the step kernel's cell mapping, addressing and launch
 * shape, then advances the shifts as a stream would.  Its rate is the ceiling
 * of the periodic-shift in-place read-modify-write on the device, the figure
 * the step kernel is priced against beside the vendor HBM peak.  The lattice
 * is left freely streamed (no collision, no overlap refresh): use it on a
 * lattice whose state is no longer needed.  Asynchronous on `stream`. */
int lbmh_pattern_ceiling(lbmh_lattice* lattice, int64_t n_steps, void* stream);

/* ---- Ensembles: many independent lattices advanced by one launch per step ----
 *
 * No single OpenLB counterpart.  The reference runs uncertainty-quantification
 * samples one SuperLattice at a time (examples/uncertaintyQuantification/
 * cavity3d/cavity3d.cpp:110-148); an ensemble holds them all on one GPU.  Each
 * member is one block lattice (same semantics as lbmh_lattice: padded PS
 * columns, dynamics ids 0..15, per-member bulk parameters and walls).  Members
 * share one descriptor, storage type and bulk kind; extents, periodicity,
 * parameters and step budgets differ.  Device memory comes from two arenas
 * sized at creation, so adding and removing members never allocates or frees
 * device memory.  lbmh_ensemble_add / _remove / _set_steps wait for all work
 * queued on the device before they change the member table. */
typedef struct lbmh_ensemble lbmh_ensemble;

typedef struct lbmh_ensemble_desc {
  int q;                    /* 19 or 27                                        */
  int storage;              /* LBMH_F64 / LBMH_F32 / LBMH_F16                  */
  int bulk_kind;            /* LBMH_BGK ... LBMH_SMAGORINSKY                   */
  int device;
  int max_members;          /* member slots                                    */
  size_t population_bytes;  /* population arena                                */
  size_t cell_capacity;     /* id arena (cells over all members)               */
  size_t boundary_capacity; /* local-boundary cells over all members
                               (0: cell_capacity / 4)                          */
} lbmh_ensemble_desc;

int lbmh_ensemble_create(const lbmh_ensemble_desc* desc, lbmh_ensemble** out);
int lbmh_ensemble_destroy(lbmh_ensemble* ensemble);

/* Device bytes one member of this extent takes from the arenas (populations,
 * ids); the admission check of a serving loop. */
int lbmh_ensemble_member_bytes(const lbmh_ensemble* ensemble, const int extent[3], size_t* population_bytes,
                               size_t* id_bytes);

/* Add a member: extent, periodicity, ids [nx][ny][nz] (values 0..15), bulk
 * parameters, and the number of steps it should run (< 0: until removed).
 * Populations start at zero (f_i = t_i).  Returns LBMH_E_MEMORY when the
 * arenas or the slots are full. */
int lbmh_ensemble_add(lbmh_ensemble* ensemble, const int extent[3], const int periodic[3], const uint8_t* ids,
                      const lbmh_bulk_params* params, int64_t steps, int* member);
int lbmh_ensemble_remove(lbmh_ensemble* ensemble, int member);
int lbmh_ensemble_set_wall(lbmh_ensemble* ensemble, int member, int id, double rho, const double u[3]);
/* Local velocity / pressure boundary on the member's cells carrying `id`, as
 * lbmh_set_local_boundary does for a lattice (every kind: LocalVelocity,
 * LocalPressure, InterpolatedVelocity, InterpolatedPressure); the per-type
 * ids come from the member's free ids 3..63.  Any bulk kind: the boundary
 * cells keep their BGK mixin, as on a lattice.
 * The first InterpolatedPressure member allocates the ensemble's plane
 * workspace (boundary capacity x Q values); later members reuse it. */
int lbmh_ensemble_set_local_boundary(lbmh_ensemble* ensemble, int member, int id, int kind, double rho,
                                     const double u[3]);
/* initialize() of one member: its edge and corner post-processors once. */
int lbmh_ensemble_initialize(lbmh_ensemble* ensemble, int member);
int lbmh_ensemble_set_steps(lbmh_ensemble* ensemble, int member, int64_t steps);
/* steps completed and steps still to run (< 0: unbounded) */
int lbmh_ensemble_status(lbmh_ensemble* ensemble, int member, int64_t* steps_done, int64_t* steps_left);
/* the same for every slot at once (arrays of max_members entries; inactive
 * slots report 0 / 0); one device round trip */
int lbmh_ensemble_status_all(lbmh_ensemble* ensemble, int64_t* steps_done, int64_t* steps_left);

int lbmh_ensemble_ini_equilibrium(lbmh_ensemble* ensemble, int member, const double* rho_u);
int lbmh_ensemble_upload(lbmh_ensemble* ensemble, int member, const void* host, int host_dtype);
int lbmh_ensemble_download(lbmh_ensemble* ensemble, int member, void* host, int host_dtype);
int lbmh_ensemble_moments(lbmh_ensemble* ensemble, int member, double* rho_u);
/* OpenLB checkpoint of one member, the same bytes lbmh_checkpoint_save
 * writes for a lattice of the member's set-up (ConcreteBlockLattice
 * serialization, src/core/blockLattice.hh:597-626).  A load replaces the
 * member's populations and shifts; its step counters are left as they are.
 * With remove and add, this preempts a member to host memory and resumes it
 * later in any free slot. */
/* lbmh_field_export for one member. */
int lbmh_ensemble_field_export(lbmh_ensemble* ensemble, int member, int field, int overlap,
                               const double conversion[3], float* host, size_t floats);
int lbmh_ensemble_checkpoint_size(lbmh_ensemble* ensemble, int member, int real, size_t* bytes);
int lbmh_ensemble_checkpoint_save(lbmh_ensemble* ensemble, int member, int real, void* buffer, size_t bytes);
int lbmh_ensemble_checkpoint_load(lbmh_ensemble* ensemble, int member, int real, const void* buffer, size_t bytes);

/* Advance every member that still has steps by up to n_steps steps: one
 * launch of the fused step kernel over all members' cells and one launch of
 * the stream/count kernel per step, asynchronously on `stream`.  The _graph
 * variant captures the step loop in a HIP graph (rebuilt when membership
 * changes) and replays it. */
int lbmh_ensemble_step(lbmh_ensemble* ensemble, int64_t n_steps, void* stream);
int lbmh_ensemble_step_graph(lbmh_ensemble* ensemble, int64_t n_steps, void* stream);

/* Launch shape of the ensemble step kernel (see lbmh_launch_config); the
 * defaults are the single-lattice ones for the storage type, except that an
 * FP64/FP32 ensemble with local-boundary members reserves 24 KB of LDS (their
 * flat faces are collided in the step kernel).  A shape set here is kept. */
int lbmh_ensemble_set_launch_config(lbmh_ensemble* ensemble, const lbmh_launch_config* config);

/* Time the candidate launch shapes (LDS reservation x XCD group) on the
 * ensemble's own members, trial_steps timed steps each, keep the fastest and
 * return it in *chosen (may be NULL).  No OpenLB counterpart.  The members'
 * populations and step state are copied aside and restored after every
 * candidate, so the ensemble ends as it began; this needs device memory for
 * one copy of the populations in use. */
int lbmh_ensemble_autotune(lbmh_ensemble* ensemble, int trial_steps, lbmh_launch_config* chosen);
/* members, cells in the launch, arena bytes in use and capacity */
int lbmh_ensemble_usage(const lbmh_ensemble* ensemble, int* members, size_t* cells, size_t* bytes_used,
                        size_t* bytes_capacity);

/* Bytes of device memory the lattice holds. */
int lbmh_lattice_footprint(const lbmh_lattice* lattice, size_t* bytes);

/* Library and device queries.  Replaces checkPlatform<GPU_CUDA>
 * (src/core/platform/gpu/cuda/communicator.hh:45-81). */
int lbmh_abi_version(void);
int lbmh_device_count(int* count);
int lbmh_device_arch(int device, char* name, size_t len);
const char* lbmh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
